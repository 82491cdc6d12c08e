"""Benchmark: Langevin steps/s of the MI355X JIMWLK solver on one GPU.

Workload (the configuration the headline metric is quoted on): N = 2048 periodic
lattice, SU(3) Wilson lines in FP64, MV initial condition (Ny = 50 sheets,
g^2 mu a = 0.25, m a = 1e-4; Philox charges), then momentum-space Langevin steps
with running coupling and per-mode correlated noise (alpha frozen at
alpha_max = 0.7 in the IR, Lambda a = 0.02), SIN kernel, Delta s = 5e-5.
Synthetic data: all noise comes from the Philox stream (seed 1).

One step = noise generation, 2-D FFTs, kernel multiplies, U xi U^dagger and
the SU(3) update; measurement is outside the timed region, as the reference's
"Evolution time" excludes it (main_optimized.cpp:236-311).

Prints one JSON line (driver contract) with `roofline` for the dominant kernel
(HIP-event timing on the library's stream) and `cpu_baseline` (the CPU oracle
timed on a bounded sample of the same workload on this host).
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "softx-d-21-00092_amd"))
sys.path.insert(0, REPO)

METRIC = "Langevin steps/s at N=2048 SU(3): mixed+batched vs FP64; S(r) err vs FP64"
HBM_PEAK_GBS = 8000.0
ALPHA_MAX = 0.7
LAMBDA_A = 0.02
# compulsory HBM bytes per site per step of the FP64 form (reals x 8 B):
# 182 for per-mode correlated noise, 132 for the white-noise couplings
REALS_PER_SITE = {"correlated": 182, "white": 132}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--N", type=int, default=2048)
    p.add_argument("--coupling", default="correlated", choices=["correlated", "fixed", "noise", "sqrt"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--calibrate", action="store_true",
                   help="also run a plain copy of 8 planes (k_calibration_copy) to scale traffic counters")
    p.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    p.add_argument("--form", default="fp64", choices=["fp64", "mixed"],
                   help="plain FP64 form, or the mixed-precision form (FP32 fields and FFTs)")
    p.add_argument("--reunit-every", type=int, default=10)
    p.add_argument("--fused-mode", type=int, default=1,
                   help="jw_set_fused mode: 1 default, 2/3/4 row-major/4-row/8-row blocked intermediates")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the lattice would shard as x-slabs with an all-to-all per column pass, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import jimwlk as jw

    N = a.N
    coupling = {"correlated": jw.CORRELATED_NOISE_COUPLING, "fixed": jw.FIXED_COUPLING,
                "noise": jw.NOISE_COUPLING_CONSTANT, "sqrt": jw.SQRT_COUPLING_CONSTANT}[a.coupling]
    step = 5e-5 if a.coupling == "correlated" else 1e-4
    al = jw.alpha(lambda_L=LAMBDA_A * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * ALPHA_MAX)))
    if a.coupling != "correlated":
        al = None

    ctx = jw.Context(N, 0)
    ctx.set_fused(a.fused_mode)
    U = ctx.field()
    ctx.philox(1, 0, 0)
    g2mu = 0.25
    t0 = time.perf_counter()
    ctx.mv_init(U, 50, 1.0, g2mu, 1e-4 / g2mu)
    ctx.sync()
    t_mv = time.perf_counter() - t0
    print(f"bench: N={N} MV init (50 sheets) {t_mv * 1e3:.1f} ms", file=sys.stderr, flush=True)

    if a.calibrate:
        ctx.calibration_copy(8)
        ctx.sync()
    def run(n):
        if a.form == "mixed":
            ctx.evolve_mixed(U, coupling, jw.SIN_KERNEL, step, al, nsteps=n, reunit_every=a.reunit_every)
        else:
            ctx.evolve(U, coupling, jw.SIN_KERNEL, step, al, nsteps=n)

    run(a.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    run(a.steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    sps = a.steps / dt
    print(f"bench: {a.steps} steps in {dt:.3f} s -> {sps:.1f} steps/s", file=sys.stderr, flush=True)

    # per-kernel durations with HIP events on the library stream (separate pass)
    ctx.timing(True)
    run(min(a.steps, 10))
    kt = ctx.kernel_times()
    ctx.timing(False)
    # Algorithmic HBM bytes per site per step of each kernel class of the fused
    # step (complex planes of 16 B each way; U is 9 complex planes):
    #   row_noise_fwd  W 8 planes                                  128
    #   col_kmult      A: R 8 W 4 (+W 8 correlated), B: R 8 W 4    384 (512)
    #   row_uxiu_fwd   R U 9 (+ xi~ 8 correlated), W 8             272 (400)
    #   row_update     R A 4 + B 4 + U 9, W U 9                    416
    # Correlated: 1456 B = 182 reals per site, the compulsory figure.  White
    # noise: 1200 B = 150 reals (U is read by both row_uxiu_fwd and row_update).
    kind = "correlated" if a.coupling == "correlated" else "white"
    corr_ = kind == "correlated"
    alg = {"row_noise_fwd": 128, "col_kmult": 512 if corr_ else 384, "row_uxiu_fwd": 400 if corr_ else 272,
           "row_update": 416}
    V = N * N
    nk = min(a.steps, 10)
    per_step_ms = {k: ms * n / nk for k, (ms, n) in kt.items()}
    dom = max(per_step_ms, key=per_step_ms.get)
    classes = {}
    for k, t in per_step_ms.items():
        if k in alg:
            classes[k] = {"ms_per_step": round(t, 5), "GBs": round(alg[k] * V / (t * 1e-3) / 1e9, 1),
                          "launch_ms": round(kt[k][0], 5), "launches_per_step": kt[k][1] / nk}
    roof = None
    if dom in alg:
        dom_ms, dom_calls = kt[dom]
        bytes_per_launch = alg[dom] * V * nk / dom_calls
        ach = bytes_per_launch / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(dom_ms, 5),
                "alg_bytes_per_launch": bytes_per_launch}
    step_bytes = REALS_PER_SITE[kind] * 8.0 * V
    step_frac = step_bytes * sps / (HBM_PEAK_GBS * 1e9)
    print(f"bench: kernel classes {json.dumps({k: [round(v[0], 4), v[1]] for k, v in kt.items()})}",
          file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        orc = oracle.Oracle()
        if a.cpu_threads > 0:
            orc.set_num_threads(a.cpu_threads)
        nthr = orc.num_threads()
        import numpy as np
        rng = np.random.default_rng(0)
        U0 = U.download()
        z = rng.normal(size=V * 16)
        oal = oracle.alpha(lambda_L=LAMBDA_A * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * ALPHA_MAX))) \
            if a.coupling == "correlated" else None
        oc = {"correlated": oracle.CORRELATED, "fixed": oracle.FIXED, "noise": oracle.NOISE,
              "sqrt": oracle.SQRT}[a.coupling]
        t0 = time.perf_counter()
        orc.langevin_step(U0, z, N, oc, oracle.SIN, step, oal)
        tc = time.perf_counter() - t0
        cpu = {"value": round(1.0 / tc, 5), "unit": "steps/s", "cores": nthr, "kind": "port",
               "sample": f"1 Langevin step at N={N} ({a.coupling} coupling) by the CPU oracle "
                         f"(reference algorithm: 9-component FFTs, eigen-decomposition exponential), "
                         f"{tc:.1f} s with {nthr} OpenMP threads"}
        print(f"bench: cpu oracle step {tc:.2f} s ({nthr} threads)", file=sys.stderr, flush=True)

    out = {"metric": METRIC, "value": round(sps, 3), "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(1e3 / sps, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (Philox noise, MV initial condition)",
           "config": {"workload": f"N={N} FP64 Langevin step, {a.coupling} coupling"
                                  + (f" (alpha_max={ALPHA_MAX}, Lambda a={LAMBDA_A})" if a.coupling == "correlated" else "")
                                  + f", SIN kernel, ds={step}", "N": N, "sites": V,
                      "form": "plain FP64", "parallelism": "single GPU"},
           "step_roofline": {"bytes_per_step": step_bytes, "achieved_GBs": round(step_bytes * sps / 1e9, 1),
                             "frac": round(step_frac, 4), "reals_per_site": REALS_PER_SITE[kind]},
           "roofline": roof, "kernels": classes, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
