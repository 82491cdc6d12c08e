"""Benchmark: Langevin steps/s of the MI355X JIMWLK solver on one GPU, mixed
precision form against the plain FP64 form, with the dipole S(r) error of
the mixed form against FP64.

Workload (the configuration the headline metric is quoted on): N = 2048 periodic
lattice of SU(3) Wilson lines, MV initial condition (Ny = 50 sheets,
g^2 mu a = 0.25, m a = 1e-4; Philox charges), then momentum-space Langevin
steps with running coupling and per-mode correlated noise (alpha frozen at
alpha_max = 0.7 in the IR, Lambda a = 0.02), SIN kernel, Delta s = 5e-5.
Synthetic data: all noise comes from the Philox stream.

Forms, each timed over the same K steps from the same start:
  * fp64  -- jw_evolve: FP64 throughout (the oracle-parity path);
  * mixed -- jw_evolve_mixed: FP32 fields, FFTs and products, FP64 exponentials, U
             projected onto SU(3) every 10 steps;
  * fp64_k, mixed_k -- the same with momentum-space noise (jw_set_noise_space:
             the noise spectra drawn per mode, the same distribution as the
             per-site normals, no noise pass);
  * mixed+batched -- jw_batch_evolve with --batch N = 2048 trajectories in one
             batch (FP32, captured step graph, tuned layout).
`value` is the mixed+batched form's trajectory-steps/s (whole job) with the noise
of --noise-space (default momentum); the FP64 form's steps/s with the same
noise, the ratio and S(r) after --sr-steps steps of the FP64 and mixed forms
from the same MV start and Philox stream (max |S_mixed(r) - S_fp64(r)| over all
lattice separations r) sit beside it, and every form's steps/s in
`forms_steps_per_s`.  `ensemble` times
the uneven batched ensemble 48 x 256^2 + 32 x 512^2 + 16 x 1024^2 (fixed
coupling, Delta s = 1e-4) in both precisions, in site-updates/s; `production`
the FP64 batch of 16 seeds of N = 1024 (fixed coupling, Delta s = 1e-4);
`large` the N = 8192 single lattice (correlated coupling) in both forms.

One step = noise generation, 2-D FFTs, kernel multiplies, U xi U^dagger and
the SU(3) update; measurement is outside the timed region, as the reference's
"Evolution time" excludes it (main_optimized.cpp:236-311).

Also printed: `roofline` for the dominant kernel of the mixed form (HIP-event
timing on the library's stream; `roofline_fp64` likewise), per-kernel tables,
and `cpu_baseline` (the CPU oracle timed on steps of the same workload on
this host).
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
# Algorithmic HBM traffic of each kernel class of the fused step, in complex
# planes per launch (read + written); the fused passes read and write u = 6
# planes of U (rows 0 and 1, row 2 rebuilt on load):
#   row_noise_fwd  W 8                                        8
#   col_kmult      A: R 8 W 4 (+W 8 correlated), B: R 8 W 4   12 and 12 (20 and 12)
#   row_uxiu_fwd   R U u (+ xi~ 8 correlated), W 8            u + 8 (u + 16)
#   row_update     R A 4 + B 4 + U u, W U u                   8 + 2u
#                  (FP64: + W 8 when it also writes the next step's xi)
# Per step: correlated 64 + 3u planes (82 = 164 reals per site); white noise
# 48 + 3u (U is read by both row passes).
# Momentum-space noise (forms *_k): no noise pass; column pass A draws the
# spectra and writes A (4) and the noise's row spectra (8), the row pass reads
# them back as in the correlated step: 48 + 3u planes = 132 reals per site for
# every coupling.
def planes_of(kind, u):
    t = {"correlated": {"row_noise_fwd": 8, "col_kmult": (20, 12), "row_uxiu_fwd": u + 16, "row_update": 8 + 2 * u},
         "white": {"row_noise_fwd": 8, "col_kmult": (12, 12), "row_uxiu_fwd": u + 8, "row_update": 8 + 2 * u},
         "kspace": {"row_noise_fwd": 0, "col_kmult": (12, 12), "row_uxiu_fwd": u + 16, "row_update": 8 + 2 * u}}[kind]
    return t


def reals_of(kind, u):
    t = planes_of(kind, u)
    return 2 * (t["row_noise_fwd"] + sum(t["col_kmult"]) + t["row_uxiu_fwd"] + t["row_update"])


# measured HBM traffic per launch of the dominant kernels (rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, scaled by the same counters on a plain copy)
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_N2048.json")
# matrix-core use of the split-BF16 column stages (rocprofv3 --pmc, mixed_mfma_k)
MFMA_FILE = os.path.join(REPO, "profiles", "mfma_N2048.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--N", type=int, default=2048)
    p.add_argument("--coupling", default="correlated", choices=["correlated", "fixed", "noise", "sqrt"])
    p.add_argument("--sr-steps", type=int, default=1000,
                   help="steps of both forms before comparing S(r) (0: skip)")
    p.add_argument("--reunit-every", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--calibrate", action="store_true",
                   help="also run a plain copy of 8 planes (k_calibration_copy) to scale traffic counters")
    p.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    p.add_argument("--cpu-steps-shipped", type=int, default=2,
                   help="CPU baseline steps with the reference's threading (serial uxiulocal/update_uf)")
    p.add_argument("--cpu-steps-omp", type=int, default=2, help="CPU baseline steps with every loop threaded")
    p.add_argument("--fused-mode", type=int, default=1,
                   help="jw_set_fused mode: 1 default, 2/3/4 row-major/4-row/8-row blocked intermediates")
    p.add_argument("--forms", default="fp64,mixed,fp64_k,mixed_k,mixed_mfma_k",
                   help="forms to time (profiling runs may pick one); *_k: momentum-space noise")
    p.add_argument("--noise-space", default="momentum", choices=["position", "momentum"],
                   help="noise of the headline forms, the batched form and the S(r) comparison")
    p.add_argument("--batch", type=int, default=2, help="N = 2048 trajectories of the batched form (0: skip)")
    p.add_argument("--no-ensemble", action="store_true", help="skip the uneven-size ensemble")
    p.add_argument("--no-production", action="store_true",
                   help="skip the FP64 production batch (16 trajectories of N = 1024, fixed coupling)")
    p.add_argument("--large-steps", type=int, default=200,
                   help="steps of the N = 8192 single-lattice configuration in both forms (0: skip)")
    return p.parse_args()


def kernel_table(kt, nk, V, planes, bytes_per_real):
    """Per-class time and algorithmic bytes per step (from each class's launch
    counts), algorithmic GB/s, and the dominant class's roofline."""
    calls = {k: n for k, (ms, n) in kt.items() if n}
    n_noise = calls.get("row_noise_fwd", 0)
    plane_b = 2 * bytes_per_real * V
    launches_planes = {}
    if "row_noise_fwd" in calls:
        launches_planes["row_noise_fwd"] = planes["row_noise_fwd"] * n_noise
    if "col_kmult" in calls:
        a_, b_ = planes["col_kmult"]
        launches_planes["col_kmult"] = (a_ + b_) * calls["col_kmult"] / 2
    if "row_uxiu_fwd" in calls:
        launches_planes["row_uxiu_fwd"] = planes["row_uxiu_fwd"] * calls["row_uxiu_fwd"]
    if "row_update" in calls:
        u = calls["row_update"]
        # updates that also wrote the next step's xi: every step whose xi the noise pass did not write
        launches_planes["row_update"] = planes["row_update"] * u + planes["row_noise_fwd"] * (u - n_noise)
    classes = {}
    for k, pl in launches_planes.items():
        ms, n = kt[k]
        b = pl * plane_b / nk
        t = ms * n / nk
        classes[k] = {"ms_per_step": round(t, 5), "GBs": round(b / (t * 1e-3) / 1e9, 1),
                      "launch_ms": round(ms, 5), "launches_per_step": n / nk, "alg_bytes_per_step": b}
    dom = max(classes, key=lambda k: classes[k]["ms_per_step"]) if classes else None
    roof = None
    if dom:
        dom_ms, dom_calls = kt[dom]
        bytes_per_launch = classes[dom]["alg_bytes_per_step"] * nk / dom_calls
        ach = bytes_per_launch / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(dom_ms, 5),
                "alg_bytes_per_launch": bytes_per_launch}
    return classes, roof


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this solver runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the lattice would shard as x-slabs with an all-to-all per column pass)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import jimwlk as jw

    N, V = a.N, a.N * a.N
    forms = [f for f in a.forms.split(",") if f]
    coupling = {"correlated": jw.CORRELATED_NOISE_COUPLING, "fixed": jw.FIXED_COUPLING,
                "noise": jw.NOISE_COUPLING_CONSTANT, "sqrt": jw.SQRT_COUPLING_CONSTANT}[a.coupling]
    step = 5e-5 if a.coupling == "correlated" else 1e-4
    al = jw.alpha(lambda_L=LAMBDA_A * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * ALPHA_MAX))) \
        if a.coupling in ("correlated", "noise", "sqrt") else None
    kind = "correlated" if a.coupling == "correlated" else "white"

    ctx = jw.Context(N, 0)
    ctx.set_fused(a.fused_mode)
    U = ctx.field()
    ctx.philox(1, 0, 0)
    g2mu = 0.25
    t0 = time.perf_counter()
    ctx.mv_init(U, 50, 1.0, g2mu, 1e-4 / g2mu)
    ctx.sync()
    print(f"bench: N={N} MV init (50 sheets) {(time.perf_counter() - t0) * 1e3:.1f} ms", file=sys.stderr, flush=True)
    U0 = U.download()
    if a.calibrate:
        ctx.calibration_copy(8)
        ctx.sync()

    def run(form, F, n):
        base = form[:-2] if form.endswith("_k") else form
        ctx.set_noise_space(jw.NOISE_MOMENTUM if form.endswith("_k") else jw.NOISE_POSITION)
        if base in ("mixed", "mixed_mfma"):
            ctx.set_mixed_fft(1 if base == "mixed_mfma" else 0)
            ctx.evolve_mixed(F, coupling, jw.SIN_KERNEL, step, al, nsteps=n, reunit_every=a.reunit_every)
            ctx.set_mixed_fft(0)
        else:
            ctx.evolve(F, coupling, jw.SIN_KERNEL, step, al, nsteps=n)
        ctx.set_noise_space(jw.NOISE_POSITION)

    ks = a.noise_space == "momentum"
    f64, fmix = ("fp64_k", "mixed_k") if ks else ("fp64", "mixed")

    res = {}
    for form in forms:
        F = ctx.field(9, U0)
        ctx.philox(2, 0, 0)
        run(form, F, a.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        run(form, F, a.steps)
        ctx.sync()
        dt = time.perf_counter() - t0
        sps = a.steps / dt
        print(f"bench: {form}: {a.steps} steps in {dt:.4f} s -> {sps:.1f} steps/s", file=sys.stderr, flush=True)
        # per-kernel durations with HIP events on the library stream (separate pass)
        nk = min(a.steps, 10)
        ctx.timing(True)
        run(form, F, nk)
        kt = ctx.kernel_times()
        ctx.timing(False)
        bpr = 4 if form.startswith("mixed") else 8
        kind_f = "kspace" if form.endswith("_k") else kind
        u_pl = 6
        classes, roof = kernel_table(kt, nk, V, planes_of(kind_f, u_pl), bpr)
        print(f"bench: {form} kernel classes {json.dumps({k: v['ms_per_step'] for k, v in classes.items()})}",
              file=sys.stderr, flush=True)
        step_bytes = reals_of(kind_f, u_pl) * bpr * V
        res[form] = {"steps_per_s": round(sps, 3), "ms_per_step": round(1e3 / sps, 4), "kernels": classes,
                     "roofline": roof,
                     "step_roofline": {"bytes_per_step": step_bytes,
                                       "achieved_GBs": round(step_bytes * sps / 1e9, 1),
                                       "frac": round(step_bytes * sps / (HBM_PEAK_GBS * 1e9), 4),
                                       "reals_per_site": reals_of(kind_f, u_pl), "bytes_per_real": bpr}}
        F.close()

    def time_batch(groups, fp32, coupling_b, step_b, al_b, nsteps, warm, seeds0, kspace=False, launch=-1):
        """Aggregate rate of a batch of MV starts (own seed each): (ensemble steps/s, layouts,
        launch mode: 0 per group, 1 one launch per kernel class)."""
        b = jw.Batch(groups, fp32=fp32)
        if launch >= 0:
            b.set_launch(launch)
        if kspace:
            b.set_noise_space(jw.NOISE_MOMENTUM)
        ctxs = {}
        t = 0
        for N_, count in groups:
            c_ = ctxs.setdefault(N_, jw.Context(N_))
            F_ = c_.field()
            for _ in range(count):
                c_.philox(seeds0 + t, 0, 0)
                c_.mv_init(F_, 50, 1.0, g2mu, 1e-4 / g2mu)
                b.set(t, F_)
                t += 1
            F_.close()
        b.philox(4, 0, 0)
        # warm-up of at least one graph period: the layout tuning and the capture
        # of the step graph happen here, outside the timed region
        period = a.reunit_every if a.reunit_every > 0 else 10
        b.evolve(coupling_b, jw.SIN_KERNEL, step_b, al_b, nsteps=max(warm, period), reunit_every=a.reunit_every)
        b.sync()
        t0 = time.perf_counter()
        b.evolve(coupling_b, jw.SIN_KERNEL, step_b, al_b, nsteps=nsteps, reunit_every=a.reunit_every)
        b.sync()
        dt = time.perf_counter() - t0
        lays = [b.layout(i) for i in range(len(groups))]
        lays = [(l, "matrix-core" if b.fft_mode(i) == 1 else "butterflies") if fp32 else l for i, l in enumerate(lays)]
        mode = b.launch()
        b.close()
        for c_ in ctxs.values():
            c_.close()
        return nsteps / dt, lays, mode

    batched = None
    if a.batch > 0:
        sps_b, lays, mode = time_batch([(N, a.batch)], True, coupling, step, al, a.steps, a.warmup, 1000, ks)
        batched = {"trajectories": a.batch, "N": N, "ensemble_steps_per_s": round(sps_b, 3),
                   "traj_steps_per_s": round(sps_b * a.batch, 3), "layout": lays[0], "launch": mode,
                   "noise_space": a.noise_space}
        print(f"bench: mixed+batched: {a.batch} x N={N}: {sps_b * a.batch:.1f} trajectory-steps/s "
              f"(layout {lays[0]})", file=sys.stderr, flush=True)
    ensemble = None
    if not a.no_ensemble:
        groups = [(256, 48), (512, 32), (1024, 16)]
        sites = sum(n * n * c for n, c in groups)
        ensemble = {"groups": [{"N": n, "count": c} for n, c in groups], "sites": sites,
                    "coupling": "fixed", "step": 1e-4, "noise_space": a.noise_space}
        for fp32, name in ((True, "mixed"), (False, "fp64")):
            sps_e, lays, mode = time_batch(groups, fp32, jw.FIXED_COUPLING, 1e-4, None, 20, 10, 2000, ks)
            ensemble[name] = {"steps_per_s": round(sps_e, 3), "site_updates_per_s": round(sps_e * sites, 1),
                              "layouts": lays, "launch": mode}
            print(f"bench: ensemble {name}: {sps_e:.1f} steps/s = {sps_e * sites:.3e} site-updates/s "
                  f"(layouts {lays}, launch mode {mode})", file=sys.stderr, flush=True)
        ensemble["mixed_over_fp64"] = round(ensemble["mixed"]["steps_per_s"] / ensemble["fp64"]["steps_per_s"], 3)

    # the FP64 production run: 16 seeds of N = 1024 (MV start Ny = 50, g^2 mu a = 0.25,
    # m a = 1e-4), fixed coupling, Delta s = 1e-4, advanced together as one batch
    production = None
    if not a.no_production:
        groups = [(1024, 16)]
        sites = sum(n * n * c for n, c in groups)
        sps_p, lays, mode = time_batch(groups, False, jw.FIXED_COUPLING, 1e-4, None, 20, 10, 1, ks)
        kind_p = "kspace" if ks else "white"
        gbs = reals_of(kind_p, 6) * 8 * sites * sps_p / 1e9
        production = {"groups": [{"N": n, "count": c} for n, c in groups], "sites": sites, "form": "fp64",
                      "coupling": "fixed", "step": 1e-4, "noise_space": a.noise_space,
                      "steps_per_s": round(sps_p, 3), "traj_steps_per_s": round(sps_p * 16, 2),
                      "site_updates_per_s": round(sps_p * sites, 1), "alg_GBs": round(gbs, 1),
                      "alg_frac": round(gbs / HBM_PEAK_GBS, 4), "reals_per_site": reals_of(kind_p, 6),
                      "layouts": lays, "launch": mode}
        print(f"bench: production fp64 16 x N=1024: {sps_p * 16:.1f} trajectory-steps/s = "
              f"{sps_p * sites:.3e} site-updates/s, {gbs:.0f} GB/s algorithmic (layouts {lays}, launch {mode})",
              file=sys.stderr, flush=True)

    # the large single lattice: N = 8192, MV start Ny = 100, g^2 mu a = 0.05,
    # correlated coupling, Delta s = 5e-5; both forms from the same start and
    # noise, timed over all their steps; S(r) and Q_s of the mixed form against FP64
    large = None
    if a.large_steps > 0:
        NL = 8192
        cl = jw.Context(NL, 0)
        al_l = jw.alpha(lambda_L=LAMBDA_A * NL, mu0_over_lambda=math.exp(2 * math.pi / (9 * ALPHA_MAX)))
        UL = cl.field()
        cl.philox(11, 0, 0)
        t0 = time.perf_counter()
        cl.mv_init(UL, 100, 1.0, 0.05, 1e-4 / 0.05)
        cl.sync()
        t_mv = time.perf_counter() - t0
        UL0 = UL.download()
        UL.close()
        cl.set_noise_space(jw.NOISE_MOMENTUM if ks else jw.NOISE_POSITION)
        large = {"N": NL, "sites": NL * NL, "mv": {"Ny": 100, "g2mu_a": 0.05, "m_a": 1e-4, "ms": round(t_mv * 1e3, 1)},
                 "coupling": "correlated", "step": 5e-5, "steps": a.large_steps, "noise_space": a.noise_space}
        SL, QL = {}, {}
        for name in ("fp64", "mixed"):
            F = cl.field(9, UL0)
            cl.philox(12, 0, 100)
            cl.sync()
            t0 = time.perf_counter()
            if name == "mixed":
                cl.evolve_mixed(F, coupling, jw.SIN_KERNEL, 5e-5, al_l, nsteps=a.large_steps, reunit_every=10)
            else:
                cl.evolve(F, coupling, jw.SIN_KERNEL, 5e-5, al_l, nsteps=a.large_steps)
            cl.sync()
            dt = time.perf_counter() - t0
            SL[name] = cl.dipole_from_corr(cl.measure_corr(F))
            r_, s_, c_ = jw.radial_bins(SL[name], NL, NL // 2)
            QL[name] = jw.saturation_scale(r_, s_, c_)
            F.close()
            bpr = 4 if name == "mixed" else 8
            comp = reals_of("kspace" if ks else kind, 6) * bpr * NL * NL
            large[name] = {"steps_per_s": round(a.large_steps / dt, 3),
                           "alg_GBs": round(comp * a.large_steps / dt / 1e9, 1),
                           "alg_frac": round(comp * a.large_steps / dt / (HBM_PEAK_GBS * 1e9), 4), "Qs": QL[name]}
            print(f"bench: N={NL} {name}: {a.large_steps} steps in {dt:.2f} s -> {a.large_steps / dt:.2f} steps/s, "
                  f"Q_s = {QL[name]:.5f}", file=sys.stderr, flush=True)
        # the matrix-core column stages: a short timed run
        F = cl.field(9, UL0)
        cl.philox(12, 0, 100)
        cl.set_mixed_fft(1)
        cl.evolve_mixed(F, coupling, jw.SIN_KERNEL, 5e-5, al_l, nsteps=2, reunit_every=10)
        cl.sync()
        t0 = time.perf_counter()
        nm = min(20, a.large_steps)
        cl.evolve_mixed(F, coupling, jw.SIN_KERNEL, 5e-5, al_l, nsteps=nm, reunit_every=10)
        cl.sync()
        large["mixed_mfma_steps_per_s"] = round(nm / (time.perf_counter() - t0), 3)
        cl.set_mixed_fft(0)
        F.close()
        large["sr_err_vs_fp64"] = float(np.abs(SL["fp64"] - SL["mixed"]).max())
        large["Qs_rel_err_vs_fp64"] = abs(QL["mixed"] - QL["fp64"]) / QL["fp64"] if QL["fp64"] else None
        large["bytes_per_step_note"] = ("alg_GBs: the compulsory bytes of the fused step (%d reals per site); the "
                                        "large-lattice step moves %d reals per site in FP32 and %d in FP64 (row "
                                        "transforms of A and B apart; FP64: also those of the noise and V)"
                                        % (reals_of("kspace" if ks else kind, 6), 176 if ks else 240,
                                           228 if ks else 292))
        print(f"bench: N={NL} S(r) mixed vs fp64 {large['sr_err_vs_fp64']:.3e}, Q_s rel err "
              f"{large['Qs_rel_err_vs_fp64']}", file=sys.stderr, flush=True)
        cl.close()

    mfma = None
    if "mixed_mfma_k" in res:
        mfma = {"form": "mixed_mfma_k: radix-16 column stages as split-BF16 products on v_mfma_f32_32x32x16_bf16",
                "steps_per_s": res["mixed_mfma_k"]["steps_per_s"],
                "column_ms_per_step": res["mixed_mfma_k"]["kernels"].get("col_kmult", {}).get("ms_per_step")}
        if os.path.exists(MFMA_FILE):
            with open(MFMA_FILE) as f:
                mfma["profile"] = json.load(f)

    # measured traffic of the dominant kernels, when a profile summary is present
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f)
        for form, r in res.items():
            roof = r["roofline"]
            key = f"{form}:{a.coupling}:{roof['kernel']}" if roof else None
            if key in traffic:
                roof["traffic"] = traffic[key]["bytes_per_launch"]
                roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, REPO)

    # S(r) of the mixed form against FP64 after the same sr-steps from the same start
    sr = None
    if a.sr_steps > 0 and f64 in forms and fmix in forms:
        S = {}
        for form in (f64, fmix):
            F = ctx.field(9, U0)
            ctx.philox(3, 0, 0)
            run(form, F, a.sr_steps)
            S[form] = ctx.dipole_from_corr(ctx.measure_corr(F))
            F.close()
        if "mixed_mfma_k" in forms and ks:       # the matrix-core form on the same stream
            F = ctx.field(9, U0)
            ctx.philox(3, 0, 0)
            run("mixed_mfma_k", F, a.sr_steps)
            S["mixed_mfma_k"] = ctx.dipole_from_corr(ctx.measure_corr(F))
            F.close()
        S0 = ctx.dipole_from_corr(ctx.measure_corr(ctx.field(9, U0)))
        r, s64, cnt = jw.radial_bins(S[f64], N, N // 2)
        _, s32, _ = jw.radial_bins(S[fmix], N, N // 2)
        sr = {"steps": a.sr_steps, "noise_space": a.noise_space, "max_abs_err": float(np.abs(S[f64] - S[fmix]).max()),
              "max_abs_err_radial": float(np.abs(s64 - s32)[cnt > 0].max()),
              "max_abs_change_from_start": float(np.abs(S[f64] - S0).max()), "bound": 2e-5}
        if "mixed_mfma_k" in S:
            sr["max_abs_err_mfma"] = float(np.abs(S[f64] - S["mixed_mfma_k"]).max())
        print(f"bench: S(r) after {a.sr_steps} steps: max |mixed - fp64| = {sr['max_abs_err']:.3e} "
              f"(evolution moved S by {sr['max_abs_change_from_start']:.3e})", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        orc = oracle.Oracle()
        if a.cpu_threads > 0:
            orc.set_num_threads(a.cpu_threads)
        nthr = orc.num_threads()
        rng = np.random.default_rng(0)
        oal = oracle.alpha(lambda_L=LAMBDA_A * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * ALPHA_MAX))) \
            if al is not None else None
        oc = {"correlated": oracle.CORRELATED, "fixed": oracle.FIXED, "noise": oracle.NOISE,
              "sqrt": oracle.SQRT}[a.coupling]

        def cpu_steps(n, as_shipped):
            orc.set_as_shipped(as_shipped)
            Uc = U0
            t0 = time.perf_counter()
            for _ in range(n):
                Uc = orc.langevin_step(Uc, rng.normal(size=V * 16), N, oc, oracle.SIN, step, oal)
            orc.set_as_shipped(0)
            return (time.perf_counter() - t0) / n

        model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        t_ship = cpu_steps(a.cpu_steps_shipped, 1)
        t_omp = cpu_steps(a.cpu_steps_omp, 0)
        cpu = {"value": round(1.0 / t_ship, 5), "unit": "steps/s", "cores": nthr, "kind": "port",
               "sample": f"{a.cpu_steps_shipped} Langevin step(s) at N={N} ({a.coupling} coupling) by the CPU oracle "
                         f"(the reference's algorithm in FP64: 9-component radix-2 FFTs, eigen-decomposition "
                         f"exponential) with the reference's threading: FFTs and noise on {nthr} OpenMP threads, "
                         f"uxiulocal/update_uf on one (field.h:1901, 1974); {t_ship:.1f} s per step",
               "all_omp": {"value": round(1.0 / t_omp, 5), "unit": "steps/s", "cores": nthr,
                           "sample": f"{a.cpu_steps_omp} step(s), every site loop on {nthr} threads, "
                                     f"{t_omp:.1f} s per step"},
               "cpu_model": model, "host_cpus": os.cpu_count(), "compiler": "g++ -O3 -fopenmp (oracle/Makefile)"}
        print(f"bench: cpu oracle step {t_ship:.2f} s as shipped, {t_omp:.2f} s all-OMP ({nthr} threads, {model})",
              file=sys.stderr, flush=True)

    head = fmix if fmix in res else forms[0]
    h = res[head]
    value = batched["traj_steps_per_s"] if batched else h["steps_per_s"]
    form_txt = (f"mixed+batched: {a.batch} trajectories of N={N} in one batch, FP32 fields, FFTs and products, FP64 "
                f"exponentials, SU(3) projection every {a.reunit_every} steps" if batched else
                ("mixed: FP32 fields, FFTs and products, FP64 exponentials, SU(3) projection every "
                 f"{a.reunit_every} steps; one trajectory") if head == fmix else "plain FP64") \
        + f"; {a.noise_space}-space noise"
    out = {"metric": METRIC, "value": value, "unit": "trajectory-steps/s" if batched else "steps/s",
           "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(1e3 / value, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32" if head == fmix else "f64",
           "data": "synthetic (Philox noise, MV initial condition)",
           "config": {"workload": f"N={N} Langevin step, {a.coupling} coupling"
                                  + (f" (alpha_max={ALPHA_MAX}, Lambda a={LAMBDA_A})" if a.coupling == "correlated" else "")
                                  + f", SIN kernel, ds={step}, MV start Ny=50 g2mu a=0.25 m a=1e-4",
                      "N": N, "sites": V,
                      "form": form_txt, "parallelism": "single GPU"},
           "noise_space": a.noise_space,
           "fp64_steps_per_s": res[f64]["steps_per_s"] if f64 in res else None,
           "mixed_steps_per_s": res[fmix]["steps_per_s"] if fmix in res else None,
           "forms_steps_per_s": {f: r["steps_per_s"] for f, r in res.items()},
           "mixed_batched": batched,
           "value_over_fp64": round(value / res[f64]["steps_per_s"], 3) if f64 in res else None,
           "mixed_over_fp64": round(res[fmix]["steps_per_s"] / res[f64]["steps_per_s"], 3)
           if fmix in res and f64 in res else None,
           "ensemble": ensemble,
           "production": production,
           "large": large,
           "mfma": mfma,
           "sr_err_vs_fp64": sr,
           "roofline": h["roofline"], "step_roofline": h["step_roofline"], "kernels": h["kernels"],
           "roofline_fp64": res[f64]["roofline"] if f64 in res else None,
           "step_roofline_fp64": res[f64]["step_roofline"] if f64 in res else None,
           "kernels_fp64": res[f64]["kernels"] if f64 in res else None,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
