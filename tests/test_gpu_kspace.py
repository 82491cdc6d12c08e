"""GPU tests of the momentum-space noise (jw_set_noise_space, NOISE_MOMENTUM).

The fused step then draws the spectra of the noise planes per mode in its first
column pass and never forms the position-space normals.  The oracle forms them
(or_kspace_normals: the same Philox words, Box-Muller, a unitary inverse FFT)
and runs its ordinary step on them, so the two agree to the rounding of the
transforms:

  * FP64, N = 64 and 256, SQRT / FIXED / correlated coupling: three steps
    against the oracle, per-site Frobenius <= 1e-12;
  * every intermediate layout of the fused step at N = 256 against the oracle;
  * N = 2048, correlated coupling (the benchmark's configuration): one FP64
    step against the oracle within 1e-12 or the spread of two FP64 CPU
    implementations (the infrared conditioning at this size);
  * the mixed form against FP64 on the same stream (FP32 rounding), and the
    batched form against single lattices (same arithmetic: 1e-13);
  * the unsupported cases fail with an error: NOISE coupling, the unfused
    path, a pending injection.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(4242)


def frob_max(a, b):
    d = (a - b).reshape(-1, 9)
    return np.sqrt((abs(d) ** 2).sum(axis=1)).max()


def alphas(jw, N, alpha_max=0.7, lam=0.02):
    import oracle
    kw = dict(lambda_L=lam * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * alpha_max)))
    return jw.alpha(**kw), oracle.alpha(**kw)


def mv_start(c, Ny=4, g2mu=0.25, seed=1):
    U = c.field()
    c.philox(seed, 0, 0)
    c.mv_init(U, Ny, 1.0, g2mu, 1e-4 / g2mu)
    return U


def oracle_steps(jw, orc, U0, N, coupling, ds, al_o, seed, traj, step0, nsteps):
    U = U0
    for s in range(step0, step0 + nsteps):
        z = orc.kspace_normals(N, seed, traj, s, jw.TAG_KSPACE)
        U = orc.langevin_step(U, z, N, coupling, 1, ds, al_o)
    return U


@pytest.mark.parametrize("N", [64, 256])
@pytest.mark.parametrize("coupling", [0, 3, 4])
def test_kspace_steps_against_oracle(jw, orc, N, coupling):
    print(f"kspace: N={N} coupling={coupling} FP64 three steps", flush=True)
    ds = 5e-5 if coupling == 4 else 1e-4
    al_g, al_o = alphas(jw, N) if coupling != 3 else (None, None)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    c.set_noise_space(jw.NOISE_MOMENTUM)
    c.philox(21, 3, 5)
    c.evolve(U, coupling, 1, ds, al_g, nsteps=3)
    got = U.download()
    c.close()
    ref = oracle_steps(jw, orc, U0, N, coupling, ds, al_o, 21, 3, 5, 3)
    err = frob_max(got, ref)
    print(f"kspace: |hip - oracle| = {err:.2e}", flush=True)
    assert err <= 1e-12, err
    assert frob_max(ref, U0) > 1e-3


# jw_set_fused modes: 2 row-major, 3 and 4 blocked rows, 5 one plane of 64-byte
# strips, 6 pair-interleaved, 7 one plane of 32-byte strips
@pytest.mark.parametrize("mode", [2, 3, 4, 5, 6, 7])
def test_kspace_layouts_against_oracle(jw, orc, mode):
    N = 256
    print(f"kspace: N={N} fused mode {mode}", flush=True)
    al_g, al_o = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    c.set_fused(mode)
    c.set_noise_space(jw.NOISE_MOMENTUM)
    c.philox(8, 1, 0)
    c.evolve(U, 4, 1, 5e-5, al_g, nsteps=2)
    got = U.download()
    c.close()
    ref = oracle_steps(jw, orc, U0, N, 4, 5e-5, al_o, 8, 1, 0, 2)
    assert frob_max(got, ref) <= 1e-12


def test_kspace_correlated_step_N2048_against_oracle(jw, orc):
    """One FP64 step at the benchmark's configuration.  Bound as for the
    position-space step at this size: 1e-12, or the spread between the oracle
    and a second FP64 implementation (numpy's FFT) on the same normals."""
    from test_gpu_production import numpy_step
    N = 2048
    print(f"kspace: N={N} correlated FP64 step", flush=True)
    al_g, al_o = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    c.set_noise_space(jw.NOISE_MOMENTUM)
    c.philox(5, 0, 17)
    c.evolve(U, 4, 1, 5e-5, al_g, nsteps=1)
    got = U.download()
    c.close()
    z = orc.kspace_normals(N, 5, 0, 17, jw.TAG_KSPACE)
    ref = orc.langevin_step(U0, z, N, 4, 1, 5e-5, al_o)
    spread = frob_max(ref, numpy_step(orc, U0, z, N, 4, 5e-5, al_o))
    err = frob_max(got, ref)
    print(f"kspace: |hip - oracle| = {err:.2e}, |numpy - oracle| = {spread:.2e}", flush=True)
    assert err <= max(1e-12, spread), (err, spread)
    assert frob_max(ref, U0) > 1e-3


@pytest.mark.parametrize("N", [256, 2048])
def test_kspace_mixed_against_fp64(jw, N):
    """The FP32 form draws the same words (FP32 Box-Muller): ten steps agree with
    FP64 to FP32 rounding, as the position-space streams do."""
    print(f"kspace: N={N} mixed vs FP64, 10 steps", flush=True)
    al_g, _ = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    c.set_noise_space(jw.NOISE_MOMENTUM)
    out = []
    for mixed in (0, 1):
        W = c.field(9, U0)
        c.philox(9, 2, 0)
        if mixed:
            c.evolve_mixed(W, 4, 1, 5e-5, al_g, nsteps=10, reunit_every=10)
        else:
            c.evolve(W, 4, 1, 5e-5, al_g, nsteps=10)
        out.append(W.download())
    c.close()
    err = frob_max(out[0], out[1])
    print(f"kspace: |mixed - fp64| = {err:.2e}", flush=True)
    assert err <= 2e-5, err
    assert frob_max(out[0], U0) > 1e-3


@pytest.mark.parametrize("fp32", [1, 0])
def test_kspace_batch_matches_single(jw, fp32):
    groups = [(64, 2), (128, 1)]
    print(f"kspace: batch {groups} fp32={fp32}", flush=True)
    ctxs = {N: jw.Context(N) for N, _ in groups}
    states = []
    for N, count in groups:
        for _ in range(count):
            states.append((N, mv_start(ctxs[N], seed=100 + len(states)).download()))
    b = jw.Batch(groups, fp32=fp32)
    for g in range(len(groups)):
        b.set_layout(g, 0)
    b.set_noise_space(jw.NOISE_MOMENTUM)
    for t, (N, U0) in enumerate(states):
        b.set(t, ctxs[N].field(9, U0))
    b.philox(7, 5, 0)
    b.evolve(4 if fp32 else 3, 1, 5e-5 if fp32 else 1e-4, alphas(jw, 64)[0] if fp32 else None, nsteps=12,
             reunit_every=10)
    b.sync()
    for t, (N, U0) in enumerate(states):
        c = ctxs[N]
        c.set_noise_space(jw.NOISE_MOMENTUM)
        got = c.field()
        b.get(t, got)
        U = c.field(9, U0)
        c.philox(7, 5 + t, 0)
        if fp32:
            c.evolve_mixed(U, 4, 1, 5e-5, alphas(jw, 64)[0], nsteps=12, reunit_every=10)
        else:
            c.evolve(U, 3, 1, 1e-4, nsteps=12)
        assert frob_max(got.download(), U.download()) <= 1e-13, t
    b.close()
    for c in ctxs.values():
        c.close()


def test_kspace_unsupported_cases(jw):
    c = jw.Context(64)
    U = mv_start(c)
    c.set_noise_space(jw.NOISE_MOMENTUM)
    al_g, _ = alphas(jw, 64)
    with pytest.raises(jw.JwError):
        c.evolve(U, jw.NOISE_COUPLING_CONSTANT, 1, 1e-4, al_g, nsteps=1)
    c.inject(np.zeros(64 * 64 * 16))
    with pytest.raises(jw.JwError):
        c.evolve(U, 3, 1, 1e-4, nsteps=1)
    c.set_noise_space(jw.NOISE_POSITION)
    c.evolve(U, 3, 1, 1e-4, nsteps=1)         # the injection is consumed in position space
    c.set_noise_space(jw.NOISE_MOMENTUM)
    c.set_fused(0)
    with pytest.raises(jw.JwError):
        c.evolve(U, 3, 1, 1e-4, nsteps=1)
    with pytest.raises(jw.JwError):
        c.evolve(U, 3, 0, 1e-4, nsteps=1)     # LINEAR kernel
    c.close()
