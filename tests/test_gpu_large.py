"""GPU tests of the large-lattice step (N = 4096 and 8192): the fused column
passes with separate row transforms and site passes (jw_packed.hpp
step_large_n), which jw_evolve and jw_evolve_mixed run at these sizes.

  * FP64 against the unfused kernel sequence (jw_set_fused(ctx, 0): full 2-D
    transforms, one kernel per operation) on the same Philox stream, two steps
    of correlated noise: two GPU implementations of the step, per-site
    Frobenius <= 1e-11 (the 1/|k| infrared weight grows with N);
  * N = 4096, one injected FP64 step of correlated noise against the oracle,
    within 1e-12 or the spread of two FP64 CPU implementations;
  * momentum-space noise at N = 4096 against the oracle fed the same stream,
    and at N = 8192 the mixed form against FP64 on it;
  * unitarity and det U = 1 after the steps.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(8192)


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


def su3_defect(U, n=4096):
    idx = rng.choice(U.shape[0], size=n, replace=False)
    M = U[idx].reshape(-1, 3, 3)
    uni = np.abs(M @ M.conj().transpose(0, 2, 1) - np.eye(3)).max()
    det = np.abs(np.linalg.det(M) - 1).max()
    return uni, det


@pytest.mark.parametrize("N,mode", [(4096, 1), (4096, 8), (8192, 1)])
def test_large_step_against_unfused(jw, N, mode):
    print(f"large: N={N} fused mode {mode} vs unfused, FP64 correlated, 2 steps", flush=True)
    al_g, _ = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    out = {}
    for m in (mode, 0):
        W = c.field(9, U0)
        c.set_fused(m)
        c.philox(3, 1, 0)
        c.evolve(W, 4, 1, 5e-5, al_g, nsteps=2)
        out[m] = W.download()
        W.close()
    c.set_fused(1)
    c.close()
    err = frob_max(out[mode], out[0])
    print(f"large: |large - unfused| = {err:.2e}", flush=True)
    assert err <= 1e-11, err
    assert frob_max(out[0], U0) > 1e-3
    uni, det = su3_defect(out[mode])
    assert uni < 1e-12 and det < 1e-12, (uni, det)


def test_large_correlated_step_N4096_against_oracle(jw, orc):
    from test_gpu_production import numpy_step
    N = 4096
    print(f"large: N={N} correlated FP64 injected step vs oracle", flush=True)
    al_g, al_o = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    z = rng.normal(size=N * N * 16)
    c.inject(z)
    c.evolve(U, 4, 1, 5e-5, al_g, nsteps=1)
    got = U.download()
    c.close()
    ref = orc.langevin_step(U0, z, N, 4, 1, 5e-5, al_o)
    spread = frob_max(ref, numpy_step(orc, U0, z, N, 4, 5e-5, al_o))
    err = frob_max(got, ref)
    print(f"large: |hip - oracle| = {err:.2e}, |numpy - oracle| = {spread:.2e}", flush=True)
    assert err <= max(1e-12, spread), (err, spread)
    assert frob_max(ref, U0) > 1e-3


def test_large_kspace_N4096_against_oracle(jw, orc):
    N = 4096
    print(f"large: N={N} momentum-space noise, FP64 step vs oracle", flush=True)
    from test_gpu_production import numpy_step
    al_g, al_o = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c)
    U0 = U.download()
    c.set_noise_space(jw.NOISE_MOMENTUM)
    c.philox(4, 2, 9)
    c.evolve(U, 4, 1, 5e-5, al_g, nsteps=1)
    got = U.download()
    c.close()
    z = orc.kspace_normals(N, 4, 2, 9, jw.TAG_KSPACE)
    ref = orc.langevin_step(U0, z, N, 4, 1, 5e-5, al_o)
    spread = frob_max(ref, numpy_step(orc, U0, z, N, 4, 5e-5, al_o))
    err = frob_max(got, ref)
    print(f"large: |hip - oracle| = {err:.2e}, |numpy - oracle| = {spread:.2e}", flush=True)
    assert err <= max(1e-12, spread), (err, spread)


def test_large_kspace_N8192_mixed_against_fp64(jw):
    N = 8192
    print(f"large: N={N} momentum-space noise, mixed vs FP64, 3 steps", flush=True)
    al_g, _ = alphas(jw, N)
    c = jw.Context(N)
    U = mv_start(c, Ny=2)
    U0 = U.download()
    c.set_noise_space(jw.NOISE_MOMENTUM)
    out = []
    for mixed in (0, 1):
        W = c.field(9, U0)
        c.philox(6, 0, 0)
        if mixed:
            c.evolve_mixed(W, 4, 1, 5e-5, al_g, nsteps=3, reunit_every=10)
        else:
            c.evolve(W, 4, 1, 5e-5, al_g, nsteps=3)
        out.append(W.download())
        W.close()
    c.close()
    err = frob_max(out[0], out[1])
    print(f"large: |mixed - fp64| = {err:.2e}", flush=True)
    assert err <= 2e-5, err
    assert frob_max(out[0], U0) > 1e-3
    uni, det = su3_defect(out[0])
    assert uni < 1e-12 and det < 1e-12, (uni, det)
