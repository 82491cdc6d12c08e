"""jw_evolve_noalloc (the allocation-free evolution on the library's per-GPU
workspace) against jw_evolve on identical inputs: the same kernels on other
buffers, so the results are bit-identical.  Calls follow one another on one
GPU, each on a smaller lattice than the one before (the workspace keeps the
previous call's contents), through the C ABI."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _start(jw, c, seed):
    U = c.field()
    c.philox(seed, 0, 0)
    c.mv_init(U, 4, 1.0, 0.25, 1.0)
    U0 = U.download()
    U.close()
    return U0


def _pair(jw, c, U0, coupling, al, nsteps, z=None):
    out = []
    for noalloc in (False, True):
        F = c.field(9, U0)
        if z is not None:
            c.inject(z)
        else:
            c.philox(11, 3, 7)
        fn = c.evolve_noalloc if noalloc else c.evolve
        fn(F, coupling, jw.SIN_KERNEL, 5e-5, al, nsteps=nsteps)
        out.append(F.download())
        F.close()
    return out


def test_noalloc_matches_evolve_successive_sizes(jw):
    rng = np.random.default_rng(7)
    for N in (512, 256, 64, 32):
        print(f"noalloc: N={N}", flush=True)
        c = jw.Context(N)
        U0 = _start(jw, c, 100 + N)
        al = jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))
        for coupling, a in ((jw.CORRELATED_NOISE_COUPLING, al), (jw.FIXED_COUPLING, None),
                            (jw.SQRT_COUPLING_CONSTANT, al), (jw.NOISE_COUPLING_CONSTANT, al)):
            ref, got = _pair(jw, c, U0, coupling, a, 3)
            assert np.array_equal(ref, got), (N, coupling)
            assert np.abs(ref - U0).max() > 1e-4
        # one injected step
        z = rng.normal(size=N * N * 16)
        ref, got = _pair(jw, c, U0, jw.CORRELATED_NOISE_COUPLING, al, 1, z=z)
        assert np.array_equal(ref, got), N
        c.close()


def test_noalloc_blocked_layout_and_limits(jw):
    N = 128
    c = jw.Context(N)
    U0 = _start(jw, c, 5)
    al = jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))
    c.set_fused(c.FUSED_BLOCK8)          # column pass reads the 16-row-blocked table
    ref, got = _pair(jw, c, U0, jw.CORRELATED_NOISE_COUPLING, al, 2)
    assert np.array_equal(ref, got)
    F = c.field(9, U0)
    with pytest.raises(jw.JwError) as e:
        c.evolve_noalloc(F, jw.FIXED_COUPLING, jw.LINEAR_KERNEL, 1e-4, None, 1)
    assert e.value.code == -4
    F.close()
    c.close()
    big = jw.Context(1024)
    F = big.field()
    with pytest.raises(jw.JwError) as e:
        big.evolve_noalloc(F, jw.FIXED_COUPLING, jw.SIN_KERNEL, 1e-4, None, 1)
    assert e.value.code == -1
    F.close()
    big.close()
