"""GPU tests of the mixed-precision form (jw_evolve_mixed): FP32 fields and
FFTs, FP64 exponentials and SU(3) projection, against the FP64 form on the
same noise.

Tolerances:
  * one step from the same U and injected noise, or the same Philox stream
    (FP32 Box-Muller from the same Philox words): per-site Frobenius <= 1e-6
    (FP32 storage of U alone is ~1e-7);
  * 1000 steps from the same MV start and Philox stream: dipole S(r) within
    2e-5 absolute of the FP64 form (the acceptance bound of the mixed form);
  * the intermediate layouts agree with each other to FP32 rounding;
  * with projection, U is unitary with unit determinant to FP32 rounding.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(11)


def frob_max(a, b):
    d = (a - b).reshape(-1, 9)
    return np.sqrt((abs(d) ** 2).sum(axis=1)).max()


def mv_start(jw, c, Ny=10):
    U = c.field()
    c.philox(3, 0, 0)
    c.mv_init(U, Ny, 1.0, 0.25, 1e-4 / 0.25)
    return U.download()


def alpha_for(jw, N, coupling):
    if coupling not in (0, 1, 4):
        return None
    return jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))


@pytest.mark.parametrize("coupling", [3, 1, 0, 4])
def test_mixed_one_step_injected(jw, coupling):
    N = 64
    c = jw.Context(N)
    U0 = mv_start(jw, c)
    z = rng.normal(size=N * N * 16)
    al = alpha_for(jw, N, coupling)
    ds = 5e-5 if coupling == 4 else 1e-4
    U64, U32 = c.field(9, U0), c.field(9, U0)
    c.inject(z)
    c.evolve(U64, coupling, 1, ds, al, nsteps=1)
    c.inject(z)
    c.evolve_mixed(U32, coupling, 1, ds, al, nsteps=1, reunit_every=0)
    a, b = U64.download(), U32.download()
    assert frob_max(a, b) <= 1e-6, frob_max(a, b)
    assert frob_max(a, U0) > 1e-3
    c.close()


@pytest.mark.parametrize("N,coupling", [(64, 3), (512, 4), (2048, 4)])
def test_mixed_one_step_philox(jw, N, coupling):
    """One step from the same U with the Philox stream: the FP32 Box-Muller
    normals (same Philox words) and the fused FP32 passes against FP64."""
    c = jw.Context(N)
    U0 = mv_start(jw, c, Ny=2)
    al = alpha_for(jw, N, coupling)
    ds = 5e-5 if coupling == 4 else 1e-4
    U64, U32 = c.field(9, U0), c.field(9, U0)
    c.philox(17, 3, 40)
    c.evolve(U64, coupling, 1, ds, al, nsteps=1)
    c.philox(17, 3, 40)
    c.evolve_mixed(U32, coupling, 1, ds, al, nsteps=1, reunit_every=0)
    a, b = U64.download(), U32.download()
    assert frob_max(a, b) <= 1e-6, frob_max(a, b)
    assert frob_max(a, U0) > 1e-3
    c.close()


@pytest.mark.parametrize("N,coupling", [(64, 3), (64, 4), (256, 3), (256, 4)])
def test_mixed_dipole_after_1000_steps(jw, N, coupling):
    """Same MV start, same Philox stream: S(r) of the two forms after 1000 steps."""
    c = jw.Context(N)
    U0 = mv_start(jw, c)
    al = alpha_for(jw, N, coupling)
    ds = 5e-5 if coupling == 4 else 1e-4
    U64, U32 = c.field(9, U0), c.field(9, U0)
    c.philox(21, 4, 0)
    c.evolve(U64, coupling, 1, ds, al, nsteps=1000)
    c.philox(21, 4, 0)
    c.evolve_mixed(U32, coupling, 1, ds, al, nsteps=1000, reunit_every=10)
    S64 = c.dipole_from_corr(c.measure_corr(U64))
    S32 = c.dipole_from_corr(c.measure_corr(U32))
    S0 = c.dipole_from_corr(c.measure_corr(c.field(9, U0)))
    err = np.abs(S64 - S32).max()
    assert err <= 2e-5, err
    assert np.abs(S64 - S0).max() > 1e-3          # the evolution moved S(r)
    c.close()


@pytest.mark.parametrize("N", [64, 512, 2048])
@pytest.mark.parametrize("coupling,space", [(3, 0), (4, 0), (3, 1), (4, 1)])
def test_mixed_layouts_agree(jw, N, coupling, space):
    """Every intermediate layout of the mixed form gives the row-major result to FP32
    rounding, with position- and momentum-space noise (the pair-interleaved layout
    runs its own row pass for the correlated and drawn noise)."""
    c = jw.Context(N)
    U0 = mv_start(jw, c, Ny=2)
    al = alpha_for(jw, N, coupling)
    c.set_noise_space(space)
    out = {}
    for mode in (c.FUSED_ROWMAJOR, c.FUSED_BLOCK4, c.FUSED_BLOCK8, c.FUSED_ROWMAJOR_WIDE, c.FUSED_INTERLEAVED,
                 c.FUSED_ROWMAJOR_NARROW, c.FUSED):
        c.set_fused(mode)
        U = c.field(9, U0)
        c.philox(8, 0, 0)
        c.evolve_mixed(U, coupling, 1, 1e-4, al, nsteps=3, reunit_every=0)
        out[mode] = U.download()
    ref = out[c.FUSED_ROWMAJOR]
    for mode in (c.FUSED_BLOCK4, c.FUSED_BLOCK8, c.FUSED_ROWMAJOR_WIDE, c.FUSED_INTERLEAVED,
                 c.FUSED_ROWMAJOR_NARROW, c.FUSED):
        assert frob_max(ref, out[mode]) <= 2e-6, mode
    assert frob_max(ref, U0) > 1e-3
    c.set_fused(c.FUSED)
    c.close()


def test_mixed_projection_keeps_su3(jw):
    N = 128
    c = jw.Context(N)
    U0 = mv_start(jw, c)
    for reunit, tol in ((10, 3e-7), (0, None)):
        U = c.field(9, U0)
        c.philox(5, 0, 0)
        c.evolve_mixed(U, 3, 1, 1e-4, None, nsteps=40, reunit_every=reunit)
        m = U.download().reshape(-1, 3, 3)
        uni = np.abs(np.einsum("sij,skj->sik", m, m.conj()) - np.eye(3)).max()
        det = np.abs(np.linalg.det(m) - 1).max()
        if tol is not None:
            assert uni <= tol and det <= tol, (uni, det)
        else:
            assert uni < 1e-4      # without projection: FP32 drift, still small over 40 steps
    c.close()


@pytest.mark.parametrize("N", [32, 64])
def test_mixed_unfused_passes(jw, N):
    """The unfused FP32 passes (N outside the fused sizes, or jw_set_fused(0)):
    one injected step against FP64, and against the fused FP32 passes."""
    c = jw.Context(N)
    U0 = mv_start(jw, c, Ny=4)
    z = rng.normal(size=N * N * 16)
    U64 = c.field(9, U0)
    c.inject(z)
    c.evolve(U64, 4, 1, 5e-5, alpha_for(jw, N, 4), nsteps=1)
    out = {}
    for mode in ((c.UNFUSED, c.FUSED) if N >= 64 else (c.FUSED,)):
        c.set_fused(mode)
        U = c.field(9, U0)
        c.inject(z)
        c.evolve_mixed(U, 4, 1, 5e-5, alpha_for(jw, N, 4), nsteps=1, reunit_every=1)
        out[mode] = U.download()
        assert frob_max(out[mode], U64.download()) <= 1e-6, mode
    if len(out) == 2:
        assert frob_max(out[c.UNFUSED], out[c.FUSED]) <= 1e-6
    c.set_fused(c.FUSED)
    c.close()


def test_mixed_large_lattice(jw):
    """N = 4096 (unfused FP32 passes): S(r) after 20 steps against FP64 from the
    same MV start and Philox stream, compared on the device's correlator."""
    N = 4096
    c = jw.Context(N)
    U0 = c.field()
    c.philox(3, 0, 0)
    c.mv_init(U0, 4, 1.0, 0.05, 1.0)
    al = alpha_for(jw, N, 4)
    U64, U32 = c.field(), c.field()
    jw.lib().jw_field_copy(U64.h, U0.h)
    jw.lib().jw_field_copy(U32.h, U0.h)
    c.philox(9, 0, 0)
    c.evolve(U64, 4, 1, 5e-5, al, nsteps=20)
    c.philox(9, 0, 0)
    c.evolve_mixed(U32, 4, 1, 5e-5, al, nsteps=20, reunit_every=10)
    S64 = c.dipole_from_corr(c.measure_corr(U64))
    S32 = c.dipole_from_corr(c.measure_corr(U32))
    S0 = c.dipole_from_corr(c.measure_corr(U0))
    assert np.abs(S64 - S32).max() <= 2e-6
    assert np.abs(S64 - S0).max() > 1e-6
    c.close()


def test_mixed_rejects_unsupported(jw):
    c = jw.Context(64)
    U = c.field()
    with pytest.raises(jw.JwError) as e:
        c.evolve_mixed(U, 3, 0, 1e-4, None, 1)     # LINEAR kernel
    assert e.value.code == -4
    with pytest.raises(jw.JwError):
        c.evolve_mixed(U, 2, 1, 1e-4, None, 1)     # HATTA
    c.close()


# ------------------------------------------------------------ split-BF16 matrix-core stages
@pytest.mark.parametrize("N,coupling", [(64, 3), (256, 4), (2048, 4)])
def test_mixed_mfma_fft_one_step(jw, N, coupling):
    """The column transforms' radix-16 stages on the matrix cores (split BF16, three
    products, FP32 accumulation; jw_set_mixed_fft(1)): one step from the same U and
    Philox stream against FP64.  Each such stage carries about 2^-17 relative error,
    so the bound is 2e-5 per site (the FP32 butterflies: 1e-6); the FP32-butterfly
    result must lie closer to FP64 than this one does."""
    c = jw.Context(N)
    U0 = mv_start(jw, c, Ny=2)
    al = alpha_for(jw, N, coupling)
    ds = 5e-5 if coupling == 4 else 1e-4
    res = {}
    for name in ("fp64", "fp32", "mfma"):
        F = c.field(9, U0)
        c.philox(17, 3, 40)
        c.set_mixed_fft(1 if name == "mfma" else 0)
        if name == "fp64":
            c.evolve(F, coupling, 1, ds, al, nsteps=1)
        else:
            c.evolve_mixed(F, coupling, 1, ds, al, nsteps=1, reunit_every=0)
        res[name] = F.download()
        F.close()
    c.set_mixed_fft(0)
    e_mfma, e_fp32 = frob_max(res["fp64"], res["mfma"]), frob_max(res["fp64"], res["fp32"])
    print(f"mfma one step N={N}: |mfma - fp64| {e_mfma:.2e}, |fp32 - fp64| {e_fp32:.2e}", flush=True)
    assert e_mfma <= 2e-5, e_mfma
    assert e_fp32 < e_mfma          # the matrix-core stages did run
    assert frob_max(res["fp64"], U0) > 1e-3
    c.close()


@pytest.mark.parametrize("N", [256, 2048])
def test_mixed_mfma_fft_dipole_1000_steps(jw, N):
    """1000 correlated steps (alpha_max = 0.7, Lambda a = 0.02, ds = 5e-5) of the
    mixed form with matrix-core column stages from the same MV start and Philox
    stream as FP64: S(r) within 2e-5 absolute."""
    c = jw.Context(N)
    U0 = mv_start(jw, c, Ny=10)
    al = alpha_for(jw, N, 4)
    S = {}
    for name in ("fp64", "mfma"):
        F = c.field(9, U0)
        c.philox(21, 4, 0)
        if name == "fp64":
            c.evolve(F, 4, 1, 5e-5, al, nsteps=1000)
        else:
            c.set_mixed_fft(1)
            c.evolve_mixed(F, 4, 1, 5e-5, al, nsteps=1000, reunit_every=10)
            c.set_mixed_fft(0)
        S[name] = c.dipole_from_corr(c.measure_corr(F))
        F.close()
    S0 = c.dipole_from_corr(c.measure_corr(c.field(9, U0)))
    err = np.abs(S["fp64"] - S["mfma"]).max()
    print(f"mfma 1000 steps N={N}: max |S_mfma - S_fp64| = {err:.2e}", flush=True)
    assert err <= 2e-5, err
    assert np.abs(S["fp64"] - S0).max() > 1e-3
    c.close()
