"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on the
same injected inputs.

Tolerances (per-site max-abs over matrix entries unless stated):
  * bit-exact (==) for the per-function kernels whose arithmetic is reproduced
    operation for operation: lambda-basis noise, prepare_A_local (NOISE/FIXED),
    uxiulocal, setMVModel, solvePoisson, trace, average_and_symmetrize, reduce;
  * 1e-14 relative for the 2-D FFT (two correct FP64 transforms);
  * 5e-15 x max(1, |X|) for the SU(3) exponential (Cayley-Hamilton here, the
    reference's eigen-decomposition in the oracle);
  * per-site Frobenius norm of U_gpu - U_ref <= 1e-12 after one Langevin step
    and <= 1e-9 after 1000 steps (N = 64, injected identical noise).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(7)


def frob_max(a, b):
    d = (a - b).reshape(-1, 9)
    return np.sqrt((abs(d) ** 2).sum(axis=1)).max()


def rand_c(shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.fixture(scope="module")
def ctx64(jw):
    c = jw.Context(64)
    yield c
    c.close()


def random_su3(orc, n, scale=2.0):
    out = np.zeros((n, 9), complex)
    for i in range(n):
        a = rand_c((3, 3))
        h = (a + a.conj().T) / 2
        h -= np.trace(h) / 3 * np.eye(3)
        out[i] = orc.su3_exponentiate(h * rng.uniform(0.05, scale), 1.0).reshape(9)
    return out


# ------------------------------------------------------------------ SU(3) exp
LT = None


def gen_matrix(h):
    X = np.zeros((3, 3), complex)
    d8 = h[7] / np.sqrt(3)
    X[0, 0], X[1, 1], X[2, 2] = h[2] + d8, -h[2] + d8, -2 * d8
    X[0, 1], X[0, 2], X[1, 2] = h[0] + 1j * h[1], h[3] + 1j * h[4], h[5] + 1j * h[6]
    X[1, 0], X[2, 0], X[2, 1] = np.conj(X[0, 1]), np.conj(X[0, 2]), np.conj(X[1, 2])
    return X


def test_su3_exp_against_oracle(jw, orc, ctx64):
    hs = [rng.normal(size=8) * s for s in np.geomspace(1e-9, 50, 400)]
    hs += [np.zeros(8), np.eye(8)[2] * 0.7, np.eye(8)[7] * 1.3, np.eye(8)[7] * 1e-7,
           np.eye(8)[0] * 2.0, np.eye(8)[2] * 1.0 + np.eye(8)[7] * np.sqrt(3) * 1.0]   # degenerate spectra
    hs = np.array(hs)
    for s in (1.0, -0.37):
        E = ctx64.su3_exp_gen(hs, s)
        for h, e in zip(hs, E):
            X = gen_matrix(h) * s
            ref = orc.su3_exponentiate(X, 1.0)
            tol = 5e-15 * max(1.0, np.abs(X).max()) * 4
            assert np.abs(e - ref).max() <= tol, (h, np.abs(e - ref).max())
            assert np.abs(e @ e.conj().T - np.eye(3)).max() < 1e-13 * max(1.0, np.abs(X).max())


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("N", [4, 8, 16, 64, 256, 1024])
def test_fft2d_against_oracle(jw, orc, N):
    c = jw.Context(N)
    f = rand_c((N * N, 9))
    F = c.field(9, f)
    for d in (1, 0):
        c.execute2D(F, d)
        f = orc.fft2d(f, N, d)
        g = F.download()
        assert np.abs(g - f).max() <= 1e-14 * np.abs(f).max() * np.log2(N)
    c.close()


def test_fft_delta_known_answer(jw):
    # tests/test_fftw_execute.cpp:59-72: delta -> constant 1/sqrt(V)
    N = 64
    c = jw.Context(N)
    f = np.zeros((N * N, 1), complex)
    f[0] = 1
    F = c.field(1, f)
    c.execute2D(F, 0)
    assert np.abs(F.download() - 1.0 / N).max() < 1e-17
    c.close()


# ------------------------------------------------------------------ per-function kernels
def test_generate_gaussian_bit_exact(jw, orc, ctx64):
    N = 64
    z = rng.normal(size=N * N * 16)
    xx, xy = ctx64.field(), ctx64.field()
    ctx64.inject(z)
    ctx64.generate_gaussian(xx, xy)
    ox, oy = orc.generate_gaussian(N, z)
    assert np.array_equal(xx.download(), ox) and np.array_equal(xy.download(), oy)
    ctx64.inject(z)
    ctx64.generate_gaussian_with_noise_coupling_constant(xx, xy)
    ox, oy = orc.generate_gaussian_noise_coupling(N, z)
    # sqrt(alpha) through the device's log/pow: a few ulp
    assert np.allclose(xx.download(), ox, rtol=1e-14, atol=0)
    assert np.allclose(xy.download(), oy, rtol=1e-14, atol=0)


def test_generate_requires_matching_injection(jw, ctx64):
    xx, xy = ctx64.field(), ctx64.field()
    ctx64.inject(np.zeros(64 * 64 * 8), per_site=8)
    with pytest.raises(jw.JwError):
        ctx64.generate_gaussian(xx, xy)
    ctx64.philox(1)


@pytest.mark.parametrize("coupling", [0, 1, 3])
@pytest.mark.parametrize("kernel", [0, 1])
def test_prepare_A_local(jw, orc, ctx64, coupling, kernel):
    N = 64
    fx, fy = rand_c((N * N, 9)), rand_c((N * N, 9))
    A, X, Y = ctx64.field(), ctx64.field(9, fx), ctx64.field(9, fy)
    ctx64.prepare_A_local(A, X, Y, coupling, kernel)
    ref = orc.prepare_A_local(fx, fy, N, coupling, kernel)
    if coupling == 0:   # alpha through the device's log/pow: last-bit differences
        assert np.allclose(A.download(), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    else:
        assert np.array_equal(A.download(), ref)


def test_prepare_A_local_rejects_hatta(jw, ctx64):
    A = ctx64.field()
    with pytest.raises(jw.JwError) as e:
        ctx64.prepare_A_local(A, A, A, 2, 1)
    assert e.value.code == -4


def test_uxiulocal_bit_exact(jw, orc, ctx64):
    N = 64
    U = random_su3(orc, N * N)
    z = rng.normal(size=N * N * 16)
    ox, oy = orc.generate_gaussian(N, z)
    vx, vy = ctx64.field(), ctx64.field()
    ctx64.uxiulocal(vx, vy, ctx64.field(9, U), ctx64.field(9, ox), ctx64.field(9, oy))
    rx, ry = orc.uxiulocal(U, ox, oy, N)
    assert np.array_equal(vx.download(), rx) and np.array_equal(vy.download(), ry)


def test_update_uf(jw, orc, ctx64):
    N = 64
    U = random_su3(orc, N * N)
    z = rng.normal(size=N * N * 16)
    ox, oy = orc.generate_gaussian(N, z)
    A = orc.fft2d(orc.prepare_A_local(orc.fft2d(ox, N, 1), orc.fft2d(oy, N, 1), N, 1, 1), N, 0)
    B = orc.fft2d(orc.prepare_A_local(orc.fft2d(oy, N, 1), orc.fft2d(ox, N, 1), N, 1, 1), N, 0)
    # Exponent norms ~0.1 (ds = 1e-4) and ~2 (ds = 0.05).  The reference's
    # eigenvector exponential (restated by the oracle) is itself off by up to
    # ~1e-14 per entry on near-degenerate spectra (against 40-digit mpmath); the
    # device's Cayley-Hamilton form stays at ~2e-16, checked below on the sites
    # where the two disagree most.
    for step, tol in ((1e-4, 5e-14), (0.05, 1e-13)):
        Uf = ctx64.field(9, U)
        ctx64.update_uf(Uf, ctx64.field(9, B), ctx64.field(9, A), step)
        got, ref = Uf.download(), orc.update_uf(U, B, A, step, N)
        assert frob_max(got, ref) < tol
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    worst = np.argsort(np.abs(got - ref).max(axis=1))[-4:]
    ss = np.sqrt(0.05)

    def herm_upper(M):
        H = np.triu(M, 1)
        return H + H.conj().T + np.diag(M.diagonal().real)

    for i in worst:
        Xb = herm_upper(B[i].reshape(3, 3)) * ss
        Xa = herm_upper(A[i].reshape(3, 3)) * (-ss)
        Xb -= np.trace(Xb) / 3 * np.eye(3)
        Xa -= np.trace(Xa) / 3 * np.eye(3)
        e = mp.expm(mp.matrix(Xb.tolist()) * mp.mpc(0, 1)) * mp.matrix(U[i].reshape(3, 3).tolist()) * \
            mp.expm(mp.matrix(Xa.tolist()) * mp.mpc(0, 1))
        e = np.array([[complex(e[p, q]) for q in range(3)] for p in range(3)]).reshape(9)
        assert np.abs(got[i] - e).max() < 3e-15, (i, np.abs(got[i] - e).max(), np.abs(ref[i] - e).max())


def test_mv_pieces(jw, orc, ctx64):
    N = 64
    z = rng.normal(size=N * N * 8)
    f = ctx64.field()
    ctx64.inject(z, per_site=8)
    ctx64.set_mv_model(f, 1.0, 0.48, 50)
    ref = orc.set_mv_model(N, z, 1.0 * 0.48 / np.sqrt(50))
    assert np.array_equal(f.download(), ref)
    fk = orc.fft2d(ref, N, 1)
    F = ctx64.field(9, fk)
    ctx64.solve_poisson(F, 0.3, 1.0)
    assert np.array_equal(F.download(), orc.solve_poisson(fk, N, 0.3, 1.0))
    U = random_su3(orc, N * N)
    A = orc.fft2d(orc.solve_poisson(fk, N, 0.3, 1.0), N, 0)
    Uf = ctx64.field(9, U)
    ctx64.mul_exp(Uf, ctx64.field(9, A))
    assert frob_max(Uf.download(), orc.mul_exp(U, A, N)) < 1e-13


def test_measurement_pieces_bit_exact(jw, orc, ctx64):
    N = 64
    Uk = orc.fft2d(random_su3(orc, N * N), N, 1)
    c = ctx64.field(1)
    ctx64.trace(ctx64.field(9, Uk), c)
    rc = orc.trace(Uk, N)
    assert np.array_equal(c.download(), rc)
    ctx64.average_and_symmetrize(c)
    rs = orc.average_and_symmetrize(rc, N)
    assert np.array_equal(c.download(), rs)
    s, e = ctx64.field(1), ctx64.field(1)
    for _ in range(2):
        ctx64.reduce(c, s, e)
    z = np.zeros_like(rs)
    s1, e1 = orc.reduce(rs, z, z, N)
    s2, e2 = orc.reduce(rs, s1, e1, N)
    assert np.array_equal(s.download(), s2) and np.array_equal(e.download(), e2)


# ------------------------------------------------------------------ Philox
def test_philox_stream_matches_oracle(jw, orc, ctx64):
    N = 64
    ctx64.philox(1234, 5, 17)
    xx, xy = ctx64.field(), ctx64.field()
    ctx64.generate_gaussian(xx, xy)
    z = orc.philox_normals(N, 1234, 5, 17, jw.TAG_LANGEVIN, 16)
    ox, oy = orc.generate_gaussian(N, z)
    assert np.abs(xx.download() - ox).max() < 1e-14
    assert np.abs(xy.download() - oy).max() < 1e-14


# ------------------------------------------------------------------ Langevin step (engine)
@pytest.mark.parametrize("coupling,kernel", [(3, 1), (1, 1), (0, 1), (4, 1), (1, 0), (3, 0), (0, 0)])
def test_evolve_one_step(jw, orc, ctx64, coupling, kernel):
    N = 64
    U0 = random_su3(orc, N * N, scale=3.0)
    z = rng.normal(size=N * N * 16)
    U = ctx64.field(9, U0)
    ctx64.inject(z)
    ctx64.evolve(U, coupling, kernel, step=1e-4, nsteps=1)
    ref = orc.langevin_step(U0, z, N, coupling, kernel, 1e-4)
    err = frob_max(U.download(), ref)
    assert err <= 1e-12, err
    assert frob_max(ref, U0) > 1e-3


@pytest.mark.parametrize("N", [64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("coupling", [3, 1, 0, 4])
def test_fused_matches_unfused(jw, orc, N, coupling):
    """The fused row/column passes, with each layout of their intermediate planes
    (row-major, 4-row and 8-row blocks), against the unfused kernel sequence;
    Philox noise, two steps."""
    c = jw.Context(N)
    U0 = c.field()
    c.philox(5, 1, 0)
    c.mv_init(U0, 2, 1.0, 0.25, 1.0)
    ref = U0.download()
    al = jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=3.0) if coupling in (0, 4) else None
    results = {}
    for mode in (c.UNFUSED, c.FUSED_ROWMAJOR, c.FUSED_BLOCK4, c.FUSED_BLOCK8, c.FUSED_ROWMAJOR_WIDE, c.FUSED_INTERLEAVED,
                 c.FUSED_ROWMAJOR_NARROW, c.FUSED):
        c.set_fused(mode)
        U = c.field(9, ref)
        c.philox(5, 1, 100)
        c.evolve(U, coupling, 1, 1e-4, al, nsteps=2)
        results[mode] = U.download()
    for mode in (c.FUSED_ROWMAJOR, c.FUSED_BLOCK4, c.FUSED_BLOCK8, c.FUSED_ROWMAJOR_WIDE, c.FUSED_INTERLEAVED,
                 c.FUSED_ROWMAJOR_NARROW, c.FUSED):
        assert frob_max(results[c.UNFUSED], results[mode]) < 1e-13, mode
    assert frob_max(results[c.UNFUSED], ref) > 1e-3
    c.close()


def test_evolve_zero_noise_is_identity(jw, orc, ctx64):
    """Zero noise leaves rows 0 and 1 of U as they were (exp(0) = 1 exactly); the fused
    passes keep only those rows and rebuild row 2 as conj(row 0 x row 1), which differs
    from the stored row 2 by U0's own distance from SU(3) (rounding)."""
    N = 64
    U0 = random_su3(orc, N * N)
    U = ctx64.field(9, U0)
    ctx64.inject(np.zeros(N * N * 16))
    ctx64.evolve(U, 3, 1, 1e-4, nsteps=1)
    got = U.download().reshape(-1, 3, 3)
    u0 = U0.reshape(-1, 3, 3)
    assert np.abs(got[:, :2] - u0[:, :2]).max() == 0.0
    row2 = np.conj(np.cross(u0[:, 0], u0[:, 1]))
    assert np.abs(got[:, 2] - row2).max() < 1e-15
    assert np.abs(got[:, 2] - u0[:, 2]).max() < 1e-13


def test_evolve_multi_step_injected(jw, orc, ctx64):
    """200 steps, FIXED coupling, identical injected noise each step."""
    N = 64
    U0 = random_su3(orc, N * N, scale=1.0)
    Ug, Uo = ctx64.field(9, U0), U0.copy()
    for s in range(200):
        z = orc.philox_normals(N, 99, 0, s, jw.TAG_LANGEVIN, 16)
        ctx64.inject(z)
        ctx64.evolve(Ug, 3, 1, 1e-4, nsteps=1)
        Uo = orc.langevin_step(Uo, z, N, 3, 1, 1e-4)
    assert frob_max(Ug.download(), Uo) <= 1e-10


@pytest.mark.parametrize("coupling", [3, 4])
def test_evolve_1000_steps(jw, orc, ctx64, coupling):
    """1000 steps from an MV-like start, FIXED and per-mode correlated noise
    (alpha frozen at 0.7, Lambda a = 0.02), Philox noise drawn on the device
    and by the oracle's restatement of the stream: per-site Frobenius <= 1e-9."""
    import math
    N, steps, ds = 64, 1000, 1e-4 if coupling == 3 else 5e-5
    al_g = jw.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7))) if coupling == 4 else None
    al_o = None
    if coupling == 4:
        import oracle
        al_o = oracle.alpha(lambda_L=0.02 * N, mu0_over_lambda=math.exp(2 * math.pi / (9 * 0.7)))
    U0 = random_su3(orc, N * N, scale=1.0)
    U = ctx64.field(9, U0)
    ctx64.philox(11, 2, 0)
    ctx64.evolve(U, coupling, 1, ds, al_g, nsteps=steps)
    Uo = U0
    for s in range(steps):
        Uo = orc.langevin_step(Uo, orc.philox_normals(N, 11, 2, s, jw.TAG_LANGEVIN, 16), N, coupling, 1, ds, al_o)
    err = frob_max(U.download(), Uo)
    assert err <= 1e-9, err
    assert frob_max(Uo, U0) > 1e-2


def test_evolve_philox_matches_oracle(jw, orc, ctx64):
    """Philox noise drawn on the device vs the oracle's restatement of the stream."""
    N = 64
    U0 = random_su3(orc, N * N, scale=1.0)
    U = ctx64.field(9, U0)
    ctx64.philox(42, 3, 0)
    ctx64.evolve(U, 1, 1, 1e-4, nsteps=5)
    Uo = U0
    for s in range(5):
        Uo = orc.langevin_step(Uo, orc.philox_normals(N, 42, 3, s, jw.TAG_LANGEVIN, 16), N, 1, 1, 1e-4)
    assert frob_max(U.download(), Uo) <= 1e-12


# ------------------------------------------------------------------ MV init, observables
def test_mv_init_against_oracle(jw, orc, ctx64):
    N, Ny, g = 64, 10, 1.0
    mu = 30.72 / N
    z = rng.normal(size=Ny * N * N * 8)
    U = ctx64.field()
    # IR mass m.a = 0.25 (mass in units of g^2 mu): away from the zero-mode floor
    ctx64.mv_init(U, Ny, g, mu, 0.25 / (g * g * mu), z)
    ref = orc.mv_init(N, z, Ny, g, mu, 0.25 / (g * g * mu))
    assert frob_max(U.download(), ref) <= 1e-12
    # the configured m.a = 1e-4: the 1/m^2 zero mode puts a constant colour
    # matrix of norm ~1e4 into every sheet; two FP64 paths then agree to ~1e-9
    ctx64.mv_init(U, Ny, g, mu, 1e-4 / (g * g * mu), z)
    ref = orc.mv_init(N, z, Ny, g, mu, 1e-4 / (g * g * mu))
    assert frob_max(U.download(), ref) <= 1e-7


@pytest.mark.parametrize("N,tol", [(64, 1e-12), (256, 3e-12)])
def test_mv_init_zero_mode_projected_against_oracle(jw, orc, N, tol):
    """m a = 1e-4 (Ny = 50, g^2 mu a = 0.25) with the zero mode projected out of each
    sheet: the start matches the oracle to 1e-12 per site at N = 64.  The remaining
    infrared modes are amplified by 1/phat2 ~ (N / 2 pi)^2, so the rounding of two
    correct FP64 paths grows with N: the oracle and a numpy-FFT composition of its
    per-site functions differ by 1.3e-13 at N = 64 and 7.5e-13 at N = 256 (50 sheets),
    hence 3e-12 there."""
    Ny, g, mu = 50, 1.0, 0.25
    c = jw.Context(N)
    c.set_mv_zero_mode(1)
    z = rng.normal(size=Ny * N * N * 8)
    U = c.field()
    c.mv_init(U, Ny, g, mu, 1e-4 / (g * g * mu), z)
    ref = orc.mv_init(N, z, Ny, g, mu, 1e-4 / (g * g * mu), zero_mode=1)
    got = U.download()
    assert frob_max(got, ref) <= tol
    # the per-function path honours the option too: one sheet through
    # set_mv_model -> execute2D -> solvePoisson -> execute2D -> operator*=
    c.inject(z[: 8 * N * N], per_site=8)
    f, W = c.field(), c.field()
    c.set_unit(W)
    c.set_mv_model(f, g, mu, Ny)
    c.execute2D(f, 1)
    c.solve_poisson(f, 1e-4 / (g * g * mu) * g * g * mu, g)
    c.execute2D(f, 0)
    c.mul_exp(W, f)
    ref1 = orc.mv_init(N, z[: 8 * N * N], 1, g, mu / math.sqrt(Ny), 1e-4 / (g * g * mu) * math.sqrt(Ny),
                       zero_mode=1)
    assert frob_max(W.download(), ref1) <= 1e-12
    c.set_mv_zero_mode(0)
    c.mv_init(U, Ny, g, mu, 1e-4 / (g * g * mu), z)
    assert frob_max(U.download(), got) > 1e-3      # the option changed the start
    c.close()


# ------------------------------------------------------------------ POSITION_EVOLUTION
@pytest.mark.parametrize("N", [16, 64])
@pytest.mark.parametrize("coupling", [0, 1, 3])
@pytest.mark.parametrize("kernel", [0, 1])
def test_prepare_A_and_B_local_against_oracle(jw, orc, N, coupling, kernel):
    """The O(V^2) position-space sums (field.h:2368-2519 at every site) against the
    oracle's restatement (pinned to the reference build); the summation orders
    differ, hence the relative bound."""
    c = jw.Context(N)
    U0 = random_su3(orc, N * N)
    xx, xy = orc.generate_gaussian(N, rng.normal(size=N * N * 16))
    A, B = c.field(), c.field()
    c.prepare_A_and_B_local(A, B, c.field(9, xx), c.field(9, xy), c.field(9, U0), coupling, kernel)
    Ao, Bo = orc.prepare_A_and_B_local(xx, xy, U0, N, coupling, kernel)
    assert np.abs(A.download() - Ao).max() <= 1e-13 * np.abs(Ao).max()
    assert np.abs(B.download() - Bo).max() <= 1e-13 * np.abs(Bo).max()
    c.close()


def test_position_evolution_steps_against_oracle(jw, orc):
    """Two POSITION_EVOLUTION steps (main_optimized.cpp:243-300) through the
    per-function calls on injected noise, NOISE coupling: white noise, per-mode
    correlation (jw_correlate_noise), the position-space sums, update_uf."""
    N, ds = 32, 1e-3
    c = jw.Context(N)
    U0 = random_su3(orc, N * N)
    U = c.field(9, U0)
    xx, xy, A, B = c.field(), c.field(), c.field(), c.field()
    Uo = U0
    for _ in range(2):
        z = rng.normal(size=N * N * 16)
        c.inject(z)
        c.generate_gaussian(xx, xy)
        c.correlate_noise(xx)
        c.correlate_noise(xy)
        c.prepare_A_and_B_local(A, B, xx, xy, U, 1, 1)
        c.update_uf(U, B, A, ds)
        ox, oy = orc.generate_gaussian(N, z)
        ox, oy = orc.correlate_noise(ox, N), orc.correlate_noise(oy, N)
        Ao, Bo = orc.prepare_A_and_B_local(ox, oy, Uo, N, 1, 1)
        Uo = orc.update_uf(Uo, Bo, Ao, ds, N)
    got = U.download()
    assert frob_max(got, Uo) <= 1e-12
    assert frob_max(Uo, U0) > 1e-3
    c.close()


def test_position_evolution_limits(jw):
    c = jw.Context(512)
    F = c.field()
    with pytest.raises(jw.JwError) as e:
        c.prepare_A_and_B_local(F, c.field(), F, F, F, 3, 1)
    assert e.value.code in (-1, -4)
    c.close()
    c = jw.Context(16)
    F, G = c.field(), c.field()
    with pytest.raises(jw.JwError) as e:
        c.prepare_A_and_B_local(F, G, c.field(), c.field(), c.field(), 2, 1)
    assert e.value.code == -4
    c.close()


def test_measure_and_dipole(jw, orc):
    N = 16
    c = jw.Context(N)
    U0 = random_su3(orc, N * N)
    corr = c.measure_corr(c.field(9, U0))
    ref = orc.measure_corr(U0, N)[:, 0]
    assert np.abs(corr - ref.real).max() <= 1e-14 * np.abs(ref).max()
    S = c.dipole_from_corr(corr)
    assert np.abs(S - orc.dipole_from_corr_bruteforce(ref, N)).max() < 1e-13
    assert abs(S[0] - 1.0) < 1e-13
    r, s, cnt = jw.radial_bins(S, N, 8)
    ro, so, co = orc.radial_bins(S, N, 8)
    assert np.array_equal(cnt, co) and np.allclose(s, so, atol=1e-15) and np.allclose(r, ro, atol=1e-15)
    c.close()


# ------------------------------------------------------------------ full-size properties (N = 2048)
def test_full_size_properties(jw):
    N = 2048
    c = jw.Context(N)
    U = c.field()
    c.philox(1, 0, 0)
    c.mv_init(U, 4, 1.0, 0.25, 1.0)
    c.evolve(U, 3, 1, 1e-4, nsteps=3)
    c.sync()
    u = U.download().reshape(-1, 3, 3)
    idx = rng.integers(0, N * N, 20000)
    m = u[idx]
    assert np.abs(np.einsum("sij,skj->sik", m, m.conj()) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(m) - 1).max() < 1e-12
    corr = c.measure_corr(U)
    assert abs(corr.sum() * (N * N) / 3.0 - 1.0) < 1e-11          # Parseval: sum_k corr = 3/V
    S = c.dipole_from_corr(corr)
    assert abs(S[0] - 1) < 1e-11
    c.close()
