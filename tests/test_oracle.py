"""Pins of the CPU oracle (oracle/jw_oracle.cpp), CPU only.

Each check is chosen so that a plausible mistake in the oracle fails it:
  * against the reference's own per-site code built from its sources
    (oracle/_ref, N = 64): bit-level agreement;
  * against numpy.fft / brute-force DFT for the transform conventions;
  * against mpmath's 50-digit matrix exponential and closed forms for exp(iH);
  * against the textbook Gell-Mann matrices for the noise colour basis;
  * against published Philox4x32-10 known-answer vectors;
  * the full Langevin step / MV init / measurement against the reference's
    per-site functions composed as main_optimized.cpp:178-191, 243-333 does,
    with numpy's FFT in the reference's unitary convention.
"""
import numpy as np
import pytest

import oracle as O_mod

N64 = 64
rng = np.random.default_rng(20261202)

# Textbook Gell-Mann matrices (standard basis, independent of the code under test)
L = np.zeros((8, 3, 3), complex)
L[0][0, 1] = L[0][1, 0] = 1
L[1][0, 1], L[1][1, 0] = -1j, 1j
L[2][0, 0], L[2][1, 1] = 1, -1
L[3][0, 2] = L[3][2, 0] = 1
L[4][0, 2], L[4][2, 0] = -1j, 1j
L[5][1, 2] = L[5][2, 1] = 1
L[6][1, 2], L[6][2, 1] = -1j, 1j
L[7] = np.diag([1, 1, -2]) / np.sqrt(3)
# the reference's basis flips the antisymmetric ones (field.h:2183-2186, 2204-2207, 2218-2221)
LT = L * np.array([1, -1, 1, 1, -1, 1, -1, 1])[:, None, None]


def herm(scale=1.0, traceless=True):
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    h = (a + a.conj().T) / 2
    if traceless:
        h -= np.trace(h) / 3 * np.eye(3)
    return h * scale


def random_su3_field(orc, N):
    return np.array([orc.su3_exponentiate(herm(rng.uniform(0.1, 3)), 1.0).reshape(9) for _ in range(N * N)])


def npfft(f, N, d):
    """Reference convention (mpi_fftw_class.h:191-315) via numpy: dir 1 = e^{-i}, both /sqrt(V)."""
    a = f.reshape(N, N, -1)
    out = np.fft.fft2(a, axes=(0, 1)) / N if d == 1 else np.fft.ifft2(a, axes=(0, 1)) * N
    return out.reshape(N * N, -1)


# ---------------------------------------------------------------- exponential
def test_exp_matches_reference_build(orc, ref):
    worst = 0.0
    for _ in range(3000):
        h = herm(rng.uniform(1e-4, 8.0), traceless=bool(rng.integers(2)))
        # non-Hermitian input: only the upper triangle and real diagonal count
        h = h + (rng.normal(size=(3, 3)) * 1e-3 * np.tril(np.ones((3, 3)), -1))
        for s in (1.0, -1.0):
            worst = max(worst, np.abs(orc.su3_exponentiate(h, s) - ref.su3_exponentiate(h, s)).max())
    assert worst <= 4e-15


@pytest.mark.parametrize("h", [
    np.zeros((3, 3)), np.diag([1.0, 1.0, -2.0]), np.diag([0.3, -0.1, -0.2]),
    np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], complex), 0.7 * L[7], 1e-9 * L[0] + 2 * L[2],
])
def test_exp_degenerate_fallback_matches_reference(orc, ref, h):
    for s in (1.0, -1.0):
        assert np.abs(orc.su3_exponentiate(h, s) - ref.su3_exponentiate(h, s)).max() <= 1e-15


def test_exp_against_mpmath_and_closed_forms(orc):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    for _ in range(40):
        h = herm(rng.uniform(0.01, 6.0))
        ex = mp.expm(mp.matrix(h.tolist()) * mp.mpc(0, 1))
        ex = np.array([[complex(ex[i, j]) for j in range(3)] for i in range(3)])
        assert np.abs(orc.su3_exponentiate(h, 1.0) - ex).max() < 1e-13
    for th in (0.0, 0.3, 2.0):
        e = orc.su3_exponentiate(th * L[2], 1.0)
        assert np.abs(e - np.diag([np.exp(1j * th), np.exp(-1j * th), 1])).max() < 1e-15
    u = orc.su3_exponentiate(herm(3.0), 1.0)
    assert abs(np.linalg.det(u) - 1) < 1e-13
    assert np.abs(u @ u.conj().T - np.eye(3)).max() < 1e-14


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [2, 4, 8, 16])
def test_fft_bruteforce(orc, N):
    f = rng.normal(size=(N * N, 3)) + 1j * rng.normal(size=(N * N, 3))
    for d in (0, 1):
        assert np.abs(orc.fft2d(f, N, d) - orc.dft2d_bruteforce(f, N, d)).max() < 1e-14


@pytest.mark.parametrize("N", [4, 32, 64, 256])
def test_fft_numpy_conventions(orc, N):
    f = rng.normal(size=(N * N, 9)) + 1j * rng.normal(size=(N * N, 9))
    for d in (0, 1):
        assert np.abs(orc.fft2d(f, N, d) - npfft(f, N, d)).max() < 1e-13
    # Parseval and round trip
    g = orc.fft2d(f, N, 1)
    assert abs(np.sum(abs(g) ** 2) / np.sum(abs(f) ** 2) - 1) < 1e-13
    assert np.abs(orc.fft2d(g, N, 0) - f).max() < 1e-13


def test_fft_delta_known_answer(orc):
    # tests/test_fftw_execute.cpp:59-72: delta at the origin -> constant 1/sqrt(V)
    N = 64
    f = np.zeros((N * N, 1), complex)
    f[0] = 1.0
    g = orc.fft2d(f, N, 0)
    assert np.abs(g - 1.0 / N).max() < 1e-16


# ---------------------------------------------------------------- noise basis
def test_lambda_basis(orc):
    N = 4
    z = rng.normal(size=N * N * 16)
    xx, xy = orc.generate_gaussian(N, z)
    for s in range(N * N):
        for field, off in ((xx, 0), (xy, 1)):
            m = field[s].reshape(3, 3)
            n = z[16 * s + off::2][:8]
            assert np.abs(m - np.einsum("a,aij->ij", n, LT)).max() < 1e-15
            assert np.abs(m - m.conj().T).max() == 0.0
            assert abs(np.trace(m)) < 1e-15
            c = np.array([np.trace(m @ LT[a]).real / 2 for a in range(8)])
            assert np.abs(c - n).max() < 1e-14


def test_noise_coupling_scaling(orc):
    # field.h:2273: sqrt(alpha) evaluated at phat2 of the POSITION index
    N = 8
    z = rng.normal(size=N * N * 16)
    a, _ = orc.generate_gaussian(N, z)
    b, _ = orc.generate_gaussian_noise_coupling(N, z)
    ph, _, _ = orc.momenta(N)
    q = ph * N * N / 36.0
    alpha = 4 * np.pi / (9 * 0.2 * np.log(6.25 ** 5 + q ** 5))   # same function, other arrangement
    assert np.allclose(b, a * np.sqrt(alpha)[:, None], rtol=1e-14, atol=0)
    # alpha freezes at 4 pi / (9 ln 6.25) in the IR
    assert abs(orc.alpha_kernel(0.0, N) - 4 * np.pi / (9 * np.log(6.25))) < 1e-15


# ---------------------------------------------------------------- kernel multiply
@pytest.mark.parametrize("coupling", [O_mod.SQRT, O_mod.NOISE, O_mod.FIXED])
@pytest.mark.parametrize("kernel", [O_mod.LINEAR, O_mod.SIN])
def test_prepare_A_local_closed_form(orc, coupling, kernel):
    N = 16
    fx = rng.normal(size=(N * N, 9)) + 1j * rng.normal(size=(N * N, 9))
    fy = rng.normal(size=(N * N, 9)) + 1j * rng.normal(size=(N * N, 9))
    A = orc.prepare_A_local(fx, fy, N, coupling, kernel)
    k = np.fft.fftfreq(N, 1.0 / N)               # wrapped mode numbers
    KX, KY = np.meshgrid(k, k, indexing="ij")
    if kernel == O_mod.SIN:
        q2 = 4 * np.sin(np.pi * np.arange(N) / N)[:, None] ** 2 + 4 * np.sin(np.pi * np.arange(N) / N)[None, :] ** 2
        bx = np.sin(2 * np.pi * np.arange(N) / N)[:, None] * np.ones((1, N))
        by = np.sin(2 * np.pi * np.arange(N) / N)[None, :] * np.ones((N, 1))
    else:
        px, py = 2 * np.pi * KX / N, 2 * np.pi * KY / N
        q2, bx, by = px ** 2 + py ** 2, px, py
    c = 4 * np.pi / (9 * 0.2 * np.log(6.25 ** 5 + (q2 * N * N / 36) ** 5)) if coupling == O_mod.SQRT else 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        kx = -2j * np.pi * np.sqrt(c) * bx / q2
        ky = -2j * np.pi * np.sqrt(c) * by / q2
    kx[0, 0] = ky[0, 0] = 0
    expect = kx.reshape(-1, 1) * fx + ky.reshape(-1, 1) * fy
    assert np.allclose(A, expect, rtol=1e-13, atol=1e-13)
    assert np.all(A[0] == 0)


def test_prepare_A_local_rejects_hatta(orc):
    N = 4
    f = np.zeros((N * N, 9), complex)
    with pytest.raises(ValueError):
        orc.prepare_A_local(f, f, N, O_mod.HATTA, O_mod.SIN)


# ---------------------------------------------------------------- oracle vs reference build, N = 64
def test_momenta_match_reference_build(orc, ref):
    for a, b in zip(orc.momenta(N64), ref.momenta()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("coupling", [O_mod.SQRT, O_mod.NOISE])
@pytest.mark.parametrize("kernel", [O_mod.LINEAR, O_mod.SIN])
def test_prepare_A_local_matches_reference_build(orc, ref, coupling, kernel):
    fx = rng.normal(size=(N64 * N64, 9)) + 1j * rng.normal(size=(N64 * N64, 9))
    fy = rng.normal(size=(N64 * N64, 9)) + 1j * rng.normal(size=(N64 * N64, 9))
    assert np.array_equal(orc.prepare_A_local(fx, fy, N64, coupling, kernel),
                          ref.prepare_A_local(fx, fy, coupling, kernel))


def test_site_functions_match_reference_build(orc, ref):
    N = N64
    U = random_su3_field(orc, N)
    z = rng.normal(size=N * N * 16)
    xx, xy = orc.generate_gaussian(N, z)
    vo = orc.uxiulocal(U, xx, xy, N)
    vr = ref.uxiulocal(U, xx, xy)
    assert np.array_equal(vo[0], vr[0]) and np.array_equal(vo[1], vr[1])
    A = npfft(orc.prepare_A_local(npfft(xx, N, 1), npfft(xy, N, 1), N, 1, 1), N, 0)
    B = npfft(orc.prepare_A_local(npfft(vo[0], N, 1), npfft(vo[1], N, 1), N, 1, 1), N, 0)
    assert np.abs(orc.update_uf(U, B, A, 1e-4, N) - ref.update_uf(U, B, A, 1e-4)).max() <= 1e-15
    assert np.abs(orc.mul_exp(U, A, N) - ref.mul_exp(U, A)).max() <= 1e-15
    assert np.array_equal(orc.set_unit(N), ref.set_unit())
    f = npfft(orc.set_mv_model(N, rng.normal(size=N * N * 8), 0.0035), N, 1)
    for mass in (1e-4, 0.3):
        assert np.array_equal(orc.solve_poisson(f, N, mass, 1.0), ref.solve_poisson(f, mass, 1.0))
    Uk = npfft(U, N, 1)
    c = orc.trace(Uk, N)
    assert np.array_equal(c, ref.trace(Uk))
    s = orc.average_and_symmetrize(c, N)
    assert np.array_equal(s, ref.average_and_symmetrize(c))
    zero = np.zeros_like(s)
    for a, b in zip(orc.reduce(s, zero, zero, N), ref.reduce(s, zero, zero)):
        assert np.array_equal(a, b)


def _step_composed_from_reference(ref, orc, U, z, coupling, step):
    """main_optimized.cpp:243-304 composed from the reference's per-site code + numpy FFT."""
    N = N64
    if coupling == O_mod.NOISE:
        xx, xy = orc.generate_gaussian_noise_coupling(N, z)
    else:
        xx, xy = orc.generate_gaussian(N, z)
    fx, fy = npfft(xx, N, 1), npfft(xy, N, 1)
    A = npfft(ref.prepare_A_local(fx, fy, coupling, O_mod.SIN), N, 0)
    xx, xy = npfft(fx, N, 0), npfft(fy, N, 0)
    vx, vy = ref.uxiulocal(U, xx, xy)
    B = npfft(ref.prepare_A_local(npfft(vx, N, 1), npfft(vy, N, 1), coupling, O_mod.SIN), N, 0)
    return ref.update_uf(U, B, A, step)


@pytest.mark.parametrize("coupling", [O_mod.SQRT, O_mod.NOISE])
def test_langevin_step_matches_reference_composition(orc, ref, coupling):
    U = random_su3_field(orc, N64)
    for _ in range(2):
        z = rng.normal(size=N64 * N64 * 16)
        a = orc.langevin_step(U, z, N64, coupling, O_mod.SIN, 1e-4)
        b = _step_composed_from_reference(ref, orc, U, z, coupling, 1e-4)
        # two correct FP64 FFTs (numpy's, the oracle's) differ at ~1e-16 relative;
        # the 1/phat2 kernel lifts that to ~2e-13 per entry after one step at N = 64
        assert np.abs(a - b).max() < 1e-12
        U = a


# ---------------------------------------------------------------- correlated noise (coupling 4)
def _sqrt_alpha_from_reference(ref):
    """sqrt(alpha(phat2(k))) from the reference's own setCorrelationsForCouplingConstant
    (field.h:1633-1651), which stores alpha(k) / sqrt(V)."""
    c = ref.correlations()[:, 0]
    assert np.abs(c.imag).max() == 0.0
    return np.sqrt(c.real * N64)


def test_correlated_kernel_equals_sqrt_coupling_kernel(orc, ref):
    # K sqrt(alpha) applied to F xi (coupling 4: noise scaled per mode, kernel factor 1)
    # is the SQRT coupling's kernel sqrt(alpha) K (field.h:2045-2053) on the same F xi
    N = N64
    xx, xy = orc.generate_gaussian(N, rng.normal(size=N * N * 16))
    fx, fy = npfft(xx, N, 1), npfft(xy, N, 1)
    a_corr = orc.prepare_A_local(orc.scale_by_sqrt_alpha(fx, N), orc.scale_by_sqrt_alpha(fy, N), N,
                                 O_mod.CORRELATED, O_mod.SIN)
    a_sqrt = ref.prepare_A_local(fx, fy, O_mod.SQRT, O_mod.SIN)
    assert np.abs(a_corr - a_sqrt).max() <= 1e-14 * np.abs(a_sqrt).max()
    # the per-mode factor is the square root of the reference's alpha(k)
    s = _sqrt_alpha_from_reference(ref)
    assert np.allclose(orc.scale_by_sqrt_alpha(np.ones((N * N, 1)), N)[:, 0].real, s, rtol=1e-15, atol=0)


def test_correlated_noise_covariance_is_reference_cholesky_target(orc, ref):
    """The oracle's map T: xi -> xi~ = F^-1[sqrt(alpha) F xi] (one colour component; the map
    is colour-blind) satisfies T T^T = M, M the dense position-space covariance whose Cholesky
    factor the reference builds: alpha(k)/sqrt(V) (field.h:1633-1651), transformed backward in
    the unitary convention (main_optimized.cpp:145-151), laid out as M[i, j] =
    C(|xi - xj|, |yi - yj|) (matrix.h:151-200)."""
    N = N64
    V = N * N
    x = np.arange(V) // N
    y = np.arange(V) % N

    def probe(j):
        f = np.zeros((V, 1), complex)
        f[j] = 1.0
        t = orc.correlate_noise(f, N)[:, 0]
        assert np.abs(t.imag).max() < 1e-16
        return t.real

    t0 = probe(0)
    # T is circulant (translation invariant): column j is column 0 shifted by j
    for j in (1, 77, 2080, V - 1):
        shift = ((x - x[j]) % N) * N + (y - y[j]) % N
        assert np.abs(probe(j) - t0[shift]).max() < 1e-15
    T = t0[((x[:, None] - x[None, :]) % N) * N + (y[:, None] - y[None, :]) % N]
    K = T @ T.T
    C = npfft(ref.correlations(), N, 0)[:, 0]
    assert np.abs(C.imag).max() < 1e-16
    M = C.real[np.abs(x[:, None] - x[None, :]) * N + np.abs(y[:, None] - y[None, :])]
    assert np.abs(K - M).max() < 1e-13
    # a mistake of alpha for sqrt(alpha) gives a visibly different covariance
    assert np.abs((T @ T) @ (T @ T).T - M).max() > 1e-3


def test_correlated_step_matches_reference_composition(orc, ref):
    """or_langevin_step, coupling 4, against main_optimized.cpp:243-304 composed from the
    reference's per-site code with the per-mode correlated noise: A with the SQRT kernel,
    xi~ = F^-1[sqrt(alpha) F xi] with sqrt(alpha) from the reference's own
    setCorrelationsForCouplingConstant, V = U xi~ U^dagger, B with kernel factor 1."""
    N = N64
    s = _sqrt_alpha_from_reference(ref)[:, None]
    U = random_su3_field(orc, N)
    for _ in range(2):
        z = rng.normal(size=N * N * 16)
        xx, xy = orc.generate_gaussian(N, z)
        fx, fy = npfft(xx, N, 1), npfft(xy, N, 1)
        A = npfft(ref.prepare_A_local(fx, fy, O_mod.SQRT, O_mod.SIN), N, 0)
        vx, vy = ref.uxiulocal(U, npfft(s * fx, N, 0), npfft(s * fy, N, 0))
        B = npfft(ref.prepare_A_local(npfft(vx, N, 1), npfft(vy, N, 1), O_mod.NOISE, O_mod.SIN), N, 0)
        expect = ref.update_uf(U, B, A, 5e-5)
        got = orc.langevin_step(U, z, N, O_mod.CORRELATED, O_mod.SIN, 5e-5)
        assert np.abs(got - expect).max() < 1e-12
        assert np.abs(got - U).max() > 1e-4
        U = got


def test_mv_init_matches_reference_composition(orc, ref):
    N, Ny, g, mu, mass = N64, 3, 1.0, 30.72 / 64, 0.5
    z = rng.normal(size=Ny * N * N * 8)
    Uo = orc.mv_init(N, z, Ny, g, mu, mass)
    U = ref.set_unit()
    for s in range(Ny):
        f = orc.set_mv_model(N, z[s * N * N * 8:(s + 1) * N * N * 8], g * mu / np.sqrt(Ny))
        f = npfft(ref.solve_poisson(npfft(f, N, 1), mass * g ** 2 * mu, g), N, 0)
        U = ref.mul_exp(U, f)
    assert np.abs(Uo - U).max() < 1e-12


@pytest.mark.parametrize("mass", [0.5, 4e-4])
def test_mv_init_zero_mode_projection(orc, ref, mass):
    """The projection option sets each sheet's k = 0 mode to zero after solvePoisson:
    the reference's composition with the k = 0 charge zeroed before its solvePoisson
    (field.h:867-877, whose factor then multiplies zero).  At m a = 1e-4 (mass 4e-4 in
    units of g^2 mu = 0.25) the projected start is well conditioned: the oracle's
    radix-2 FFT and numpy's agree to 1e-12 per site."""
    N, Ny, g, mu = N64, 3, 1.0, 0.25
    z = rng.normal(size=Ny * N * N * 8)
    Uo = orc.mv_init(N, z, Ny, g, mu, mass, zero_mode=1)
    U = ref.set_unit()
    for s in range(Ny):
        f = npfft(orc.set_mv_model(N, z[s * N * N * 8:(s + 1) * N * N * 8], g * mu / np.sqrt(Ny)), N, 1)
        f[0, :] = 0.0
        f = npfft(ref.solve_poisson(f, mass * g ** 2 * mu, g), N, 0)
        U = ref.mul_exp(U, f)
    assert np.abs(Uo - U).max() < 1e-12
    m = Uo.reshape(-1, 3, 3)
    assert np.abs(np.einsum("sij,skj->sik", m, m.conj()) - np.eye(3)).max() < 1e-13
    # the option changes the start (a plausible no-op implementation fails here)
    assert np.abs(orc.mv_init(N, z, Ny, g, mu, mass) - Uo).max() > 1e-3


def test_measurement_matches_reference_composition(orc, ref):
    U = random_su3_field(orc, N64)
    c = orc.measure_corr(U, N64)
    expect = ref.average_and_symmetrize(ref.trace(npfft(U, N64, 1)))
    assert np.abs(c - expect).max() < 1e-17


# ---------------------------------------------------------------- invariants of the step
def test_step_invariants(orc):
    N = 16
    U = random_su3_field(orc, N)
    # zero noise leaves U unchanged (A = B = 0, exp(0) = 1 exactly)
    assert np.array_equal(orc.langevin_step(U, np.zeros(N * N * 16), N, O_mod.FIXED, O_mod.SIN, 1e-4), U)
    U2 = orc.langevin_step(U, rng.normal(size=N * N * 16), N, O_mod.FIXED, O_mod.SIN, 1e-4)
    m = U2.reshape(-1, 3, 3)
    assert np.abs(np.einsum("sij,skj->sik", m, m.conj()) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(m) - 1).max() < 1e-13
    assert np.abs(U2 - U).max() > 1e-3


# ---------------------------------------------------------------- dipole / Wiener-Khinchin
def test_dipole_wiener_khinchin(orc):
    N = 8
    U = random_su3_field(orc, N)
    corr = orc.measure_corr(U, N)
    S_k = orc.dipole_from_corr_bruteforce(corr, N)
    S_x = orc.dipole_position_space(U, N)
    S_x_sym = np.zeros_like(S_x).reshape(N, N)
    Sx = S_x.reshape(N, N)
    for x in range(N):
        for y in range(N):
            t = [Sx[x, y], Sx[y, x]]
            t = t + [Sx[(-x) % N, y], Sx[y, (-x) % N], Sx[x, (-y) % N], Sx[(-y) % N, x],
                     Sx[(-x) % N, (-y) % N], Sx[(-y) % N, (-x) % N]]
            S_x_sym[x, y] = np.mean(t)
    assert np.abs(S_k - S_x_sym.reshape(-1)).max() < 1e-13
    assert abs(S_k[0] - 1) < 1e-13
    assert abs(corr[:, 0].real.sum() - 3.0 / (N * N)) < 1e-16


def test_radial_bins(orc):
    N = 8
    S = np.arange(N * N, dtype=float)
    r, s, c = orc.radial_bins(S, N, 6)
    assert c[0] == 1 and s[0] == 0 and r[0] == 0
    assert c.sum() == sum(1 for x in range(N) for y in range(N)
                          if np.hypot(min(x, N - x) if x != N // 2 else N // 2,
                                      min(y, N - y) if y != N // 2 else N // 2) < 6)


# ---------------------------------------------------------------- Philox
def test_philox_known_answers(orc):
    # Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11)
    assert orc.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert orc.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert orc.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_normals_statistics(orc):
    N = 64
    z = orc.philox_normals(N, 7, 0, 0, 0x4C414E47, 16)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    z2 = orc.philox_normals(N, 7, 0, 1, 0x4C414E47, 16)
    assert abs(np.corrcoef(z, z2)[0, 1]) < 0.02


TAG_KSPACE = 0x4B535043


def test_kspace_noise_spectrum_round_trip(orc):
    """The position-space normals of the momentum-space stream transform back
    (numpy's forward FFT, e^{-2 pi i k.x/N}, unnormalised) to N times the drawn
    spectra, plane pair by plane pair and direction by direction: pins the sign
    of the exponent, the 1/N normalisation and the draw-order slots."""
    N = 32
    z, spec = orc.kspace_normals(N, 5, 1, 3, TAG_KSPACE, spectra=True)
    z = z.reshape(N * N, 16)
    for p in range(4):
        wx = (z[:, 4 * p] + 1j * z[:, 4 * p + 2]).reshape(N, N)
        wy = (z[:, 4 * p + 1] + 1j * z[:, 4 * p + 3]).reshape(N, N)
        px = (spec[:, p, 0] + 1j * spec[:, p, 1]).reshape(N, N)
        py = (spec[:, p, 2] + 1j * spec[:, p, 3]).reshape(N, N)
        assert np.abs(np.fft.fft2(wx) - N * px).max() < 1e-12 * N
        assert np.abs(np.fft.fft2(wy) - N * py).max() < 1e-12 * N


def test_kspace_noise_words_are_philox(orc):
    """Mode k, pair p: the four Philox words of counter {k, step, tag, p} give the
    spectra through Box-Muller on 32-bit uniforms (checked on a few modes against
    a numpy evaluation of the published transform)."""
    N = 16
    seed, traj, step = 11, 2, 9
    _, spec = orc.kspace_normals(N, seed, traj, step, TAG_KSPACE, spectra=True)
    for k in (0, 1, 17, 255):
        for p in range(4):
            w = orc.philox4x32_10([k, step, TAG_KSPACE, p], [seed, traj])
            u = (np.array(w, dtype=np.float64) + 0.5) * 2.0 ** -32
            r0, r1 = np.sqrt(-2 * np.log(u[0])), np.sqrt(-2 * np.log(u[2]))
            want = [r0 * np.cos(2 * np.pi * u[1]), r0 * np.sin(2 * np.pi * u[1]),
                    r1 * np.cos(2 * np.pi * u[3]), r1 * np.sin(2 * np.pi * u[3])]
            assert np.allclose(spec[k, p], want, rtol=1e-13, atol=1e-13)


def test_kspace_noise_is_white_unit_gaussian(orc):
    """In distribution the momentum-space stream is the position-space one: 16
    independent unit normals per site, uncorrelated between sites and steps.
    Bounds are about 5 standard errors of each estimate."""
    N = 64
    V = N * N
    z = orc.kspace_normals(N, 7, 0, 0, TAG_KSPACE).reshape(V, 16)
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / n)
    # component covariance: identity (off-diagonal standard error 1/sqrt(V))
    c = np.cov(z.T)
    assert np.abs(np.diag(c) - 1).max() < 6 * np.sqrt(2 / V)
    assert np.abs(c - np.diag(np.diag(c))).max() < 6 / np.sqrt(V)
    # neighbours in x and y, and the next step: uncorrelated
    f = z.reshape(N, N, 16)
    for g in (np.roll(f, 1, 0), np.roll(f, 1, 1), orc.kspace_normals(N, 7, 0, 1, TAG_KSPACE).reshape(N, N, 16)):
        assert abs(np.mean(f * g)) < 5 / np.sqrt(n)
    # Gaussian tails: fourth moment 3, and the KS distance to N(0, 1)
    assert abs(np.mean(z ** 4) - 3) < 5 * np.sqrt(96 / n)
    from scipy import stats
    assert stats.kstest(z.ravel(), "norm").pvalue > 1e-4


# ---------------------------------------------------------------- POSITION_EVOLUTION
@pytest.mark.parametrize("coupling", [0, 1])
@pytest.mark.parametrize("kernel", [0, 1])
def test_prepare_A_and_B_local_matches_reference(orc, ref, coupling, kernel):
    """The brute-force position-space sums (field.h:2368-2519 at every site) against the
    reference's own function; its OpenMP reduction fixes no summation order, hence the
    relative tolerance."""
    N = N64
    U = random_su3_field(orc, N)
    xx, xy = orc.generate_gaussian(N, rng.normal(size=N * N * 16))
    Ar, Br = ref.prepare_A_and_B_local(xx, xy, U, coupling, kernel)
    Ao, Bo = orc.prepare_A_and_B_local(xx, xy, U, N, coupling, kernel)
    assert np.abs(Ao - Ar).max() <= 5e-14 * np.abs(Ar).max()
    assert np.abs(Bo - Br).max() <= 5e-14 * np.abs(Br).max()


@pytest.mark.parametrize("kernel", [0, 1])
def test_prepare_A_and_B_local_is_a_convolution(orc, kernel):
    """A = w * xi and B = w * (U xi U^dagger) as cyclic convolutions over the
    displacement x - y (convolution theorem with numpy FFTs), and the SIN weights are
    the positions.h:131-141 tables: a sign or index slip (y - x for x - y) fails."""
    N = 16
    U = random_su3_field(orc, N)
    xx, xy = orc.generate_gaussian(N, rng.normal(size=N * N * 16))
    A, B = orc.prepare_A_and_B_local(xx, xy, U, N, 3, kernel)
    wa, wb = orc.position_weights(N, 3, kernel)
    vx, vy = orc.uxiulocal(U, xx, xy, N)

    def conv(w, f):
        W = np.fft.fft2(w.reshape(N, N))
        F = np.fft.fft2(f.reshape(N, N, 9), axes=(0, 1))
        return np.fft.ifft2(W[:, :, None] * F, axes=(0, 1)).reshape(N * N, 9)

    assert np.abs(A - (conv(wa, xx) + conv(wb, xy))).max() < 1e-13 * np.abs(A).max()
    assert np.abs(B - (conv(wa, vx) + conv(wb, vy))).max() < 1e-13 * np.abs(B).max()
    # the reversed displacement gives a different (sign-flipped, odd kernel) A
    rev = lambda w: w.reshape(N, N)[(-np.arange(N)) % N][:, (-np.arange(N)) % N].reshape(-1)
    assert np.abs(A - (conv(rev(wa), xx) + conv(rev(wb), xy))).max() > 1e-2 * np.abs(A).max()
    if kernel == 1:
        i = np.arange(N)
        xh = 0.5 * N * np.sin(2 * np.pi * i / N) / np.pi
        xb = N * np.sin(np.pi * i / N) / np.pi
        r2 = xb[:, None] ** 2 + xb[None, :] ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            expect = np.where(r2 > 1e-8, xh[:, None] / r2, 0.0).reshape(-1)
        assert np.allclose(wa, expect, rtol=1e-14, atol=0)


def test_position_kernel_against_momentum_kernel():
    """Cross-check of the two discretisations of the same kernel (reference facts, no
    oracle call): the momentum path's kernel -2 pi i pbar/phat2 (field.h:2052-2053)
    acting through the unitary transforms equals, in position space, the convolution
    with (1/V) sum_k K(k) e^{ik.r}; the position path uses xhat(r)/xbar2(r)
    (positions.h:131-141).  Their spectra agree within 16% at every nonzero mode of
    wavelength 8 or more lattice spacings (|k| <= N/8 per axis), exactly at none: the
    momentum and position evolutions are different lattice discretisations (they are
    not expected to match step for step)."""
    for N in (16, 64):
        k = np.arange(N)
        q4 = 4 * np.sin(np.pi * k / N) ** 2
        sb = np.sin(2 * np.pi * k / N)
        ph = q4[:, None] + q4[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            Kx = np.where(ph > 1e-8, -2 * np.pi * 1j * sb[:, None] / ph, 0)
            xh = 0.5 * N * np.sin(2 * np.pi * k / N) / np.pi
            xb = N * np.sin(np.pi * k / N) / np.pi
            r2 = xb[:, None] ** 2 + xb[None, :] ** 2
            w = np.where(r2 > 1e-8, xh[:, None] / r2, 0.0)
        Kpos = np.fft.fft2(w)
        kw = np.abs(np.where(k < N // 2, k, k - N))
        m = (np.abs(Kx) > 1e-12) & (kw[:, None] <= N // 8) & (kw[None, :] <= N // 8)
        ratio = Kpos[m] / Kx[m]
        assert np.abs(ratio.imag).max() < 1e-12
        assert 0.84 < ratio.real.min() and ratio.real.max() < 1.15
        assert np.abs(ratio.real - 1).max() > 0.05
