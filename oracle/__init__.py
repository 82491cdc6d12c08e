"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
per-site code (_ref/libjw_ref.so, when built).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker, never as the thing measured
or shipped.

Fields are numpy complex128 arrays of shape (N*N, t) in the reference's AoS
layout (site index x*N + y, component k = 3*row + col), field.h:199-209.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libjw_ref.so")

SQRT, NOISE, HATTA, FIXED, CORRELATED = 0, 1, 2, 3, 4
LINEAR, SIN = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)


class Alpha(ctypes.Structure):
    """Running-coupling parameters (field.h:2046, 2257-2273 hard-code 6, 2.5, 0.2, 9)."""
    _fields_ = [("lambda_L", ctypes.c_double), ("mu0_over_lambda", ctypes.c_double),
                ("freeze_c", ctypes.c_double), ("beta0", ctypes.c_double)]


def alpha(lambda_L=6.0, mu0_over_lambda=2.5, freeze_c=0.2, beta0=9.0):
    return Alpha(lambda_L, mu0_over_lambda, freeze_c, beta0)


def build(ref=True):
    """Compile liboracle.so (and _ref/libjw_ref.so when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "oracle"], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def _p(a):
    return a.ctypes.data_as(_dp)


def _field(N, t=9):
    return np.zeros((N * N, t), dtype=np.complex128)


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run oracle.build())")
        L = ctypes.CDLL(path)
        self.L = L
        L.or_sqrt_alpha_noise.restype = ctypes.c_double
        L.or_alpha_kernel.restype = ctypes.c_double
        L.or_sqrt_alpha_noise.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(Alpha)]
        L.or_alpha_kernel.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(Alpha)]
        L.or_philox4x32_10.restype = None
        L.or_philox_normals.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_int, _dp]
        L.or_kspace_normals.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, _dp, ctypes.c_void_p]

    def _ok(self, rc):
        if rc != 0:
            raise ValueError(f"oracle call failed ({rc})")

    def set_num_threads(self, n):
        self.L.or_set_num_threads(int(n))

    def num_threads(self):
        return self.L.or_num_threads()

    def set_as_shipped(self, on):
        """1: uxiulocal/update_uf on one thread, as the reference ships them (field.h:1901, 1974)."""
        self.L.or_set_as_shipped(int(on))

    # --- SU(3) ---
    def su3_exponentiate(self, m, sign):
        m = np.ascontiguousarray(m, dtype=np.complex128).reshape(9)
        out = np.zeros(9, np.complex128)
        self._ok(self.L.or_su3_exponentiate(_p(m), ctypes.c_double(sign), _p(out)))
        return out.reshape(3, 3)

    def su3_eigen(self, m):
        m = np.ascontiguousarray(m, dtype=np.complex128).reshape(9)
        Q = np.zeros(9, np.complex128)
        w = np.zeros(3)
        self._ok(self.L.or_su3_eigen(_p(m), _p(Q), _p(w)))
        return Q.reshape(3, 3), w

    # --- tables ---
    def momenta(self, N):
        a, b, c = (np.zeros(N * N) for _ in range(3))
        self._ok(self.L.or_momenta(N, _p(a), _p(b), _p(c)))
        return a, b, c

    def sqrt_alpha_noise(self, q2, N, al=None):
        return self.L.or_sqrt_alpha_noise(q2, N, ctypes.byref(al) if al else None)

    def alpha_kernel(self, q2, N, al=None):
        return self.L.or_alpha_kernel(q2, N, ctypes.byref(al) if al else None)

    # --- field ops (in place on copies) ---
    def fft2d(self, f, N, dir):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        t = f.size // (N * N)
        self._ok(self.L.or_fft2d(N, t, _p(f), int(dir)))
        return f

    def dft2d_bruteforce(self, f, N, dir):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        t = f.size // (N * N)
        self._ok(self.L.or_dft2d_bruteforce(N, t, _p(f), int(dir)))
        return f

    def set_unit(self, N):
        u = _field(N)
        self._ok(self.L.or_set_unit(N, _p(u)))
        return u

    def generate_gaussian(self, N, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        x, y = _field(N), _field(N)
        self._ok(self.L.or_generate_gaussian(N, _p(z), _p(x), _p(y)))
        return x, y

    def generate_gaussian_noise_coupling(self, N, z, al=None):
        z = np.ascontiguousarray(z, dtype=np.float64)
        x, y = _field(N), _field(N)
        self._ok(self.L.or_generate_gaussian_noise_coupling(N, _p(z), ctypes.byref(al) if al else None,
                                                            _p(x), _p(y)))
        return x, y

    def set_mv_model(self, N, z, sigma):
        z = np.ascontiguousarray(z, dtype=np.float64)
        f = _field(N)
        self._ok(self.L.or_set_mv_model(N, _p(z), ctypes.c_double(sigma), _p(f)))
        return f

    def solve_poisson(self, f, N, mass, g):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        self._ok(self.L.or_solve_poisson(N, _p(f), ctypes.c_double(mass), ctypes.c_double(g)))
        return f

    def mul_exp(self, U, f, N):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        f = np.ascontiguousarray(f, dtype=np.complex128)
        self._ok(self.L.or_mul_exp(N, _p(U), _p(f)))
        return U

    def prepare_A_local(self, xix, xiy, N, coupling, kernel, al=None):
        A = _field(N)
        xix = np.ascontiguousarray(xix, dtype=np.complex128)
        xiy = np.ascontiguousarray(xiy, dtype=np.complex128)
        self._ok(self.L.or_prepare_A_local(N, _p(A), _p(xix), _p(xiy), coupling, kernel,
                                           ctypes.byref(al) if al else None))
        return A

    def uxiulocal(self, U, xix, xiy, N):
        Vx, Vy = _field(N), _field(N)
        U, xix, xiy = (np.ascontiguousarray(a, dtype=np.complex128) for a in (U, xix, xiy))
        self._ok(self.L.or_uxiulocal(N, _p(Vx), _p(Vy), _p(U), _p(xix), _p(xiy)))
        return Vx, Vy

    def position_weights(self, N, coupling, kernel):
        wa, wb = np.zeros(N * N), np.zeros(N * N)
        self._ok(self.L.or_position_weights(N, coupling, kernel, _p(wa), _p(wb)))
        return wa, wb

    def prepare_A_and_B_local(self, xix, xiy, U, N, coupling, kernel):
        A, B = _field(N), _field(N)
        xix, xiy, U = (np.ascontiguousarray(a, dtype=np.complex128) for a in (xix, xiy, U))
        self._ok(self.L.or_prepare_A_and_B_local(N, _p(A), _p(B), _p(xix), _p(xiy), _p(U), coupling, kernel))
        return A, B

    def update_uf(self, U, B, A, step, N):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        B, A = (np.ascontiguousarray(a, dtype=np.complex128) for a in (B, A))
        self._ok(self.L.or_update_uf(N, _p(U), _p(B), _p(A), ctypes.c_double(step)))
        return U

    def trace(self, U, N):
        c = _field(N, 1)
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self._ok(self.L.or_trace(N, _p(U), _p(c)))
        return c

    def average_and_symmetrize(self, c, N):
        c = np.ascontiguousarray(c, dtype=np.complex128).copy()
        self._ok(self.L.or_average_and_symmetrize(N, _p(c)))
        return c

    def reduce(self, c, s, e, N):
        s = np.ascontiguousarray(s, dtype=np.complex128).copy()
        e = np.ascontiguousarray(e, dtype=np.complex128).copy()
        c = np.ascontiguousarray(c, dtype=np.complex128)
        self._ok(self.L.or_reduce(N, _p(c), _p(s), _p(e)))
        return s, e

    def scale_by_sqrt_alpha(self, f, N, al=None):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        t = f.size // (N * N)
        self._ok(self.L.or_scale_by_sqrt_alpha(N, t, _p(f), ctypes.byref(al) if al else None))
        return f

    def correlate_noise(self, f, N, al=None):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        t = f.size // (N * N)
        self._ok(self.L.or_correlate_noise(N, t, _p(f), ctypes.byref(al) if al else None))
        return f

    def langevin_step(self, U, z, N, coupling, kernel, step, al=None):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        z = np.ascontiguousarray(z, dtype=np.float64)
        self._ok(self.L.or_langevin_step(N, _p(U), _p(z), coupling, kernel, ctypes.c_double(step),
                                         ctypes.byref(al) if al else None))
        return U

    def mv_init(self, N, z, Ny, g, mu, mass, zero_mode=0):
        U = _field(N)
        z = np.ascontiguousarray(z, dtype=np.float64)
        self._ok(self.L.or_mv_init_opt(N, _p(U), _p(z), Ny, ctypes.c_double(g), ctypes.c_double(mu),
                                       ctypes.c_double(mass), int(zero_mode)))
        return U

    def measure_corr(self, U, N):
        c = _field(N, 1)
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self._ok(self.L.or_measure_corr(N, _p(U), _p(c)))
        return c

    def dipole_from_corr_bruteforce(self, corr, N):
        corr = np.ascontiguousarray(corr, dtype=np.complex128)
        S = np.zeros(N * N)
        self._ok(self.L.or_dipole_from_corr_bruteforce(N, _p(corr), _p(S)))
        return S

    def dipole_position_space(self, U, N):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        S = np.zeros(N * N)
        self._ok(self.L.or_dipole_position_space(N, _p(U), _p(S)))
        return S

    def radial_bins(self, S, N, nbins):
        S = np.ascontiguousarray(S, dtype=np.float64)
        r, s, c = np.zeros(nbins), np.zeros(nbins), np.zeros(nbins)
        self._ok(self.L.or_radial_bins(N, _p(S), nbins, _p(r), _p(s), _p(c)))
        return r, s, c

    # --- Philox ---
    def philox4x32_10(self, ctr, key):
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        self.L.or_philox4x32_10(c, k, o)
        return list(o)

    def philox_normals(self, N, seed, traj, step, tag, per_site):
        z = np.zeros(N * N * per_site)
        self._ok(self.L.or_philox_normals(N, seed, traj, step, tag, per_site, _p(z)))
        return z

    def kspace_normals(self, N, seed, traj, step, tag, spectra=False):
        """Position-space normals (16 per site, draw order n0, m0, n1, m1, ...) of
        the momentum-space noise stream; with spectra=True also the drawn
        unit-variance spectra, (V, 4 pairs, [Px re, Px im, Py re, Py im])."""
        z = np.zeros(N * N * 16)
        spec = np.zeros(N * N * 16) if spectra else None
        self._ok(self.L.or_kspace_normals(N, seed, traj, step, tag, _p(z),
                                          None if spec is None else spec.ctypes.data))
        return (z, spec.reshape(N * N, 4, 4)) if spectra else z


class Reference:
    """The reference's own per-site code at N = 64 (oracle/_ref/libjw_ref.so)."""
    N = 64

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        self.L = ctypes.CDLL(path)
        assert self.L.ref_lattice_size() == self.N

    def su3_exponentiate(self, m, sign):
        m = np.ascontiguousarray(m, dtype=np.complex128).reshape(9)
        out = np.zeros(9, np.complex128)
        self.L.ref_su3_exponentiate(_p(m), ctypes.c_double(sign), _p(out))
        return out.reshape(3, 3)

    def su3_eigen(self, m):
        m = np.ascontiguousarray(m, dtype=np.complex128).reshape(9)
        Q = np.zeros(9, np.complex128)
        w = np.zeros(3)
        rc = self.L.ref_su3_eigen(_p(m), _p(Q), _p(w))
        return Q.reshape(3, 3), w, rc

    def momenta(self):
        N = self.N
        a, b, c = (np.zeros(N * N) for _ in range(3))
        self.L.ref_momenta(_p(a), _p(b), _p(c))
        return a, b, c

    def prepare_A_local(self, xix, xiy, coupling, kernel):
        A = _field(self.N)
        xix = np.ascontiguousarray(xix, dtype=np.complex128)
        xiy = np.ascontiguousarray(xiy, dtype=np.complex128)
        assert self.L.ref_prepare_A_local(_p(A), _p(xix), _p(xiy), coupling, kernel) == 0
        return A

    def uxiulocal(self, U, xix, xiy):
        Vx, Vy = _field(self.N), _field(self.N)
        U, xix, xiy = (np.ascontiguousarray(a, dtype=np.complex128) for a in (U, xix, xiy))
        self.L.ref_uxiulocal(_p(Vx), _p(Vy), _p(U), _p(xix), _p(xiy))
        return Vx, Vy

    def prepare_A_and_B_local(self, xix, xiy, U, coupling, kernel):
        A, B = _field(self.N), _field(self.N)
        xix, xiy, U = (np.ascontiguousarray(a, dtype=np.complex128) for a in (xix, xiy, U))
        assert self.L.ref_prepare_A_and_B_local(_p(A), _p(B), _p(xix), _p(xiy), _p(U), coupling, kernel) == 0
        return A, B

    def update_uf(self, U, B, A, step):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        B, A = (np.ascontiguousarray(a, dtype=np.complex128) for a in (B, A))
        self.L.ref_update_uf(_p(U), _p(B), _p(A), ctypes.c_double(step))
        return U

    def mul_exp(self, U, f):
        U = np.ascontiguousarray(U, dtype=np.complex128).copy()
        f = np.ascontiguousarray(f, dtype=np.complex128)
        self.L.ref_mul_exp(_p(U), _p(f))
        return U

    def solve_poisson(self, f, mass, g):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        self.L.ref_solve_poisson(_p(f), ctypes.c_double(mass), ctypes.c_double(g))
        return f

    def set_unit(self):
        u = _field(self.N)
        self.L.ref_set_unit(_p(u))
        return u

    def trace(self, U):
        c = _field(self.N, 1)
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self.L.ref_trace(_p(U), _p(c))
        return c

    def average_and_symmetrize(self, c):
        c = np.ascontiguousarray(c, dtype=np.complex128).copy()
        self.L.ref_average_and_symmetrize(_p(c))
        return c

    def reduce(self, c, s, e):
        s = np.ascontiguousarray(s, dtype=np.complex128).copy()
        e = np.ascontiguousarray(e, dtype=np.complex128).copy()
        c = np.ascontiguousarray(c, dtype=np.complex128)
        self.L.ref_reduce(_p(c), _p(s), _p(e))
        return s, e

    def correlations(self):
        """setCorrelationsForCouplingConstant(momenta*): alpha(phat2(k)) / sqrt(V)."""
        c = _field(self.N, 1)
        self.L.ref_correlations(_p(c))
        return c
