"""Python binding of the MI355X JIMWLK library (libjimwlk_hip.so, C ABI in
include/jimwlk.h), for tests and benchmarks.

Host fields are numpy complex128 arrays of shape (N*N, t) in the reference's
AoS layout (site x*N + y, entry k = 3*row + col; field.h:199-209).  There is
no CPU fallback: loading fails when the HIP library is missing.
"""
import ctypes
import os
import subprocess
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjimwlk_hip.so")

SQRT_COUPLING_CONSTANT, NOISE_COUPLING_CONSTANT, HATTA_COUPLING_CONSTANT = 0, 1, 2
FIXED_COUPLING, CORRELATED_NOISE_COUPLING = 3, 4
LINEAR_KERNEL, SIN_KERNEL = 0, 1
TAG_LANGEVIN, TAG_MV, TAG_KSPACE = 0x4C414E47, 0x4D565348, 0x4B535043
NOISE_POSITION, NOISE_MOMENTUM = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)


class JwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"jimwlk error {code}: {msg}")
        self.code = code


class Alpha(ctypes.Structure):
    _fields_ = [("lambda_L", ctypes.c_double), ("mu0_over_lambda", ctypes.c_double),
                ("freeze_c", ctypes.c_double), ("beta0", ctypes.c_double)]


def alpha(lambda_L=6.0, mu0_over_lambda=2.5, freeze_c=0.2, beta0=9.0):
    return Alpha(lambda_L, mu0_over_lambda, freeze_c, beta0)


def build():
    subprocess.run(["make", "-s", "-j8", "-C", HERE, "-B", "lib"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.jw_last_error.restype = ctypes.c_char_p
        L.jw_version.restype = ctypes.c_char_p
        L.jw_kernel_name.restype = ctypes.c_char_p
        L.jw_destroy.restype = None
        L.jw_field_destroy.restype = None
        L.jw_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
        L.jw_field_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in ("jw_destroy", "jw_sync", "jw_lattice_size"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.jw_field_destroy.argtypes = [ctypes.c_void_p]
        L.jw_field_upload.argtypes = [ctypes.c_void_p, _dp]
        L.jw_field_download.argtypes = [ctypes.c_void_p, _dp]
        L.jw_field_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.jw_noise_inject.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.jw_noise_philox.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.jw_noise_get_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.jw_checkpoint_write.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p]
        L.jw_checkpoint_read.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p]
        L.jw_evolve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.POINTER(Alpha), ctypes.c_int]
        L.jw_evolve_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.POINTER(Alpha), ctypes.c_int]
        L.jw_evolve_mixed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.POINTER(Alpha), ctypes.c_int, ctypes.c_int]
        L.jw_mv_init.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _dp]
        L.jw_update_uf.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double]
        L.jw_set_mv_model.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.jw_solve_poisson.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.jw_set_mv_zero_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_prepare_A_and_B_local.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int, ctypes.c_int]
        L.jw_correlate_noise.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Alpha)]
        L.jw_prepare_A_local.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Alpha)]
        L.jw_generate_gaussian_with_noise_coupling_constant.argtypes = [ctypes.c_void_p] * 3 + [ctypes.POINTER(Alpha)]
        L.jw_measure_corr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp]
        L.jw_dipole_from_corr.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.jw_radial_bins.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.jw_saturation_scale.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.jw_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long)]
        L.jw_su3_exp_gen.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, ctypes.c_int]
        L.jw_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_set_fused.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_set_mixed_fft.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_set_noise_space.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_batch_set_noise_space.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_calibration_copy.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.jw_batch_destroy.restype = None
        L.jw_batch_destroy.argtypes = [ctypes.c_void_p]
        L.jw_batch_size.argtypes = [ctypes.c_void_p]
        L.jw_batch_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.jw_batch_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.jw_batch_philox.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.jw_batch_set_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.jw_batch_layout.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_batch_fft_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_batch_set_launch.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.jw_batch_launch.argtypes = [ctypes.c_void_p]
        L.jw_batch_evolve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.POINTER(Alpha), ctypes.c_int, ctypes.c_int]
        L.jw_batch_sync.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise JwError(rc, lib().jw_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_dp)


class Field:
    def __init__(self, ctx, t=9):
        self.ctx, self.t = ctx, t
        h = ctypes.c_void_p()
        _check(lib().jw_field_create(ctx.h, t, ctypes.byref(h)))
        self.h = h
        ctx._fields.add(self)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.ctx.V, self.t)
        _check(lib().jw_field_upload(self.h, _p(a)))
        return self

    def download(self):
        a = np.zeros((self.ctx.V, self.t), np.complex128)
        _check(lib().jw_field_download(self.h, _p(a)))
        return a

    def close(self):
        if self.h:
            lib().jw_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One lattice (N x N) on one GPU (jw_create)."""

    def __init__(self, N, device=0):
        h = ctypes.c_void_p()
        _check(lib().jw_create(ctypes.byref(h), N, device))
        self.h, self.N, self.V = h, N, N * N
        self._fields = weakref.WeakSet()

    def close(self):
        """Destroys the context after its remaining fields (a field must not outlive
        its context: jw_field_destroy uses the context's stream)."""
        if self.h:
            for f in list(self._fields):
                f.close()
            lib().jw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, t=9, data=None):
        f = Field(self, t)
        if data is not None:
            f.upload(data)
        return f

    def sync(self):
        _check(lib().jw_sync(self.h))

    def copy(self, dst, src):
        """lfield::operator= (jw_field_copy)."""
        _check(lib().jw_field_copy(dst.h, src.h))

    # noise
    def inject(self, z, per_site=16):
        z = np.ascontiguousarray(z, dtype=np.float64)
        assert z.size == per_site * self.V
        _check(lib().jw_noise_inject(self.h, _p(z), per_site))

    def philox(self, seed, traj=0, step=0):
        _check(lib().jw_noise_philox(self.h, seed, traj, step))

    def noise_step(self):
        s = ctypes.c_uint32()
        _check(lib().jw_noise_get_step(self.h, ctypes.byref(s)))
        return s.value

    def checkpoint_write(self, U, path):
        """jw_checkpoint_write: U and the Philox state to a file."""
        _check(lib().jw_checkpoint_write(self.h, U.h, os.fsencode(path)))

    def checkpoint_read(self, U, path):
        """jw_checkpoint_read: U and the Philox state from a file."""
        _check(lib().jw_checkpoint_read(self.h, U.h, os.fsencode(path)))

    # reference per-function entry points
    def execute2D(self, f, dir):
        _check(lib().jw_execute2D(self.h, f.h, int(dir)))

    def generate_gaussian(self, xx, xy):
        _check(lib().jw_generate_gaussian(self.h, xx.h, xy.h))

    def generate_gaussian_with_noise_coupling_constant(self, xx, xy, al=None):
        _check(lib().jw_generate_gaussian_with_noise_coupling_constant(self.h, xx.h, xy.h,
                                                                        ctypes.byref(al) if al else None))

    def prepare_A_local(self, A, xx, xy, coupling, kernel, al=None):
        _check(lib().jw_prepare_A_local(self.h, A.h, xx.h, xy.h, coupling, kernel, ctypes.byref(al) if al else None))

    def uxiulocal(self, vx, vy, U, xx, xy):
        _check(lib().jw_uxiulocal(self.h, vx.h, vy.h, U.h, xx.h, xy.h))

    def update_uf(self, U, B, A, step):
        _check(lib().jw_update_uf(self.h, U.h, B.h, A.h, step))

    def prepare_A_and_B_local(self, A, B, xx, xy, U, coupling, kernel):
        """POSITION_EVOLUTION kernel sums at every site (O(V^2), N <= 256)."""
        _check(lib().jw_prepare_A_and_B_local(self.h, A.h, B.h, xx.h, xy.h, U.h, coupling, kernel))

    def correlate_noise(self, f, al=None):
        """f <- F^-1[sqrt(alpha(k)) F f] (the noise correlation of the position path)."""
        _check(lib().jw_correlate_noise(self.h, f.h, ctypes.byref(al) if al else None))

    def set_mv_model(self, f, g, mu, Ny):
        _check(lib().jw_set_mv_model(self.h, f.h, g, mu, Ny))

    def solve_poisson(self, f, mass, g):
        _check(lib().jw_solve_poisson(self.h, f.h, mass, g))

    def set_mv_zero_mode(self, mode):
        """0: the reference's k = 0 factor -g/m^2 in solvePoisson; 1: the k = 0 mode set to zero."""
        _check(lib().jw_set_mv_zero_mode(self.h, int(mode)))

    def mul_exp(self, U, f):
        _check(lib().jw_mul_exp(self.h, U.h, f.h))

    def set_unit(self, f):
        _check(lib().jw_set_to_unit(f.h))

    def trace(self, U, corr):
        _check(lib().jw_trace(self.h, U.h, corr.h))

    def average_and_symmetrize(self, corr):
        _check(lib().jw_average_and_symmetrize(self.h, corr.h))

    def reduce(self, corr, s, e):
        _check(lib().jw_reduce(self.h, corr.h, s.h, e.h))

    # engine
    def evolve(self, U, coupling=FIXED_COUPLING, kernel=SIN_KERNEL, step=1e-4, al=None, nsteps=1):
        _check(lib().jw_evolve(self.h, U.h, coupling, kernel, step, ctypes.byref(al) if al else None, nsteps))

    def evolve_noalloc(self, U, coupling=FIXED_COUPLING, kernel=SIN_KERNEL, step=1e-4, al=None, nsteps=1):
        """jw_evolve_noalloc: jw_evolve on the library's per-GPU workspace (N <= 512, no allocation)."""
        _check(lib().jw_evolve_noalloc(self.h, U.h, coupling, kernel, step, ctypes.byref(al) if al else None,
                                       nsteps))

    def evolve_mixed(self, U, coupling=FIXED_COUPLING, kernel=SIN_KERNEL, step=1e-4, al=None, nsteps=1,
                     reunit_every=10):
        """jw_evolve_mixed: FP32 fields and FFTs, FP64 exponentials and SU(3) projection."""
        _check(lib().jw_evolve_mixed(self.h, U.h, coupling, kernel, step, ctypes.byref(al) if al else None, nsteps,
                                     reunit_every))

    def mv_init(self, U, Ny, g, mu, mass, z=None):
        if z is not None:
            z = np.ascontiguousarray(z, dtype=np.float64)
            assert z.size == 8 * self.V * Ny
        _check(lib().jw_mv_init(self.h, U.h, Ny, g, mu, mass, _p(z) if z is not None else None))

    # observables
    def measure_corr(self, U):
        c = np.zeros(self.V)
        _check(lib().jw_measure_corr(self.h, U.h, _p(c)))
        return c

    def dipole_from_corr(self, corr):
        corr = np.ascontiguousarray(corr, dtype=np.float64)
        S = np.zeros(self.V)
        _check(lib().jw_dipole_from_corr(self.h, _p(corr), _p(S)))
        return S

    # instrumentation
    def timing(self, enable=True):
        _check(lib().jw_timing(self.h, 1 if enable else 0))

    def calibration_copy(self, nplanes=8):
        _check(lib().jw_calibration_copy(self.h, nplanes))

    # jw_set_fused modes
    (UNFUSED, FUSED, FUSED_ROWMAJOR, FUSED_BLOCK4, FUSED_BLOCK8, FUSED_ROWMAJOR_WIDE, FUSED_INTERLEAVED,
     FUSED_ROWMAJOR_NARROW) = range(8)

    def set_fused(self, mode=1):
        """mode: False/0 unfused kernel sequence, True/1 fused (default layout),
        2/3/4 fused with row-major / 4-row / 8-row blocked intermediate planes,
        5 row-major with the column pass over one plane of 64-byte strips, 6 the
        two directions of each plane pair interleaved site by site, 7 row-major with
        the column pass over one plane of 32-byte strips."""
        _check(lib().jw_set_fused(self.h, int(mode)))

    def set_mixed_fft(self, mode):
        """0: FP32 butterflies; 1: radix-16 stages of the FP32 column passes as split-BF16
        DFT-matrix products on the matrix cores."""
        _check(lib().jw_set_mixed_fft(self.h, int(mode)))

    def set_noise_space(self, space):
        """NOISE_POSITION (default): 16 Philox normals per site; NOISE_MOMENTUM: the
        noise spectra drawn per mode in the first column pass (same distribution)."""
        _check(lib().jw_set_noise_space(self.h, int(space)))

    def kernel_times(self):
        out = {}
        for k in range(13):
            ms, n = ctypes.c_double(), ctypes.c_long()
            _check(lib().jw_kernel_time(self.h, k, ctypes.byref(ms), ctypes.byref(n)))
            if n.value:
                out[lib().jw_kernel_name(k).decode()] = (ms.value, n.value)
        return out

    def su3_exp_gen(self, h8, s):
        h8 = np.ascontiguousarray(h8, dtype=np.float64).reshape(-1, 8)
        out = np.zeros((h8.shape[0], 9), np.complex128)
        _check(lib().jw_su3_exp_gen(self.h, _p(h8), s, _p(out), h8.shape[0]))
        return out.reshape(-1, 3, 3)


class Batch:
    """Independent trajectories in size groups (jw_batch_*): groups = [(N, count), ...]."""

    def __init__(self, groups, fp32=True, device=0):
        n = len(groups)
        Ns = (ctypes.c_int * n)(*[g[0] for g in groups])
        cs = (ctypes.c_int * n)(*[g[1] for g in groups])
        h = ctypes.c_void_p()
        _check(lib().jw_batch_create(ctypes.byref(h), device, n, Ns, cs, 1 if fp32 else 0))
        self.h, self.groups = h, list(groups)

    def close(self):
        if self.h:
            lib().jw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return lib().jw_batch_size(self.h)

    def set(self, t, U):
        _check(lib().jw_batch_set(self.h, t, U.h))

    def get(self, t, U):
        _check(lib().jw_batch_get(self.h, t, U.h))

    def philox(self, seed, traj0=0, step=0):
        _check(lib().jw_batch_philox(self.h, seed, traj0, step))

    def set_layout(self, group, layout):
        _check(lib().jw_batch_set_layout(self.h, group, layout))

    def set_noise_space(self, space):
        _check(lib().jw_batch_set_noise_space(self.h, int(space)))

    def layout(self, group):
        return lib().jw_batch_layout(self.h, group)

    def set_launch(self, mode):
        """0: per-group launches on their own streams; 1: one launch per kernel class over
        all groups; -1: timed on the next evolution."""
        _check(lib().jw_batch_set_launch(self.h, int(mode)))

    def launch(self):
        return lib().jw_batch_launch(self.h)

    def fft_mode(self, group):
        """1 when the tuner put the group's FP32 radix-16 column stages on the matrix cores."""
        return lib().jw_batch_fft_mode(self.h, group)

    def evolve(self, coupling=FIXED_COUPLING, kernel=SIN_KERNEL, step=1e-4, al=None, nsteps=1, reunit_every=10):
        _check(lib().jw_batch_evolve(self.h, coupling, kernel, step, ctypes.byref(al) if al else None, nsteps,
                                     reunit_every))

    def sync(self):
        _check(lib().jw_batch_sync(self.h))


def radial_bins(S, N, nbins):
    S = np.ascontiguousarray(S, dtype=np.float64)
    r, s, c = np.zeros(nbins), np.zeros(nbins), np.zeros(nbins)
    _check(lib().jw_radial_bins(N, _p(S), nbins, _p(r), _p(s), _p(c)))
    return r, s, c


def saturation_scale(r, s, c):
    q = ctypes.c_double()
    r, s, c = (np.ascontiguousarray(a, dtype=np.float64) for a in (r, s, c))
    _check(lib().jw_saturation_scale(len(r), _p(r), _p(s), _p(c), ctypes.byref(q)))
    return q.value
