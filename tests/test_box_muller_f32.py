"""The FP32 Box-Muller radius of the mixed form (box_muller_pair_f in
jw_device.hpp): one logarithm for all lanes, log u0 below 1/2 and above it
log1p(-t) = log(a) t / (1 - a) with t = 1 - u0 from the 53-bit integer and
a = 1 - t rounded.  Emulated here in numpy float32 (correctly rounded log,
exact reciprocal) against the FP64 radius sqrt(-2 log u0): the method keeps
every radius within a few FP32 ulps of max(r, 1), including u0 next to 1
(small radii, where log u0 computed from a rounded u0 would lose them) and
next to 1/2.  The GPU form's own check is tests/test_gpu_mixed.py (one step on
the Philox stream against FP64)."""
import numpy as np

F = np.float32


def radius_f32(m0):
    lo = m0 < (1 << 52)
    t = (((1 << 53) - 1 - m0).astype(np.float64).astype(F) + F(0.5)) * F(2.0 ** -53)
    a = np.where(lo, (m0.astype(np.float64).astype(F) + F(0.5)) * F(2.0 ** -53), F(1) - t)
    lu = np.log(a).astype(F)
    with np.errstate(divide="ignore", invalid="ignore"):
        hi = np.where(a == F(1), -t, (lu * (t * (F(1) / (F(1) - a)).astype(F))).astype(F))
    lu = np.where(lo, lu, hi).astype(F)
    return np.sqrt(F(-2) * lu).astype(F)


def test_radius_within_few_ulps():
    rng = np.random.default_rng(5)
    m0 = np.concatenate([
        rng.integers(0, 1 << 53, 200_000, dtype=np.uint64),
        ((1 << 53) - 1 - rng.integers(0, 1 << 30, 20_000, dtype=np.uint64)).astype(np.uint64),
        ((1 << 52) + rng.integers(-1000, 1000, 2000)).astype(np.uint64),
        np.array([(1 << 53) - 1, 1 << 52, (1 << 52) - 1, 0, 1], dtype=np.uint64)])
    u0 = (m0.astype(np.float64) + 0.5) * 2.0 ** -53
    ref = np.sqrt(-2.0 * np.log(u0))
    err = np.abs(radius_f32(m0).astype(np.float64) - ref) / np.maximum(ref, 1.0)
    assert err.max() <= 4 * 2.0 ** -24, err.max() / 2.0 ** -24


def test_naive_log_loses_small_radii():
    """The case the log1p form exists for: u0 within 2^-30 of 1."""
    m0 = np.array([(1 << 53) - 1 - k for k in (0, 1, 7, 1000, 1 << 20)], dtype=np.uint64)
    u0 = (m0.astype(np.float64) + 0.5) * 2.0 ** -53
    ref = np.sqrt(-2.0 * np.log(u0))
    naive = np.sqrt(F(-2) * np.log(u0.astype(F))).astype(np.float64)
    assert np.abs(radius_f32(m0) - ref).max() <= 4 * 2.0 ** -24
    assert np.abs(naive - ref).max() > 1e-5


def radius32_f32(w0):
    """box_muller32's FP32 radius (momentum-space noise): the same method on one
    32-bit word, u0 = (w0 + 1/2) 2^-32."""
    lo = w0 < (1 << 31)
    t = (((0xFFFFFFFF - w0).astype(np.float64) + 0.5).astype(F)) * F(2.0 ** -32)
    a = np.where(lo, ((w0.astype(np.float64) + 0.5).astype(F)) * F(2.0 ** -32), F(1) - t)
    lu = np.log(a).astype(F)
    with np.errstate(divide="ignore", invalid="ignore"):
        hi = np.where(a == F(1), -t, (lu * (t * (F(1) / (F(1) - a)).astype(F))).astype(F))
    lu = np.where(lo, lu, hi).astype(F)
    return np.sqrt(F(-2) * lu).astype(F)


def test_radius32_within_few_ulps():
    rng = np.random.default_rng(6)
    w0 = np.concatenate([
        rng.integers(0, 1 << 32, 200_000, dtype=np.uint64),
        (0xFFFFFFFF - rng.integers(0, 1 << 12, 20_000, dtype=np.uint64)).astype(np.uint64),
        ((1 << 31) + rng.integers(-1000, 1000, 2000)).astype(np.uint64),
        np.array([0xFFFFFFFF, 1 << 31, (1 << 31) - 1, 0, 1], dtype=np.uint64)])
    u0 = (w0.astype(np.float64) + 0.5) * 2.0 ** -32
    ref = np.sqrt(-2.0 * np.log(u0))
    err = np.abs(radius32_f32(w0).astype(np.float64) - ref) / np.maximum(ref, 1.0)
    assert err.max() <= 4 * 2.0 ** -24, err.max() / 2.0 ** -24
