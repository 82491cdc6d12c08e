// jw_device.hpp -- device-side building blocks of the MI355X JIMWLK solver:
// complex FP64 arithmetic, SU(3) / su(3) algebra in the reference's colour basis,
// the SU(3) exponential (Cayley-Hamilton with series coefficients, scaling and
// squaring), the Philox4x32-10 noise stream and the LDS Stockham FFT.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jw {

// ---------------------------------------------------------------------------
// complex double
// ---------------------------------------------------------------------------
typedef double2 cplx;

__device__ __forceinline__ cplx mk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }  // a*conj(b)
__device__ __forceinline__ cplx conjg(cplx a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ cplx cmuli(cplx a) { return mk(-a.y, a.x); }       // i*a
__device__ __forceinline__ cplx cmulmi(cplx a) { return mk(a.y, -a.x); }      // -i*a

// ---------------------------------------------------------------------------
// complex float (storage and FFTs of the mixed-precision form): the same
// operations, overloaded on float2
// ---------------------------------------------------------------------------
typedef float2 cplxf;

__device__ __forceinline__ cplxf mkf(float r, float i) { return make_float2(r, i); }
__device__ __forceinline__ cplxf cadd(cplxf a, cplxf b) { return mkf(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplxf csub(cplxf a, cplxf b) { return mkf(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplxf cmul(cplxf a, cplxf b) { return mkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cplxf cscale(cplxf a, float s) { return mkf(a.x * s, a.y * s); }
__device__ __forceinline__ cplxf cmuli(cplxf a) { return mkf(-a.y, a.x); }
__device__ __forceinline__ cplxf cmulmi(cplxf a) { return mkf(a.y, -a.x); }
__device__ __forceinline__ cplxf cmulc(cplxf a, cplxf b) { return mkf(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ cplxf conjg(cplxf a) { return mkf(a.x, -a.y); }

// complex value of type C (double2 or float2) from two doubles, and back
template <typename C>
__device__ __forceinline__ C mkc(double r, double i) {
    if constexpr (sizeof(C) == sizeof(double2)) return make_double2(r, i);
    else return make_float2((float)r, (float)i);
}
__device__ __forceinline__ cplx to_d(cplx a) { return a; }
__device__ __forceinline__ cplx to_d(cplxf a) { return make_double2((double)a.x, (double)a.y); }

// ---------------------------------------------------------------------------
// 3x3 complex matrices, row-major (entry k = 3*row + col, as field.h:199-209)
// ---------------------------------------------------------------------------
struct m3 { cplx m[9]; };

__device__ __forceinline__ m3 mmul(const m3& a, const m3& b) {
    m3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx s = cmul(a.m[3 * i], b.m[j]);
            s = cadd(s, cmul(a.m[3 * i + 1], b.m[3 + j]));
            s = cadd(s, cmul(a.m[3 * i + 2], b.m[6 + j]));
            c.m[3 * i + j] = s;
        }
    return c;
}

// the same in FP32 (the mixed-precision form's products of exponentials and
// links: its fields carry FP32 rounding anyway)
struct m3f { cplxf m[9]; };

__device__ __forceinline__ m3f mmul(const m3f& a, const m3f& b) {
    m3f c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplxf s = cmul(a.m[3 * i], b.m[j]);
            s = cadd(s, cmul(a.m[3 * i + 1], b.m[3 + j]));
            s = cadd(s, cmul(a.m[3 * i + 2], b.m[6 + j]));
            c.m[3 * i + j] = s;
        }
    return c;
}
__device__ __forceinline__ m3f to_f(const m3& a) {
    m3f c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.m[k] = mkf((float)a.m[k].x, (float)a.m[k].y);
    return c;
}
__device__ __forceinline__ m3 to_d(const m3f& a) {
    m3 c;
#pragma unroll
    for (int k = 0; k < 9; k++) c.m[k] = make_double2((double)a.m[k].x, (double)a.m[k].y);
    return c;
}

// ---------------------------------------------------------------------------
// su(3) generator coefficients in the reference's colour basis
//   {l1, -l2, l3, l4, -l5, l6, -l7, l8}   (field.h:2169-2231)
// h[0..7] <-> n[0..7]:  H01 = h0 + i h1, H02 = h3 + i h4, H12 = h5 + i h6,
//                       H00 = h2 + h7/sqrt3, H11 = -h2 + h7/sqrt3, H22 = -2 h7/sqrt3.
// ---------------------------------------------------------------------------
#define JW_SQRT3 1.7320508075688772935274463415059
#define JW_INV_SQRT3 0.57735026918962576450914878050196

// Hermitian traceless matrix s*H(h)
__device__ __forceinline__ m3 herm_from_gen(const double h[8], double s) {
    m3 X;
    const double d8 = h[7] * JW_INV_SQRT3;
    X.m[0] = mk(s * (h[2] + d8), 0.0);
    X.m[4] = mk(s * (-h[2] + d8), 0.0);
    X.m[8] = mk(s * (-2.0 * d8), 0.0);
    X.m[1] = mk(s * h[0], s * h[1]);
    X.m[2] = mk(s * h[3], s * h[4]);
    X.m[5] = mk(s * h[5], s * h[6]);
    X.m[3] = conjg(X.m[1]);
    X.m[6] = conjg(X.m[2]);
    X.m[7] = conjg(X.m[5]);
    return X;
}

// Generator coefficients of a (nearly) Hermitian matrix, read the way the
// reference's exponential reads its argument: upper triangle and real diagonal
// (su3_matrix.h:263-264).  The trace part of the diagonal is dropped.
__device__ __forceinline__ void gen_from_upper(const m3& M, double h[8]) {
    h[0] = M.m[1].x; h[1] = M.m[1].y;
    h[3] = M.m[2].x; h[4] = M.m[2].y;
    h[5] = M.m[5].x; h[6] = M.m[5].y;
    h[2] = 0.5 * (M.m[0].x - M.m[4].x);
    h[7] = (M.m[0].x + M.m[4].x - 2.0 * M.m[8].x) * (0.5 * JW_INV_SQRT3);
}

// ---------------------------------------------------------------------------
// exp(i X) for X Hermitian traceless 3x3 (Cayley-Hamilton form).
// With X^3 = c1 X + c0 (c1 = tr(X^2)/2, c0 = det X), X^n = a_n + b_n X + d_n X^2
// where (a, b, d)_{n+1} = (c0 d_n, a_n + c1 d_n, b_n); hence
//   exp(iX) = f0 + f1 X + f2 X^2,  f = sum_n i^n (a_n, b_n, d_n) / n!.
// The sum runs to at most 24 terms for spectral radius rho <= 2 sqrt(c1/3)
// <= 1.6 (c1 <= 1.92): the remainder is below rho^25/25! ~ 1e-20; it stops
// earlier once a term's |a| + |b| + |d| is below 1e-17 (small X: a Langevin
// step's exponent converges in about 10 terms).  Larger X are scaled by 2^-k
// (exact) into that range and the result squared k times.
// X = 0 gives exactly 1.
// ---------------------------------------------------------------------------
#define JW_EXP_SERIES_C1 1.92
#define JW_EXP_SERIES_TERMS 24
#define JW_EXP_SERIES_TAIL 1e-17

__device__ __forceinline__ m3 su3_project(const m3& M);

__device__ __forceinline__ m3 exp_i_gen(const double h[8], double s) {
    double c1 = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) c1 += h[a] * h[a];
    c1 *= s * s;
    m3 E;
    if (!(c1 > 0.0)) {   // X == 0 (or NaN input: propagate it)
#pragma unroll
        for (int k = 0; k < 9; k++) E.m[k] = mk((k % 4 == 0) ? 1.0 : 0.0, 0.0);
        if (c1 != c1) E.m[0] = mk(c1, c1);
        return E;
    }
    // squarings needed: c1 / 4^k <= JW_EXP_SERIES_C1
    int k = 0;
    while (c1 > JW_EXP_SERIES_C1 && k < 60) { c1 *= 0.25; k++; }
    const m3 X = herm_from_gen(h, ldexp(s, -k));
    // det X (Hermitian: real)
    const double d0 = X.m[0].x, d1 = X.m[4].x, d2 = X.m[8].x;
    const cplx x = X.m[1], y = X.m[2], z = X.m[5];
    const cplx xz = cmul(x, z);
    const double c0 = d0 * d1 * d2 + 2.0 * (xz.x * y.x + xz.y * y.y)
                      - d0 * (z.x * z.x + z.y * z.y) - d1 * (y.x * y.x + y.y * y.y)
                      - d2 * (x.x * x.x + x.y * x.y);
    // scaled coefficients A = a_n/n!, B = b_n/n!, D = d_n/n!; n = 0
    double A = 1.0, B = 0.0, D = 0.0;
    double f0r = 1.0, f0i = 0.0, f1r = 0.0, f1i = 0.0, f2r = 0.0, f2i = 0.0;
#pragma unroll
    for (int n = 1; n <= JW_EXP_SERIES_TERMS; n++) {
        const double inv = 1.0 / n;
        const double An = c0 * D * inv, Bn = (A + c1 * D) * inv, Dn = B * inv;
        A = An; B = Bn; D = Dn;
        // i^n: n%4 = 0 -> +1, 1 -> +i, 2 -> -1, 3 -> -i
        if ((n & 3) == 0) { f0r += A; f1r += B; f2r += D; }
        else if ((n & 3) == 1) { f0i += A; f1i += B; f2i += D; }
        else if ((n & 3) == 2) { f0r -= A; f1r -= B; f2r -= D; }
        else { f0i -= A; f1i -= B; f2i -= D; }
        // the rest of the series is below 4e-17 in E (terms fall at least 5x per
        // step from here, |X| <= 1.6): stop (Langevin steps converge in ~10 terms)
        if (fabs(A) + fabs(B) + fabs(D) < JW_EXP_SERIES_TAIL) break;
    }
    const cplx f0 = mk(f0r, f0i), f1 = mk(f1r, f1i), f2 = mk(f2r, f2i);
    const m3 X2 = mmul(X, X);
#pragma unroll
    for (int q = 0; q < 9; q++) E.m[q] = cadd(cmul(f1, X.m[q]), cmul(f2, X2.m[q]));
    E.m[0] = cadd(E.m[0], f0);
    E.m[4] = cadd(E.m[4], f0);
    E.m[8] = cadd(E.m[8], f0);
    for (int j = 0; j < k; j++) E = mmul(E, E);
    // k squarings multiply the rounding of E by up to 2^k; the part that breaks
    // unitarity is taken out again (an MV sheet at m a = 1e-4 carries a zero
    // mode of norm ~1e4: k ~ 15, |E^dagger E - 1| ~ 1e-11 without this)
    if (k > 0) E = su3_project(E);
    return E;
}

// exp(i s H(h)) as exp_i_gen, rounded to FP32 for the mixed-precision form:
// the same FP64 series coefficients, X and X^2 formed from their upper
// triangles (both are Hermitian), the squarings (c1 > 1.92 only) in FP32.
// Fewer live registers than the FP64 matrices of exp_i_gen.
template <typename H>
__device__ __forceinline__ m3f exp_i_gen_f(const H h[8], double s) {
    double c1 = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) c1 += (double)h[a] * (double)h[a];
    c1 *= s * s;
    m3f E;
    if (!(c1 > 0.0)) {
#pragma unroll
        for (int k = 0; k < 9; k++) E.m[k] = mkf((k % 4 == 0) ? 1.f : 0.f, 0.f);
        if (c1 != c1) E.m[0] = mkf((float)c1, (float)c1);
        return E;
    }
    int k = 0;
    while (c1 > JW_EXP_SERIES_C1 && k < 60) { c1 *= 0.25; k++; }
    const double sk = ldexp(s, -k);
    // X = [[d0, x, y], [x*, d1, z], [y*, z*, d2]]
    const double d8 = (double)h[7] * JW_INV_SQRT3;
    const double d0 = sk * ((double)h[2] + d8), d1 = sk * (-(double)h[2] + d8), d2 = sk * (-2.0 * d8);
    const cplx x = mk(sk * (double)h[0], sk * (double)h[1]), y = mk(sk * (double)h[3], sk * (double)h[4]),
               z = mk(sk * (double)h[5], sk * (double)h[6]);
    const double ax = x.x * x.x + x.y * x.y, ay = y.x * y.x + y.y * y.y, az = z.x * z.x + z.y * z.y;
    const cplx xz = cmul(x, z);
    const double c0 = d0 * d1 * d2 + 2.0 * (xz.x * y.x + xz.y * y.y) - d0 * az - d1 * ay - d2 * ax;
    double A = 1.0, B = 0.0, D = 0.0;
    double f0r = 1.0, f0i = 0.0, f1r = 0.0, f1i = 0.0, f2r = 0.0, f2i = 0.0;
#pragma unroll
    for (int n = 1; n <= JW_EXP_SERIES_TERMS; n++) {
        const double inv = 1.0 / n;
        const double An = c0 * D * inv, Bn = (A + c1 * D) * inv, Dn = B * inv;
        A = An; B = Bn; D = Dn;
        if ((n & 3) == 0) { f0r += A; f1r += B; f2r += D; }
        else if ((n & 3) == 1) { f0i += A; f1i += B; f2i += D; }
        else if ((n & 3) == 2) { f0r -= A; f1r -= B; f2r -= D; }
        else { f0i -= A; f1i -= B; f2i -= D; }
        // the result is rounded to FP32: a tail below 1e-12 cannot show
        if (fabs(A) + fabs(B) + fabs(D) < 1e-13) break;
    }
    const cplx f1 = mk(f1r, f1i), f2 = mk(f2r, f2i);
    // X^2: real diagonal and upper triangle
    const double q0 = d0 * d0 + ax + ay, q1 = ax + d1 * d1 + az, q2 = ay + az + d2 * d2;
    const cplx q01 = cadd(cscale(x, d0 + d1), cmulc(y, z));      // x (d0 + d1) + y z*
    const cplx q02 = cadd(cscale(y, d0 + d2), xz);               // y (d0 + d2) + x z
    const cplx q12 = cadd(cmulc(y, x), cscale(z, d1 + d2));      // x* y + z (d1 + d2)
    // E = f0 + f1 X + f2 X^2; the lower triangle from conj(X), conj(X^2)
    E.m[0] = mkf((float)(f0r + f1r * d0 + f2r * q0), (float)(f0i + f1i * d0 + f2i * q0));
    E.m[4] = mkf((float)(f0r + f1r * d1 + f2r * q1), (float)(f0i + f1i * d1 + f2i * q1));
    E.m[8] = mkf((float)(f0r + f1r * d2 + f2r * q2), (float)(f0i + f1i * d2 + f2i * q2));
    auto off = [&](cplx a, cplx b, cplxf& up, cplxf& lo) {
        const cplx u = cadd(cmul(f1, a), cmul(f2, b));
        const cplx l = cadd(cmul(f1, conjg(a)), cmul(f2, conjg(b)));
        up = mkf((float)u.x, (float)u.y);
        lo = mkf((float)l.x, (float)l.y);
    };
    off(x, q01, E.m[1], E.m[3]);
    off(y, q02, E.m[2], E.m[6]);
    off(z, q12, E.m[5], E.m[7]);
    for (int j = 0; j < k; j++) E = mmul(E, E);
    return E;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the library's noise
// stream: key {seed, traj}, counter {site, step, tag, call}; each call gives
// one Box-Muller pair from two 53-bit uniforms in (0,1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32x32 -> 64 bit product each (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ void box_muller_pair(uint32_t site, uint32_t step, uint32_t tag, uint32_t call,
                                                uint32_t seed, uint32_t traj, double& z0, double& z1) {
    uint32_t c[4] = {site, step, tag, call};
    philox4x32_10(c, seed, traj);
    const uint64_t a = ((uint64_t)c[0] << 32) | c[1];
    const uint64_t b = ((uint64_t)c[2] << 32) | c[3];
    const double u0 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u1 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
    const double rr = sqrt(-2.0 * log(u0));
    double s, co;
    sincospi(2.0 * u1, &s, &co);
    z0 = rr * co;
    z1 = rr * s;
}

// The same pair in FP32 from the same Philox words (mixed-precision form):
// log u0 from the 53-bit integer (as log1p of -(1 - u0) above 1/2, so that
// small |z| keep their relative precision) and the angle as pi x with
// x = 2 u1 - 1 in (-1, 1): sincospi(x + 1) = -sincospi(x), on the hardware
// sine and cosine.  Each normal agrees with the FP64 pair to a few FP32 ulps
// of max(|z|, 1).
__device__ __forceinline__ void box_muller_pair_f(uint32_t site, uint32_t step, uint32_t tag, uint32_t call,
                                                  uint32_t seed, uint32_t traj, float& z0, float& z1) {
    uint32_t c[4] = {site, step, tag, call};
    philox4x32_10(c, seed, traj);
    const uint64_t m0 = (((uint64_t)c[0] << 32) | c[1]) >> 11;     // u0 = (m0 + 1/2) 2^-53
    const uint64_t m1 = (((uint64_t)c[2] << 32) | c[3]) >> 11;     // u1 = (m1 + 1/2) 2^-53
    // 53-bit integers to FP32 through FP64 (exact there, then one rounding: the
    // same value as a direct conversion, in fewer instructions)
    // One logarithm for both halves: u0 < 1/2 takes log u0; above, with
    // t = 1 - u0 from the integer and a = 1 - t rounded (1 - a exact),
    // log1p(-t) = log(a) t / (1 - a) (Goldberg's log1p), or -t when a = 1.
    const bool lo = m0 < (1ull << 52);
    const float t = ((float)(double)((1ull << 53) - 1 - m0) + 0.5f) * 0x1.0p-53f;
    const float a = lo ? ((float)(double)m0 + 0.5f) * 0x1.0p-53f : 1.0f - t;
    float lu = logf(a);
    if (!lo) lu = a == 1.0f ? -t : lu * (t * __builtin_amdgcn_rcpf(1.0f - a));
    const float rr = sqrtf(-2.0f * lu);
    const int64_t d = 2 * (int64_t)m1 + 1 - (1ll << 53);            // x = d 2^-53
    // sin and cos of pi x on the hardware's revolution-based units (x / 2
    // revolutions, in (-1/2, 1/2)): within 1.9e-7 of FP64 sincospi on all inputs,
    // against 1.2e-7 for the library's sincospif, in far fewer instructions
    const float rev = (float)(double)d * 0x1.0p-54f;
    z0 = -rr * __builtin_amdgcn_cosf(rev);
    z1 = -rr * __builtin_amdgcn_sinf(rev);
}

// ---------------------------------------------------------------------------
// Momentum-space noise.  The spectrum of a pair of independent white fields,
// F[n_a + i n_b](k), is itself white circular complex Gaussian: real and
// imaginary parts N(0, V), independent across k.  So the xi spectra can be
// drawn per mode instead of per site and transformed: Philox call
// {k = kx N + ky, step, tag, p} (key {seed, traj}) gives four 32-bit uniforms
// u_j = (w_j + 1/2) 2^-32 and two Box-Muller pairs
//   (z0, z1) = sqrt(-2 ln u0) (cos 2 pi u1, sin 2 pi u1),  (z2, z3) from u2, u3,
// with X^_p(k) = sqrt(V) (z0 + i z1), Y^_p(k) = sqrt(V) (z2 + i z3) for the x-
// and y-direction plane pairs p.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void box_muller32(uint32_t w0, uint32_t w1, double& z0, double& z1) {
    const double u0 = ((double)w0 + 0.5) * 0x1.0p-32;
    const double u1 = ((double)w1 + 0.5) * 0x1.0p-32;
    const double rr = sqrt(-2.0 * log(u0));
    double s, co;
    sincospi(2.0 * u1, &s, &co);
    z0 = rr * co;
    z1 = rr * s;
}
// FP32 from the same words, as box_muller_pair_f: log u0 below 1/2, above it
// log1p(-t) with t = 1 - u0 taken from the integer; the angle pi x with
// x = 2 u1 - 1 in (-1, 1)
__device__ __forceinline__ void box_muller32(uint32_t w0, uint32_t w1, float& z0, float& z1) {
    const bool lo = w0 < 0x80000000u;
    const float t = (float)((double)(0xFFFFFFFFu - w0) + 0.5) * 0x1.0p-32f;
    const float a = lo ? (float)((double)w0 + 0.5) * 0x1.0p-32f : 1.0f - t;
    float lu = logf(a);
    if (!lo) lu = a == 1.0f ? -t : lu * (t * __builtin_amdgcn_rcpf(1.0f - a));
    const float rr = sqrtf(-2.0f * lu);
    const int64_t d = 2 * (int64_t)w1 + 1 - (1ll << 32);            // x = d 2^-32
    const float rev = (float)(double)d * 0x1.0p-33f;                 // x / 2 revolutions, as above
    z0 = -rr * __builtin_amdgcn_cosf(rev);
    z1 = -rr * __builtin_amdgcn_sinf(rev);
}
template <typename T>
__device__ __forceinline__ void kspace_normals4(uint32_t k, uint32_t step, uint32_t tag, uint32_t p, uint32_t seed,
                                                uint32_t traj, T z[4]) {
    uint32_t c[4] = {k, step, tag, p};
    philox4x32_10(c, seed, traj);
    box_muller32(c[0], c[1], z[0], z[1]);
    box_muller32(c[2], c[3], z[2], z[3]);
}

// Projection of a near-unitary 3x3 matrix onto SU(3): Gram-Schmidt on rows 0
// and 1, row 2 = conj(row0 x row1) (unit determinant).  The periodic
// reunitarisation of the mixed-precision form.
__device__ __forceinline__ m3 su3_project(const m3& M) {
    m3 R;
    double n0 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) n0 += M.m[k].x * M.m[k].x + M.m[k].y * M.m[k].y;
    n0 = 1.0 / sqrt(n0);
#pragma unroll
    for (int k = 0; k < 3; k++) R.m[k] = cscale(M.m[k], n0);
    cplx p = mk(0.0, 0.0);     // <row0, row1>
#pragma unroll
    for (int k = 0; k < 3; k++) p = cadd(p, cmulc(M.m[3 + k], R.m[k]));
    double n1 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        R.m[3 + k] = csub(M.m[3 + k], cmul(p, R.m[k]));
        n1 += R.m[3 + k].x * R.m[3 + k].x + R.m[3 + k].y * R.m[3 + k].y;
    }
    n1 = 1.0 / sqrt(n1);
#pragma unroll
    for (int k = 0; k < 3; k++) R.m[3 + k] = cscale(R.m[3 + k], n1);
    R.m[6] = conjg(csub(cmul(R.m[1], R.m[5]), cmul(R.m[2], R.m[4])));
    R.m[7] = conjg(csub(cmul(R.m[2], R.m[3]), cmul(R.m[0], R.m[5])));
    R.m[8] = conjg(csub(cmul(R.m[0], R.m[4]), cmul(R.m[1], R.m[3])));
    return R;
}

// ---------------------------------------------------------------------------
// Radix-2/4/8 DFT kernels in registers.  DIR = -1: exp(-2 pi i jk/R) (forward),
// DIR = +1: exp(+2 pi i jk/R) (backward).
// ---------------------------------------------------------------------------
template <int DIR, typename C>
__device__ __forceinline__ C rot_q(C a) { return DIR < 0 ? cmulmi(a) : cmuli(a); }  // a * e^{DIR i pi/2}

template <int R, int DIR, typename C>
__device__ __forceinline__ void dft_small(C* v) {
    typedef decltype(v[0].x) T;
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const C t2 = cadd(v[1], v[3]), t3 = rot_q<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else {   // R == 8: two radix-4 halves and the e^{DIR i pi k/4} twiddles
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4, DIR>(e);
        dft_small<4, DIR>(o);
        const T h = (T)0.70710678118654752440084436210485;
        // o1 *= (1 + DIR i)/sqrt2 ; o2 *= DIR i ; o3 *= (-1 + DIR i)/sqrt2
        C o1, o3;
        if (DIR < 0) {
            o1.x = (o[1].x + o[1].y) * h; o1.y = (o[1].y - o[1].x) * h;
            o3.x = (-o[3].x + o[3].y) * h; o3.y = (-o[3].y - o[3].x) * h;
        } else {
            o1.x = (o[1].x - o[1].y) * h; o1.y = (o[1].y + o[1].x) * h;
            o3.x = (-o[3].x - o[3].y) * h; o3.y = (-o[3].y + o[3].x) * h;
        }
        const C o2 = rot_q<DIR>(o[2]);
        v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
    }
}

// ---------------------------------------------------------------------------
// In-LDS Stockham autosort FFT (mixed radix 8/4/2), out-of-order-free:
// stage with radix R after a product P of earlier radices maps butterfly j
// (k = j mod P) from inputs x[j + r N/R] to outputs x[(j-k) R + k + r P],
// after twiddling input r by w_N^{r k N/(P R)}.  Every thread reads all its
// inputs, the block synchronises, then writes, so one buffer suffices.
// NSEQ sequences of length N; element n of sequence s sits at
// lpos<ES, SS, PS>(n, s) = (n + (n >> PS)) ES + s SS: one padding slot every
// 2^PS elements (PS = 0: none) keeps the strided Stockham writes off common
// LDS banks (PS = 4 for rows, ES = 1; PS = 3 for column tiles, ES > 1).
// tw[m] = exp(-2 pi i m / N), in the element type.  NT threads take part.
// ---------------------------------------------------------------------------
template <int ES, int SS, int PS>
__device__ __forceinline__ int lpos(int n, int s) {
    return (n + (PS ? (n >> PS) : 0)) * ES + s * SS;
}

// padded length of N elements
template <int N, int PS>
struct padded { static constexpr int value = N + (PS ? (N >> PS) : 0); };

template <int REM>
struct radix_of { static constexpr int value = (REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2); };

template <int N, int P, int R, int NSEQ, int ES, int SS, int NT, int DIR, int PS, typename C, typename TW>
__device__ __forceinline__ void fft_stage(C* buf, TW tw, int tid) {
    constexpr int S = N / R;
    constexpr int NB = S * NSEQ;
    constexpr int PER = (NB + NT - 1) / NT;
    // With power-of-two N, P, R and padding, lpos(j + r S) = lpos(j) + off(r S) and
    // lpos(b + r P) = lpos(b) + off(r P), off(m) = (m + (m >> PS)) ES: one computed
    // base per butterfly, compile-time offsets per input/output.
    C v[PER][R];
    int jj[PER], ss[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * NT;
        int j, s;
        if constexpr (ES == 1) { j = b % S; s = b / S; }       // rows: j fastest
        else { s = b % NSEQ; j = b / NSEQ; }                    // column tiles: sequence fastest
        jj[q] = j;
        ss[q] = s;
        if (NB % NT == 0 || b < NB) {
            const C* rb = buf + lpos<ES, SS, PS>(j, s);
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = rb[lpos<ES, 0, PS>(r * S, 0)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * NT;
        if (NB % NT == 0 || b < NB) {
            const int j = jj[q], s = ss[q];
            const int k = j % P;
            if constexpr (P > 1) {
                constexpr int STRIDE = N / (P * R);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C w = tw[r * k * STRIDE];
                    if (DIR > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dft_small<R, DIR>(v[q]);
            C* wb = buf + lpos<ES, SS, PS>((j - k) * R + k, s);
#pragma unroll
            for (int r = 0; r < R; r++) wb[lpos<ES, 0, PS>(r * P, 0)] = v[q][r];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Split-BF16 radix-16 stage on the matrix cores (FP32 data, the mixed form).
// The radix-16 DFT of 32 butterflies at a time is one real product D = A B
// with v_mfma_f32_32x32x16_bf16, in two K = 16 halves:
//   B = [Re v; Im v] (32 x 32: rows the 16 twiddled inputs' real then
//   imaginary parts, columns the butterflies),
//   A = [[Fr, -Fi], [Fi, Fr]] (32 x 32), F[i][r] = exp(DIR 2 pi i i r / 16),
// so that rows 0..15 of D are Re y and rows 16..31 Im y.  Every FP32 operand x
// is split into hi = bf16(x) and lo = bf16(x - hi); the products lo.hi +
// hi.lo, then hi.hi, accumulate in FP32 (the dropped lo.lo and the split leave
// about 2^-17 relative per stage, against 2^-24 for the FP32 butterflies).
// Lane l of the 32x32x16 instruction (wave64) holds A row l mod 32, columns
// 8 (l / 32) + j; B rows 8 (l / 32) + j, column l mod 32 (j < 8); D rows
// 8 (r / 4) + 4 (l / 32) + r mod 4, column l mod 32 (r < 16).  So lane l loads
// inputs 8 (l / 32) .. + 7 of butterfly l mod 32 once, and ends with eight of
// its complex outputs: y[o] = (D[r], D[r + 8]), o = 8 (r / 4) + 4 (l / 32) + r mod 4.
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// cos(2 pi m / 16), m = 0..15
__device__ __forceinline__ float cos16(int m) {
    const float c1 = 0.92387953251128675613f, c2 = 0.70710678118654752440f, c3 = 0.38268343236508977173f;
    const float t[16] = {1.f, c1, c2, c3, 0.f, -c3, -c2, -c1, -1.f, -c1, -c2, -c3, 0.f, c3, c2, c1};
    return t[m & 15];
}

// element t of (hi, lo) <- the split of x
__device__ __forceinline__ void split_bf16(float x, bf16x8& hi, bf16x8& lo, int t) {
    const __bf16 h = (__bf16)x;
    hi[t] = h;
    lo[t] = (__bf16)(x - (float)h);
}

// this lane's slices of the two K halves of A for direction DIR
struct Dft16Mfma {
    bf16x8 a0_hi, a0_lo, a1_hi, a1_lo;
};
template <int DIR>
__device__ __forceinline__ Dft16Mfma dft16_operands(int lane) {
    const int i = lane & 31, o = i & 15, half = lane >> 5;
    Dft16Mfma a;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int r = 8 * half + j, m = (o * r) & 15;
        const float c = cos16(m), s = DIR * cos16(m + 12);       // sin(x) = cos(x - pi/2)
        // K half 0 multiplies Re v: Fr (rows 0..15), Fi (rows 16..31);
        // K half 1 multiplies Im v: -Fi, Fr
        split_bf16(i < 16 ? c : s, a.a0_hi, a.a0_lo, j);
        split_bf16(i < 16 ? -s : c, a.a1_hi, a.a1_lo, j);
    }
    return a;
}

// One radix-16 Stockham stage (the indexing of fft_stage with R = 16) on the
// matrix cores: butterflies in groups of 32, one group per wave at a time;
// reads, a barrier, MFMAs and writes, a barrier.
template <int N, int P, int NSEQ, int ES, int SS, int NT, int DIR, int PS, typename TW>
__device__ __forceinline__ void fft_stage_mfma16(float2* buf, TW tw, int tid) {
    constexpr int R = 16, S = N / R, NB = S * NSEQ;
    constexpr int NG = (NB + 31) / 32, NW = NT / 64, GPW = (NG + NW - 1) / NW;
    static_assert(NT % 64 == 0, "the matrix-core stage needs whole waves");
    const int lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    bf16x8 rh[GPW], rl[GPW], ih[GPW], il[GPW];
    int jj[GPW], ss[GPW];
#pragma unroll
    for (int q = 0; q < GPW; q++) {
        const int g = wave + q * NW, b = 32 * g + col;
        int j, s;
        if constexpr (ES == 1) { j = b % S; s = b / S; }
        else { s = b % NSEQ; j = b / NSEQ; }
        jj[q] = j;
        ss[q] = s;
        if (g < NG && b < NB) {
            const int k = j % P;
            const float2* rb = buf + lpos<ES, SS, PS>(j, s);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int r = 8 * half + t;
                float2 v = rb[lpos<ES, 0, PS>(r * S, 0)];
                if constexpr (P > 1) {
                    constexpr int STRIDE = N / (P * R);
                    if (r > 0) {
                        float2 w = tw[r * k * STRIDE];
                        if (DIR > 0) w.y = -w.y;
                        v = cmul(v, w);
                    }
                }
                split_bf16(v.x, rh[q], rl[q], t);
                split_bf16(v.y, ih[q], il[q], t);
            }
        } else {
#pragma unroll
            for (int t = 0; t < 8; t++) rh[q][t] = rl[q][t] = ih[q][t] = il[q][t] = (__bf16)0.f;
        }
    }
    __syncthreads();
    // the DFT matrix slices, formed after the reads (the lane index is made
    // opaque so that they are not computed early and kept live across the reads)
    int lane_v = lane;
    asm volatile("" : "+v"(lane_v));
    const Dft16Mfma A = dft16_operands<DIR>(lane_v);
#pragma unroll
    for (int q = 0; q < GPW; q++) {
        const int g = wave + q * NW;
        if (g >= NG) continue;            // uniform over the wave
        f32x16 d;
#pragma unroll
        for (int r = 0; r < 16; r++) d[r] = 0.f;
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a0_lo, rh[q], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a0_hi, rl[q], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a1_lo, ih[q], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a1_hi, il[q], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a0_hi, rh[q], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.a1_hi, ih[q], d, 0, 0, 0);
        if (32 * g + col < NB) {
            const int j = jj[q], s = ss[q], k = j % P;
            float2* wb = buf + lpos<ES, SS, PS>((j - k) * R + k, s);
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int o = 8 * (r / 4) + 4 * half + (r % 4);
                wb[lpos<ES, 0, PS>(o * P, 0)] = make_float2(d[r], d[r + 8]);
            }
        }
    }
    __syncthreads();
}

// MF = 1: radix-16 stages on the matrix cores (float2 data only) where the
// remaining length allows, the others as usual
template <int N, int P, int NSEQ, int ES, int SS, int NT, int DIR, int PS, int MF = 0>
struct fft_stages {
    template <typename C, typename TW>
    static __device__ __forceinline__ void run(C* buf, TW tw, int tid) {
        if constexpr (MF && sizeof(C) == sizeof(float2) && (N / P) % 16 == 0) {
            fft_stage_mfma16<N, P, NSEQ, ES, SS, NT, DIR, PS>(buf, tw, tid);
            fft_stages<N, P * 16, NSEQ, ES, SS, NT, DIR, PS, MF>::run(buf, tw, tid);
        } else {
            constexpr int R = radix_of<N / P>::value;
            fft_stage<N, P, R, NSEQ, ES, SS, NT, DIR, PS>(buf, tw, tid);
            fft_stages<N, P * R, NSEQ, ES, SS, NT, DIR, PS, MF>::run(buf, tw, tid);
        }
    }
};
template <int N, int NSEQ, int ES, int SS, int NT, int DIR, int PS, int MF>
struct fft_stages<N, N, NSEQ, ES, SS, NT, DIR, PS, MF> {
    template <typename C, typename TW>
    static __device__ __forceinline__ void run(C*, TW, int) {}
};

// C = double2 (FP64 form) or float2 (mixed form), deduced from buf; tw is the
// global twiddle table (const C*) or an LDS two-level table (TwL); MF = 1 puts
// the radix-16 stages of a float2 transform on the matrix cores
template <int N, int NSEQ, int ES, int SS, int NT, int DIR, int PS = 0, int MF = 0, typename C, typename TW>
__device__ __forceinline__ void lds_fft(C* buf, TW tw, int tid) {
    fft_stages<N, 1, NSEQ, ES, SS, NT, DIR, PS, MF>::run(buf, tw, tid);
}

// Twiddles w^m = exp(-2 pi i m / N) from two small LDS tables holding entries
// of the global table: lo[i] = w^i (i < L), hi[j] = w^(j L) (j < N / L), and
// w^m = hi[m / L] lo[m mod L] (one complex product, about 2 ulp).  A stage then
// waits on LDS (tens of cycles) instead of a global load that the streamed
// tiles keep evicting from L1.
constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
template <int N>
struct tw_split {
    static constexpr int LB = (ilog2c(N) + 1) / 2, L = 1 << LB, H = N / L;
};
template <typename C, int N>
struct TwL {
    const C* t;     // lo (L entries), then hi (H entries)
    __device__ __forceinline__ C operator[](int m) const {
        return cmul(t[tw_split<N>::L + (m >> tw_split<N>::LB)], t[m & (tw_split<N>::L - 1)]);
    }
};
// the twiddle source of a kernel: the LDS tables (USE_LDS) or the global table
template <int N, bool USE_LDS, typename C>
__device__ __forceinline__ auto tw_source(const C* __restrict__ tw);
// as passed to one FFT: a global table pointer made opaque (its loads are not
// kept live from one FFT to the next), LDS tables as they are
template <typename C>
__device__ __forceinline__ const C* tw_fft(const C* p) {
    // through a global-memory pointer, so that loads after the empty asm stay
    // global loads whatever the compiler knew of p (flat loads would also count
    // against the LDS wait counter)
    typedef const __attribute__((address_space(1))) C GC;
    GC* g = (GC*)p;
    asm volatile("" : "+s"(g));
    return (const C*)g;
}
template <typename C, int N>
__device__ __forceinline__ TwL<C, N> tw_fft(TwL<C, N> t) { return t; }

// fills the workgroup's tables from the global table (one barrier)
template <int N, typename C>
__device__ __forceinline__ TwL<C, N> lds_twiddles(const C* __restrict__ tw) {
    constexpr int L = tw_split<N>::L, H = tw_split<N>::H;
    __shared__ C t[L + H];
    for (int i = threadIdx.x; i < L + H; i += blockDim.x) t[i] = i < L ? tw[i] : tw[(i - L) * L];
    __syncthreads();
    return TwL<C, N>{t};
}
template <int N, bool USE_LDS, typename C>
__device__ __forceinline__ auto tw_source(const C* __restrict__ tw) {
    if constexpr (USE_LDS) return lds_twiddles<N>(tw);
    else return tw;
}

}  // namespace jw
