// jw_common.hpp -- helpers shared by the HIP translation units: kernel-class
// timing, launch geometry, bit-matching complex arithmetic, the noise readers
// and the lambda-basis map, SoA 3x3 loads/stores and U xi U^dagger in the
// generator basis.
#pragma once
#include <type_traits>

#include "jw_device.hpp"
#include "jw_internal.hpp"

namespace jw {

// ---------------------------------------------------------------------------
// timing of kernel classes with HIP events on the context stream
// ---------------------------------------------------------------------------
enum KClass { K_FFT_ROWS = 0, K_FFT_COLS, K_GEN_NOISE, K_KMULT, K_UXIU, K_UPDATE, K_MV, K_OBS, K_DROPIN,
              K_ROW_NOISE, K_COL_KMULT, K_ROW_UXIU, K_ROW_UPDATE, K_NCLASS };

struct KTimer {
    jw_ctx* c;
    int cls;
    hipEvent_t a, b;
    KTimer(jw_ctx* c_, int cls_) : c(c_), cls(cls_) {
        if (c->timing) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    ~KTimer() {
        if (c->timing) {
            hipEventRecord(b, c->stream);
            hipEventSynchronize(b);
            float ms = 0.f;
            hipEventElapsedTime(&ms, a, b);
            c->kernel_ms[cls] += ms;
            c->kernel_calls[cls] += 1;
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    }
};

static inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)(g ? g : 1);
}

#define GRID_STRIDE(i, n) for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); \
                               i += (size_t)gridDim.x * blockDim.x)

// non-contracting complex helpers for the bit-matching kernels
__device__ __forceinline__ cplx xmul(cplx a, cplx b) {
#pragma clang fp contract(off)
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx xadd(cplx a, cplx b) {
#pragma clang fp contract(off)
    return mk(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ m3 xmmul(const m3& a, const m3& b) {
    m3 c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx s = mk(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < 3; k++) s = xadd(s, xmul(a.m[3 * i + k], b.m[3 * k + j]));
            c.m[3 * i + j] = s;
        }
    return c;
}

// ===========================================================================
// noise
// ===========================================================================
__device__ __forceinline__ void normals16(const NoiseSrc& ns, size_t site, double z[16]) {
    if (ns.inj) {
        const double2* p = reinterpret_cast<const double2*>(ns.inj + 16 * site);
#pragma unroll
        for (int q = 0; q < 8; q++) { const double2 v = p[q]; z[2 * q] = v.x; z[2 * q + 1] = v.y; }
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++)
            box_muller_pair((uint32_t)site, ns.step, ns.tag, (uint32_t)q, ns.seed, ns.traj, z[2 * q], z[2 * q + 1]);
    }
}

__device__ __forceinline__ void normals8(const NoiseSrc& ns, size_t site, double z[8]) {
    if (ns.inj) {
        const double2* p = reinterpret_cast<const double2*>(ns.inj + 8 * site);
#pragma unroll
        for (int q = 0; q < 4; q++) { const double2 v = p[q]; z[2 * q] = v.x; z[2 * q + 1] = v.y; }
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            box_muller_pair((uint32_t)site, ns.step, ns.tag, (uint32_t)q, ns.seed, ns.traj, z[2 * q], z[2 * q + 1]);
    }
}

// 16 normals of a site in the storage precision: FP64 pairs, or the FP32 pairs
// of the same Philox words (box_muller_pair_f); injected normals are rounded
template <typename T>
__device__ __forceinline__ void normals16_t(const NoiseSrc& ns, size_t site, T z[16]) {
    if constexpr (sizeof(T) == sizeof(double)) {
        normals16(ns, site, z);
    } else if (ns.inj) {
        const double2* p = reinterpret_cast<const double2*>(ns.inj + 16 * site);
#pragma unroll
        for (int q = 0; q < 8; q++) { const double2 v = p[q]; z[2 * q] = (T)v.x; z[2 * q + 1] = (T)v.y; }
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++)
            box_muller_pair_f((uint32_t)site, ns.step, ns.tag, (uint32_t)q, ns.seed, ns.traj, z[2 * q], z[2 * q + 1]);
    }
}

// scalar type of a complex storage type
template <typename C> struct scalar_of;
template <> struct scalar_of<double2> { typedef double type; };
template <> struct scalar_of<float2> { typedef float type; };

// the reference's lambda-basis site map (field.h:2169-2231; setMVModel field.h:714-745),
// operation for operation
__device__ __forceinline__ void lambda_entries(const double n[8], cplx u[9]) {
#pragma clang fp contract(off)
    const double s3 = sqrt(3.0);
    const double d = n[7] / s3;
    u[0] = mk(n[2] + d, 0.0);
    u[1] = mk(n[0], n[1]);
    u[2] = mk(n[3], n[4]);
    u[3] = mk(n[0], -n[1]);
    u[4] = mk(-n[2] + d, 0.0);
    u[5] = mk(n[5], n[6]);
    u[6] = mk(n[3], -n[4]);
    u[7] = mk(n[5], -n[6]);
    u[8] = mk(-2.0 * n[7] / s3, 0.0);
}

__device__ __forceinline__ m3 load9(const cplx* __restrict__ f, size_t V, size_t i) {
    m3 a;
#pragma unroll
    for (int k = 0; k < 9; k++) a.m[k] = f[k * V + i];
    return a;
}
__device__ __forceinline__ void store9(cplx* __restrict__ f, size_t V, size_t i, const m3& a) {
#pragma unroll
    for (int k = 0; k < 9; k++) f[k * V + i] = a.m[k];
}

// V_i = U xi_i U^dagger in generator coefficients (upper triangle + real diagonal)
// Computed one row r of W = U X at a time, then the entries V_rc (c >= r) that
// gen_from_upper reads: keeps few values live.  M = m3 with S = double (FP64
// form), or m3f with S = float (the mixed form's products).
template <typename M, typename S>
__device__ __forceinline__ void uxiu_gen(const M& u, const S h[8], S out[8]) {
    typedef decltype(u.m[0]) CR;
    typedef typename std::remove_const<typename std::remove_reference<CR>::type>::type Cx;
    const S d8 = h[7] * (S)JW_INV_SQRT3;
    const S x00 = h[2] + d8, x11 = -h[2] + d8, x22 = (S)-2.0 * d8;
    const Cx x01 = mkc<Cx>(h[0], h[1]), x02 = mkc<Cx>(h[3], h[4]), x12 = mkc<Cx>(h[5], h[6]);
    S d[3];
    Cx up[3];   // V01, V02, V12
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const Cx a = u.m[3 * r], b = u.m[3 * r + 1], c = u.m[3 * r + 2];
        // row r of U X, X Hermitian: X10 = conj(x01), X20 = conj(x02), X21 = conj(x12)
        const Cx w0 = cadd(cadd(cscale(a, x00), cmulc(b, x01)), cmulc(c, x02));
        const Cx w1 = cadd(cadd(cmul(a, x01), cscale(b, x11)), cmulc(c, x12));
        const Cx w2 = cadd(cadd(cmul(a, x02), cmul(b, x12)), cscale(c, x22));
#pragma unroll
        for (int cc = r; cc < 3; cc++) {
            Cx s = cmulc(w0, u.m[3 * cc]);
            s = cadd(s, cmulc(w1, u.m[3 * cc + 1]));
            s = cadd(s, cmulc(w2, u.m[3 * cc + 2]));
            if (cc == r) d[r] = s.x;
            else up[r + cc - 1] = s;
        }
    }
    out[0] = up[0].x; out[1] = up[0].y;
    out[3] = up[1].x; out[4] = up[1].y;
    out[5] = up[2].x; out[6] = up[2].y;
    out[2] = (S)0.5 * (d[0] - d[1]);
    out[7] = (d[0] + d[1] - (S)2.0 * d[2]) * (S)(0.5 * JW_INV_SQRT3);
}

}  // namespace jw
