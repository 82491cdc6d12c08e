// jw_packed.hpp -- the unfused passes of the generator-basis step, templated on
// the storage type C (double2: FP64 form, float2: mixed form): the 2-D FFT
// row and column passes and the per-site kernels (noise, kernel multiply,
// U xi U^dagger, update).  Site algebra is FP64 in both forms.  Used where the
// fused passes are not (N outside 64..2048, the LINEAR kernel's 9-component
// transforms, jw_set_fused(ctx, 0)).
#pragma once

#include "jw_common.hpp"
#include "jw_engine.hpp"

namespace jw {

// ===========================================================================
// FFT passes
// ===========================================================================
// Rows per workgroup (at most N: a workgroup never reaches past its plane) and
// columns per workgroup (at most N), each a divisor of N.  Column strips are
// as wide as about 150 KB of LDS allows (at most 16 columns).
template <int N, typename C = double2>
struct RowCfg {
    static constexpr int RPW0 = (N >= 1024) ? 1 : (1024 / N);
    static constexpr int RPW = RPW0 < N ? RPW0 : N;
    static constexpr int ELEMS = RPW * N;
    static constexpr int NT = (ELEMS / 8 < 64) ? 64 : ((ELEMS / 8 > 1024) ? 1024 : ELEMS / 8);
    static_assert(RPW >= 1 && RPW <= N && N % RPW == 0, "row tiling must divide the plane");
};
template <int N, typename C = double2>
struct ColCfg {
    static constexpr int BUDGET = 150 * 1024;
    static constexpr int W1 = BUDGET / (padded<N, 3>::value * (int)sizeof(C));
    static constexpr int W0 = W1 >= 16 ? 16 : (W1 >= 8 ? 8 : (W1 >= 4 ? 4 : (W1 >= 2 ? 2 : 1)));
    static constexpr int W = W0 < N ? W0 : N;
    static constexpr int ELEMS = W * N;
    static constexpr int NT = (ELEMS / 8 < 64) ? 64 : ((ELEMS / 8 > 1024) ? 1024 : ELEMS / 8);
    static_assert(W >= 1 && W <= N && N % W == 0, "column tiling must divide the plane");
};

template <int N, int DIR, typename C>
__global__ __launch_bounds__((RowCfg<N, C>::NT)) void k_fft_rows(C* __restrict__ data, size_t plane_stride,
                                                              const C* __restrict__ tw, double scale) {
    typedef typename scalar_of<C>::type T;
    using G = RowCfg<N, C>;
    constexpr int PN = padded<N, 4>::value;            // padded row length in LDS
    __shared__ C lds[G::RPW * PN];
    const int tid = threadIdx.x;
    C* base = data + blockIdx.y * plane_stride + (size_t)blockIdx.x * G::ELEMS;
    for (int i = tid; i < G::ELEMS; i += G::NT) lds[lpos<1, PN, 4>(i % N, i / N)] = base[i];
    __syncthreads();
    lds_fft<N, G::RPW, 1, PN, G::NT, DIR, 4>(lds, tw, tid);
    for (int i = tid; i < G::ELEMS; i += G::NT) base[i] = cscale(lds[lpos<1, PN, 4>(i % N, i / N)], (T)scale);
}

template <int N, int DIR, typename C>
__global__ __launch_bounds__((ColCfg<N, C>::NT)) void k_fft_cols(C* __restrict__ data, size_t plane_stride,
                                                              const C* __restrict__ tw, double scale) {
    typedef typename scalar_of<C>::type T;
    using G = ColCfg<N, C>;
    __shared__ C lds[padded<N, 3>::value * G::W];
    const int tid = threadIdx.x;
    C* base = data + blockIdx.y * plane_stride + (size_t)blockIdx.x * G::W;
    for (int i = tid; i < G::ELEMS; i += G::NT)
        lds[lpos<G::W, 1, 3>(i / G::W, i % G::W)] = base[(size_t)(i / G::W) * N + (i % G::W)];
    __syncthreads();
    lds_fft<N, G::W, G::W, 1, G::NT, DIR, 3>(lds, tw, tid);
    for (int i = tid; i < G::ELEMS; i += G::NT)
        base[(size_t)(i / G::W) * N + (i % G::W)] = cscale(lds[lpos<G::W, 1, 3>(i / G::W, i % G::W)], (T)scale);
}

template <int N, typename C>
static hipError_t launch_rows(jw_ctx* c, C* p, int np, int dir, double scale, const C* tw) {
    using G = RowCfg<N, C>;
    dim3 grid(N / G::RPW, np);
    KTimer t(c, K_FFT_ROWS);
    if (dir) hipLaunchKernelGGL((k_fft_rows<N, -1, C>), grid, dim3(G::NT), 0, c->stream, p, c->V, tw, scale);
    else hipLaunchKernelGGL((k_fft_rows<N, +1, C>), grid, dim3(G::NT), 0, c->stream, p, c->V, tw, scale);
    return hipGetLastError();
}

template <int N, typename C>
static hipError_t launch_cols(jw_ctx* c, C* p, int np, int dir, double scale, const C* tw) {
    using G = ColCfg<N, C>;
    dim3 grid(N / G::W, np);
    KTimer t(c, K_FFT_COLS);
    if (dir) hipLaunchKernelGGL((k_fft_cols<N, -1, C>), grid, dim3(G::NT), 0, c->stream, p, c->V, tw, scale);
    else hipLaunchKernelGGL((k_fft_cols<N, +1, C>), grid, dim3(G::NT), 0, c->stream, p, c->V, tw, scale);
    return hipGetLastError();
}

#define JW_SIZE_SWITCH(N, CALL)                         \
    switch (N) {                                        \
        case 4: return CALL(4);                         \
        case 8: return CALL(8);                         \
        case 16: return CALL(16);                       \
        case 32: return CALL(32);                       \
        case 64: return CALL(64);                       \
        case 128: return CALL(128);                     \
        case 256: return CALL(256);                     \
        case 512: return CALL(512);                     \
        case 1024: return CALL(1024);                   \
        case 2048: return CALL(2048);                   \
        case 4096: return CALL(4096);                   \
        case 8192: return CALL(8192);                   \
        default: return hipErrorInvalidValue;           \
    }

template <typename C>
static const C* twiddles_of(const jw_ctx* c) {
    if constexpr (sizeof(C) == sizeof(double2)) return c->tw;
    else return c->twf;
}

template <typename C>
static hipError_t fft_rows_t(jw_ctx* c, C* planes, int nplanes, int dir, double scale) {
#define CALL_ROWS(n) launch_rows<n, C>(c, planes, nplanes, dir, scale, twiddles_of<C>(c))
    JW_SIZE_SWITCH(c->N, CALL_ROWS)
#undef CALL_ROWS
}

template <typename C>
static hipError_t fft_cols_t(jw_ctx* c, C* planes, int nplanes, int dir, double scale) {
#define CALL_COLS(n) launch_cols<n, C>(c, planes, nplanes, dir, scale, twiddles_of<C>(c))
    JW_SIZE_SWITCH(c->N, CALL_COLS)
#undef CALL_COLS
}

// unitary 2-D transform of nplanes planes: dir 1 -> exp(-i k.x), dir 0 -> exp(+i k.x),
// scaled by `scale` once (1/N for the reference's 1/sqrt(V))
template <typename C>
static hipError_t fft2d_t(jw_ctx* c, C* planes, int nplanes, int dir, double scale) {
    hipError_t e = fft_rows_t(c, planes, nplanes, dir, 1.0);
    if (e != hipSuccess) return e;
    return fft_cols_t(c, planes, nplanes, dir, scale);
}

// ===========================================================================
// generator-basis engine (packed planes)
//   Xi planes 0..3: (n0,n1),(n2,n3),(n4,n5),(n6,n7) of xi_x; planes 4..7: same for xi_y
//   Vv planes: the same layout for V_x, V_y;  ZA, ZB: 4 planes (A, B)
// ===========================================================================
template <typename C>
__global__ __launch_bounds__(256) void k_gen_noise_packed(C* __restrict__ Xi, size_t V, NoiseSrc ns,
                                                          const double* __restrict__ sa_pos) {
    typedef typename scalar_of<C>::type T;
    GRID_STRIDE(i, V) {
        T z[16];
        normals16_t(ns, i, z);
        const double s = sa_pos ? sa_pos[i] : 1.0;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            Xi[p * V + i] = mkc<C>(z[4 * p] * s, z[4 * p + 2] * s);           // n_{2p}, n_{2p+1}
            Xi[(4 + p) * V + i] = mkc<C>(z[4 * p + 1] * s, z[4 * p + 3] * s); // m_{2p}, m_{2p+1}
        }
    }
}

// Z[p] = i (kx(k) X[p] + ky(k) Y[p]),  kx = -2 pi c pbarX / phat2, zero at k = 0;
// scale_inputs: X, Y *= sa(k) first (and written back) for per-mode correlated noise.
template <typename C>
__global__ __launch_bounds__(256) void k_kernel_mult_packed(C* __restrict__ Z, C* __restrict__ X, C* __restrict__ Y,
                                                            int N, int coupling, const double* __restrict__ q4,
                                                            const double* __restrict__ sb,
                                                            const double* __restrict__ sa, int scale_inputs) {
    typedef typename scalar_of<C>::type T;
    const size_t V = (size_t)N * N;
    GRID_STRIDE(i, V) {
        const int ix = (int)(i / N), iy = (int)(i % N);
        const double ph = q4[ix] + q4[iy];
        const double sc = (coupling == 0) ? sa[i] : 1.0;
        const double kxv = (i == 0) ? 0.0 : -2.0 * M_PI * sc * sb[ix] / ph;
        const double kyv = (i == 0) ? 0.0 : -2.0 * M_PI * sc * sb[iy] / ph;
        const double si = scale_inputs ? sa[i] : 1.0;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            C x = X[p * V + i], y = Y[p * V + i];
            if (scale_inputs) {
                x = cscale(x, (T)si);
                y = cscale(y, (T)si);
                X[p * V + i] = x;
                Y[p * V + i] = y;
            }
            Z[p * V + i] = mkc<C>(-(kxv * x.y + kyv * y.y), kxv * x.x + kyv * y.x);   // i (kx x + ky y)
        }
    }
}

// U for the FP64 site algebra of the packed kernels: all nine planes, or (U6,
// the large-lattice step) rows 0 and 1 with row 2 rebuilt as conj(row 0 x row 1)
template <int U6, typename C>
__device__ __forceinline__ void load_u_d(m3& u, const C* __restrict__ U, size_t V, size_t i) {
#pragma unroll
    for (int k = 0; k < (U6 ? 6 : 9); k++) u.m[k] = to_d(U[k * V + i]);
    if constexpr (U6 != 0) {
        u.m[6] = conjg(csub(cmul(u.m[1], u.m[5]), cmul(u.m[2], u.m[4])));
        u.m[7] = conjg(csub(cmul(u.m[2], u.m[3]), cmul(u.m[0], u.m[5])));
        u.m[8] = conjg(csub(cmul(u.m[0], u.m[4]), cmul(u.m[1], u.m[3])));
    }
}

template <typename C, int U6 = 0>
__global__ __launch_bounds__(256) void k_uxiu_packed(C* __restrict__ Vv, const C* __restrict__ U,
                                                     const C* __restrict__ Xi_pos, size_t V, NoiseSrc ns,
                                                     const double* __restrict__ sa_pos) {
    typedef typename scalar_of<C>::type T;
    GRID_STRIDE(i, V) {
        double hx[8], hy[8];
        if (Xi_pos) {
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const C a = Xi_pos[p * V + i], b = Xi_pos[(4 + p) * V + i];
                hx[2 * p] = a.x; hx[2 * p + 1] = a.y;
                hy[2 * p] = b.x; hy[2 * p + 1] = b.y;
            }
        } else {
            T z[16];
            normals16_t(ns, i, z);
            const double s = sa_pos ? sa_pos[i] : 1.0;
#pragma unroll
            for (int k = 0; k < 8; k++) { hx[k] = z[2 * k] * s; hy[k] = z[2 * k + 1] * s; }
        }
        m3 u;
        load_u_d<U6>(u, U, V, i);
        double vx[8], vy[8];
        uxiu_gen(u, hx, vx);
        uxiu_gen(u, hy, vy);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            Vv[p * V + i] = mkc<C>(vx[2 * p], vx[2 * p + 1]);
            Vv[(4 + p) * V + i] = mkc<C>(vy[2 * p], vy[2 * p + 1]);
        }
    }
}

template <typename C, int U6 = 0>
__global__ __launch_bounds__(256) void k_update_packed(C* __restrict__ U, const C* __restrict__ ZA,
                                                       const C* __restrict__ ZB, size_t V, double ss, int reunit) {
    GRID_STRIDE(i, V) {
        double ha[8], hb[8];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const C a = ZA[p * V + i], b = ZB[p * V + i];
            ha[2 * p] = a.x; ha[2 * p + 1] = a.y;
            hb[2 * p] = b.x; hb[2 * p + 1] = b.y;
        }
        const m3 eb = exp_i_gen(hb, ss);
        const m3 ea = exp_i_gen(ha, -ss);
        m3 u;
        load_u_d<U6>(u, U, V, i);
        m3 r = mmul(mmul(eb, u), ea);
        if (reunit) r = su3_project(r);
#pragma unroll
        for (int k = 0; k < (U6 ? 6 : 9); k++) U[k * V + i] = mkc<C>(r.m[k].x, r.m[k].y);
    }
}

// the unfused step in either precision: noise, 2-D FFTs, kernel multiplies,
// U xi U^dagger, update (reunit: project U onto SU(3) after it)
template <typename C>
static hipError_t step_unfused_t(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const NoiseSrc& ns,
                                 const double* sa_pos, const double* sa_k, int kc, int corr, double ss, int reunit) {
    const size_t V = c->V;
    const double s = 1.0 / c->N;
    const unsigned g = grid_for(V, 256);
    hipError_t e;
#define JW_TRY(x) if ((e = (x)) != hipSuccess) return e
    {
        KTimer t(c, K_GEN_NOISE);
        hipLaunchKernelGGL(k_gen_noise_packed<C>, dim3(g), dim3(256), 0, c->stream, Xi, V, ns, sa_pos);
        JW_TRY(hipGetLastError());
    }
    JW_TRY(fft2d_t(c, Xi, 8, 1, s));
    {
        KTimer t(c, K_KMULT);
        hipLaunchKernelGGL(k_kernel_mult_packed<C>, dim3(g), dim3(256), 0, c->stream, ZA, Xi, Xi + 4 * V, c->N, kc,
                           c->q4, c->sb, sa_k, corr);
        JW_TRY(hipGetLastError());
    }
    JW_TRY(fft2d_t(c, ZA, 4, 0, s));
    if (corr) JW_TRY(fft2d_t(c, Xi, 8, 0, s));
    {
        KTimer t(c, K_UXIU);
        hipLaunchKernelGGL(k_uxiu_packed<C>, dim3(g), dim3(256), 0, c->stream, Vv, U, corr ? Xi : nullptr, V, ns,
                           sa_pos);
        JW_TRY(hipGetLastError());
    }
    JW_TRY(fft2d_t(c, Vv, 8, 1, s));
    {
        KTimer t(c, K_KMULT);
        hipLaunchKernelGGL(k_kernel_mult_packed<C>, dim3(g), dim3(256), 0, c->stream, ZB, Vv, Vv + 4 * V, c->N, kc,
                           c->q4, c->sb, sa_k, 0);
        JW_TRY(hipGetLastError());
    }
    JW_TRY(fft2d_t(c, ZB, 4, 0, s));
    {
        KTimer t(c, K_UPDATE);
        hipLaunchKernelGGL(k_update_packed<C>, dim3(g), dim3(256), 0, c->stream, U, ZA, ZB, V, ss, reunit);
        JW_TRY(hipGetLastError());
    }
#undef JW_TRY
    return hipSuccess;
}


// ===========================================================================
// Large lattices (N = 4096, 8192): the fused column passes of jw_engine.hpp
// (forward column FFT, kernel multiply, inverse column FFT in LDS per strip;
// the momentum-space noise drawn there) with the row transforms and the site
// algebra as separate passes, since one lattice row of four planes no longer
// fits LDS (FP64 at N = 4096 needs 256 KB).  U is held as rows 0 and 1 (row 2
// rebuilt on load), as in the fused step.  Per step, in complex planes moved:
//   [position-space noise: noise W 8, forward row FFT R 8 W 8]
//   column A                     R 8 (0 drawn) W 4 (+8 correlated or drawn)
//   inverse row FFT of A         R 4 W 4
//   correlated or drawn noise, k_row_uxiu_1p (FP32 at N = 8192):
//     inverse row FFT of xi~, U xi~ U^dagger, forward row FFT of V  R 20 W 8
//   otherwise: [inverse row FFT of xi~ R 8 W 8]
//     U xi U^dagger              R 6 (+8) W 8
//     forward row FFT of V       R 8 W 8
//   column B                     R 8 W 4
//   inverse row FFT of B         R 4 W 4
//   update                       R 14 W 6
// 88 planes with the noise drawn per mode in FP32 at N = 8192 (114 otherwise),
// against 66 for the fused step of N <= 2048.  Column strips: BF = 32 (16-byte
// row segments) at N = 8192, BF = 0 or 32 at N = 4096 (the intermediate planes
// stay row-major).
// ===========================================================================
// U xi~ U^dagger of the large-lattice step as one pass over a lattice row (FP32
// at N = 8192; FP32 at 4096 and FP64 spill registers), where four planes of a row do not fit
// LDS: each plane of xi~ is inverse-transformed alone in LDS and its values at
// the thread's SPT sites are kept in registers; with the four planes of a
// direction held, V = U xi~ U^dagger per site (U as rows 0 and 1, read once per
// direction), and the four planes of V go through LDS one at a time for the
// forward row transform.  In place of the inverse row FFT of xi~ (R 8 W 8), the
// site pass (R 14 W 8) and the forward row FFT of V (R 8 W 8): R 20 W 8.
template <int N, typename C>
struct Row1P {
    static constexpr int NT = 1024, SPT = N / NT;
    static constexpr bool ok = N == 8192 && sizeof(C) == sizeof(float2);
};
template <int N, typename C>
__global__ __launch_bounds__(1024, 1) void k_row_uxiu_1p(C* __restrict__ Vv, const C* __restrict__ U,
                                                          const C* __restrict__ Xs, size_t V,
                                                          const C* __restrict__ tw) {
    typedef typename scalar_of<C>::type T;
    constexpr int NT = Row1P<N, C>::NT, SPT = Row1P<N, C>::SPT;
    constexpr int PN = padded<N, 4>::value;
    static_assert(SPT * NT == N, "row tiling");
    __shared__ C lds[PN];
    const int tid = threadIdx.x;
    const size_t row0 = (size_t)blockIdx.x * N;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
        C h[4][SPT];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const C* src = Xs + (size_t)(4 * half + p) * V + row0;
            const int tl = opaque_v(tid);
#pragma unroll
            for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tl + q * NT, 0)] = src[tl + q * NT];
            __syncthreads();
            lds_fft<N, 1, 1, PN, NT, +1, 4>(lds, tw_fft(tw), opaque_v(tid));
            const int tr = opaque_v(tid);
#pragma unroll
            for (int q = 0; q < SPT; q++) h[p][q] = lds[lpos<1, PN, 4>(tr + q * NT, 0)];
            __syncthreads();
        }
        const int tu = opaque_v(tid);
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const size_t i = row0 + tu + q * NT;
            T hh[8];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                hh[2 * p] = h[p][q].x;
                hh[2 * p + 1] = h[p][q].y;
            }
            typename site_mat<C>::type u;
            load_u(u, U, V, i);
            T v[8];
            uxiu_gen(u, hh, v);
#pragma unroll
            for (int p = 0; p < 4; p++) h[p][q] = mkc<C>(v[2 * p], v[2 * p + 1]);
        }
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int tl = opaque_v(tid);
#pragma unroll
            for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tl + q * NT, 0)] = h[p][q];
            __syncthreads();
            lds_fft<N, 1, 1, PN, NT, -1, 4>(lds, tw_fft(tw), opaque_v(tid));
            C* dst = Vv + (size_t)(4 * half + p) * V + row0;
            const int ts = opaque_v(tid);
#pragma unroll
            for (int q = 0; q < SPT; q++) dst[ts + q * NT] = lds[lpos<1, PN, 4>(ts + q * NT, 0)];
            __syncthreads();
        }
    }
}

template <int N, int BF, typename C>
static hipError_t step_large_n(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const StepArgs& a) {
    const size_t V = c->V;
    const C* tw = twiddles_of<C>(c);
    const unsigned g = grid_for(V, 256);
    const int wb = a.corr || a.kspace;      // xi~ written back by column A
    hipError_t e;
#define JW_TRY(x) if ((e = (x)) != hipSuccess) return e
    if (!a.kspace) {
        {
            KTimer t(c, K_GEN_NOISE);
            hipLaunchKernelGGL(k_gen_noise_packed<C>, dim3(g), dim3(256), 0, c->stream, Xi, V, a.ns, a.sa_pos);
            JW_TRY(hipGetLastError());
        }
        JW_TRY((launch_rows<N, C>(c, Xi, 8, 1, 1.0, tw)));
    }
    {
        KTimer t(c, K_COL_KMULT);
        JW_TRY((col_pass<N, BF>(c, ZA, Xi, Xi + 4 * V, a.sa_k, a.sqrt_coupling, a.corr, 1, a.ns, a.kspace, a.corr)));
    }
    JW_TRY((launch_rows<N, C>(c, ZA, 4, 0, 1.0, tw)));
    bool one_pass = false;
    if constexpr (Row1P<N, C>::ok) {
        if (wb) {
            KTimer t(c, K_UXIU);
            hipLaunchKernelGGL((k_row_uxiu_1p<N, C>), dim3(N), dim3(Row1P<N, C>::NT), 0, c->stream, Vv, U, Xi, V, tw);
            JW_TRY(hipGetLastError());
            one_pass = true;
        }
    }
    if (!one_pass) {
        if (wb) JW_TRY((launch_rows<N, C>(c, Xi, 8, 0, 1.0, tw)));
        {
            KTimer t(c, K_UXIU);
            hipLaunchKernelGGL((k_uxiu_packed<C, 1>), dim3(g), dim3(256), 0, c->stream, Vv, U, wb ? Xi : nullptr, V,
                               a.ns, a.sa_pos);
            JW_TRY(hipGetLastError());
        }
        JW_TRY((launch_rows<N, C>(c, Vv, 8, 1, 1.0, tw)));
    }
    {
        KTimer t(c, K_COL_KMULT);
        JW_TRY((col_pass<N, BF>(c, ZB, Vv, Vv + 4 * V, a.sa_k, a.sqrt_coupling, 0, 1, a.ns, 0, 0)));
    }
    JW_TRY((launch_rows<N, C>(c, ZB, 4, 0, 1.0, tw)));
    {
        KTimer t(c, K_UPDATE);
        hipLaunchKernelGGL((k_update_packed<C, 1>), dim3(g), dim3(256), 0, c->stream, U, ZA, ZB, V, a.sqrt_step,
                           a.reunit);
        JW_TRY(hipGetLastError());
    }
#undef JW_TRY
    return hipSuccess;
}

// the column strips of a large lattice: 16-byte segments one plane at a time
// at N = 8192 (the only strip that fits LDS); at N = 4096 the X + Y tiles of
// one column (layout 0, default) or one plane at a time over 16-byte segments
// (layout 32)
template <typename C>
static hipError_t step_large_t(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const StepArgs& a) {
    if (a.nbatch != 1) return hipErrorInvalidValue;
    if (c->N == 8192) return step_large_n<8192, 32>(c, U, Xi, Vv, ZA, ZB, a);
    if (c->N == 4096) {
        if (c->layout == 32) return step_large_n<4096, 32>(c, U, Xi, Vv, ZA, ZB, a);
        return step_large_n<4096, 0>(c, U, Xi, Vv, ZA, ZB, a);
    }
    return hipErrorInvalidValue;
}

}  // namespace jw
