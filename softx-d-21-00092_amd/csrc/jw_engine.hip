// jw_engine.hip -- fused generator-basis Langevin step for 64 <= N <= 2048.
//
// One step (main_optimized.cpp:243-304 in the generator basis) as six passes
// over HBM instead of the reference's 8 nine-component FFTs and 5 site loops:
//   row_noise_fwd  : Philox/injected noise -> xi planes, forward row FFT    W 8
//   col_kmult (A)  : forward column FFT, kernel multiply, inverse column FFT R 8 W 4
//                    (correlated noise: also writes F_col^-1[sqrt(alpha) F xi])
//   row_uxiu       : V = U xi U^dagger, forward row FFT                      R 9 W 8
//                    (correlated noise: inverse row FFT of xi~ first)     (R 17 W 8)
//   col_kmult (B)  : as for A                                                R 8 W 4
//   row_update     : inverse row FFTs of A, B; U <- exp(i s B) U exp(-i s A) R 17 W 9
// (units: complex planes of V doubles pairs, 16 B per site each).
//
// Row passes: one workgroup per lattice row, four planes at a time in LDS.
// Column passes: W adjacent columns of one plane per workgroup, workgroups
// mapped so that strips sharing cache lines run on the same XCD.
#include "jw_common.hpp"

namespace jw {

template <int N>
struct RowE {
    static constexpr int NT = (N / 4 < 64) ? 64 : ((N / 4 > 512) ? 512 : N / 4);
    static constexpr int SPT = N / NT;                  // sites per thread
    static_assert(SPT >= 1 && SPT * NT == N, "row pass needs N a multiple of its thread count");
};
// Column passes hold the X and Y tiles of one W-column strip together in LDS
// (2 W N complex values plus one padding row in eight, 144 KB at most):
// W = min(16, 4096 / N).
template <int N>
struct ColE {
    static constexpr int W = (4096 / N > 16) ? 16 : 4096 / N;
    static constexpr int ELEMS = 2 * W * N;
    static constexpr int NT = (ELEMS / 8 < 64) ? 64 : ((ELEMS / 8 > 1024) ? 1024 : ELEMS / 8);
    static constexpr int PER = ELEMS / NT;          // elements per thread
    static constexpr int PAIRS = (W * N) / NT;      // (X, Y) element pairs per thread
    static_assert(W >= 1 && W <= N && N % W == 0 && PER * NT == ELEMS && PAIRS * NT == W * N, "column tiling");
};

// strip index for block b so that the strips on one XCD (blocks b, b+8, ...) are adjacent
__device__ __forceinline__ int xcd_strip(int b, int nstrips) {
    if (nstrips % 8) return b;
    return (b % 8) * (nstrips / 8) + b / 8;
}

// ---------------------------------------------------------------------------
// row pass 1: noise -> xi row spectra (planes 0..3 xi_x, 4..7 xi_y)
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(RowE<N>::NT) void k_row_noise_fwd(cplx* __restrict__ Xi, size_t V, NoiseSrc ns,
                                                              const double* __restrict__ sa_pos,
                                                              const cplx* __restrict__ tw) {
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    __shared__ cplx lds[4 * PN];
    const int tid = threadIdx.x;
    const size_t row0 = (size_t)blockIdx.x * N;
    double m[SPT][8];
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int y = tid + q * NT;
        double z[16];
        normals16(ns, row0 + y, z);
        const double s = sa_pos ? sa_pos[row0 + y] : 1.0;
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = mk(z[4 * p] * s, z[4 * p + 2] * s);
#pragma unroll
        for (int k = 0; k < 8; k++) m[q][k] = z[2 * k + 1] * s;
    }
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Xi[p * V + row0 + tid + q * NT] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = mk(m[q][2 * p], m[q][2 * p + 1]);
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Xi[(4 + p) * V + row0 + tid + q * NT] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
}

// ---------------------------------------------------------------------------
// row pass 2: V_i = U xi_i U^dagger -> V row spectra (planes 0..3 V_x, 4..7 V_y)
// xi from the noise source, or from position-space planes Xi_pos (correlated)
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(RowE<N>::NT) void k_row_uxiu_fwd(cplx* __restrict__ Vv, const cplx* __restrict__ U,
                                                             const cplx* __restrict__ Xi_pos, size_t V, NoiseSrc ns,
                                                             const double* __restrict__ sa_pos,
                                                             const cplx* __restrict__ tw) {
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    __shared__ cplx lds[4 * PN];
    const int tid = threadIdx.x;
    const size_t row0 = (size_t)blockIdx.x * N;
    double vy[SPT][8];
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int y = tid + q * NT;
        const size_t i = row0 + y;
        double hx[8], hy[8];
        if (Xi_pos) {
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const cplx a = Xi_pos[p * V + i], b = Xi_pos[(4 + p) * V + i];
                hx[2 * p] = a.x; hx[2 * p + 1] = a.y;
                hy[2 * p] = b.x; hy[2 * p + 1] = b.y;
            }
        } else {
            double z[16];
            normals16(ns, i, z);
            const double s = sa_pos ? sa_pos[i] : 1.0;
#pragma unroll
            for (int k = 0; k < 8; k++) { hx[k] = z[2 * k] * s; hy[k] = z[2 * k + 1] * s; }
        }
        m3 u;
#pragma unroll
        for (int k = 0; k < 9; k++) u.m[k] = U[k * V + i];
        double vx[8];
        uxiu_gen(u, hx, vx);
        uxiu_gen(u, hy, vy[q]);
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = mk(vx[2 * p], vx[2 * p + 1]);
    }
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Vv[p * V + row0 + tid + q * NT] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = mk(vy[q][2 * p], vy[q][2 * p + 1]);
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Vv[(4 + p) * V + row0 + tid + q * NT] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
}

// ---------------------------------------------------------------------------
// row pass 2, correlated noise: reads the column-inverse spectra of xi~,
// finishes their inverse transform along the row, and forms V_i = U xi~_i U^dagger
// in two batches (i = x, then y).  In each batch a site's V coefficients
// overwrite its own xi~ slots in LDS before the forward row FFT; U is read once
// per batch (the second read is served by the caches).
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(RowE<N>::NT) void k_row_uxiu_corr(cplx* __restrict__ Vv, const cplx* __restrict__ U,
                                                              const cplx* __restrict__ Xs, size_t V,
                                                              const cplx* __restrict__ tw) {
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;
    __shared__ cplx lds[4 * PN];
    const int tid = threadIdx.x;
    const size_t row0 = (size_t)blockIdx.x * N;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < SPT; q++)
                lds[lpos<1, PN, 4>(tid + q * NT, p)] = Xs[(4 * half + p) * V + row0 + tid + q * NT];
        __syncthreads();
        lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, tw, tid);
        // one site at a time: keeps a single U (9 complex) live
#pragma unroll 1
        for (int q = 0; q < SPT; q++) {
            const int y = tid + q * NT;
            const size_t i = row0 + y;
            double h[8];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const cplx a = lds[lpos<1, PN, 4>(y, p)];
                h[2 * p] = a.x;
                h[2 * p + 1] = a.y;
            }
            m3 u;
#pragma unroll
            for (int k = 0; k < 9; k++) u.m[k] = U[k * V + i];
            double v[8];
            uxiu_gen(u, h, v);
#pragma unroll
            for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = mk(v[2 * p], v[2 * p + 1]);
        }
        __syncthreads();
        lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < SPT; q++)
                Vv[(4 * half + p) * V + row0 + tid + q * NT] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// column pass: Z[p] = F_col^-1[ i (kx X^[p] + ky Y^[p]) / V ], X^ = F_col X
//   kx = -2 pi c pbarX / phat2 (c = sa(k) for SQRT, else 1), zero at k = 0.
//   CORR: X^, Y^ are first multiplied by sa(k) (per-mode correlated noise) and
//   F_col^-1[sa X^ / V] is written back over X (likewise Y).
// X, Y: row spectra of plane p = blockIdx.y; strips of W columns.
// LDS row x holds [X(x, c0..c0+W-1), Y(x, c0..c0+W-1)]: 2W sequences of length N
// with element stride 2W.
// ---------------------------------------------------------------------------
template <int N, int CORR>
__global__ __launch_bounds__(ColE<N>::NT) void k_col_kmult(cplx* __restrict__ Z, cplx* __restrict__ X,
                                                          cplx* __restrict__ Y, size_t V, int sqrt_coupling,
                                                          const double* __restrict__ q4,
                                                          const double* __restrict__ sb,
                                                          const double* __restrict__ sa,
                                                          const cplx* __restrict__ tw) {
    constexpr int W = ColE<N>::W, NT = ColE<N>::NT, PER = ColE<N>::PER, PAIRS = ColE<N>::PAIRS;
    constexpr int S2 = 2 * W;
    __shared__ cplx lds[padded<N, 3>::value * S2];
    const int tid = threadIdx.x;
    const int strip = xcd_strip(blockIdx.x, N / W);
    const int col0 = strip * W;
    const size_t pofs = (size_t)blockIdx.y * V;
    const double invV = 1.0 / ((double)N * (double)N);
    // load both tiles: element e -> row e / 2W, slot e % 2W (X for slot < W)
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int e = tid + q * NT;
        const int x = e / S2, s = e % S2;
        const size_t g = pofs + (size_t)x * N + col0 + (s % W);
        lds[lpos<S2, 1, 3>(x, s)] = s < W ? X[g] : Y[g];
    }
    __syncthreads();
    lds_fft<N, S2, S2, 1, NT, -1, 3>(lds, tw, tid);
    // pair (x, c): X^ in slot c of LDS row x, Y^ in slot W + c
    cplx zreg[CORR ? PAIRS : 1];
#pragma unroll
    for (int q = 0; q < PAIRS; q++) {
        const int pr = tid + q * NT;
        const int kx = pr / W, c = pr % W, ky = col0 + c;
        const size_t k = (size_t)kx * N + ky;
        cplx vx = lds[lpos<S2, 1, 3>(kx, c)], vy = lds[lpos<S2, 1, 3>(kx, W + c)];
        double fx = 0.0, fy = 0.0;
        if (kx != 0 || ky != 0) {
            const double ph = q4[kx] + q4[ky];
            const double cc = sqrt_coupling ? sa[k] : 1.0;
            fx = -2.0 * M_PI * cc * sb[kx] / ph * invV;
            fy = -2.0 * M_PI * cc * sb[ky] / ph * invV;
        }
        if (CORR) {
            const double s = sa[k];
            vx = cscale(vx, s);
            vy = cscale(vy, s);
            lds[lpos<S2, 1, 3>(kx, c)] = cscale(vx, invV);
            lds[lpos<S2, 1, 3>(kx, W + c)] = cscale(vy, invV);
        }
        const cplx z = mk(-(fx * vx.y + fy * vy.y), fx * vx.x + fy * vy.x);   // i (fx vx + fy vy)
        if (CORR) zreg[q] = z;
        else lds[lpos<S2, 1, 3>(kx, c)] = z;
    }
    __syncthreads();
    if (CORR) {
        // xi~ = F_col^-1 of the scaled tiles, written back over X and Y
        lds_fft<N, S2, S2, 1, NT, +1, 3>(lds, tw, tid);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = tid + q * NT;
            const int x = e / S2, s = e % S2;
            const size_t g = pofs + (size_t)x * N + col0 + (s % W);
            if (s < W) X[g] = lds[lpos<S2, 1, 3>(x, s)];
            else Y[g] = lds[lpos<S2, 1, 3>(x, s)];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PAIRS; q++) {
            const int pr = tid + q * NT;
            lds[lpos<S2, 1, 3>(pr / W, pr % W)] = zreg[q];
        }
        __syncthreads();
    }
    // Z: the W sequences in the X slots
    lds_fft<N, W, S2, 1, NT, +1, 3>(lds, tw, tid);
#pragma unroll
    for (int q = 0; q < PAIRS; q++) {
        const int pr = tid + q * NT;
        const int x = pr / W, c = pr % W;
        Z[pofs + (size_t)x * N + col0 + c] = lds[lpos<S2, 1, 3>(x, c)];
    }
}

// ---------------------------------------------------------------------------
// row pass 3: inverse row FFTs of the A and B column-inverse spectra and the
// update U <- exp(i s B) U exp(-i s A) (field.h:1966-1996) in one pass.
// A's coefficients are kept in registers, B's stay in LDS.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(RowE<N>::NT) void k_row_update(cplx* __restrict__ U, const cplx* __restrict__ ZA,
                                                           const cplx* __restrict__ ZB, size_t V, double ss,
                                                           const cplx* __restrict__ tw) {
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    __shared__ cplx lds[4 * PN];
    const int tid = threadIdx.x;
    const size_t row0 = (size_t)blockIdx.x * N;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = ZA[p * V + row0 + tid + q * NT];
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, tw, tid);
    double ha[SPT][8];
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const cplx a = lds[lpos<1, PN, 4>(tid + q * NT, p)];
            ha[q][2 * p] = a.x;
            ha[q][2 * p + 1] = a.y;
        }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = ZB[p * V + row0 + tid + q * NT];
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, tw, tid);
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int y = tid + q * NT;
        const size_t i = row0 + y;
        double hb[8];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const cplx b = lds[lpos<1, PN, 4>(y, p)];
            hb[2 * p] = b.x;
            hb[2 * p + 1] = b.y;
        }
        m3 u;
#pragma unroll
        for (int k = 0; k < 9; k++) u.m[k] = U[k * V + i];
        const m3 w = mmul(exp_i_gen(hb, ss), u);
        const m3 r = mmul(w, exp_i_gen(ha[q], -ss));
#pragma unroll
        for (int k = 0; k < 9; k++) U[k * V + i] = r.m[k];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N>
static hipError_t step_fused_n(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const NoiseSrc& ns,
                               const double* sa_pos, const double* sa_k, int sqrt_coupling, int corr,
                               double sqrt_step, int* kcls) {
    const size_t V = c->V;
    constexpr int W = ColE<N>::W;
    hipError_t e;
    {
        KTimer t(c, kcls[0]);
        hipLaunchKernelGGL((k_row_noise_fwd<N>), dim3(N), dim3(RowE<N>::NT), 0, c->stream, Xi, V, ns, sa_pos, c->tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[1]);
        if (corr)
            hipLaunchKernelGGL((k_col_kmult<N, 1>), dim3(N / W, 4), dim3(ColE<N>::NT), 0, c->stream, ZA, Xi,
                               Xi + 4 * V, V, sqrt_coupling, c->q4, c->sb, sa_k, c->tw);
        else
            hipLaunchKernelGGL((k_col_kmult<N, 0>), dim3(N / W, 4), dim3(ColE<N>::NT), 0, c->stream, ZA, Xi,
                               Xi + 4 * V, V, sqrt_coupling, c->q4, c->sb, sa_k, c->tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[2]);
        if (corr)
            hipLaunchKernelGGL((k_row_uxiu_corr<N>), dim3(N), dim3(RowE<N>::NT), 0, c->stream, Vv, U, Xi, V, c->tw);
        else
            hipLaunchKernelGGL((k_row_uxiu_fwd<N>), dim3(N), dim3(RowE<N>::NT), 0, c->stream, Vv, U, nullptr, V, ns,
                               sa_pos, c->tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[1]);
        hipLaunchKernelGGL((k_col_kmult<N, 0>), dim3(N / W, 4), dim3(ColE<N>::NT), 0, c->stream, ZB, Vv, Vv + 4 * V,
                           V, sqrt_coupling, c->q4, c->sb, sa_k, c->tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[3]);
        hipLaunchKernelGGL((k_row_update<N>), dim3(N), dim3(RowE<N>::NT), 0, c->stream, U, ZA, ZB, V, sqrt_step,
                           c->tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

bool fused_supported(int N) { return N >= 64 && N <= 2048; }

hipError_t step_fused(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const NoiseSrc& ns,
                      const double* sa_pos, const double* sa_k, int sqrt_coupling, int corr, double sqrt_step) {
    int kcls[4] = {K_ROW_NOISE, K_COL_KMULT, K_ROW_UXIU, K_ROW_UPDATE};
    switch (c->N) {
        case 64: return step_fused_n<64>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        case 128: return step_fused_n<128>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        case 256: return step_fused_n<256>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        case 512: return step_fused_n<512>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        case 1024: return step_fused_n<1024>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        case 2048: return step_fused_n<2048>(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, sqrt_coupling, corr, sqrt_step, kcls);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace jw
