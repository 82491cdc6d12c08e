// jw_engine.hpp -- fused generator-basis Langevin step for 64 <= N <= 2048,
// templated on the storage type C (double2: FP64 form, float2: mixed form);
// instantiated by jw_step_f64.hip and jw_step_f32.hip.
//
// One step (main_optimized.cpp:243-304 in the generator basis) as five passes
// over HBM instead of the reference's 8 nine-component FFTs and 5 site loops:
//   row_noise_fwd  : Philox/injected noise -> xi planes, forward row FFT    W 8
//   col_kmult (A)  : forward column FFT, kernel multiply, inverse column FFT R 8 W 4
//                    (correlated noise: also writes F_col^-1[sqrt(alpha) F xi])
//   row_uxiu       : V = U xi U^dagger, forward row FFT                      R 6 W 8
//                    (correlated noise: inverse row FFT of xi~ first)     (R 14 W 8)
//   col_kmult (B)  : as for A                                                R 8 W 4
//   row_update     : inverse row FFTs of A, B; U <- exp(i s B) U exp(-i s A) R 14 W 6
// (units: complex planes of V elements, 16 B (FP64) or 8 B (FP32) per site each;
// U as its rows 0 and 1, UPlanes below).  With the spectra drawn per mode the
// noise pass goes and column A reads nothing and writes 12 planes.
//
// Row passes: one workgroup per lattice row, four planes at a time in LDS.
//
// The intermediate planes (xi, V, A, B spectra) between a row pass and a
// column pass are stored in one of two layouts (template parameter BF):
//   BF = 0   row-major, site (x, y) at x N + y.  Column passes then take strips
//            of W adjacent columns (X and Y tiles together in LDS), workgroups
//            mapped so that strips sharing cache lines run on the same XCD.
//   BF = 4, 8  blocks of BF rows: (x, y) at ((x / BF) N + y) BF + x mod BF, so
//            that one 128-byte segment holds 8 / BF adjacent columns of BF
//            consecutive rows.  A column pass then reads and writes whole
//            segments (8 / BF columns per workgroup, one plane at a time, all
//            four plane pairs in one workgroup); row passes touch BF-row
//            segments whose rows are processed on one XCD at neighbouring times.
// U always stays row-major.
#pragma once

#include "jw_common.hpp"

namespace jw {

// Row passes: N / DIV threads (64..1024), SPT sites per thread.  At N = 2048
// the four padded rows in LDS (139 KB) allow one workgroup per CU; 512 threads
// (DIV = 4) measured faster than 1024 for these passes.
template <int N, int DIV = 4>
struct RowE {
    static constexpr int NT = (N / DIV < 64) ? 64 : ((N / DIV > 1024) ? 1024 : N / DIV);
    static constexpr int SPT = N / NT;                  // sites per thread
    static_assert(SPT >= 1 && SPT * NT == N, "row pass needs N a multiple of its thread count");
};
// Row-major column passes hold the X and Y tiles of one W-column strip together
// in LDS (2 W N complex values plus one padding row in eight, 144 KB at most in
// FP64): W = min(16, 4096 / N) for both forms (wider FP32 strips measured
// slower: one workgroup per CU and spills in the correlated variant).
template <int N, typename C = double2, int CORR = 0>
struct ColE {
    static constexpr int NTMAX = 1024;
    static constexpr int W = (4096 / N > 16) ? 16 : 4096 / N;
    static constexpr int ELEMS = 2 * W * N;
    static constexpr int NT = (ELEMS / 8 < 64) ? 64 : ((ELEMS / 8 > NTMAX) ? NTMAX : ELEMS / 8);
    static constexpr int PER = ELEMS / NT;          // elements per thread
    static constexpr int PAIRS = (W * N) / NT;      // (X, Y) element pairs per thread
    static_assert(W >= 1 && W <= N && N % W == 0 && PER * NT == ELEMS && PAIRS * NT == W * N, "column tiling");
};

// minimum waves per SIMD requested of the row passes: the FP32 rows (4 padded
// rows, 70 KB of LDS at N = 2048) fit two 512-thread workgroups per CU when
// they stay within 128 VGPRs
template <typename C>
struct RowWaves { static constexpr int value = sizeof(C) == sizeof(double2) ? 1 : 4; };

// strip index for block b so that the strips on one XCD (blocks b, b+8, ...) are adjacent
__device__ __forceinline__ int xcd_strip(int b, int nstrips) {
    if (nstrips % 8) return b;
    return (b % 8) * (nstrips / 8) + b / 8;
}

// offset of site (x, y) in an intermediate plane of layout BF (0 and 2
// row-major, 1 row-major for the 64-byte-strip column pass, 4 and 8 row blocks)
template <int N, int BF>
__device__ __forceinline__ size_t lay(int x, int y) {
    if constexpr (BF <= 2 || BF >= 16) return (size_t)x * N + y;
    else return ((size_t)(x / BF) * N + y) * BF + (x % BF);
}

// offset of site (x, y) of plane q (0..3: the x direction's plane pairs,
// 4..7: the y direction's) in an 8-plane spectral array.  Layout 2 interleaves
// the two directions of a plane pair site by site, (p, x, y, d) at
// 2 V p + 2 (x N + y) + d, so that a column strip of W columns of both
// directions is one 2W-element row segment.
template <int N, int BF>
__device__ __forceinline__ size_t pat(int q, int x, int y, size_t V) {
    if constexpr (BF == 2) return (size_t)(q & 3) * 2 * V + 2 * ((size_t)x * N + y) + (q >> 2);
    else return (size_t)q * V + lay<N, BF>(x, y);
}

// lattice row of row-pass block b.  Blocks are dealt round-robin to the 8 XCDs;
// with a blocked layout the BF rows sharing segments go to one XCD in
// consecutive turns (a bijection on [0, N) for N a multiple of 8 BF).
template <int N, int BF>
__device__ __forceinline__ int row_of_block(int b) {
    if constexpr (BF <= 2 || BF >= 16) return b;
    else {
        static_assert(N % (8 * BF) == 0, "blocked layout needs N a multiple of 8 BF");
        const int t = b >> 3, xcd = b & 7;
        return (t / BF) * (8 * BF) + xcd * BF + (t % BF);
    }
}

// Trajectory blockIdx.z of a batched launch: its planes sit at z times the
// per-trajectory size of each array, and its noise key is traj + z (the
// Philox step is read from the device counter when one is given).
__device__ __forceinline__ NoiseSrc batch_noise(NoiseSrc ns, size_t V) {
    ns.traj += blockIdx.z;
    if (ns.inj) ns.inj += (size_t)blockIdx.z * 16 * V;
    if (ns.stepp) ns.step = *ns.stepp;
    return ns;
}
__device__ __forceinline__ size_t zofs(int planes, size_t V) { return (size_t)blockIdx.z * planes * V; }
// the same for an explicit trajectory z (kernel bodies run by one workgroup or,
// in a launch over several size groups, by a part of one)
__device__ __forceinline__ NoiseSrc batch_noise_z(NoiseSrc ns, size_t V, int z) {
    ns.traj += z;
    if (ns.inj) ns.inj += (size_t)z * 16 * V;
    if (ns.stepp) ns.step = *ns.stepp;
    return ns;
}
__device__ __forceinline__ size_t zofs_z(int z, int planes, size_t V) { return (size_t)z * planes * V; }
// where a kernel body runs: its row, strip or segment (bx), plane pair (by),
// trajectory (bz) and thread index among the threads that run it (tid)
struct BlockIx {
    int bx, by, bz, tid;
};
__device__ __forceinline__ BlockIx own_block() {
    return BlockIx{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)threadIdx.x};
}

// matrix type of the site algebra's products: FP64 in the FP64 form, FP32 in
// the mixed form (whose exponentials and projection stay FP64)
template <typename C>
struct site_mat { typedef m3 type; };
template <>
struct site_mat<float2> { typedef m3f type; };

// U planes read and written by the fused passes: rows 0 and 1 (six planes), row
// 2 rebuilt on load as conj(row 0 x row 1) -- U is in SU(3) to the rounding of
// its form (and the mixed form projects it every few steps, rebuilding row 2
// the same way) -- so that the row passes move a third less U.  The FP64 field
// keeps its nine planes; row 2 is rebuilt into it after the fused steps
// (u_rebuild).
template <typename C>
struct UPlanes { static constexpr int value = 6; };
template <typename C, typename M>
__device__ __forceinline__ void load_u(M& u, const C* __restrict__ U, size_t V, size_t i) {
#pragma unroll
    for (int k = 0; k < UPlanes<C>::value; k++) u.m[k] = U[k * V + i];
    if constexpr (UPlanes<C>::value == 6) {
        u.m[6] = conjg(csub(cmul(u.m[1], u.m[5]), cmul(u.m[2], u.m[4])));
        u.m[7] = conjg(csub(cmul(u.m[2], u.m[3]), cmul(u.m[0], u.m[5])));
        u.m[8] = conjg(csub(cmul(u.m[0], u.m[4]), cmul(u.m[1], u.m[3])));
    }
}
template <typename C, typename M>
__device__ __forceinline__ void store_u(C* __restrict__ U, size_t V, size_t i, const M& r) {
#pragma unroll
    for (int k = 0; k < UPlanes<C>::value; k++) U[k * V + i] = r.m[k];
}

// ordering fences: values passed through them are final at the fence and
// values read after it are defined there (empty asm statements keep their order)
__device__ __forceinline__ void fence(double& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void fence(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void fence(float2& v) { asm volatile("" : "+v"(v.x), "+v"(v.y)); }
__device__ __forceinline__ void fence(double2& v) { asm volatile("" : "+v"(v.x), "+v"(v.y)); }

// the same pointer, opaque to the optimiser: loads through it are not merged
// with earlier loads of the same addresses (which would keep every twiddle of
// one FFT live in registers until the next)
template <typename T>
__device__ __forceinline__ const T* opaque(const T* p) {
    typedef const __attribute__((address_space(1))) T GT;     // a global-memory table
    GT* g = (GT*)p;
    asm volatile("" : "+s"(g));
    return (const T*)g;
}
// likewise for a per-thread index: values derived from it are recomputed
// rather than kept live from one phase of a kernel to the next
__device__ __forceinline__ int opaque_v(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// inverse row FFT of 4 planes in LDS; in FP32 through opaque twiddles and
// thread index (fewer live registers around the site algebra)
template <int N, int NT, int PN, typename C, typename TW>
__device__ __forceinline__ void row_ifft4(C* lds, TW tw, int tid) {
    if constexpr (sizeof(C) == sizeof(float2) || N <= 1024) lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, tw_fft(tw), opaque_v(tid));
    else lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, tw, tid);
}

// a thread's site index in a row pass's site loop, opaque, so that the compiler
// runs the sites one after another instead of interleaving their site algebra
// (FP32: registers for two workgroups per CU; FP64: fewer spills, the same
// arithmetic)
template <typename C>
__device__ __forceinline__ int site_of(int y) {
    return opaque_v(y);
}

// ---------------------------------------------------------------------------
// row pass 1: noise -> xi row spectra (planes 0..3 xi_x, 4..7 xi_y) of row x.
// The body is shared by the noise pass and by the update pass, which runs it
// for the next step after its own sites (the noise is compute-bound, the
// update memory-bound).  lds: 4 padded rows.
// ---------------------------------------------------------------------------
template <int N, int BF, typename C, typename TW>
__device__ __forceinline__ void row_noise_body(C* lds, C* __restrict__ Xi, size_t V, const NoiseSrc& ns,
                                               const double* __restrict__ sa_pos, TW tw, int x, int tid) {
    typedef typename scalar_of<C>::type T;
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    const size_t row0 = (size_t)x * N;
    T m[SPT][8];
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        // FP32: sites one after another (no spills at two workgroups per CU)
        const int y = sizeof(C) == sizeof(float2) ? site_of<C>(tid + q * NT) : tid + q * NT;
        T z[16];
        normals16_t(ns, row0 + y, z);
        const T s = sa_pos ? (T)sa_pos[row0 + y] : (T)1;      // in the element type: no FP64 in FP32
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = C{z[4 * p] * s, z[4 * p + 2] * s};
#pragma unroll
        for (int k = 0; k < 8; k++) m[q][k] = z[2 * k + 1] * s;
    }
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Xi[pat<N, BF>(p, x, tid + q * NT, V)] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = mkc<C>(m[q][2 * p], m[q][2 * p + 1]);
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, tw, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Xi[pat<N, BF>(4 + p, x, tid + q * NT, V)] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
}

// The noise pass proper: one plane pair at a time.  Plane p of xi_x holds
// (n_2p, n_2p+1) and plane p of xi_y (m_2p, m_2p+1), the normals of Philox
// calls 2p and 2p+1 (draw order n0, m0, n1, m1, ...: call q gives n_q, m_q),
// so each pair is generated, transformed and stored on its own and nothing is
// held across pairs: two padded rows in LDS (70 KB in FP64 at N = 2048) fit
// two workgroups per CU (FP32: 35 KB, four); the waves-per-SIMD bound asks
// the compiler for the registers of that many.
template <int N, typename C>
struct NoiseWaves { static constexpr int value = sizeof(C) == sizeof(double2) ? 4 : 8; };

template <int N, int BF, typename C>
__device__ __forceinline__ void row_noise_fwd_body(C* lds, BlockIx bi, C* __restrict__ Xi, size_t V, NoiseSrc ns,
                                                              const double* __restrict__ sa_pos,
                                                              const C* __restrict__ tw) {
    typedef typename scalar_of<C>::type T;
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;
    const auto twl = tw_source<N, true>(tw);
    Xi += zofs_z(bi.bz, 8, V);
    ns = batch_noise_z(ns, V, bi.bz);
    const int x = row_of_block<N, BF>(bi.bx);
    const size_t row0 = (size_t)x * N;
    const int tid = bi.tid;
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int y = site_of<C>(tid + q * NT);
            const size_t i = row0 + y;
            T z[4];
            if (ns.inj) {
                const double* zi = ns.inj + 16 * i + 4 * p;
#pragma unroll
                for (int k = 0; k < 4; k++) z[k] = (T)zi[k];
            } else if constexpr (sizeof(T) == sizeof(double)) {
                box_muller_pair((uint32_t)i, ns.step, ns.tag, (uint32_t)(2 * p), ns.seed, ns.traj, z[0], z[1]);
                box_muller_pair((uint32_t)i, ns.step, ns.tag, (uint32_t)(2 * p + 1), ns.seed, ns.traj, z[2], z[3]);
            } else {
                box_muller_pair_f((uint32_t)i, ns.step, ns.tag, (uint32_t)(2 * p), ns.seed, ns.traj, z[0], z[1]);
                box_muller_pair_f((uint32_t)i, ns.step, ns.tag, (uint32_t)(2 * p + 1), ns.seed, ns.traj, z[2], z[3]);
            }
            const T s = sa_pos ? (T)sa_pos[i] : (T)1;
            lds[lpos<1, PN, 4>(y, 0)] = C{z[0] * s, z[2] * s};      // xi_x plane p: (n_2p, n_2p+1)
            lds[lpos<1, PN, 4>(y, 1)] = C{z[1] * s, z[3] * s};      // xi_y plane p: (m_2p, m_2p+1)
        }
        __syncthreads();
        lds_fft<N, 2, 1, PN, NT, -1, 4>(lds, tw_fft(twl), opaque_v(tid));
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int y = tid + q * NT;
            Xi[pat<N, BF>(p, x, y, V)] = lds[lpos<1, PN, 4>(y, 0)];
            Xi[pat<N, BF>(4 + p, x, y, V)] = lds[lpos<1, PN, 4>(y, 1)];
        }
        __syncthreads();
    }
}

template <int N, int BF, typename C>
__global__ __launch_bounds__(RowE<N>::NT, (NoiseWaves<N, C>::value)) void k_row_noise_fwd(C* __restrict__ Xi, size_t V, NoiseSrc ns,
                                                              const double* __restrict__ sa_pos,
                                                              const C* __restrict__ tw) {
    __shared__ C lds[2 * padded<N, 4>::value];
    row_noise_fwd_body<N, BF, C>(lds, own_block(), Xi, V, ns, sa_pos, tw);
}

// ---------------------------------------------------------------------------
// row pass 2: V_i = U xi_i U^dagger -> V row spectra (planes 0..3 V_x, 4..7 V_y),
// xi regenerated from the noise source (white noise)
// ---------------------------------------------------------------------------
template <int N, int BF, typename C>
__device__ __forceinline__ void row_uxiu_fwd_body(C* lds, BlockIx bi, C* __restrict__ Vv, const C* __restrict__ U,
                                                             size_t V, NoiseSrc ns, const double* __restrict__ sa_pos,
                                                             const C* __restrict__ tw) {
    typedef typename scalar_of<C>::type T;
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    const int tid = bi.tid;
    const auto twl = tw_source<N, true>(tw);
    Vv += zofs_z(bi.bz, 8, V);
    U += zofs_z(bi.bz, UPlanes<C>::value, V);
    ns = batch_noise_z(ns, V, bi.bz);
    const int x = row_of_block<N, BF>(bi.bx);
    const size_t row0 = (size_t)x * N;
    T vy[SPT][8];      // in the element type: they go to LDS in it
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int y = site_of<C>(tid + q * NT);
        const size_t i = row0 + y;
        T hx[8], hy[8];
        T z[16];
        normals16_t(ns, i, z);
        const T s = sa_pos ? (T)sa_pos[i] : (T)1;
#pragma unroll
        for (int k = 0; k < 8; k++) { hx[k] = z[2 * k] * s; hy[k] = z[2 * k + 1] * s; }
        // FP64 form: FP64 products; mixed form: FP32 products
        typename site_mat<C>::type u;
        load_u(u, U, V, i);
        T vx[8], v2[8];
        uxiu_gen(u, hx, vx);
        uxiu_gen(u, hy, v2);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            vy[q][k] = (T)v2[k];
            fence(vy[q][k]);
        }
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = mkc<C>(vx[2 * p], vx[2 * p + 1]);
    }
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, twl, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) Vv[pat<N, BF>(p, x, tid + q * NT, V)] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = mkc<C>(vy[q][2 * p], vy[q][2 * p + 1]);
    __syncthreads();
    lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, twl, tid);
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++)
            Vv[pat<N, BF>(4 + p, x, tid + q * NT, V)] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
}

template <int N, int BF, typename C>
__global__ __launch_bounds__(RowE<N>::NT, RowWaves<C>::value) void k_row_uxiu_fwd(C* __restrict__ Vv, const C* __restrict__ U,
                                                             size_t V, NoiseSrc ns, const double* __restrict__ sa_pos,
                                                             const C* __restrict__ tw) {
    __shared__ C lds[4 * padded<N, 4>::value];
    row_uxiu_fwd_body<N, BF, C>(lds, own_block(), Vv, U, V, ns, sa_pos, tw);
}

// ---------------------------------------------------------------------------
// row pass 2, correlated noise: reads the column-inverse spectra of xi~,
// finishes their inverse transform along the row, and forms V_i = U xi~_i U^dagger
// in two batches (i = x, then y).  In each batch a site's V coefficients
// overwrite its own xi~ slots in LDS before the forward row FFT.  U is read in
// both batches: holding xi~_x in registers to read U once needs more registers
// and measured slower (one workgroup per CU instead of two in FP32).
// ---------------------------------------------------------------------------
template <int N, int BF, typename C>
__device__ __forceinline__ void row_uxiu_corr_body(C* lds, BlockIx bi, C* __restrict__ Vv, const C* __restrict__ U,
                                                              const C* __restrict__ Xs, size_t V,
                                                              const C* __restrict__ tw) {
    typedef typename scalar_of<C>::type T;
    constexpr int NT = RowE<N>::NT, SPT = RowE<N>::SPT;
    constexpr int PN = padded<N, 4>::value;
    const int tid = bi.tid;
    const auto twl = tw_source<N, false>(tw);
    Vv += zofs_z(bi.bz, 8, V);
    U += zofs_z(bi.bz, UPlanes<C>::value, V);
    Xs += zofs_z(bi.bz, 8, V);
    const int x = row_of_block<N, BF>(bi.bx);
    const size_t row0 = (size_t)x * N;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < SPT; q++)
                lds[lpos<1, PN, 4>(tid + q * NT, p)] = Xs[pat<N, BF>(4 * half + p, x, tid + q * NT, V)];
        __syncthreads();
        lds_fft<N, 4, 1, PN, NT, +1, 4>(lds, twl, tid);
        // one site at a time: keeps a single U (9 complex, row 2 rebuilt) live
#pragma unroll 1
        for (int q = 0; q < SPT; q++) {
            const int y = tid + q * NT;
            const size_t i = row0 + y;
            T h[8];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const C a = lds[lpos<1, PN, 4>(y, p)];
                h[2 * p] = a.x;
                h[2 * p + 1] = a.y;
            }
            typename site_mat<C>::type u;
            load_u(u, U, V, i);
            T v[8];
            uxiu_gen(u, h, v);
#pragma unroll
            for (int p = 0; p < 4; p++) lds[lpos<1, PN, 4>(y, p)] = mkc<C>(v[2 * p], v[2 * p + 1]);
        }
        __syncthreads();
        lds_fft<N, 4, 1, PN, NT, -1, 4>(lds, twl, tid);
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < SPT; q++)
                Vv[pat<N, BF>(4 * half + p, x, tid + q * NT, V)] = lds[lpos<1, PN, 4>(tid + q * NT, p)];
        __syncthreads();
    }
}

template <int N, int BF, typename C>
__global__ __launch_bounds__(RowE<N>::NT, RowWaves<C>::value) void k_row_uxiu_corr(C* __restrict__ Vv, const C* __restrict__ U,
                                                              const C* __restrict__ Xs, size_t V,
                                                              const C* __restrict__ tw) {
    __shared__ C lds[4 * padded<N, 4>::value];
    row_uxiu_corr_body<N, BF, C>(lds, own_block(), Vv, U, Xs, V, tw);
}

// ---------------------------------------------------------------------------
// column pass: Z[p] = F_col^-1[ i (kx X^[p] + ky Y^[p]) / V ], X^ = F_col X
//   kx = -2 pi c pbarX / phat2 (c = sa(k) for SQRT, else 1), zero at k = 0.
//   CORR: X^, Y^ are first multiplied by sa(k) (per-mode correlated noise;
//   by 1 when noise_scale is 0) and F_col^-1[sa X^ / V] is written back over X
//   (likewise Y).
//   GEN (with CORR): X^, Y^ are not transformed from X, Y but drawn per mode
//   (kspace_normals4, plane pair p = blockIdx.y, noise key from ns); X and Y
//   then receive the row spectra of the position-space noise.
// X, Y: row spectra of plane p = blockIdx.y; strips of W columns.  PI: X is
// the pair-interleaved array (layout 2) holding both directions, Y unused.
// LDS row x holds [X(x, c0..c0+W-1), Y(x, c0..c0+W-1)]: 2W sequences of length N
// with element stride 2W.
// ---------------------------------------------------------------------------
// MF = 1, white noise (FP32 tiles of 74 KB, two per CU): registers of two
// workgroups asked for, as the FP32 butterfly version fits them by itself; the
// correlated variant needs more than that without spilling and runs one per CU
template <int N, int CORR, typename C, int PI = 0, int MF = 0, int GEN = 0>
__device__ __forceinline__ void col_kmult_body(C* lds, BlockIx bi, C* __restrict__ Z, C* __restrict__ X,
                                                          C* __restrict__ Y, size_t V, int sqrt_coupling,
                                                          const double* __restrict__ q4,
                                                          const double* __restrict__ sb,
                                                          const double* __restrict__ sa,
                                                          const C* __restrict__ tw, NoiseSrc ns, int noise_scale) {
    static_assert(!GEN || CORR, "drawn spectra are written back as position-space noise");
    typedef typename scalar_of<C>::type T;
    typedef ColE<N, C, CORR> CE;
    constexpr int W = CE::W, NT = CE::NT, PER = CE::PER, PAIRS = CE::PAIRS;
    constexpr int S2 = 2 * W;
    const int tid = bi.tid;
    const auto twl = tw_source<N, sizeof(C) == sizeof(float2)>(tw);
    const int strip = xcd_strip(bi.bx, N / W);
    const int col0 = strip * W;
    Z += zofs_z(bi.bz, 4, V);
    X += zofs_z(bi.bz, 8, V);
    Y += zofs_z(bi.bz, 8, V);
    const size_t pofs = (size_t)bi.by * V;
    const double invV = 1.0 / ((double)N * (double)N);
    // element e of the tile: row e / 2W, LDS slot s (X for s < W, Y above),
    // global offset g within X (PI: the pair-interleaved array) or Y
    auto elem = [&](int e, int& x, int& s, size_t& g) {
        x = e / S2;
        const int m = e % S2;
        if constexpr (PI) {
            s = (m & 1) * W + (m >> 1);
            g = 2 * pofs + 2 * ((size_t)x * N + col0) + m;
        } else {
            s = m;
            g = pofs + (size_t)x * N + col0 + (m % W);
        }
    };
    if constexpr (!GEN) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int x, s;
            size_t g;
            elem(tid + q * NT, x, s, g);
            lds[lpos<S2, 1, 3>(x, s)] = (PI || s < W) ? X[g] : Y[g];
        }
        __syncthreads();
        lds_fft<N, S2, S2, 1, NT, -1, 3, MF>(lds, tw_fft(twl), opaque_v(tid));
    }
    const NoiseSrc nk = batch_noise_z(ns, V, bi.bz);
    // pair (x, c): X^ in slot c of LDS row x, Y^ in slot W + c
    C zreg[CORR ? PAIRS : 1];
    // FP32 correlated: indices recomputed per phase (opaque_v), since addresses
    // kept live across the FFTs would cost the registers of a second
    // workgroup per CU (measured 0.76 -> 0.70 ms per step at N = 2048)
    constexpr bool REIDX = CORR && sizeof(C) == sizeof(float2);
    const int tk = REIDX ? opaque_v(tid) : tid;
#pragma unroll
    for (int q = 0; q < PAIRS; q++) {
        const int pr = tk + q * NT;
        const int kx = pr / W, c = pr % W, ky = col0 + c;
        const size_t k = (size_t)kx * N + ky;
        C vx, vy;
        if constexpr (GEN) {
            T g[4];
            kspace_normals4((uint32_t)k, nk.step, nk.tag, bi.by, nk.seed, nk.traj, g);
            vx = mkc<C>(g[0] * (T)N, g[1] * (T)N);          // sqrt(V) = N
            vy = mkc<C>(g[2] * (T)N, g[3] * (T)N);
        } else {
            vx = lds[lpos<S2, 1, 3>(kx, c)];
            vy = lds[lpos<S2, 1, 3>(kx, W + c)];
        }
        // kernel factors in the element type: one FP64 division per pair in FP64,
        // one FP32 division in the mixed form (whose spectra are FP32)
        T fx = 0, fy = 0;
        if (kx != 0 || ky != 0) {
            const double cc = sqrt_coupling ? sa[k] : 1.0;
            const T f = (T)(-2.0 * M_PI * invV * cc) / (T)(q4[kx] + q4[ky]);
            fx = f * (T)sb[kx];
            fy = f * (T)sb[ky];
        }
        if (CORR) {
            if (noise_scale) {
                const T s = (T)sa[k];
                vx = cscale(vx, s);
                vy = cscale(vy, s);
            }
            lds[lpos<S2, 1, 3>(kx, c)] = cscale(vx, (T)invV);
            lds[lpos<S2, 1, 3>(kx, W + c)] = cscale(vy, (T)invV);
        }
        const C z = mkc<C>(-(fx * vx.y + fy * vy.y), fx * vx.x + fy * vy.x);   // i (fx vx + fy vy)
        if (CORR) zreg[q] = z;
        else lds[lpos<S2, 1, 3>(kx, c)] = z;
    }
    __syncthreads();
    if (CORR) {
        // xi~ = F_col^-1 of the scaled tiles, written back over X and Y
        lds_fft<N, S2, S2, 1, NT, +1, 3, MF>(lds, tw_fft(twl), opaque_v(tid));
        const int ts = REIDX ? opaque_v(tid) : tid;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int x, s;
            size_t g;
            elem(ts + q * NT, x, s, g);
            if (PI || s < W) X[g] = lds[lpos<S2, 1, 3>(x, s)];
            else Y[g] = lds[lpos<S2, 1, 3>(x, s)];
        }
        __syncthreads();
        const int tz = REIDX ? opaque_v(tid) : tid;
#pragma unroll
        for (int q = 0; q < PAIRS; q++) {
            const int pr = tz + q * NT;
            lds[lpos<S2, 1, 3>(pr / W, pr % W)] = zreg[q];
        }
        __syncthreads();
    }
    // Z: the W sequences in the X slots
    lds_fft<N, W, S2, 1, NT, +1, 3, MF>(lds, tw_fft(twl), opaque_v(tid));
    const int to = REIDX ? opaque_v(tid) : tid;
#pragma unroll
    for (int q = 0; q < PAIRS; q++) {
        const int pr = to + q * NT;
        const int x = pr / W, c = pr % W;
        Z[pofs + (size_t)x * N + col0 + c] = lds[lpos<S2, 1, 3>(x, c)];
    }
}

template <int N, int CORR, typename C, int PI = 0, int MF = 0, int GEN = 0>
__global__ __launch_bounds__((ColE<N, C, CORR>::NT), (MF && !CORR ? 2 * ColE<N, C, CORR>::NT / 256 : 1)) void k_col_kmult(C* __restrict__ Z, C* __restrict__ X,
                                                          C* __restrict__ Y, size_t V, int sqrt_coupling,
                                                          const double* __restrict__ q4,
                                                          const double* __restrict__ sb,
                                                          const double* __restrict__ sa,
                                                          const C* __restrict__ tw, NoiseSrc ns, int noise_scale) {
    __shared__ C lds[padded<N, 3>::value * 2 * ColE<N, C, CORR>::W];
    col_kmult_body<N, CORR, C, PI, MF, GEN>(lds, own_block(), Z, X, Y, V, sqrt_coupling, q4, sb, sa, tw, ns, noise_scale);
}

// ---------------------------------------------------------------------------
// column pass on the blocked layout: one workgroup per 128-byte column segment
// (CPB = 8 / BF adjacent columns) and plane pair p, one plane at a time in LDS:
//   X^ = F_col X[p]  ->  acc = i fx X^       (CORR: X^ *= sa, X[p] <- F_col^-1[X^ / V])
//   Y^ = F_col Y[p]  ->  acc += i fy Y^      (CORR: likewise for Y[p])
//   Z[p] = F_col^-1 acc
// with fx, fy, sa as in k_col_kmult.  Element e of a loaded segment set is
// x-block e / 8, slot e mod 8 = c BF + x mod BF: consecutive threads read
// consecutive 16-byte words.  sa16 holds sa(k) in 16-row blocks.
// ---------------------------------------------------------------------------
template <int N, int BF, typename C>
struct ColB {
    // BF = 4, 8: one 128-byte segment (SEG elements) holds CPB columns of BF rows;
    // BF = 1 (row-major): 64-byte row segments of CPB adjacent columns;
    // BF = 16 (row-major): 32-byte row segments (two workgroups per CU in FP64
    // at N = 2048); BF = 32 (row-major): 16-byte row segments, the strip that
    // fits LDS at N = 8192 (one FP64 or two FP32 columns, 139 KB)
    static constexpr int RB = (BF == 1 || BF >= 16) ? 1 : BF;     // rows per block of the layout
    static constexpr int SEG = BF == 1    ? 64 / (int)sizeof(C)
                               : BF == 16 ? 32 / (int)sizeof(C)
                               : BF == 32 ? 16 / (int)sizeof(C)
                                          : 128 / (int)sizeof(C);
    static constexpr int CPB = SEG / RB;               // columns per workgroup
    static constexpr int ELEMS = CPB * N;
    static constexpr int NTMAX = (BF == 1 || BF == 32) ? 1024 : 512;
    static constexpr int NT = (ELEMS / 8 < 64) ? 64 : ((ELEMS / 8 > NTMAX) ? NTMAX : ELEMS / 8);
    static constexpr int PER = ELEMS / NT;             // elements per thread
    static constexpr int LDS_BYTES = CPB * padded<N, 4>::value * (int)sizeof(C);
    // the row-major narrow strips ask for the registers of two workgroups per
    // CU where their LDS allows two
    static constexpr int WAVES = (BF >= 16 && 2 * LDS_BYTES <= 160 * 1024) ? (NT * 2 + 255) / 256 : 1;
    static_assert((BF == 1 || BF == 4 || BF == 8 || BF == 16 || BF == 32) && CPB >= 1 && CPB <= N &&
                      PER * NT == ELEMS && N % 16 == 0 && N % CPB == 0,
                  "blocked column tiling");
};

template <int N, int BF, int CORR, typename C, int GEN = 0, int MF = 0>
__global__ __launch_bounds__((ColB<N, BF, C>::NT), (ColB<N, BF, C>::WAVES)) void k_col_kmult_blk(C* __restrict__ Z, C* __restrict__ X,
                                                                     C* __restrict__ Y, size_t V, int sqrt_coupling,
                                                                     const double* __restrict__ q4,
                                                                     const double* __restrict__ sb,
                                                                     const double* __restrict__ sa16,
                                                                     const C* __restrict__ tw, NoiseSrc ns,
                                                                     int noise_scale) {
    static_assert(!GEN || CORR, "drawn spectra are written back as position-space noise");
    typedef typename scalar_of<C>::type T;
    constexpr int SEG = ColB<N, BF, C>::SEG, CPB = ColB<N, BF, C>::CPB, RB = ColB<N, BF, C>::RB;
    constexpr int NT = ColB<N, BF, C>::NT, PER = ColB<N, BF, C>::PER;
    constexpr int PN = padded<N, 4>::value;
    __shared__ C lds[CPB * PN];
    const int tid = threadIdx.x;
    const auto twl = tw_source<N, true>(tw);
    Z += zofs(4, V);
    X += zofs(8, V);
    Y += zofs(8, V);
    // 32-byte strips: the four strips of one 128-byte line on one XCD at neighbouring times
    const int y0 = (BF >= 16 ? xcd_strip(blockIdx.x, N / CPB) : (int)blockIdx.x) * CPB;
    const double invV = 1.0 / ((double)N * (double)N);
    // global offset within a plane / LDS slot of loaded element e: x-block
    // e / SEG, slot e mod SEG = c BF + x mod BF (row-major, BF = 1: row e / SEG,
    // column y0 + e mod SEG)
    auto goff = [&](int e) -> size_t { return (size_t)(e / SEG) * ((size_t)N * RB) + (size_t)y0 * RB + (e % SEG); };
    auto lofs = [&](int e) -> int {
        const int sl = e % SEG;
        return lpos<1, PN, 4>((e / SEG) * RB + sl % RB, sl / RB);
    };
    auto load = [&](const C* __restrict__ P) {
        const int t = opaque_v(tid);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = t + q * NT;
            lds[lofs(e)] = P[goff(e)];
        }
    };
    auto store = [&](C* __restrict__ P) {
        const int t = opaque_v(tid);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = t + q * NT;
            P[goff(e)] = lds[lofs(e)];
        }
    };
    // one direction d (0: x, 1: y): forward column FFT of P, accumulate
    // i f_d P^ into acc for this thread's pairs (kx = pr mod N, column c = pr / N)
    // with f_d = -2 pi cc sb[k_d] / phat^2 / V, zero at k = 0; CORR: scale P^
    // by sa (noise_scale) and write F_col^-1[sa P^ / V] back over P.  GEN: P^
    // drawn per mode instead; the x direction's draw also gives the y spectrum
    // (held in yg until then).
    const NoiseSrc nk = batch_noise(ns, V);
    C yg[GEN ? PER : 1];
    auto direction = [&](C* __restrict__ P, C acc[PER], int d) {
        const double* q4d = opaque(q4);
        const double* sbd = opaque(sb);
        const double* sad = opaque(sa16);
        if constexpr (!GEN) {
            load(P);
            __syncthreads();
            lds_fft<N, CPB, 1, PN, NT, -1, 4, MF>(lds, tw_fft(twl), opaque_v(tid));
        }
        const int t = opaque_v(tid);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int pr = t + q * NT;
            const int kx = pr % N, c = pr / N, ky = y0 + c;
            const int l = lpos<1, PN, 4>(kx, c);
            const size_t k16 = ((size_t)(kx >> 4) * N + ky) * 16 + (kx & 15);
            C v;
            if constexpr (GEN) {
                if (d == 0) {
                    T g[4];
                    kspace_normals4((uint32_t)((size_t)kx * N + ky), nk.step, nk.tag, blockIdx.y, nk.seed, nk.traj, g);
                    v = mkc<C>(g[0] * (T)N, g[1] * (T)N);      // sqrt(V) = N
                    yg[q] = mkc<C>(g[2] * (T)N, g[3] * (T)N);
                } else {
                    v = yg[q];
                }
            } else {
                v = lds[l];
            }
            if (CORR) {
                if (noise_scale) v = cscale(v, (T)sad[k16]);
                lds[l] = cscale(v, (T)invV);
            }
            T f = 0;
            if (kx != 0 || ky != 0) {
                const double cc = sqrt_coupling ? sad[k16] : 1.0;
                f = (T)(-2.0 * M_PI * invV * cc) / (T)(q4d[kx] + q4d[ky]) * (T)sbd[d ? ky : kx];
            }
            if (d == 0) acc[q] = mkc<C>(-(f * v.y), f * v.x);
            else acc[q] = mkc<C>(acc[q].x - f * v.y, acc[q].y + f * v.x);
        }
        __syncthreads();
        if (CORR) {
            lds_fft<N, CPB, 1, PN, NT, +1, 4, MF>(lds, tw_fft(twl), opaque_v(tid));
            store(P);
            __syncthreads();
        }
    };
    // plane pair p = blockIdx.y (one pass over the pairs would let the compiler
    // hoist every twiddle and kernel factor out of the loop into registers)
    const size_t po = (size_t)blockIdx.y * V;
    C acc[PER];
    direction(X + po, acc, 0);
    direction(Y + po, acc, 1);
    const int t = opaque_v(tid);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int pr = t + q * NT;
        lds[lpos<1, PN, 4>(pr % N, pr / N)] = acc[q];
    }
    __syncthreads();
    lds_fft<N, CPB, 1, PN, NT, +1, 4, MF>(lds, tw_fft(twl), opaque_v(tid));
    store(Z + po);
}


// ---------------------------------------------------------------------------
// row pass 3: inverse row FFTs of the A and B column-inverse spectra and the
// update U <- exp(i s B) U exp(-i s A) (field.h:1966-1996) in one pass.
// A's coefficients are kept in registers, B's stay in LDS.  The site algebra
// (exponentials, products, the optional projection onto SU(3)) is FP64 in both
// forms; U is stored in the element type C.
// ---------------------------------------------------------------------------
template <int N, int BF, typename C>
__device__ __forceinline__ void row_update_body(C* lds, BlockIx bi, C* __restrict__ U, const C* __restrict__ ZA,
                                                              const C* __restrict__ ZB, size_t V, double ss,
                                                              int reunit, const C* __restrict__ tw, int next_noise,
                                                              C* __restrict__ Xi, NoiseSrc ns,
                                                              const double* __restrict__ sa_pos) {
    static_assert(RowE<N, 4>::NT == RowE<N>::NT, "the update and noise bodies share the thread count");
    typedef typename scalar_of<C>::type T;
    constexpr int NT = RowE<N, 4>::NT, SPT = RowE<N, 4>::SPT;
    constexpr int PN = padded<N, 4>::value;     // padded row length in LDS
    const int tid = bi.tid;
    const auto twl = tw_source<N, sizeof(C) == sizeof(float2)>(tw);
    U += zofs_z(bi.bz, UPlanes<C>::value, V);
    ZA += zofs_z(bi.bz, 4, V);
    ZB += zofs_z(bi.bz, 4, V);
    const int x = row_of_block<N, BF>(bi.bx);
    const size_t row0 = (size_t)x * N;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = ZA[p * V + lay<N, BF>(x, tid + q * NT)];
    __syncthreads();
    row_ifft4<N, NT, PN>(lds, twl, tid);
    T ha[SPT][8];
#pragma unroll
    for (int q = 0; q < SPT; q++)
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const C a = lds[lpos<1, PN, 4>(tid + q * NT, p)];
            ha[q][2 * p] = a.x;
            ha[q][2 * p + 1] = a.y;
        }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < SPT; q++) lds[lpos<1, PN, 4>(tid + q * NT, p)] = ZB[p * V + lay<N, BF>(x, tid + q * NT)];
    __syncthreads();
    row_ifft4<N, NT, PN>(lds, twl, tid);
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int y = site_of<C>(tid + q * NT);
        const size_t i = row0 + y;
        T hb[8];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const C b = lds[lpos<1, PN, 4>(y, p)];
            hb[2 * p] = b.x;
            hb[2 * p + 1] = b.y;
        }
        if constexpr (sizeof(C) == sizeof(double2)) {
            m3 u;
            load_u(u, U, V, i);
            m3 w = mmul(exp_i_gen(hb, ss), u);
            double hd[8];
#pragma unroll
            for (int k = 0; k < UPlanes<C>::value; k++) fence(w.m[k]);      // the rows that are stored
#pragma unroll
            for (int k = 0; k < 8; k++) {
                hd[k] = ha[q][k];
                fence(hd[k]);
            }
            m3 r = mmul(w, exp_i_gen(hd, -ss));
            if (reunit) r = su3_project(r);
            store_u(U, V, i, r);
        } else {
            // FP64 exponentials, FP32 products; the projection in FP64.  The
            // fences keep the two exponentials apart (fewer live registers).
            const m3f eb = exp_i_gen_f(hb, ss);
            m3f u;
            load_u(u, U, V, i);
            m3f w = mmul(eb, u);
            T hd[8];
#pragma unroll
            for (int k = 0; k < UPlanes<C>::value; k++) fence(w.m[k]);      // the rows that are stored
#pragma unroll
            for (int k = 0; k < 8; k++) {
                hd[k] = ha[q][k];
                fence(hd[k]);
            }
            m3f r = mmul(w, exp_i_gen_f(hd, -ss));
            if (reunit) r = to_f(su3_project(to_d(r)));
            store_u(U, V, i, r);
        }
    }
    if (next_noise) {        // the next step's xi row spectra
        __syncthreads();
        NoiseSrc nn = batch_noise_z(ns, V, bi.bz);
        nn.step += 1;
        row_noise_body<N, BF, C>(lds, Xi + zofs_z(bi.bz, 8, V), V, nn, sa_pos, tw_fft(twl), x, bi.tid);
    }
}

template <int N, int BF, typename C>
__global__ __launch_bounds__((RowE<N, 4>::NT), RowWaves<C>::value) void k_row_update(C* __restrict__ U, const C* __restrict__ ZA,
                                                              const C* __restrict__ ZB, size_t V, double ss,
                                                              int reunit, const C* __restrict__ tw, int next_noise,
                                                              C* __restrict__ Xi, NoiseSrc ns,
                                                              const double* __restrict__ sa_pos) {
    __shared__ C lds[4 * padded<N, 4>::value];
    row_update_body<N, BF, C>(lds, own_block(), U, ZA, ZB, V, ss, reunit, tw, next_noise, Xi, ns, sa_pos);
}

// ---------------------------------------------------------------------------
// One launch per kernel class over every size group of a batch (row-major
// intermediate planes).  Block b of a launch belongs to the group whose block
// range holds it.  A group whose body needs NT threads, fewer than the T of
// the launch, runs T / NT bodies side by side in the block (rows, or column
// strips), each on its own thread range and its own part of LDS; all bodies
// of one block have the same N, so the barriers inside them coincide.
// ---------------------------------------------------------------------------
template <typename C>
struct GroupRun {
    int N, nblk;                      // lattice size, blocks of this group in the launch
    C *U, *Xi, *Vv, *ZA, *ZB;         // the group's trajectories, as for one grouped launch
    size_t V;
    NoiseSrc ns;                      // key of the group's first trajectory
    const double *sa_pos, *sa_k, *q4, *sb;
    const C* tw;
};
template <typename C>
struct MultiRun {
    int ng;
    GroupRun<C> g[JW_MAX_GROUPS];
};

// the group of launch block b; b becomes the block's index inside the group
template <typename C>
__device__ __forceinline__ const GroupRun<C>& group_of(const MultiRun<C>& m, int& b) {
    int g = 0;
    while (g < m.ng - 1 && b >= m.g[g].nblk) {
        b -= m.g[g].nblk;
        g++;
    }
    return m.g[g];
}

template <int N>
using int_c = std::integral_constant<int, N>;

// f(int_c<N>, group, block index in the group) for the block's group, N from
// the sizes a launch of T threads can hold (threads per body = MAXNT(N) <= T)
template <int T, template <int> class NTOF, typename C, typename F>
__device__ __forceinline__ void dispatch_group(const MultiRun<C>& m, F&& f) {
    int b = blockIdx.x;
    const GroupRun<C>& g = group_of(m, b);
    switch (g.N) {
        case 64: if constexpr (NTOF<64>::value <= T) f(int_c<64>{}, g, b); break;
        case 128: if constexpr (NTOF<128>::value <= T) f(int_c<128>{}, g, b); break;
        case 256: if constexpr (NTOF<256>::value <= T) f(int_c<256>{}, g, b); break;
        case 512: if constexpr (NTOF<512>::value <= T) f(int_c<512>{}, g, b); break;
        case 1024: if constexpr (NTOF<1024>::value <= T) f(int_c<1024>{}, g, b); break;
        case 2048: if constexpr (NTOF<2048>::value <= T) f(int_c<2048>{}, g, b); break;
        default: break;
    }
}
template <int N> struct RowNT { static constexpr int value = RowE<N>::NT; };
template <typename C>
struct ColNT { template <int N> struct of { static constexpr int value = ColE<N, C, 1>::NT; }; };

// LDS of a row launch of T threads: T / NT bodies of `planes` padded rows
template <int T, int planes, typename C>
struct RowMultiLds {
    template <int N>
    static constexpr int need() { return RowE<N>::NT <= T ? (T / RowE<N>::NT) * planes * padded<N, 4>::value : 0; }
    static constexpr int a = need<64>() > need<128>() ? need<64>() : need<128>();
    static constexpr int b = need<256>() > need<512>() ? need<256>() : need<512>();
    static constexpr int c = need<1024>() > need<2048>() ? need<1024>() : need<2048>();
    static constexpr int ab = a > b ? a : b;
    static constexpr int value = ab > c ? ab : c;
};
template <typename C>
struct ColMultiLds {
    template <int N>
    static constexpr int need() { return (1024 / ColE<N, C, 1>::NT) * padded<N, 3>::value * 2 * ColE<N, C, 1>::W; }
    static constexpr int a = need<64>() > need<128>() ? need<64>() : need<128>();
    static constexpr int b = need<256>() > need<512>() ? need<256>() : need<512>();
    static constexpr int c = need<1024>() > need<2048>() ? need<1024>() : need<2048>();
    static constexpr int ab = a > b ? a : b;
    static constexpr int value = ab > c ? ab : c;
};

// the row (x), trajectory (z) and thread of a row body in block b of its group
template <int N, int T>
__device__ __forceinline__ BlockIx row_slot(int b) {
    constexpr int NT = RowE<N>::NT, RPB = T / NT, NRB = N / RPB;
    const int s = threadIdx.x / NT;
    return BlockIx{(b % NRB) * RPB + s, 0, b / NRB, (int)threadIdx.x % NT};
}
template <int N, int T>
__device__ __forceinline__ int row_sub() { return threadIdx.x / RowE<N>::NT; }

template <typename C, int T>
__global__ __launch_bounds__(T, (NoiseWaves<64, C>::value)) void k_row_noise_fwd_multi(MultiRun<C> m) {
    __shared__ C lds[RowMultiLds<T, 2, C>::value];
    dispatch_group<T, RowNT>(m, [&](auto n, const GroupRun<C>& g, int b) {
        constexpr int N = decltype(n)::value;
        row_noise_fwd_body<N, 0, C>(lds + row_sub<N, T>() * 2 * padded<N, 4>::value, row_slot<N, T>(b), g.Xi, g.V,
                                    g.ns, g.sa_pos, g.tw);
    });
}
template <typename C, int T>
__global__ __launch_bounds__(T, RowWaves<C>::value) void k_row_uxiu_fwd_multi(MultiRun<C> m) {
    __shared__ C lds[RowMultiLds<T, 4, C>::value];
    dispatch_group<T, RowNT>(m, [&](auto n, const GroupRun<C>& g, int b) {
        constexpr int N = decltype(n)::value;
        row_uxiu_fwd_body<N, 0, C>(lds + row_sub<N, T>() * 4 * padded<N, 4>::value, row_slot<N, T>(b), g.Vv, g.U, g.V,
                                   g.ns, g.sa_pos, g.tw);
    });
}
template <typename C, int T>
__global__ __launch_bounds__(T, RowWaves<C>::value) void k_row_uxiu_corr_multi(MultiRun<C> m) {
    __shared__ C lds[RowMultiLds<T, 4, C>::value];
    dispatch_group<T, RowNT>(m, [&](auto n, const GroupRun<C>& g, int b) {
        constexpr int N = decltype(n)::value;
        row_uxiu_corr_body<N, 0, C>(lds + row_sub<N, T>() * 4 * padded<N, 4>::value, row_slot<N, T>(b), g.Vv, g.U, g.Xi,
                                    g.V, g.tw);
    });
}
template <typename C, int T>
__global__ __launch_bounds__(T, RowWaves<C>::value) void k_row_update_multi(MultiRun<C> m, double ss, int reunit) {
    __shared__ C lds[RowMultiLds<T, 4, C>::value];
    dispatch_group<T, RowNT>(m, [&](auto n, const GroupRun<C>& g, int b) {
        constexpr int N = decltype(n)::value;
        row_update_body<N, 0, C>(lds + row_sub<N, T>() * 4 * padded<N, 4>::value, row_slot<N, T>(b), g.U, g.ZA, g.ZB,
                                 g.V, ss, reunit, g.tw, 0, g.Xi, g.ns, g.sa_pos);
    });
}
// column passes: T = 1024; block b of a group: strip block, plane pair, trajectory
// in that order (the order of the grouped launch's grid)
template <typename C, int CORR, int GEN>
__global__ __launch_bounds__(1024, 1) void k_col_kmult_multi(MultiRun<C> m, int use_v, int sqrt_coupling,
                                                             int noise_scale) {
    __shared__ C lds[ColMultiLds<C>::value];
    dispatch_group<1024, ColNT<C>::template of>(m, [&](auto n, const GroupRun<C>& g, int b) {
        constexpr int N = decltype(n)::value;
        typedef ColE<N, C, CORR> CE;
        constexpr int SPB = 1024 / CE::NT, NSB = N / CE::W / SPB;
        const int s = threadIdx.x / CE::NT;
        const BlockIx bi{(b % NSB) * SPB + s, (b / NSB) % 4, b / (NSB * 4), (int)threadIdx.x % CE::NT};
        C* X = use_v ? g.Vv : g.Xi;
        col_kmult_body<N, CORR, C, 0, 0, GEN>(lds + s * padded<N, 3>::value * 2 * CE::W, bi, use_v ? g.ZB : g.ZA, X,
                                              X + 4 * g.V, g.V, sqrt_coupling, g.q4, g.sb, g.sa_k, g.tw, g.ns,
                                              noise_scale);
    });
}

// blocks of group g in a row launch of T threads / in a column launch
template <typename C>
inline int row_blocks(int N, int count, int T) {
    const int nt = N / 4 < 64 ? 64 : (N / 4 > 1024 ? 1024 : N / 4);      // RowE<N>::NT
    return count * N / (T / nt);
}
template <typename C>
inline int col_blocks(int N, int count) {
    const int W = 4096 / N > 16 ? 16 : 4096 / N;                          // ColE<N>::W
    const int elems = 2 * W * N, nt = elems / 8 < 64 ? 64 : (elems / 8 > 1024 ? 1024 : elems / 8);
    return count * 4 * (N / W) / (1024 / nt);
}

// one step of every group of a batch, one launch per kernel class (row-major
// planes; the noise pass separate from the update).  counts[g]: trajectories.
template <typename C>
static hipError_t step_multi_t(hipStream_t st, MultiRun<C> m, const int* counts, int maxN, const StepArgs& a) {
    if (m.ng < 1 || m.ng > JW_MAX_GROUPS) return hipErrorInvalidValue;
    const int T = maxN > 1024 ? 512 : 256;
    int nrow = 0, ncol = 0;
    for (int g = 0; g < m.ng; g++) {
        if (m.g[g].N < 64 || m.g[g].N > 2048) return hipErrorInvalidValue;
        nrow += row_blocks<C>(m.g[g].N, counts[g], T);
        ncol += col_blocks<C>(m.g[g].N, counts[g]);
    }
    auto set_blocks = [&](bool col) {
        for (int g = 0; g < m.ng; g++)
            m.g[g].nblk = col ? col_blocks<C>(m.g[g].N, counts[g]) : row_blocks<C>(m.g[g].N, counts[g], T);
    };
    const int wb = a.corr || a.kspace;
    hipError_t e;
    if (!a.kspace) {
        set_blocks(false);
        if (T == 256) hipLaunchKernelGGL((k_row_noise_fwd_multi<C, 256>), dim3(nrow), dim3(256), 0, st, m);
        else hipLaunchKernelGGL((k_row_noise_fwd_multi<C, 512>), dim3(nrow), dim3(512), 0, st, m);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    set_blocks(true);
    if (a.kspace)
        hipLaunchKernelGGL((k_col_kmult_multi<C, 1, 1>), dim3(ncol), dim3(1024), 0, st, m, 0, a.sqrt_coupling, a.corr);
    else if (a.corr)
        hipLaunchKernelGGL((k_col_kmult_multi<C, 1, 0>), dim3(ncol), dim3(1024), 0, st, m, 0, a.sqrt_coupling, 1);
    else
        hipLaunchKernelGGL((k_col_kmult_multi<C, 0, 0>), dim3(ncol), dim3(1024), 0, st, m, 0, a.sqrt_coupling, 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    set_blocks(false);
    if (wb) {
        if (T == 256) hipLaunchKernelGGL((k_row_uxiu_corr_multi<C, 256>), dim3(nrow), dim3(256), 0, st, m);
        else hipLaunchKernelGGL((k_row_uxiu_corr_multi<C, 512>), dim3(nrow), dim3(512), 0, st, m);
    } else {
        if (T == 256) hipLaunchKernelGGL((k_row_uxiu_fwd_multi<C, 256>), dim3(nrow), dim3(256), 0, st, m);
        else hipLaunchKernelGGL((k_row_uxiu_fwd_multi<C, 512>), dim3(nrow), dim3(512), 0, st, m);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    set_blocks(true);
    hipLaunchKernelGGL((k_col_kmult_multi<C, 0, 0>), dim3(ncol), dim3(1024), 0, st, m, 1, a.sqrt_coupling, 0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    set_blocks(false);
    if (T == 256)
        hipLaunchKernelGGL((k_row_update_multi<C, 256>), dim3(nrow), dim3(256), 0, st, m, a.sqrt_step, a.reunit);
    else
        hipLaunchKernelGGL((k_row_update_multi<C, 512>), dim3(nrow), dim3(512), 0, st, m, a.sqrt_step, a.reunit);
    return hipGetLastError();
}

// the batch's groups as one MultiRun (at most JW_MAX_GROUPS)
template <typename C>
static hipError_t step_multi_from(hipStream_t st, const GroupArgs* ga, int ng, const StepArgs& a) {
    if (ng < 1 || ng > JW_MAX_GROUPS) return hipErrorInvalidValue;
    MultiRun<C> m{};
    int counts[JW_MAX_GROUPS], maxN = 0;
    m.ng = ng;
    for (int g = 0; g < ng; g++) {
        const GroupArgs& s = ga[g];
        m.g[g] = GroupRun<C>{s.N, 0, (C*)s.U, (C*)s.Xi, (C*)s.Vv, (C*)s.ZA, (C*)s.ZB, s.V, s.ns,
                             s.sa_pos, s.sa_k, s.q4, s.sb, (const C*)s.tw};
        counts[g] = s.count;
        maxN = s.N > maxN ? s.N : maxN;
    }
    return step_multi_t<C>(st, m, counts, maxN, a);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename C>
static const C* twiddles(const jw_ctx* c) {
    if constexpr (sizeof(C) == sizeof(cplx)) return c->tw;
    else return c->twf;
}

// corr: the correlated-noise variant (xi~ written back); gen: the xi spectra
// drawn per mode from ns (implies the write-back); noise_scale: the drawn or
// transformed spectra are multiplied by sa (correlated coupling)
template <int N, int BF, typename C>
static hipError_t col_pass(jw_ctx* c, C* Z, C* X, C* Y, const double* sa_k, int sqrt_coupling, int corr,
                           int nbatch, const NoiseSrc& ns, int gen, int noise_scale) {
    const size_t V = c->V;
    const C* tw = twiddles<C>(c);
#define JW_COL_ARGS Z, X, Y, V, sqrt_coupling, c->q4, c->sb, sa_k, tw, ns, noise_scale
    if constexpr (BF == 0 || BF == 2) {
        constexpr int W = ColE<N, C>::W, PI = BF == 2;
        const dim3 grid(N / W, 4, nbatch);
        // split-BF16 matrix-core radix-16 stages: FP32 storage, row-major layout
        if constexpr (sizeof(C) == sizeof(float2) && BF == 0) {
            if (c->fft_mfma) {
                if (gen)
                    hipLaunchKernelGGL((k_col_kmult<N, 1, C, 0, 1, 1>), grid, dim3(ColE<N, C, 1>::NT), 0, c->stream,
                                       JW_COL_ARGS);
                else if (corr)
                    hipLaunchKernelGGL((k_col_kmult<N, 1, C, 0, 1>), grid, dim3(ColE<N, C, 1>::NT), 0, c->stream,
                                       JW_COL_ARGS);
                else
                    hipLaunchKernelGGL((k_col_kmult<N, 0, C, 0, 1>), grid, dim3(ColE<N, C, 0>::NT), 0, c->stream,
                                       JW_COL_ARGS);
                return hipGetLastError();
            }
        }
        if (gen)
            hipLaunchKernelGGL((k_col_kmult<N, 1, C, PI, 0, 1>), grid, dim3(ColE<N, C, 1>::NT), 0, c->stream, JW_COL_ARGS);
        else if (corr)
            hipLaunchKernelGGL((k_col_kmult<N, 1, C, PI>), grid, dim3(ColE<N, C, 1>::NT), 0, c->stream, JW_COL_ARGS);
        else
            hipLaunchKernelGGL((k_col_kmult<N, 0, C, PI>), grid, dim3(ColE<N, C, 0>::NT), 0, c->stream, JW_COL_ARGS);
    } else {
        constexpr int CPB = ColB<N, BF, C>::CPB, NT = ColB<N, BF, C>::NT;
        const dim3 grid(N / CPB, 4, nbatch);
        // the matrix-core stages in the 16-byte strips of the large lattices
        if constexpr (sizeof(C) == sizeof(float2) && BF == 32) {
            if (c->fft_mfma) {
                if (gen)
                    hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 1, C, 1, 1>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
                else if (corr)
                    hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 1, C, 0, 1>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
                else
                    hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 0, C, 0, 1>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
                return hipGetLastError();
            }
        }
        if (gen)
            hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 1, C, 1>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
        else if (corr)
            hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 1, C>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
        else
            hipLaunchKernelGGL((k_col_kmult_blk<N, BF, 0, C>), grid, dim3(NT), 0, c->stream, JW_COL_ARGS);
    }
#undef JW_COL_ARGS
    return hipGetLastError();
}

template <int N, int BF, typename C>
static hipError_t step_fused_n(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const StepArgs& a, const int* kcls) {
    const size_t V = c->V;
    const C* tw = twiddles<C>(c);
    hipError_t e;
    // The next step's noise rides on the update pass in FP64 for N <= 512,
    // where that measured faster (steps/s, fixed / correlated coupling, fused
    // against separate: N = 256 10032 / 10783 against 9785 / 10571; 512 7258 /
    // 7134 against 6670 / 6579); at 1024 and 2048 the separate pass is faster
    // (1916 / 1742 against 1746 / 1597; 471 / 426 against 463 / 420), and in
    // FP32 the separate pass, at two workgroups per CU, is faster everywhere.
    // Momentum-space noise (a.kspace): no noise pass; the first column pass
    // draws the spectra and writes the position-space noise's row spectra back
    // for the correlated row pass.
    constexpr bool fuse_noise = sizeof(C) == sizeof(double2) && N <= 512;
    const int first_noise = a.kspace ? 0 : fuse_noise ? a.first_noise : 1;
    const int next_noise = a.kspace ? 0 : fuse_noise ? a.next_noise : 0;
    if (first_noise) {
        KTimer t(c, kcls[0]);
        hipLaunchKernelGGL((k_row_noise_fwd<N, BF, C>), dim3(N, 1, a.nbatch), dim3(RowE<N>::NT), 0, c->stream, Xi, V, a.ns,
                           a.sa_pos, tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[1]);
        if ((e = col_pass<N, BF>(c, ZA, Xi, Xi + 4 * V, a.sa_k, a.sqrt_coupling, a.corr, a.nbatch, a.ns, a.kspace,
                                 a.corr)) != hipSuccess)
            return e;
    }
    {
        KTimer t(c, kcls[2]);
        if (a.corr || a.kspace)
            hipLaunchKernelGGL((k_row_uxiu_corr<N, BF, C>), dim3(N, 1, a.nbatch), dim3(RowE<N>::NT), 0, c->stream, Vv, U, Xi, V,
                               tw);
        else
            hipLaunchKernelGGL((k_row_uxiu_fwd<N, BF, C>), dim3(N, 1, a.nbatch), dim3(RowE<N>::NT), 0, c->stream, Vv, U, V, a.ns,
                               a.sa_pos, tw);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    {
        KTimer t(c, kcls[1]);
        if ((e = col_pass<N, BF>(c, ZB, Vv, Vv + 4 * V, a.sa_k, a.sqrt_coupling, 0, a.nbatch, a.ns, 0, 0)) != hipSuccess)
            return e;
    }
    {
        KTimer t(c, kcls[3]);
        hipLaunchKernelGGL((k_row_update<N, BF, C>), dim3(N, 1, a.nbatch), dim3(RowE<N, 4>::NT), 0, c->stream, U, ZA, ZB, V,
                           a.sqrt_step, a.reunit, tw, next_noise, Xi, a.ns, a.sa_pos);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int N, typename C>
static hipError_t step_fused_layout(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const StepArgs& a,
                                    const int* kcls) {
    switch (c->layout) {
        case 0: return step_fused_n<N, 0>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        case 1: return step_fused_n<N, 1>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        case 2: return step_fused_n<N, 2>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        case 4: return step_fused_n<N, 4>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        case 8: return step_fused_n<N, 8>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        case 16: return step_fused_n<N, 16>(c, U, Xi, Vv, ZA, ZB, a, kcls);
        default: return hipErrorInvalidValue;
    }
}

template <typename C>
static hipError_t step_fused_t(jw_ctx* c, C* U, C* Xi, C* Vv, C* ZA, C* ZB, const StepArgs& a) {
    static const int kcls[4] = {K_ROW_NOISE, K_COL_KMULT, K_ROW_UXIU, K_ROW_UPDATE};
#define JW_FUSED_CASE(n) \
    case n: return step_fused_layout<n>(c, U, Xi, Vv, ZA, ZB, a, kcls);
    switch (c->N) {
        JW_FUSED_CASE(64)
        JW_FUSED_CASE(128)
        JW_FUSED_CASE(256)
        JW_FUSED_CASE(512)
        JW_FUSED_CASE(1024)
        JW_FUSED_CASE(2048)
        default: return hipErrorInvalidValue;
    }
#undef JW_FUSED_CASE
}

}  // namespace jw
