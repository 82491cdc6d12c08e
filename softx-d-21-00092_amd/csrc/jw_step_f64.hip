// jw_step_f64.hip -- the fused step with FP64 storage (the plain form), its
// large-lattice variant (N = 4096, 8192), and the
// helpers shared by both forms: the 16-row-blocked sqrt(alpha) table and the
// U conversions of the mixed form.
#include "jw_engine.hpp"
#include "jw_packed.hpp"

namespace jw {

bool fused_supported(int N) { return N >= 64 && N <= 2048; }
bool large_supported(int N) { return N == 4096 || N == 8192; }

hipError_t step_fused(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const StepArgs& a) {
    return step_fused_t(c, U, Xi, Vv, ZA, ZB, a);
}

hipError_t step_multi(hipStream_t st, const GroupArgs* g, int ng, const StepArgs& a) {
    return step_multi_from<cplx>(st, g, ng, a);
}

hipError_t step_large(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const StepArgs& a) {
    return step_large_t(c, U, Xi, Vv, ZA, ZB, a);
}

__global__ __launch_bounds__(256) void k_block16(double* __restrict__ out, const double* __restrict__ in, int N) {
    const size_t V = (size_t)N * N;
    GRID_STRIDE(i, V) {
        const int kx = (int)(i / N), ky = (int)(i % N);
        out[((size_t)(kx >> 4) * N + ky) * 16 + (kx & 15)] = in[i];
    }
}

hipError_t block16(jw_ctx* c, double* out, const double* in) {
    if (c->N % 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_block16, dim3(grid_for(c->V, 256)), dim3(256), 0, c->stream, out, in, c->N);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// U between the FP64 field (9 planes) and the FP32 state of the mixed-precision
// form: 9 planes, or rows 0 and 1 (6 planes, the fused and large-lattice passes) with row 2
// rebuilt as conj(row 0 x row 1) in FP64 on the way back
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_u_to_f32(cplxf* __restrict__ out, const cplx* __restrict__ in, size_t n) {
    GRID_STRIDE(i, n) out[i] = mkf((float)in[i].x, (float)in[i].y);
}
__global__ __launch_bounds__(256) void k_u_to_f64(cplx* __restrict__ out, const cplxf* __restrict__ in, size_t n) {
    GRID_STRIDE(i, n) out[i] = to_d(in[i]);
}
__global__ __launch_bounds__(256) void k_u2_to_f64(cplx* __restrict__ out, const cplxf* __restrict__ in, size_t V) {
    GRID_STRIDE(i, V) {
        cplx m[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            m[k] = to_d(in[k * V + i]);
            out[k * V + i] = m[k];
        }
        out[6 * V + i] = conjg(csub(cmul(m[1], m[5]), cmul(m[2], m[4])));
        out[7 * V + i] = conjg(csub(cmul(m[2], m[3]), cmul(m[0], m[5])));
        out[8 * V + i] = conjg(csub(cmul(m[0], m[4]), cmul(m[1], m[3])));
    }
}

// row 2 of an FP64 field rebuilt from rows 0 and 1 (after the fused passes,
// which keep only those rows current)
__global__ __launch_bounds__(256) void k_u_rebuild(cplx* __restrict__ U, size_t V) {
    GRID_STRIDE(i, V) {
        cplx m[6];
#pragma unroll
        for (int k = 0; k < 6; k++) m[k] = U[k * V + i];
        U[6 * V + i] = conjg(csub(cmul(m[1], m[5]), cmul(m[2], m[4])));
        U[7 * V + i] = conjg(csub(cmul(m[2], m[3]), cmul(m[0], m[5])));
        U[8 * V + i] = conjg(csub(cmul(m[0], m[4]), cmul(m[1], m[3])));
    }
}

hipError_t u_rebuild(jw_ctx* c, cplx* U, hipStream_t s) {
    hipLaunchKernelGGL(k_u_rebuild, dim3(grid_for(c->V, 256)), dim3(256), 0, s, U, c->V);
    return hipGetLastError();
}

hipError_t u_convert(jw_ctx* c, cplxf* u32, cplx* u64, int to_f32, int planes) {
    if (planes != 6 && planes != 9) return hipErrorInvalidValue;
    const size_t n = (size_t)planes * c->V;
    if (to_f32) hipLaunchKernelGGL(k_u_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, u32, u64, n);
    else if (planes == 9) hipLaunchKernelGGL(k_u_to_f64, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, u64, u32, n);
    else hipLaunchKernelGGL(k_u2_to_f64, dim3(grid_for(c->V, 256)), dim3(256), 0, c->stream, u64, u32, c->V);
    return hipGetLastError();
}

// the device Philox step counter of a batch, advanced once per step (inside the
// captured step graph)
__global__ void k_step_inc(uint32_t* step) {
    if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1u;
}

hipError_t step_inc(hipStream_t s, uint32_t* step) {
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);
    return hipGetLastError();
}

}  // namespace jw
