// jw_internal.hpp -- context/field structures and kernel launchers shared by the
// kernels (jw_kernels.hip) and the C ABI (jw_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/jimwlk.h"

// intermediate-plane layout of the fused step unless jw_set_fused selects one:
// -1 picks per step kind (jw_default_layout in jw_api.hip)
#define JW_DEFAULT_LAYOUT (-1)

// Philox counter tags of the two noise streams
#define JW_TAG_LANGEVIN 0x4C414E47u   // 'LANG'
#define JW_TAG_MV 0x4D565348u         // 'MVSH'
#define JW_TAG_KSPACE 0x4B535043u     // 'KSPC': momentum-space Langevin noise

namespace jw {

typedef double2 cplx;
typedef float2 cplxf;

// Noise source for the generators: injected standard normals (AoS, `per_site`
// per site, reference draw order) or the Philox stream.
struct NoiseSrc {
    const double* inj;     // device pointer or nullptr -> Philox
    uint32_t seed, traj, step, tag;
    const uint32_t* stepp; // device step counter (read instead of step when set)
};

struct Alpha {
    double lambda_L, mu0_over_lambda, freeze_c, beta0;
};

}  // namespace jw

struct jw_field {
    jw_ctx* ctx;
    int t;                 // components per site (1 or 9)
    jw::cplx* d;           // t planes of V complex doubles (SoA)
};

struct jw_ctx {
    int N;
    size_t V;
    int device;
    hipStream_t stream;
    jw::cplx* tw;          // tw[m] = exp(-2 pi i m / N), m < N
    float2* twf;           // the same in FP32 (mixed-precision form)
    double* q4;            // q4[k] = 4 pow(sin(pi k / N), 2)   (momenta.h:111)
    double* sb;            // sb[k] = sin(2 pi k / N)           (momenta.h:119-120)
    // sqrt(alpha) table over the lattice, rebuilt when the variant/params change
    double* sa;
    int sa_variant;        // -1 none, 0 kernel formula, 1 noise formula
    jw::Alpha sa_params;
    unsigned sa_gen;       // bumped whenever sa is rebuilt
    // k-space sqrt(alpha) in 16-row blocks for the blocked column pass:
    // sa16[((kx / 16) N + ky) 16 + kx mod 16] = sa[kx N + ky]
    double* sa16;
    unsigned sa16_gen;     // sa_gen it was built from (0: never)
    // injected normals (AoS, up to 16 per site), valid for one generator call
    double* inj;
    int inj_per_site;      // 0 when no injection is pending
    // Philox stream state for the generators
    uint32_t seed, traj, step;
    // workspace: generator-basis planes (packed) and scratch complex planes
    jw::cplx* ws;
    size_t ws_planes;
    // 1: evolve through the unfused kernel sequence even where the fused passes exist
    int unfused;
    // fused step's intermediate-plane layout: 0 row-major (X and Y tiles of
    // 32-byte column strips in LDS), 1 row-major (one plane at a time over
    // 64-byte strips), 4 or 8 rows per block, -1 default
    int layout;
    // 1: the FP32 row-major column passes run their radix-16 stages as
    // split-BF16 DFT-matrix products on the matrix cores (jw_set_mixed_fft)
    int fft_mfma;
    // 1: solve_poisson sets the k = 0 mode to zero instead of multiplying it
    // by -g/mass^2 (jw_set_mv_zero_mode)
    int mv_zero_mode;
    // 1: the fused step draws the Langevin noise per momentum mode
    // (kspace_normals4) instead of per site (jw_set_noise_space)
    int noise_kspace;
    // per-kernel timing (HIP events on ctx->stream)
    int timing;
    double kernel_ms[16];
    long kernel_calls[16];
};

namespace jw {

// ---- launchers (all asynchronous on ctx->stream; return hipError_t) ----
hipError_t fft2d(jw_ctx* c, cplx* planes, int nplanes, int dir, double scale);
hipError_t fft_rows(jw_ctx* c, cplx* planes, int nplanes, int dir, double scale);
hipError_t fft_cols(jw_ctx* c, cplx* planes, int nplanes, int dir, double scale);

// sqrt(alpha) over the lattice into sa (V doubles)
hipError_t build_sqrt_alpha(jw_ctx* c, double* sa, int variant, const Alpha& a);

// reference per-function kernels on 9-component (SoA) fields
hipError_t set_unit(jw_ctx* c, cplx* U);
hipError_t generate_gaussian(jw_ctx* c, cplx* xix, cplx* xiy, const NoiseSrc& ns, const double* sa_pos);
hipError_t prepare_A_local(jw_ctx* c, cplx* A, const cplx* xix, const cplx* xiy, int coupling, int kernel,
                           const double* sa_k, const Alpha& al);
hipError_t uxiulocal(jw_ctx* c, cplx* Vx, cplx* Vy, const cplx* U, const cplx* xix, const cplx* xiy);
hipError_t update_uf(jw_ctx* c, cplx* U, const cplx* B, const cplx* A, double step);
hipError_t set_mv_model(jw_ctx* c, cplx* f, const NoiseSrc& ns, double sigma);
hipError_t solve_poisson(jw_ctx* c, cplx* f, int t, double mass, double g);
hipError_t mul_exp(jw_ctx* c, cplx* U, const cplx* f);
hipError_t trace(jw_ctx* c, const cplx* U, cplx* corr);
hipError_t average_and_symmetrize(jw_ctx* c, const cplx* in, cplx* out);
hipError_t reduce(jw_ctx* c, const cplx* corr, cplx* sum, cplx* err);
hipError_t scale_planes(jw_ctx* c, cplx* p, size_t n, double s);
hipError_t scale_by_table(jw_ctx* c, cplx* f, int t, const double* s);
// POSITION_EVOLUTION kernel sums over all site pairs (weights per displacement)
hipError_t position_sums(jw_ctx* c, cplx* A, cplx* B, const cplx* X, const cplx* Y, const cplx* VX, const cplx* VY,
                         const double* wa, const double* wb);

// generator-basis engine: the unfused step (kc: 0 for SQRT coupling, 1 otherwise)
hipError_t step_unfused(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const NoiseSrc& ns,
                        const double* sa_pos, const double* sa_k, int kc, int corr, double ss);
hipError_t step_unfused_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const NoiseSrc& ns,
                            const double* sa_pos, const double* sa_k, int kc, int corr, double ss, int reunit);
hipError_t mv_noise_packed(jw_ctx* c, cplx* R, const NoiseSrc& ns, double sigma);
hipError_t poisson_packed(jw_ctx* c, cplx* R, double mass, double g);
hipError_t mul_exp_packed(jw_ctx* c, cplx* U, const cplx* R);

// fused generator-basis step (64 <= N <= 2048), FP64 or FP32 storage
bool fused_supported(int N);
struct StepArgs {
    NoiseSrc ns;
    const double* sa_pos;    // sqrt(alpha) at positions (NOISE coupling) or nullptr
    const double* sa_k;      // sqrt(alpha) in k: row-major for layout 0, sa16 otherwise; or nullptr
    int sqrt_coupling;       // kernel factor carries sqrt(alpha) (SQRT coupling)
    int corr;                // per-mode correlated noise
    double sqrt_step;        // sqrt(Delta s)
    int reunit;              // project the updated U onto SU(3)
    int nbatch;              // trajectories per launch (grid z; 1 for a single lattice)
    int first_noise;         // generate this step's xi first (else the previous update did)
    int next_noise;          // the update pass also generates the next step's xi
    int kspace;              // xi drawn per mode in the first column pass (no noise pass)
};
hipError_t step_fused(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const StepArgs& a);
hipError_t step_fused_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const StepArgs& a);
// one step of every size group of a batch with one launch per kernel class
// (row-major intermediate planes, at most 8 groups of 64 <= N <= 2048); the
// pointers are the group's state arrays in the element type of the form
#define JW_MAX_GROUPS 8
struct GroupArgs {
    int N, count;
    void *U, *Xi, *Vv, *ZA, *ZB;
    size_t V;
    NoiseSrc ns;
    const double *sa_pos, *sa_k, *q4, *sb;
    const void* tw;
};
hipError_t step_multi(hipStream_t st, const GroupArgs* g, int ng, const StepArgs& a);
hipError_t step_multi_f32(hipStream_t st, const GroupArgs* g, int ng, const StepArgs& a);
// N = 4096, 8192: fused column passes, separate row transforms and site passes
bool large_supported(int N);
hipError_t step_large(jw_ctx* c, cplx* U, cplx* Xi, cplx* Vv, cplx* ZA, cplx* ZB, const StepArgs& a);
hipError_t step_large_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const StepArgs& a);
// U (9 FP64 planes) to / from the mixed form's FP32 state of `planes` planes
// (9, or 6: rows 0 and 1, row 2 rebuilt on the way back)
hipError_t u_convert(jw_ctx* c, cplxf* u32, cplx* u64, int to_f32, int planes);
// row 2 of a 9-plane FP64 U rebuilt from rows 0 and 1 (on stream s)
hipError_t u_rebuild(jw_ctx* c, cplx* U, hipStream_t s);
hipError_t step_inc(hipStream_t s, uint32_t* step);
hipError_t block16(jw_ctx* c, double* out, const double* in);

// observables
hipError_t corr_from_Uk(jw_ctx* c, const cplx* Uk, double* corr);
hipError_t accumulate_real(jw_ctx* c, const double* corr, double* sum, double* err);

// instrumentation and unit-test hooks
hipError_t calibration_copy(jw_ctx* c, const cplx* a, cplx* b, size_t n);
hipError_t exp_batch(jw_ctx* c, const double* h8, double s, double* out18, int n);

}  // namespace jw
