/* jimwlk.h -- C ABI of the MI355X JIMWLK Langevin library (libjimwlk_hip.so).
 *
 * The reference (ElsevierSoftwareX/SOFTX-D-21-00092) has no FFI: its boundary is
 * the C++ class and free-function surface that main_optimized.cpp:159-355 calls.
 * Each entry point below replaces one piece of that surface and cites it.  The
 * C++ classes in softx-d-21-00092_amd/host/ (lfield, gfield, fftw2D, momenta,
 * MV_class, rand_class, positions, config, mpi_class) sit on top of this ABI.
 *
 * Conventions
 *  - Lattice: N x N periodic, N a power of two in [4, 8192].  Nx == Ny as the
 *    reference assumes (field.h:1471).
 *  - Host fields are AoS arrays of complex doubles (interleaved re, im):
 *    entry k = 3*row + col of site x*N + y at [(x*N + y)*t + k]  (field.h:199-209).
 *    Device storage is opaque (SoA planes) and owned by the context.
 *  - Every call returns JW_OK (0) or a negative JW_ERR_* code; jw_last_error()
 *    returns the message of the last failure on the calling thread.  Calls are
 *    asynchronous on the context's stream unless they return host data;
 *    jw_sync() waits.  One host thread per context; not re-entrant.
 *  - FP64 throughout (the reference's arithmetic).
 */
#ifndef JIMWLK_H
#define JIMWLK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    JW_OK = 0,
    JW_ERR_ARG = -1,          /* bad argument (size, enum, null pointer)          */
    JW_ERR_HIP = -2,          /* HIP runtime error                                  */
    JW_ERR_ALLOC = -3,        /* device allocation failed                           */
    JW_ERR_UNSUPPORTED = -4,  /* valid in the reference but not supported here      */
    JW_ERR_STATE = -5         /* call out of sequence (e.g. noise not provided)     */
};

/* Coupling choices (config.h:40, field.h:2022-2050) plus two additions:
 *   JW_FIXED_COUPLING: kernel factor 1, white noise (the reference has no switch
 *     for it on the momentum path);
 *   JW_CORRELATED_NOISE_COUPLING: kernel factor 1, noise correlated per mode,
 *     xi~ = F^-1[sqrt(alpha(k)) F xi] (the per-mode Cholesky factor of the
 *     coupling covariance that field.h:1633-1651 + matrix.h:151-200 build densely).
 * JW_HATTA_COUPLING_CONSTANT is rejected: the reference leaves the coupling
 * uninitialised for it on this path (field.h:2022). */
enum jw_coupling {
    JW_SQRT_COUPLING_CONSTANT = 0,
    JW_NOISE_COUPLING_CONSTANT = 1,
    JW_HATTA_COUPLING_CONSTANT = 2,
    JW_FIXED_COUPLING = 3,
    JW_CORRELATED_NOISE_COUPLING = 4
};

/* Kernel choices (config.h:42) */
enum jw_kernel { JW_LINEAR_KERNEL = 0, JW_SIN_KERNEL = 1 };

/* Running coupling alpha(q2) = (4 pi / beta0) / ln[ (r^2)^(1/c) + (q2 N^2 / L^2)^(1/c) ]^c
 * with L = lambda_L = Lambda*N, r = mu0/Lambda.  The reference hard-codes
 * {6, 2.5, 0.2, 9} (field.h:2046, 2257-2273); a null pointer selects those. */
typedef struct jw_alpha {
    double lambda_L;
    double mu0_over_lambda;
    double freeze_c;
    double beta0;
} jw_alpha;

typedef struct jw_ctx jw_ctx;
typedef struct jw_field jw_field;

/* ---- context ------------------------------------------------------------
 * Replaces the lattice set-up of main_optimized.cpp:64-92 for one rank:
 * config sizes, momenta::set (momenta.h:90-125), fftw2D(config*) + init2D
 * (mpi_fftw_class.h:82-189).  Multi-rank MPI is not part of this library:
 * the context is one lattice on one GPU. */
int jw_create(jw_ctx** out, int N, int device);
void jw_destroy(jw_ctx* ctx);
int jw_sync(jw_ctx* ctx);
int jw_lattice_size(const jw_ctx* ctx);
const char* jw_last_error(void);
const char* jw_version(void);

/* ---- fields: lfield<double,t>(Nxl, Nyl) (field.h:152-299) ------------------
 * A field belongs to its context: destroy the fields before the context. */
int jw_field_create(jw_ctx* ctx, int t, jw_field** out);          /* zero-filled (field.h:70-80) */
void jw_field_destroy(jw_field* f);
int jw_field_upload(jw_field* f, const double* host_aos);          /* 2*t*V doubles */
int jw_field_download(const jw_field* f, double* host_aos);
int jw_field_copy(jw_field* dst, const jw_field* src);             /* lfield::operator= (field.h:301-318) */
int jw_set_to_unit(jw_field* f);                                   /* lfield::setToUnit (field.h:196-211) */
int jw_set_to_zero(jw_field* f);                                   /* lfield::setToZero (field.h:180-188) */

/* ---- noise --------------------------------------------------------------
 * The reference seeds thread-local ranlux24 generators from the clock
 * (field.h:692-697, 2131-2135) and cannot be reproduced.  Here the generators
 * take either injected standard normals (one call's worth, in the reference's
 * draw order: 16 per site n0,m0,n1,m1,... for the Langevin noise,
 * field.h:2162-2165; 8 per site n0..n7 for an MV sheet, field.h:708-710), or
 * the Philox4x32-10 stream: key {seed, traj}, counter {site, step, tag, call}.
 * rand_class (rand_class.h:47-68) is re-backed by this stream. */
int jw_noise_inject(jw_ctx* ctx, const double* host_normals, int per_site);
int jw_noise_philox(jw_ctx* ctx, uint32_t seed, uint32_t traj, uint32_t step);
int jw_noise_get_step(const jw_ctx* ctx, uint32_t* step);

/* ---- checkpoints ----------------------------------------------------------
 * No counterpart in the reference (its clock-seeded noise cannot be resumed).
 * A checkpoint holds U (FP64, the reference's AoS entry order) and the Philox
 * state {seed, traj, step} of the context; reading it restores both, so that
 * the evolution continues bit for bit as if it had not stopped.  Format: the
 * 8 bytes "JWCKPT01", six uint32 (N, seed, traj, step, 9, 0), then 18 V
 * doubles.  Reading into a context of another N is JW_ERR_ARG; writing while
 * a noise injection is pending is JW_ERR_STATE.  Both calls synchronise the
 * context's stream. */
int jw_checkpoint_write(jw_ctx* ctx, const jw_field* U, const char* path);
int jw_checkpoint_read(jw_ctx* ctx, jw_field* U, const char* path);

/* ---- reference per-function entry points (9-component fields) ------------ */
/* fftw2D::execute2D (mpi_fftw_class.h:191-315): unitary, dir 1 = FFTW_FORWARD. */
int jw_execute2D(jw_ctx* ctx, jw_field* f, int dir);
/* generate_gaussian (field.h:2116-2243) */
int jw_generate_gaussian(jw_ctx* ctx, jw_field* xi_x, jw_field* xi_y);
/* generate_gaussian_with_noise_coupling_constant (field.h:2245-2365) */
int jw_generate_gaussian_with_noise_coupling_constant(jw_ctx* ctx, jw_field* xi_x, jw_field* xi_y,
                                                      const jw_alpha* alpha);
/* prepare_A_local, 7-argument form (field.h:1998-2075) */
int jw_prepare_A_local(jw_ctx* ctx, jw_field* A, const jw_field* xi_x, const jw_field* xi_y,
                       int coupling, int kernel, const jw_alpha* alpha);
/* uxiulocal (field.h:1892-1933) */
int jw_uxiulocal(jw_ctx* ctx, jw_field* v_x, jw_field* v_y, const jw_field* U,
                 const jw_field* xi_x, const jw_field* xi_y);
/* update_uf (field.h:1966-1996) */
int jw_update_uf(jw_ctx* ctx, jw_field* U, const jw_field* B, const jw_field* A, double step);
/* ---- POSITION_EVOLUTION (main_optimized.cpp:271-297) ---------------------
 * prepare_A_and_B_local (field.h:2368-2519) at every site, the loop of
 * main_optimized.cpp:288-297: brute-force position-space sums over all site
 * pairs, A(x) = sum_y w(x-y).(xi_x(y), xi_y(y)) and B(x) the same over
 * U(y) xi(y) U(y)^dagger, with the kernel of positions.h:131-141 (SIN) or the
 * integer distances (LINEAR).  coupling: SQRT (the position-space formula of
 * field.h:2449-2456), NOISE, FIXED or CORRELATED (factor 1: for NOISE the
 * reference correlates the noise first, main_optimized.cpp:277-282, which
 * jw_correlate_noise replaces); HATTA is JW_ERR_UNSUPPORTED.  O(V^2) work:
 * N <= JW_POSITION_MAX_N, a cross-check of the momentum path at small N.
 * Synchronous. */
#define JW_POSITION_MAX_N 256
int jw_prepare_A_and_B_local(jw_ctx* ctx, jw_field* A, jw_field* B, const jw_field* xi_x, const jw_field* xi_y,
                             const jw_field* U, int coupling, int kernel);
/* gfield::multiplyByCholesky (field.h:1653-1689) with the factor of
 * main_optimized.cpp:145-157 (dense Cholesky of the coupling covariance,
 * matrix.h:151-200), replaced in distribution by the per-mode factor: in place
 * f <- F^-1[sqrt(alpha(k)) F f] on every component.  Its covariance is the
 * matrix the reference factors (alpha(k)/sqrt(V) of field.h:1633-1651,
 * transformed back); the sample-wise factor is the symmetric square root
 * instead of the Cholesky one. */
int jw_correlate_noise(jw_ctx* ctx, jw_field* f, const jw_alpha* alpha);

/* lfield::setMVModel (field.h:683-757); mu is MV_class::mu_parameter */
int jw_set_mv_model(jw_ctx* ctx, jw_field* f, double g, double mu, int Ny);
/* lfield::solvePoisson (field.h:867-877) */
int jw_solve_poisson(jw_ctx* ctx, jw_field* f, double mass, double g);
/* The k = 0 mode of solvePoisson, for jw_solve_poisson and jw_mv_init on this
 * context.  mode 0 (default): the reference's factor -g/mass^2 with no guard
 * (field.h:867-877), which at m a = 1e-4 amplifies the zero mode of each
 * sheet's charge by about 1e8; mode 1: that mode is set to zero (each sheet's
 * A loses its constant colour part).  No counterpart in the reference. */
int jw_set_mv_zero_mode(jw_ctx* ctx, int mode);
/* lfield::operator*= : U <- U exp(i f) (field.h:337-365) */
int jw_mul_exp(jw_ctx* ctx, jw_field* U, const jw_field* f);
/* lfield::trace (field.h:1375-1432): corr (t = 1) = tr[(U/V)(U/V)^dagger] */
int jw_trace(jw_ctx* ctx, const jw_field* U, jw_field* corr);
/* gfield::average_and_symmetrize (field.h:1460-1492), in place on a t = 1 field */
int jw_average_and_symmetrize(jw_ctx* ctx, jw_field* corr);
/* gfield::reduce(sum, err, mpi) (field.h:1521-1536): sum += c, err += c^2 */
int jw_reduce(jw_ctx* ctx, const jw_field* corr, jw_field* sum, jw_field* err);

/* ---- evolution engine ----------------------------------------------------
 * main_optimized.cpp:234-311, MOMENTUM_EVOLUTION branch: nsteps Langevin steps
 * U <- exp(i sqrt(ds) B) U exp(-i sqrt(ds) A) (field.h:1966-1996), computed in
 * the generator basis (8 real planes per su(3) field, two per complex FFT).
 * Noise: the pending injection (nsteps must be 1) or the Philox stream (its
 * step counter advances by one per step).  step = Delta s as the caller wants
 * it (the reference parses it through float, config.h:92-93).  The SIN kernel
 * runs the fused passes for 64 <= N <= 2048; at N = 4096 and 8192 the fused
 * column passes (forward column FFT, kernel multiply, inverse column FFT per
 * strip) with the row transforms and the site algebra as separate passes; the
 * unfused kernel sequence otherwise (and with jw_set_fused(ctx, 0)).  All of
 * them compute the same step. */
int jw_evolve(jw_ctx* ctx, jw_field* U, int coupling, int kernel, double step, const jw_alpha* alpha,
              int nsteps);

/* The same evolution in the mixed-precision form (no counterpart in the
 * reference, which is FP64 throughout; it replaces the same loop,
 * main_optimized.cpp:234-311): U and every field and FFT in FP32, the SU(3)
 * exponentials and the projection of U onto SU(3) in FP64; the products
 * around the exponentials and U xi U^dagger are FP32 in the fused passes
 * (64 <= N <= 2048; U xi U^dagger also at N = 8192) and FP64 in the others.  U is
 * converted to FP32 on entry and back on return, and projected onto SU(3)
 * every reunit_every steps and after the last step (reunit_every = 0: never).
 * The noise is the same Philox stream (normals computed in FP32 from the same
 * words) or the pending injection.  SIN kernel; the fused passes for
 * 64 <= N <= 2048, at N = 4096 and 8192 the fused column passes with separate
 * row transforms and site passes, the unfused ones otherwise. */
int jw_evolve_mixed(jw_ctx* ctx, jw_field* U, int coupling, int kernel, double step, const jw_alpha* alpha,
                    int nsteps, int reunit_every);

/* Where the Langevin noise of jw_evolve / jw_evolve_mixed / jw_evolve_noalloc
 * on this context is drawn.  JW_NOISE_POSITION (default): 16 normals per site,
 * as generate_gaussian (field.h:2116-2167) in the library's Philox stream.
 * JW_NOISE_MOMENTUM: the spectra of the noise planes drawn per mode -- the
 * Fourier transform of a pair of independent white fields is white circular
 * complex Gaussian, so the same noise in distribution -- which removes the
 * noise pass and the forward column transform of the step.  Philox call
 * {k = kx N + ky, step, 'KSPC', p} (key {seed, traj}) gives four 32-bit
 * uniforms u_j = (w_j + 1/2) 2^-32 and two Box-Muller pairs (z0, z1), (z2, z3);
 * the x- and y-direction spectra of plane pair p (generators 2p, 2p+1) are
 * N (z0 + i z1) and N (z2 + i z3).  The position-space normals are then
 * n_2p + i n_2p+1 = N^-1 sum_k e^{+2 pi i k.x/N} (z0 + i z1)_k and likewise m
 * from (z2, z3).  Needs the fused passes (64 <= N <= 8192), the SIN kernel and a
 * coupling other than NOISE_COUPLING_CONSTANT (which scales the noise per
 * site): JW_ERR_UNSUPPORTED otherwise, JW_ERR_STATE with a pending injection.
 * No counterpart in the reference. */
#define JW_NOISE_POSITION 0
#define JW_NOISE_MOMENTUM 1
int jw_set_noise_space(jw_ctx* ctx, int space);

/* ---- allocation-free evolution -------------------------------------------
 * No counterpart in the reference.  It is for callers that cannot allocate
 * device memory while they run: code under HIP stream or graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  jw_evolve_noalloc computes
 * what jw_evolve computes (same kernels, same results) and calls no hipMalloc,
 * hipFree or allocating library routine: its workspace (the 24 intermediate
 * planes and the coupling tables, 400 V bytes) is one library-owned device
 * array per GPU of JW_NOALLOC_SCRATCH_BYTES (100 MiB), held until the process
 * ends.  The context, the field and any noise injection are made beforehand
 * (those calls allocate).  Limits: N <= JW_NOALLOC_MAX_N (JW_ERR_ARG above);
 * the LINEAR kernel is JW_ERR_UNSUPPORTED (its 9-component path needs 54
 * planes).  Every call on one GPU shares that GPU's array and leaves its
 * contents behind: calls that overlap on one GPU (two contexts' streams, or
 * two threads) race, and the caller must serialise them.  Per-kernel timing
 * (jw_timing) is not recorded in this call. */
#define JW_NOALLOC_MAX_N 512
#define JW_NOALLOC_SCRATCH_BYTES (400ull * JW_NOALLOC_MAX_N * JW_NOALLOC_MAX_N)
int jw_evolve_noalloc(jw_ctx* ctx, jw_field* U, int coupling, int kernel, double step, const jw_alpha* alpha,
                      int nsteps);

/* ---- batched form ------------------------------------------------------
 * No counterpart in the reference (one lattice per run); it advances many
 * independent trajectories -- the reference's `stat` loop of main_optimized.cpp
 * and its MV + evolution body, run concurrently -- as size groups of equal N
 * (64 <= N <= 2048): group i holds count[i] trajectories of N[i]^2, numbered
 * consecutively over the groups.  fp32 = 1: the mixed-precision form (as
 * jw_evolve_mixed), 0: FP64 (as jw_evolve).  Each group's kernels take all its
 * trajectories in one launch, or every group's in one launch per kernel class
 * (jw_batch_set_launch); one step of all groups is captured once into a HIP
 * graph and replayed.
 * Trajectory t draws the Philox stream of key {seed, traj0 + t}, so it is the
 * trajectory a single lattice evolves after jw_noise_philox(seed, traj0 + t,
 * step).  The intermediate-plane layout of each group is chosen on first use
 * by timing each of layouts 0, 2, 4, 8 and 16, and in FP32 layout 0 with the
 * matrix-core column stages, on a copy of its state (jw_batch_set_layout
 * overrides). */
typedef struct jw_batch jw_batch;
int jw_batch_create(jw_batch** out, int device, int ngroups, const int* N, const int* count, int fp32);
void jw_batch_destroy(jw_batch* b);
int jw_batch_size(const jw_batch* b);
/* copy trajectory t from / to a 9-component field of a context with the same N */
int jw_batch_set(jw_batch* b, int t, const jw_field* U);
int jw_batch_get(jw_batch* b, int t, jw_field* U);
/* noise key and step of every trajectory (trajectory t: key {seed, traj0 + t}) */
int jw_batch_philox(jw_batch* b, uint32_t seed, uint32_t traj0, uint32_t step);
/* layout 0, 2, 4, 8 or 16 for group g (FP32 butterflies), or -1 to tune on the
 * next evolution; the chosen layout, and whether the tuner put the FP32
 * row-major column transforms' radix-16 stages on the matrix cores (1) or not
 * (0) (as jw_set_mixed_fft) */
int jw_batch_set_layout(jw_batch* b, int group, int layout);
int jw_batch_layout(const jw_batch* b, int group);
int jw_batch_fft_mode(const jw_batch* b, int group);
/* How a step launches its kernels: 0, each size group its own launches on its
 * own stream (the groups run concurrently; per-group layouts); 1, one launch
 * per kernel class over all groups (at most JW_MAX_GROUPS = 8; row-major
 * planes, FP32 butterflies): block b of a launch finds its group from the
 * groups' block ranges, and a group of small N runs several rows or column
 * strips side by side in one workgroup; -1 (default): timed on the first
 * evolution after the layouts, the faster kept.  jw_batch_launch: the mode. */
int jw_batch_set_launch(jw_batch* b, int mode);
int jw_batch_launch(const jw_batch* b);
/* JW_NOISE_POSITION or JW_NOISE_MOMENTUM for every trajectory, as
 * jw_set_noise_space (trajectory t keyed {seed, traj0 + t}) */
int jw_batch_set_noise_space(jw_batch* b, int space);
/* nsteps steps of every trajectory (SIN kernel); reunit_every as jw_evolve_mixed
 * (ignored in FP64).  Asynchronous: jw_batch_sync waits. */
int jw_batch_evolve(jw_batch* b, int coupling, int kernel, double step, const jw_alpha* alpha, int nsteps,
                    int reunit_every);
int jw_batch_sync(jw_batch* b);

/* MV initial condition (main_optimized.cpp:178-191): U = 1, then Ny sheets
 * U <- U exp(i A_s), A_s = F^-1[ F rho_s * (-g / (-phat2 + m^2)) ], rho_s with
 * sigma = g mu / sqrt(Ny), m = mass * g^2 * mu.  Charges: injected (8*V*Ny
 * normals, sheet-major; may be null to use Philox with the context's seed,
 * traj and sheet index as the step) */
int jw_mv_init(jw_ctx* ctx, jw_field* U, int Ny, double g, double mu, double mass,
               const double* host_normals);

/* ---- observables ----------------------------------------------------------
 * Measurement (main_optimized.cpp:317-333): F U -> trace -> symmetrise, as a
 * real array corr[V] (the reference's values are real; field.h:1375-1492). */
int jw_measure_corr(jw_ctx* ctx, const jw_field* U, double* host_corr);
/* Dipole S(r) = (V/3) Re sum_k corr(k) e^{i k.r} over the whole lattice
 * (Wiener-Khinchin; the reference prints k-space only).  host_corr: V reals. */
int jw_dipole_from_corr(jw_ctx* ctx, const double* host_corr, double* host_S);
/* Radial bins of S(r) by minimal-image distance (bin b: b <= |r| < b+1) and the
 * saturation scale Q_s = sqrt(2)/R_s, S(R_s) = exp(-1/2), linear interpolation.
 * Host-side post-processing; *Qs = 0 when S never drops below exp(-1/2). */
int jw_radial_bins(int N, const double* S, int nbins, double* rbar, double* Sbin, double* count);
int jw_saturation_scale(int nbins, const double* rbar, const double* Sbin, const double* count, double* Qs);

/* ---- instrumentation ------------------------------------------------------ */
/* Select how jw_evolve runs the generator-basis step; all modes compute the
 * same step.  mode 0: the unfused kernel sequence; 1: the fused row/column
 * passes (default, 64 <= N <= 2048) with the default layout of their
 * intermediate planes; 2, 3, 4: the fused passes with row-major, 4-row-blocked
 * or 8-row-blocked intermediate planes; 5: row-major planes with the column
 * pass taking one plane at a time over 64-byte column strips; 6: the two
 * directions of each plane pair interleaved site by site, so that a column
 * strip of both is one row segment; 7: row-major planes with the column pass
 * taking one plane at a time over 32-byte column strips (two workgroups per CU
 * in FP64); 8: one plane at a time over 16-byte strips (N = 4096 and 8192
 * only).  At N = 4096 modes 0, 1, 2 and 8 exist, at N = 8192 modes 0, 1 and 8
 * (the column strips that fit LDS); others are JW_ERR_UNSUPPORTED there
 * (2..8 for measurement). */
int jw_set_fused(jw_ctx* ctx, int mode);
/* The column transforms of the mixed-precision form (jw_evolve_mixed, fused
 * passes, row-major layout).  mode 0 (default): FP32 radix-8/4/2 butterflies;
 * 1: the radix-16 stages as DFT-matrix products on the matrix cores in split
 * BF16 (x = hi + lo, three products lo.hi + hi.lo + hi.hi accumulated in FP32,
 * v_mfma_f32_32x32x16_bf16), the remaining stage in FP32 butterflies: about
 * 2^-17 relative per matrix-core stage instead of 2^-24.  No counterpart in
 * the reference (FP64 FFTW). */
int jw_set_mixed_fft(jw_ctx* ctx, int mode);
/* Enable per-kernel timing with HIP events on the context's stream. */
int jw_timing(jw_ctx* ctx, int enable);
/* Average duration (ms) and call count of kernel class `which` (see jw_kernel_name). */
int jw_kernel_time(const jw_ctx* ctx, int which, double* avg_ms, long* calls);
const char* jw_kernel_name(int which);
/* A plain device copy of nplanes complex planes (16 V nplanes bytes read and
 * as many written) in the context's workspace, kernel k_calibration_copy: the
 * reference size for scaling hardware traffic counters. */
int jw_calibration_copy(jw_ctx* ctx, int nplanes);
/* exp(i s H(h)) for n generator vectors h (8 reals each) -> n 3x3 matrices (18 reals). */
int jw_su3_exp_gen(jw_ctx* ctx, const double* host_h8, double s, double* host_out18, int n);

#ifdef __cplusplus
}
#endif

#endif /* JIMWLK_H */
