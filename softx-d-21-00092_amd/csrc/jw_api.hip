// jw_api.hip -- the C ABI declared in include/jimwlk.h.
#include "jw_internal.hpp"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

using jw::cplx;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(JW_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

const jw::Alpha kAlphaRef = {6.0, 2.5, 0.2, 9.0};

jw::Alpha to_alpha(const jw_alpha* a) {
    if (!a) return kAlphaRef;
    return jw::Alpha{a->lambda_L, a->mu0_over_lambda, a->freeze_c, a->beta0};
}

bool valid_N(int N) { return N >= 4 && N <= 8192 && (N & (N - 1)) == 0; }

int ensure_ws(jw_ctx* c, size_t planes) {
    if (c->ws_planes >= planes) return JW_OK;
    if (c->ws) hipFree(c->ws);
    c->ws = nullptr;
    c->ws_planes = 0;
    if (hipMalloc(&c->ws, planes * c->V * sizeof(cplx)) != hipSuccess)
        return fail(JW_ERR_ALLOC, "workspace of %zu planes (%zu MB) failed", planes,
                    planes * c->V * sizeof(cplx) >> 20);
    c->ws_planes = planes;
    return JW_OK;
}

int check_alpha(const jw::Alpha& a) {
    if (!(a.lambda_L > 0) || !(a.mu0_over_lambda > 1.0) || !(a.freeze_c > 0) || !(a.beta0 > 0))
        return fail(JW_ERR_ARG, "invalid running-coupling parameters");
    return JW_OK;
}

int ensure_sa(jw_ctx* c, int variant, const jw::Alpha& a) {
    if (!c->sa) {
        if (hipMalloc(&c->sa, c->V * sizeof(double)) != hipSuccess)
            return fail(JW_ERR_ALLOC, "coupling table allocation failed");
        c->sa_variant = -1;
    }
    if (c->sa_variant == variant && memcmp(&c->sa_params, &a, sizeof a) == 0) return JW_OK;
    if (int e = check_alpha(a)) return e;
    HIP_CHECK(jw::build_sqrt_alpha(c, c->sa, variant, a));
    c->sa_variant = variant;
    c->sa_params = a;
    c->sa_gen += 1;
    return JW_OK;
}

// the current sa table in 16-row blocks (for the blocked column pass)
int ensure_sa16(jw_ctx* c) {
    if (!c->sa16 && hipMalloc(&c->sa16, c->V * sizeof(double)) != hipSuccess)
        return fail(JW_ERR_ALLOC, "coupling table allocation failed");
    if (c->sa16_gen == c->sa_gen) return JW_OK;
    HIP_CHECK(jw::block16(c, c->sa16, c->sa));
    c->sa16_gen = c->sa_gen;
    return JW_OK;
}

// noise source for the next generator call; advances the Philox step afterwards
jw::NoiseSrc take_noise(jw_ctx* c, int per_site, uint32_t tag, bool* ok) {
    jw::NoiseSrc ns{nullptr, c->seed, c->traj, c->step, tag};
    *ok = true;
    if (c->inj_per_site) {
        if (c->inj_per_site != per_site) { *ok = false; return ns; }
        ns.inj = c->inj;
        c->inj_per_site = 0;
    } else {
        c->step += 1;
    }
    return ns;
}

// momentum-space noise runs on the fused passes (64 <= N <= 8192) with a
// coupling that does not scale the noise per site, and replaces the Philox
// draws (an injection is position-space noise)
int check_kspace(const jw_ctx* c, int coupling, bool fused) {
    if (!c->noise_kspace) return JW_OK;
    if (!fused)
        return fail(JW_ERR_UNSUPPORTED, "momentum-space noise needs the fused passes (64 <= N <= 8192, fused mode on)");
    if (coupling == JW_NOISE_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "momentum-space noise: NOISE coupling scales the noise per site");
    if (c->inj_per_site) return fail(JW_ERR_STATE, "momentum-space noise: injected normals are position-space noise");
    return JW_OK;
}

int check_field(const jw_field* f, const jw_ctx* c, int t, const char* name) {
    if (!f) return fail(JW_ERR_ARG, "%s is null", name);
    if (c && f->ctx != c) return fail(JW_ERR_ARG, "%s belongs to another context", name);
    if (t && f->t != t) return fail(JW_ERR_ARG, "%s has %d components, expected %d", name, f->t, t);
    return JW_OK;
}

// the fused step's intermediate layout when none is selected.  Measured at
// N = 2048 (steps/s, FP64 / FP32 storage, correlated / fixed coupling):
// row-major 382 / 568, 418 / 653; 4-row blocks 363 / 474, 415 / 536; 8-row
// blocks 345 / 507, 407 / 559.  Row-major everywhere.
// the column passes of layouts 1, 4 and 8 read the 16-row-blocked sqrt(alpha) table
bool uses_sa16(int layout) { return layout == 1 || layout == 4 || layout == 8 || layout >= 16; }

// N = 8192: 16-byte strips (the only ones that fit LDS); N = 4096: 16-byte
// strips in FP32 (measured 143 against 125 steps/s for the X + Y column tiles,
// correlated, momentum-space noise), the X + Y tiles in FP64 (89 against 85);
// otherwise row-major X + Y column tiles
int default_layout(int N, int corr, int fp32) {
    (void)corr;
    return N == 8192 || (N == 4096 && fp32) ? 32 : 0;
}

// nsteps generator-basis steps: the fused passes or the unfused kernel sequence
int evolve_generator_basis(jw_ctx* c, jw_field* U, int nsteps, bool fused, bool large, jw::cplx* Xi, jw::cplx* Vv,
                           jw::cplx* ZA, jw::cplx* ZB, const double* sa_pos, const double* sa_k, int kc, int corr,
                           double s, double ss) {
    const size_t V = c->V;
    for (int it = 0; it < nsteps; it++) {
        bool ok;
        jw::NoiseSrc ns = take_noise(c, 16, c->noise_kspace ? JW_TAG_KSPACE : JW_TAG_LANGEVIN, &ok);
        if (!ok) return fail(JW_ERR_STATE, "pending injection has %d normals per site, need 16", c->inj_per_site);
        if (fused || large) {
            const jw::StepArgs sa{ns, sa_pos, sa_k, kc == 0, corr, ss, 0, 1, it == 0, it < nsteps - 1, c->noise_kspace};
            if (fused) HIP_CHECK(jw::step_fused(c, U->d, Xi, Vv, ZA, ZB, sa));
            else HIP_CHECK(jw::step_large(c, U->d, Xi, Vv, ZA, ZB, sa));
            continue;
        }
        HIP_CHECK(jw::step_unfused(c, U->d, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, kc, corr, ss));
    }
    // the fused and large-lattice passes keep rows 0 and 1 of U current
    if ((fused || large) && nsteps > 0) HIP_CHECK(jw::u_rebuild(c, U->d, c->stream));
    return JW_OK;
}


// The generator-basis evolution of jw_evolve for the SIN kernel on the 24
// intermediate planes at Xi (xi 8, V 8, A 4, B 4).  tabs == nullptr: the
// context's cached coupling tables (allocated on first use); otherwise two
// V-double arrays into which this call builds sqrt(alpha) and, for the blocked
// layouts, its 16-row-blocked copy (nothing allocated).
int evolve_sin(jw_ctx* c, jw_field* U, int coupling, double step, const jw_alpha* a, int nsteps, cplx* Xi,
               double* tabs) {
    const size_t V = c->V;
    cplx* Vv = Xi + 8 * V;
    cplx* ZA = Vv + 8 * V;
    cplx* ZB = ZA + 4 * V;
    const jw::Alpha al = to_alpha(a);
    const double* sa_pos = nullptr;
    const double* sa_k = nullptr;
    if (coupling == JW_NOISE_COUPLING_CONSTANT || coupling == JW_SQRT_COUPLING_CONSTANT ||
        coupling == JW_CORRELATED_NOISE_COUPLING) {
        const int variant = coupling == JW_NOISE_COUPLING_CONSTANT ? 1 : 0;
        const double* sa;
        if (tabs) {
            if (int e = check_alpha(al)) return e;
            HIP_CHECK(jw::build_sqrt_alpha(c, tabs, variant, al));
            sa = tabs;
        } else {
            if (int e = ensure_sa(c, variant, al)) return e;
            sa = c->sa;
        }
        if (variant == 1) sa_pos = sa;
        else sa_k = sa;
    }
    const int kc = coupling == JW_SQRT_COUPLING_CONSTANT ? 0 : 1;
    const int corr = coupling == JW_CORRELATED_NOISE_COUPLING;
    const double s = 1.0 / c->N, ss = sqrt(step);
    const bool fused = jw::fused_supported(c->N) && !c->unfused;
    const bool large = jw::large_supported(c->N) && !c->unfused;
    if (int e = check_kspace(c, coupling, fused || large)) return e;
    const int saved_layout = c->layout;
    if (c->layout < 0) c->layout = default_layout(c->N, corr, 0);
    if ((fused || large) && sa_k && uses_sa16(c->layout)) {
        if (tabs) {
            const hipError_t he = jw::block16(c, tabs + V, tabs);
            if (he != hipSuccess) {
                c->layout = saved_layout;
                return fail(JW_ERR_HIP, "coupling table: %s", hipGetErrorString(he));
            }
            sa_k = tabs + V;
        } else {
            if (int e = ensure_sa16(c)) {
                c->layout = saved_layout;
                return e;
            }
            sa_k = c->sa16;
        }
    }
    const int rc = evolve_generator_basis(c, U, nsteps, fused, large, Xi, Vv, ZA, ZB, sa_pos, sa_k, kc, corr, s, ss);
    c->layout = saved_layout;
    return rc;
}

// ---------------------------------------------------------------- allocation-free evolution
// Workspace of jw_evolve_noalloc: the 24 intermediate planes (384 V bytes) and
// two coupling tables of V doubles (16 V bytes), for N <= JW_NOALLOC_MAX_N.
// One copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: each device's copy is reached through the
// address hipGetSymbolAddress gives for it (scratch_setup, at context
// creation, with that device current).
constexpr size_t kScratchBytes = (size_t)400 * JW_NOALLOC_MAX_N * JW_NOALLOC_MAX_N;
static_assert(kScratchBytes == JW_NOALLOC_SCRATCH_BYTES, "scratch size");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// once per device, with `device` current: the device address of its scratch copy
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(JW_ERR_UNSUPPORTED, "device %d: at most %d devices", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return JW_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(JW_ERR_HIP, "device %d: address of the allocation-free workspace: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return JW_OK;
}

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// ---------------------------------------------------------------- batches
// One size group of a batch: `count` trajectories of one N, state arrays of
// `count` x (9, 8, 8, 4, 4) planes (U, xi, V, A, B), element type by precision.
struct BatchGroup {
    jw_ctx* ctx = nullptr;     // tables, coupling table and stream for this N
    int N = 0, count = 0, first = 0;
    void* U = nullptr;
    void* Xi = nullptr;
    void* Vv = nullptr;
    void* ZA = nullptr;
    void* ZB = nullptr;
    int layout = -1;           // -1: tune on first use
    int fft_mfma = 0;          // FP32 groups: radix-16 column stages on the matrix cores (tuned)
    hipEvent_t join = nullptr;
};

// key of the captured step graph: the arguments baked into it
struct BatchKey {
    int coupling, reunit_every;
    double step;
    jw::Alpha alpha;
    bool operator==(const BatchKey& o) const {
        return coupling == o.coupling && reunit_every == o.reunit_every && step == o.step &&
               memcmp(&alpha, &o.alpha, sizeof alpha) == 0;
    }
};

}  // namespace

struct jw_batch {
    int device = 0, fp32 = 0, ntraj = 0;
    int kspace = 0;                  // momentum-space noise (jw_batch_set_noise_space)
    int launch = -1;                 // 0: per group, 1: one launch per kernel class, -1: tune
    std::vector<BatchGroup> g;
    hipStream_t stream = nullptr;    // origin of the step graph; groups fork from it
    hipEvent_t fork = nullptr;
    uint32_t* d_step = nullptr;      // Philox step counter on the device
    uint32_t seed = 0, traj0 = 0, step = 0;
    hipGraphExec_t graph = nullptr;  // reunit_every steps (10 without projection)
    int graph_steps = 0;
    bool have_key = false;
    BatchKey key{};
};

namespace {

size_t elem_size(const jw_batch* b) { return b->fp32 ? sizeof(jw::cplxf) : sizeof(cplx); }
// U planes per trajectory: the fused passes hold rows 0 and 1
int u_planes(const jw_batch* b) {
    (void)b;
    return 6;
}

void batch_free(jw_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    for (auto& g : b->g) {
        for (void* p : {g.U, g.Xi, g.Vv, g.ZA, g.ZB})
            if (p) hipFree(p);
        if (g.join) hipEventDestroy(g.join);
        if (g.ctx) jw_destroy(g.ctx);
    }
    if (b->fork) hipEventDestroy(b->fork);
    if (b->d_step) hipFree(b->d_step);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

// locate trajectory t: its group and index inside the group
BatchGroup* batch_find(jw_batch* b, int t, int* k) {
    for (auto& g : b->g)
        if (t >= g.first && t < g.first + g.count) {
            *k = t - g.first;
            return &g;
        }
    return nullptr;
}

// coupling tables of one group for this evolution
int batch_tables(BatchGroup& g, int coupling, const jw::Alpha& al, const double** sa_pos, const double** sa_k) {
    jw_ctx* c = g.ctx;
    *sa_pos = *sa_k = nullptr;
    if (coupling == JW_NOISE_COUPLING_CONSTANT) {
        if (int e = ensure_sa(c, 1, al)) return e;
        *sa_pos = c->sa;
    } else if (coupling == JW_SQRT_COUPLING_CONSTANT || coupling == JW_CORRELATED_NOISE_COUPLING) {
        if (int e = ensure_sa(c, 0, al)) return e;
        if (uses_sa16(c->layout)) {
            if (int e = ensure_sa16(c)) return e;
            *sa_k = c->sa16;
        } else {
            *sa_k = c->sa;
        }
    }
    return JW_OK;
}

// one step of group g on its own stream, noise step from the device counter
hipError_t batch_group_step(jw_batch* b, BatchGroup& g, int coupling, double step, const double* sa_pos,
                            const double* sa_k, int reunit, int count, void* U, int first_noise, int next_noise) {
    jw::NoiseSrc ns{nullptr, b->seed, b->traj0 + (uint32_t)g.first, 0u, b->kspace ? JW_TAG_KSPACE : JW_TAG_LANGEVIN,
                    b->d_step};
    const jw::StepArgs a{ns,      sa_pos,      sa_k,  coupling == JW_SQRT_COUPLING_CONSTANT,
                         coupling == JW_CORRELATED_NOISE_COUPLING, sqrt(step), reunit, count, first_noise, next_noise,
                         b->kspace};
    if (b->fp32)
        return jw::step_fused_f32(g.ctx, (jw::cplxf*)U, (jw::cplxf*)g.Xi, (jw::cplxf*)g.Vv, (jw::cplxf*)g.ZA,
                                  (jw::cplxf*)g.ZB, a);
    return jw::step_fused(g.ctx, (cplx*)U, (cplx*)g.Xi, (cplx*)g.Vv, (cplx*)g.ZA, (cplx*)g.ZB, a);
}

// one step of every group, forked from and joined back to b->stream, then the
// step counter; capturable.  first/next: as StepArgs::first_noise/next_noise.
// Each group stream waits on the fork before anything is enqueued on it, the
// coupling-table check included (the tables are built before a capture, so
// inside one nothing is rebuilt).
//
// launch 1: all groups in one launch per kernel class on b->stream (their
// planes row-major); Uv: the groups' U arrays (nullptr: their own); inc: the
// step counter advances (not while tuning).
int batch_step_all(jw_batch* b, int coupling, double step, const jw::Alpha& al, int reunit, int first, int next,
                   void* const* Uv = nullptr, bool inc = true, int launch = -1) {
    if (launch < 0) launch = b->launch;
    if (launch == 1) {
        const int ng = (int)b->g.size();
        if (ng > JW_MAX_GROUPS) return fail(JW_ERR_UNSUPPORTED, "one launch per kernel class holds %d groups", JW_MAX_GROUPS);
        jw::GroupArgs ga[JW_MAX_GROUPS];
        for (int gi = 0; gi < ng; gi++) {
            BatchGroup& g = b->g[gi];
            const double *sa_pos, *sa_k;
            if (int e = batch_tables(g, coupling, al, &sa_pos, &sa_k)) return e;
            const jw::NoiseSrc ns{nullptr, b->seed, b->traj0 + (uint32_t)g.first, 0u,
                                  b->kspace ? JW_TAG_KSPACE : JW_TAG_LANGEVIN, b->d_step};
            ga[gi] = jw::GroupArgs{g.N,    g.count, Uv ? Uv[gi] : g.U, g.Xi,         g.Vv,     g.ZA,     g.ZB,
                                   g.ctx->V, ns,    sa_pos,           sa_k,         g.ctx->q4, g.ctx->sb,
                                   b->fp32 ? (const void*)g.ctx->twf : (const void*)g.ctx->tw};
        }
        const jw::StepArgs a{ga[0].ns, nullptr, nullptr, coupling == JW_SQRT_COUPLING_CONSTANT,
                             coupling == JW_CORRELATED_NOISE_COUPLING, sqrt(step), reunit, 1, 1, 0, b->kspace};
        HIP_CHECK(b->fp32 ? jw::step_multi_f32(b->stream, ga, ng, a) : jw::step_multi(b->stream, ga, ng, a));
    } else {
        HIP_CHECK(hipEventRecord(b->fork, b->stream));
        for (size_t gi = 0; gi < b->g.size(); gi++) {
            BatchGroup& g = b->g[gi];
            const double *sa_pos, *sa_k;
            HIP_CHECK(hipStreamWaitEvent(g.ctx->stream, b->fork, 0));
            if (int e = batch_tables(g, coupling, al, &sa_pos, &sa_k)) return e;
            HIP_CHECK(batch_group_step(b, g, coupling, step, sa_pos, sa_k, reunit, g.count, Uv ? Uv[gi] : g.U, first,
                                       next));
            HIP_CHECK(hipEventRecord(g.join, g.ctx->stream));
            HIP_CHECK(hipStreamWaitEvent(b->stream, g.join, 0));
        }
    }
    if (inc) HIP_CHECK(jw::step_inc(b->stream, b->d_step));
    return JW_OK;
}

// the row-major planes and FP32 butterflies that one launch per kernel class runs
void batch_rowmajor(jw_batch* b) {
    for (auto& g : b->g) {
        g.layout = 0;
        g.ctx->layout = 0;
        g.fft_mfma = 0;
        g.ctx->fft_mfma = 0;
    }
}

// The launch mode: time 3 steps (after one untimed) of the whole batch with the
// tuned per-group launches and with one launch per kernel class (row-major
// planes), on copies of every trajectory with a scratch noise key; keep the
// faster.  The trajectories and the step counter are left as they were.
int batch_tune_launch(jw_batch* b, int coupling, double step, const jw::Alpha& al) {
    const int ng = (int)b->g.size();
    if (ng > JW_MAX_GROUPS) {
        b->launch = 0;
        return JW_OK;
    }
    std::vector<void*> uc(ng, nullptr);
    std::vector<int> lay(ng), mf(ng);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = JW_OK;
    for (auto& g : b->g) HIP_CHECK(hipStreamSynchronize(g.ctx->stream));
    for (int gi = 0; gi < ng && rc == JW_OK; gi++) {
        const size_t bytes = (size_t)b->g[gi].count * u_planes(b) * b->g[gi].ctx->V * elem_size(b);
        if (hipMalloc(&uc[gi], bytes) != hipSuccess ||
            hipMemcpyAsync(uc[gi], b->g[gi].U, bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
            rc = fail(JW_ERR_ALLOC, "launch tuning buffer allocation failed");
        lay[gi] = b->g[gi].layout;
        mf[gi] = b->g[gi].fft_mfma;
    }
    if (rc == JW_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(JW_ERR_HIP, "launch tuning events");
    const uint32_t saved_traj0 = b->traj0;
    b->traj0 = 0xFFFFF000u;      // scratch noise key
    float ms[2] = {1e30f, 1e30f};
    for (int mode = 0; mode < 2 && rc == JW_OK; mode++) {
        if (mode == 1) batch_rowmajor(b);
        bool ok = batch_step_all(b, coupling, step, al, 0, 1, 0, uc.data(), false, mode) == JW_OK;
        ok = ok && hipEventRecord(e0, b->stream) == hipSuccess;
        for (int it = 0; it < 3 && ok; it++) ok = batch_step_all(b, coupling, step, al, 0, 1, 0, uc.data(), false, mode) == JW_OK;
        ok = ok && hipEventRecord(e1, b->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms[mode], e0, e1) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            ms[mode] = 1e30f;
        }
    }
    b->traj0 = saved_traj0;
    hipStreamSynchronize(b->stream);
    for (auto& g : b->g) hipStreamSynchronize(g.ctx->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    for (void* p : uc)
        if (p) hipFree(p);
    if (rc != JW_OK) return rc;
    if (ms[0] >= 1e30f && ms[1] >= 1e30f) return fail(JW_ERR_HIP, "launch tuning: neither launch mode ran");
    b->launch = ms[1] < ms[0] ? 1 : 0;
    if (b->launch == 0)      // back to the tuned per-group layouts
        for (int gi = 0; gi < ng; gi++) {
            b->g[gi].layout = b->g[gi].ctx->layout = lay[gi];
            b->g[gi].fft_mfma = b->g[gi].ctx->fft_mfma = mf[gi];
        }
    return JW_OK;
}

// The tuner: for each group with no layout yet, time 3 steps (after one
// untimed) of each intermediate layout (and, in FP32, of the row-major layout
// with the column transforms' radix-16 stages on the matrix cores) on a copy
// of its U with a scratch noise key, and keep the fastest.  A layout whose tables or steps fail is skipped;
// the tuning fails only when no layout ran.  The trajectories and the step
// counter are left as they were.
int batch_tune(jw_batch* b, int coupling, double step, const jw::Alpha& al) {
    for (size_t gi = 0; gi < b->g.size(); gi++) {
        BatchGroup& g = b->g[gi];
        if (g.layout >= 0) continue;
        const size_t ubytes = (size_t)g.count * u_planes(b) * g.ctx->V * elem_size(b);
        void* ucopy = nullptr;
        if (hipMalloc(&ucopy, ubytes) != hipSuccess) return fail(JW_ERR_ALLOC, "tuning buffer allocation failed");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipMemcpyAsync(ucopy, g.U, ubytes, hipMemcpyDeviceToDevice, g.ctx->stream) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            if (e0) hipEventDestroy(e0);
            hipStreamSynchronize(g.ctx->stream);
            hipFree(ucopy);
            return fail(JW_ERR_HIP, "tuning set-up failed for group %zu", gi);
        }
        const uint32_t saved_traj0 = b->traj0;
        b->traj0 = 0xFFFFF000u;      // scratch noise key
        float best = 1e30f;
        int best_layout = -1, best_mfma = 0;
        // candidates (layout, matrix-core stages); the stages exist for the FP32 row-major passes
        static const int cand[6][2] = {{0, 0}, {2, 0}, {4, 0}, {8, 0}, {16, 0}, {0, 1}};
        for (const auto& cd : cand) {
            const int lay = cd[0];
            if (cd[1] && !b->fp32) continue;
            g.ctx->layout = lay;
            g.ctx->fft_mfma = cd[1];
            const double *sa_pos, *sa_k;
            if (batch_tables(g, coupling, al, &sa_pos, &sa_k) != JW_OK) continue;
            bool ok = batch_group_step(b, g, coupling, step, sa_pos, sa_k, 0, g.count, ucopy, 1, 0) == hipSuccess;
            ok = ok && hipEventRecord(e0, g.ctx->stream) == hipSuccess;
            for (int it = 0; it < 3 && ok; it++)
                ok = batch_group_step(b, g, coupling, step, sa_pos, sa_k, 0, g.count, ucopy, 1, 0) == hipSuccess;
            ok = ok && hipEventRecord(e1, g.ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0.f;
            ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (!ok) {
                (void)hipGetLastError();
                continue;
            }
            if (ms < best) {
                best = ms;
                best_layout = lay;
                best_mfma = cd[1];
            }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        b->traj0 = saved_traj0;
        hipStreamSynchronize(g.ctx->stream);
        hipFree(ucopy);
        if (best_layout < 0) {
            g.ctx->layout = -1;
            g.ctx->fft_mfma = 0;
            return fail(JW_ERR_HIP, "tuning: no intermediate layout ran for group %zu", gi);
        }
        g.layout = best_layout;
        g.ctx->layout = best_layout;
        g.fft_mfma = best_mfma;
        g.ctx->fft_mfma = best_mfma;
    }
    return JW_OK;
}

}  // namespace

extern "C" {

const char* jw_last_error(void) { return g_err; }
const char* jw_version(void) { return "jimwlk-mi355x 0.1 (gfx950, FP64)"; }

int jw_create(jw_ctx** out, int N, int device) {
    if (!out) return fail(JW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!valid_N(N)) return fail(JW_ERR_ARG, "N = %d: must be a power of two in [4, 8192]", N);
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(JW_ERR_ARG, "device %d of %d", device, ndev);
    HIP_CHECK(hipSetDevice(device));
    if (int e = scratch_setup(device)) return e;
    jw_ctx* c = new jw_ctx();
    memset(c, 0, sizeof *c);
    c->N = N;
    c->V = (size_t)N * N;
    c->device = device;
    c->sa_variant = -1;
    c->layout = JW_DEFAULT_LAYOUT;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(JW_ERR_HIP, "stream creation failed");
    }
    // twiddles exp(-2 pi i m/N) rounded from long double (FP64 and FP32)
    std::vector<cplx> tw(N);
    std::vector<float2> twf(N);
    for (int m = 0; m < N; m++) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)N;
        tw[m] = make_double2((double)cosl(ang), (double)sinl(ang));
        twf[m] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    // momenta.h:90-125 per axis: 4 pow(sin(pi k/N), 2) and sin(2 pi k/N)
    std::vector<double> q4(N), sb(N);
    for (int k = 0; k < N; k++) {
        q4[k] = 4.0 * pow(sin(M_PI * k / (1.0 * N)), 2.0);
        sb[k] = sin(2.0 * M_PI * k / (1.0 * N));
    }
    if (hipMalloc(&c->tw, N * sizeof(cplx)) != hipSuccess || hipMalloc(&c->q4, N * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->sb, N * sizeof(double)) != hipSuccess || hipMalloc(&c->twf, N * sizeof(float2)) != hipSuccess) {
        jw_destroy(c);
        return fail(JW_ERR_ALLOC, "table allocation failed");
    }
    hipMemcpy(c->tw, tw.data(), N * sizeof(cplx), hipMemcpyHostToDevice);
    hipMemcpy(c->twf, twf.data(), N * sizeof(float2), hipMemcpyHostToDevice);
    hipMemcpy(c->q4, q4.data(), N * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(c->sb, sb.data(), N * sizeof(double), hipMemcpyHostToDevice);
    *out = c;
    return JW_OK;
}

void jw_destroy(jw_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->tw) hipFree(c->tw);
    if (c->twf) hipFree(c->twf);
    if (c->q4) hipFree(c->q4);
    if (c->sb) hipFree(c->sb);
    if (c->sa) hipFree(c->sa);
    if (c->sa16) hipFree(c->sa16);
    if (c->inj) hipFree(c->inj);
    if (c->ws) hipFree(c->ws);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int jw_sync(jw_ctx* c) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return JW_OK;
}

int jw_lattice_size(const jw_ctx* c) { return c ? c->N : 0; }

// ---------------------------------------------------------------- fields
int jw_field_create(jw_ctx* c, int t, jw_field** out) {
    if (!c || !out) return fail(JW_ERR_ARG, "null argument");
    if (t != 1 && t != 9) return fail(JW_ERR_ARG, "t = %d: fields have 1 or 9 components", t);
    jw_field* f = new jw_field{c, t, nullptr};
    if (hipMalloc(&f->d, t * c->V * sizeof(cplx)) != hipSuccess) {
        delete f;
        return fail(JW_ERR_ALLOC, "field allocation (%zu MB) failed", t * c->V * sizeof(cplx) >> 20);
    }
    HIP_CHECK(hipMemsetAsync(f->d, 0, t * c->V * sizeof(cplx), c->stream));
    *out = f;
    return JW_OK;
}

void jw_field_destroy(jw_field* f) {
    if (!f) return;
    hipStreamSynchronize(f->ctx->stream);
    hipFree(f->d);
    delete f;
}

int jw_field_upload(jw_field* f, const double* h) {
    if (!f || !h) return fail(JW_ERR_ARG, "null argument");
    jw_ctx* c = f->ctx;
    const size_t V = c->V;
    std::vector<cplx> soa(f->t * V);
    for (size_t i = 0; i < V; i++)
        for (int k = 0; k < f->t; k++) soa[k * V + i] = make_double2(h[2 * (i * f->t + k)], h[2 * (i * f->t + k) + 1]);
    HIP_CHECK(hipMemcpyAsync(f->d, soa.data(), soa.size() * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return JW_OK;
}

int jw_field_download(const jw_field* f, double* h) {
    if (!f || !h) return fail(JW_ERR_ARG, "null argument");
    jw_ctx* c = f->ctx;
    const size_t V = c->V;
    std::vector<cplx> soa(f->t * V);
    HIP_CHECK(hipMemcpyAsync(soa.data(), f->d, soa.size() * sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < V; i++)
        for (int k = 0; k < f->t; k++) {
            h[2 * (i * f->t + k)] = soa[k * V + i].x;
            h[2 * (i * f->t + k) + 1] = soa[k * V + i].y;
        }
    return JW_OK;
}

int jw_field_copy(jw_field* dst, const jw_field* src) {
    if (!dst || !src || dst->ctx != src->ctx || dst->t != src->t) return fail(JW_ERR_ARG, "incompatible fields");
    if (dst == src) return JW_OK;
    HIP_CHECK(hipMemcpyAsync(dst->d, src->d, dst->t * dst->ctx->V * sizeof(cplx), hipMemcpyDeviceToDevice,
                             dst->ctx->stream));
    return JW_OK;
}

int jw_set_to_unit(jw_field* f) {
    if (int e = check_field(f, nullptr, 9, "field")) return e;
    HIP_CHECK(jw::set_unit(f->ctx, f->d));
    return JW_OK;
}

int jw_set_to_zero(jw_field* f) {
    if (!f) return fail(JW_ERR_ARG, "null field");
    HIP_CHECK(hipMemsetAsync(f->d, 0, f->t * f->ctx->V * sizeof(cplx), f->ctx->stream));
    return JW_OK;
}

// ---------------------------------------------------------------- noise
int jw_noise_inject(jw_ctx* c, const double* h, int per_site) {
    if (!c || !h) return fail(JW_ERR_ARG, "null argument");
    if (per_site != 16 && per_site != 8) return fail(JW_ERR_ARG, "per_site = %d: 16 (Langevin) or 8 (MV)", per_site);
    if (!c->inj && hipMalloc(&c->inj, 16 * c->V * sizeof(double)) != hipSuccess)
        return fail(JW_ERR_ALLOC, "noise buffer allocation failed");
    HIP_CHECK(hipMemcpyAsync(c->inj, h, per_site * c->V * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->inj_per_site = per_site;
    return JW_OK;
}

int jw_noise_philox(jw_ctx* c, uint32_t seed, uint32_t traj, uint32_t step) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    c->seed = seed;
    c->traj = traj;
    c->step = step;
    c->inj_per_site = 0;
    return JW_OK;
}

int jw_noise_get_step(const jw_ctx* c, uint32_t* step) {
    if (!c || !step) return fail(JW_ERR_ARG, "null argument");
    *step = c->step;
    return JW_OK;
}

// ---------------------------------------------------------------- checkpoints
// File: 8-byte magic "JWCKPT01", then uint32 N, seed, traj, step (Philox
// state), uint32 t = 9, uint32 0, then 18 V doubles of U in the reference's AoS
// order (site x N + y, entry 3 row + col, re and im).
static const char kCkptMagic[8] = {'J', 'W', 'C', 'K', 'P', 'T', '0', '1'};

int jw_checkpoint_write(jw_ctx* c, const jw_field* U, const char* path) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (!path) return fail(JW_ERR_ARG, "null path");
    if (c->inj_per_site) return fail(JW_ERR_STATE, "a noise injection is pending: the stream state is not Philox");
    std::vector<double> h(18 * c->V);
    if (int e = jw_field_download(U, h.data())) return e;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(JW_ERR_ARG, "checkpoint %s: cannot open for writing", path);
    const uint32_t hdr[6] = {(uint32_t)c->N, c->seed, c->traj, c->step, 9u, 0u};
    const bool ok = fwrite(kCkptMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof hdr, 1, f) == 1 &&
                    fwrite(h.data(), sizeof(double), h.size(), f) == h.size();
    if (fclose(f) != 0 || !ok) return fail(JW_ERR_ARG, "checkpoint %s: write failed", path);
    return JW_OK;
}

int jw_checkpoint_read(jw_ctx* c, jw_field* U, const char* path) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (!path) return fail(JW_ERR_ARG, "null path");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(JW_ERR_ARG, "checkpoint %s: cannot open", path);
    char magic[8];
    uint32_t hdr[6];
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kCkptMagic, 8) == 0 && fread(hdr, sizeof hdr, 1, f) == 1;
    if (!ok) {
        fclose(f);
        return fail(JW_ERR_ARG, "checkpoint %s: not a checkpoint file", path);
    }
    if ((int)hdr[0] != c->N || hdr[4] != 9u) {
        fclose(f);
        return fail(JW_ERR_ARG, "checkpoint %s: lattice N = %u, context N = %d", path, hdr[0], c->N);
    }
    std::vector<double> h(18 * c->V);
    ok = fread(h.data(), sizeof(double), h.size(), f) == h.size();
    fclose(f);
    if (!ok) return fail(JW_ERR_ARG, "checkpoint %s: truncated", path);
    if (int e = jw_field_upload(U, h.data())) return e;
    return jw_noise_philox(c, hdr[1], hdr[2], hdr[3]);
}

// ---------------------------------------------------------------- per-function entry points
int jw_execute2D(jw_ctx* c, jw_field* f, int dir) {
    if (int e = check_field(f, c, 0, "field")) return e;
    HIP_CHECK(jw::fft2d(c, f->d, f->t, dir ? 1 : 0, 1.0 / c->N));
    return JW_OK;
}

static int generate_common(jw_ctx* c, jw_field* x, jw_field* y, const double* sa_pos) {
    bool ok;
    jw::NoiseSrc ns = take_noise(c, 16, JW_TAG_LANGEVIN, &ok);
    if (!ok) return fail(JW_ERR_STATE, "pending injection has %d normals per site, need 16", c->inj_per_site);
    HIP_CHECK(jw::generate_gaussian(c, x->d, y->d, ns, sa_pos));
    return JW_OK;
}

int jw_generate_gaussian(jw_ctx* c, jw_field* x, jw_field* y) {
    if (int e = check_field(x, c, 9, "xi_x")) return e;
    if (int e = check_field(y, c, 9, "xi_y")) return e;
    return generate_common(c, x, y, nullptr);
}

int jw_generate_gaussian_with_noise_coupling_constant(jw_ctx* c, jw_field* x, jw_field* y, const jw_alpha* a) {
    if (int e = check_field(x, c, 9, "xi_x")) return e;
    if (int e = check_field(y, c, 9, "xi_y")) return e;
    if (int e = ensure_sa(c, 1, to_alpha(a))) return e;
    return generate_common(c, x, y, c->sa);
}

int jw_prepare_A_local(jw_ctx* c, jw_field* A, const jw_field* x, const jw_field* y, int coupling, int kernel,
                       const jw_alpha* a) {
    if (int e = check_field(A, c, 9, "A")) return e;
    if (int e = check_field(x, c, 9, "xi_x")) return e;
    if (int e = check_field(y, c, 9, "xi_y")) return e;
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || (kernel != 0 && kernel != 1)) return fail(JW_ERR_ARG, "bad coupling/kernel");
    const double* sa = nullptr;
    if (coupling == JW_SQRT_COUPLING_CONSTANT && kernel == JW_SIN_KERNEL) {
        if (int e = ensure_sa(c, 0, to_alpha(a))) return e;
        sa = c->sa;
    }
    HIP_CHECK(jw::prepare_A_local(c, A->d, x->d, y->d, coupling == 0 ? 0 : 1, kernel, sa, to_alpha(a)));
    return JW_OK;
}

int jw_uxiulocal(jw_ctx* c, jw_field* vx, jw_field* vy, const jw_field* U, const jw_field* x, const jw_field* y) {
    if (int e = check_field(vx, c, 9, "v_x")) return e;
    if (int e = check_field(vy, c, 9, "v_y")) return e;
    if (int e = check_field(U, c, 9, "U")) return e;
    if (int e = check_field(x, c, 9, "xi_x")) return e;
    if (int e = check_field(y, c, 9, "xi_y")) return e;
    HIP_CHECK(jw::uxiulocal(c, vx->d, vy->d, U->d, x->d, y->d));
    return JW_OK;
}

int jw_update_uf(jw_ctx* c, jw_field* U, const jw_field* B, const jw_field* A, double step) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (int e = check_field(B, c, 9, "B")) return e;
    if (int e = check_field(A, c, 9, "A")) return e;
    if (!(step >= 0)) return fail(JW_ERR_ARG, "step must be >= 0");
    HIP_CHECK(jw::update_uf(c, U->d, B->d, A->d, step));
    return JW_OK;
}

int jw_set_mv_model(jw_ctx* c, jw_field* f, double g, double mu, int Ny) {
    if (int e = check_field(f, c, 9, "f")) return e;
    if (Ny < 1) return fail(JW_ERR_ARG, "Ny must be >= 1");
    bool ok;
    jw::NoiseSrc ns = take_noise(c, 8, JW_TAG_MV, &ok);
    if (!ok) return fail(JW_ERR_STATE, "pending injection has %d normals per site, need 8", c->inj_per_site);
    HIP_CHECK(jw::set_mv_model(c, f->d, ns, g * mu / sqrt((double)Ny)));
    return JW_OK;
}

int jw_solve_poisson(jw_ctx* c, jw_field* f, double mass, double g) {
    if (int e = check_field(f, c, 0, "f")) return e;
    HIP_CHECK(jw::solve_poisson(c, f->d, f->t, mass, g));
    return JW_OK;
}

// POSITION_EVOLUTION weights per displacement r = (rx, ry), index rx N + ry:
// SIN kernel from the positions.h:131-141 tables, LINEAR from the wrapped
// integer distances, SQRT coupling with field.h:2449-2456's position-space
// formula, else factor 1; zero where rrr <= 10e-9 (field.h:2460-2468)
static void position_weights(int N, int coupling, int kernel, double* wa, double* wb) {
    const double lambda = pow(15.0 * 15.0 / 6.0 / 6.0, 1.0 / 0.2);
    for (int rx = 0; rx < N; rx++)
        for (int ry = 0; ry < N; ry++) {
            double dx, dy, rrr;
            if (kernel == JW_SIN_KERNEL) {
                dx = 0.5 * N * sin(2.0 * M_PI * rx / (1.0 * N)) / M_PI;
                dy = 0.5 * N * sin(2.0 * M_PI * ry / (1.0 * N)) / M_PI;
                const double bx = N * sin(M_PI * rx / (1.0 * N)) / M_PI;
                const double by = N * sin(M_PI * ry / (1.0 * N)) / M_PI;
                rrr = pow(bx, 2.0) + pow(by, 2.0);
            } else {
                dx = rx >= N / 2 ? rx - N : rx;
                dy = ry >= N / 2 ? ry - N : ry;
                rrr = 1.0 * (dx * dx + dy * dy);
            }
            double cc = 1.0;
            if (coupling == JW_SQRT_COUPLING_CONSTANT)
                cc = sqrt(4.0 * M_PI /
                          ((11.0 - 2.0 * 3.0 / 3.0) * log(pow(lambda + 1.26 / pow(6.0 * 6.0 * rrr / N / N, 1.0 / 0.2), 0.2))));
            const size_t i = (size_t)rx * N + ry;
            wa[i] = rrr > 10e-9 ? cc * dx / rrr : 0.0;
            wb[i] = rrr > 10e-9 ? cc * dy / rrr : 0.0;
        }
}

int jw_prepare_A_and_B_local(jw_ctx* c, jw_field* A, jw_field* B, const jw_field* x, const jw_field* y,
                             const jw_field* U, int coupling, int kernel) {
    if (int e = check_field(A, c, 9, "A")) return e;
    if (int e = check_field(B, c, 9, "B")) return e;
    if (int e = check_field(x, c, 9, "xi_x")) return e;
    if (int e = check_field(y, c, 9, "xi_y")) return e;
    if (int e = check_field(U, c, 9, "U")) return e;
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || (kernel != 0 && kernel != 1)) return fail(JW_ERR_ARG, "bad coupling/kernel");
    if (c->N > JW_POSITION_MAX_N)
        return fail(JW_ERR_UNSUPPORTED, "N = %d: the O(V^2) position-space sums run up to N = %d", c->N,
                    JW_POSITION_MAX_N);
    if (A == B || A == x || A == y || A == U || B == x || B == y || B == U)
        return fail(JW_ERR_ARG, "A and B must be distinct from the inputs");
    // workspace: V_x (9 planes), V_y (9), the two weight tables (2 V doubles = 1 plane)
    if (int e = ensure_ws(c, 19)) return e;
    const size_t V = c->V;
    cplx* vx = c->ws;
    cplx* vy = vx + 9 * V;
    double* w = reinterpret_cast<double*>(vy + 9 * V);
    std::vector<double> hw(2 * V);
    position_weights(c->N, coupling, kernel, hw.data(), hw.data() + V);
    HIP_CHECK(hipMemcpyAsync(w, hw.data(), 2 * V * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(jw::uxiulocal(c, vx, vy, U->d, x->d, y->d));
    HIP_CHECK(jw::position_sums(c, A->d, B->d, x->d, y->d, vx, vy, w, w + V));
    HIP_CHECK(hipStreamSynchronize(c->stream));     // hw is released on return
    return JW_OK;
}

int jw_correlate_noise(jw_ctx* c, jw_field* f, const jw_alpha* a) {
    if (int e = check_field(f, c, 0, "field")) return e;
    if (int e = ensure_sa(c, 0, to_alpha(a))) return e;
    HIP_CHECK(jw::fft2d(c, f->d, f->t, 1, 1.0 / c->N));
    HIP_CHECK(jw::scale_by_table(c, f->d, f->t, c->sa));
    HIP_CHECK(jw::fft2d(c, f->d, f->t, 0, 1.0 / c->N));
    return JW_OK;
}

int jw_set_noise_space(jw_ctx* c, int space) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    if (space != JW_NOISE_POSITION && space != JW_NOISE_MOMENTUM) return fail(JW_ERR_ARG, "noise space %d", space);
    c->noise_kspace = space == JW_NOISE_MOMENTUM;
    return JW_OK;
}

int jw_set_mv_zero_mode(jw_ctx* c, int mode) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    if (mode != 0 && mode != 1) return fail(JW_ERR_ARG, "zero mode %d: 0 (reference) or 1 (projected out)", mode);
    c->mv_zero_mode = mode;
    return JW_OK;
}

int jw_mul_exp(jw_ctx* c, jw_field* U, const jw_field* f) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (int e = check_field(f, c, 9, "f")) return e;
    HIP_CHECK(jw::mul_exp(c, U->d, f->d));
    return JW_OK;
}

int jw_trace(jw_ctx* c, const jw_field* U, jw_field* corr) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (int e = check_field(corr, c, 1, "corr")) return e;
    HIP_CHECK(jw::trace(c, U->d, corr->d));
    return JW_OK;
}

int jw_average_and_symmetrize(jw_ctx* c, jw_field* corr) {
    if (int e = check_field(corr, c, 1, "corr")) return e;
    if (int e = ensure_ws(c, 24)) return e;
    HIP_CHECK(jw::average_and_symmetrize(c, corr->d, c->ws));
    HIP_CHECK(hipMemcpyAsync(corr->d, c->ws, c->V * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
    return JW_OK;
}

int jw_reduce(jw_ctx* c, const jw_field* corr, jw_field* sum, jw_field* err) {
    if (int e = check_field(corr, c, 1, "corr")) return e;
    if (int e = check_field(sum, c, 1, "sum")) return e;
    if (int e = check_field(err, c, 1, "err")) return e;
    HIP_CHECK(jw::reduce(c, corr->d, sum->d, err->d));
    return JW_OK;
}

// ---------------------------------------------------------------- evolution
// 9-component form of main_optimized.cpp:243-304 (used for the LINEAR kernel,
// whose Nyquist line is not odd in k and so does not preserve the real
// generator planes that the packed form relies on)
static int evolve_9comp(jw_ctx* c, jw_field* U, int coupling, int kernel, double step, const jw_alpha* a) {
    if (coupling == JW_CORRELATED_NOISE_COUPLING)
        return fail(JW_ERR_UNSUPPORTED, "LINEAR kernel with correlated noise is not supported");
    if (int e = ensure_ws(c, 54)) return e;
    cplx* xx = c->ws;
    cplx* xy = xx + 9 * c->V;
    cplx* A = xy + 9 * c->V;
    cplx* vx = A + 9 * c->V;
    cplx* vy = vx + 9 * c->V;
    cplx* B = vy + 9 * c->V;
    const double* sa_pos = nullptr;
    if (coupling == JW_NOISE_COUPLING_CONSTANT) {
        if (int e = ensure_sa(c, 1, to_alpha(a))) return e;
        sa_pos = c->sa;
    }
    bool ok;
    jw::NoiseSrc ns = take_noise(c, 16, JW_TAG_LANGEVIN, &ok);
    if (!ok) return fail(JW_ERR_STATE, "pending injection has %d normals per site, need 16", c->inj_per_site);
    const double s = 1.0 / c->N;
    HIP_CHECK(jw::generate_gaussian(c, xx, xy, ns, sa_pos));
    HIP_CHECK(jw::fft2d(c, xx, 18, 1, s));                       // xi_x, xi_y are adjacent
    const int kc = coupling == JW_SQRT_COUPLING_CONSTANT ? 0 : 1;
    HIP_CHECK(jw::prepare_A_local(c, A, xx, xy, kc, kernel, nullptr, to_alpha(a)));
    HIP_CHECK(jw::fft2d(c, A, 9, 0, s));
    HIP_CHECK(jw::fft2d(c, xx, 18, 0, s));
    HIP_CHECK(jw::uxiulocal(c, vx, vy, U->d, xx, xy));
    HIP_CHECK(jw::fft2d(c, vx, 18, 1, s));
    HIP_CHECK(jw::prepare_A_local(c, B, vx, vy, kc, kernel, nullptr, to_alpha(a)));
    HIP_CHECK(jw::fft2d(c, B, 9, 0, s));
    HIP_CHECK(jw::update_uf(c, U->d, B, A, step));
    return JW_OK;
}

int jw_evolve(jw_ctx* c, jw_field* U, int coupling, int kernel, double step, const jw_alpha* a, int nsteps) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || (kernel != 0 && kernel != 1)) return fail(JW_ERR_ARG, "bad coupling/kernel");
    if (!(step >= 0)) return fail(JW_ERR_ARG, "step must be >= 0");
    if (nsteps < 0) return fail(JW_ERR_ARG, "nsteps < 0");
    if (c->inj_per_site && nsteps > 1) return fail(JW_ERR_STATE, "an injection covers one step; nsteps = %d", nsteps);
    if (kernel == JW_LINEAR_KERNEL) {
        if (c->noise_kspace) return fail(JW_ERR_UNSUPPORTED, "momentum-space noise runs the SIN kernel");
        for (int s = 0; s < nsteps; s++)
            if (int e = evolve_9comp(c, U, coupling, kernel, step, a)) return e;
        return JW_OK;
    }
    if (int e = ensure_ws(c, 24)) return e;
    return evolve_sin(c, U, coupling, step, a, nsteps, c->ws, nullptr);
}

int jw_evolve_noalloc(jw_ctx* c, jw_field* U, int coupling, int kernel, double step, const jw_alpha* a,
                      int nsteps) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || (kernel != 0 && kernel != 1)) return fail(JW_ERR_ARG, "bad coupling/kernel");
    if (kernel == JW_LINEAR_KERNEL)
        return fail(JW_ERR_UNSUPPORTED, "the allocation-free evolution runs the SIN kernel");
    if (c->N > JW_NOALLOC_MAX_N)
        return fail(JW_ERR_ARG, "N = %d: the allocation-free evolution holds lattices up to N = %d", c->N,
                    JW_NOALLOC_MAX_N);
    if (!(step >= 0)) return fail(JW_ERR_ARG, "step must be >= 0");
    if (nsteps < 0) return fail(JW_ERR_ARG, "nsteps < 0");
    if (c->inj_per_site && nsteps > 1) return fail(JW_ERR_STATE, "an injection covers one step; nsteps = %d", nsteps);
    void* ws = scratch_of(c->device);
    if (!ws) return fail(JW_ERR_STATE, "device %d has no allocation-free workspace", c->device);
    cplx* Xi = static_cast<cplx*>(ws);
    double* tabs = reinterpret_cast<double*>(Xi + 24 * c->V);
    const int saved_timing = c->timing;      // timing events are not created here
    c->timing = 0;
    const int rc = evolve_sin(c, U, coupling, step, a, nsteps, Xi, tabs);
    c->timing = saved_timing;
    return rc;
}

int jw_evolve_mixed(jw_ctx* c, jw_field* U, int coupling, int kernel, double step, const jw_alpha* a, int nsteps,
                    int reunit_every) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || (kernel != 0 && kernel != 1)) return fail(JW_ERR_ARG, "bad coupling/kernel");
    if (kernel != JW_SIN_KERNEL) return fail(JW_ERR_UNSUPPORTED, "the mixed-precision form runs the SIN kernel");
    if (!(step >= 0)) return fail(JW_ERR_ARG, "step must be >= 0");
    if (nsteps < 0 || reunit_every < 0) return fail(JW_ERR_ARG, "nsteps or reunit_every < 0");
    if (c->inj_per_site && nsteps > 1) return fail(JW_ERR_STATE, "an injection covers one step; nsteps = %d", nsteps);
    if (nsteps == 0) return JW_OK;
    // FP32 state: U (9 planes of room; 6 used by the fused and large-lattice
    // passes) and the 24 intermediate planes, inside the workspace of 24 FP64
    // planes (48 FP32 planes)
    if (int e = ensure_ws(c, 24)) return e;
    const size_t V = c->V;
    jw::cplxf* U32 = reinterpret_cast<jw::cplxf*>(c->ws);
    jw::cplxf* Xi = U32 + 9 * V;
    jw::cplxf* Vv = Xi + 8 * V;
    jw::cplxf* ZA = Vv + 8 * V;
    jw::cplxf* ZB = ZA + 4 * V;
    const double* sa_pos = nullptr;
    const double* sa_k = nullptr;
    if (coupling == JW_NOISE_COUPLING_CONSTANT) {
        if (int e = ensure_sa(c, 1, to_alpha(a))) return e;
        sa_pos = c->sa;
    } else if (coupling == JW_SQRT_COUPLING_CONSTANT || coupling == JW_CORRELATED_NOISE_COUPLING) {
        if (int e = ensure_sa(c, 0, to_alpha(a))) return e;
        sa_k = c->sa;
    }
    const int corr = coupling == JW_CORRELATED_NOISE_COUPLING;
    const bool fused = jw::fused_supported(c->N) && !c->unfused;
    const bool large = jw::large_supported(c->N) && !c->unfused;
    if (int e = check_kspace(c, coupling, fused || large)) return e;
    const int saved_layout = c->layout;
    if (c->layout < 0) c->layout = default_layout(c->N, corr, 1);
    int rc = JW_OK;
    if ((fused || large) && sa_k && uses_sa16(c->layout)) {
        rc = ensure_sa16(c);
        sa_k = c->sa16;
    }
    // the fused and large-lattice passes hold rows 0 and 1 of U (row 2 rebuilt),
    // the unfused sequence all nine entries
    const int uplanes = fused || large ? 6 : 9;
    if (rc == JW_OK && jw::u_convert(c, U32, U->d, 1, uplanes) != hipSuccess) rc = fail(JW_ERR_HIP, "U conversion failed");
    for (int it = 0; rc == JW_OK && it < nsteps; it++) {
        bool ok;
        jw::NoiseSrc ns = take_noise(c, 16, c->noise_kspace ? JW_TAG_KSPACE : JW_TAG_LANGEVIN, &ok);
        if (!ok) {
            rc = fail(JW_ERR_STATE, "pending injection has %d normals per site, need 16", c->inj_per_site);
            break;
        }
        // projection every reunit_every steps and after the last one
        const int reunit = reunit_every > 0 && ((it + 1) % reunit_every == 0 || it == nsteps - 1);
        const jw::StepArgs sa{ns,      sa_pos,        sa_k,           coupling == JW_SQRT_COUPLING_CONSTANT,
                              corr,    sqrt(step),    reunit,         1,
                              it == 0, it < nsteps - 1, c->noise_kspace};
        const hipError_t he = fused ? jw::step_fused_f32(c, U32, Xi, Vv, ZA, ZB, sa)
                              : large ? jw::step_large_f32(c, U32, Xi, Vv, ZA, ZB, sa)
                                    : jw::step_unfused_f32(c, U32, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k,
                                                           coupling == JW_SQRT_COUPLING_CONSTANT ? 0 : 1, corr,
                                                           sqrt(step), reunit);
        if (he != hipSuccess) rc = fail(JW_ERR_HIP, "%s", hipGetErrorString(he));
    }
    if (rc == JW_OK && jw::u_convert(c, U32, U->d, 0, uplanes) != hipSuccess) rc = fail(JW_ERR_HIP, "U conversion failed");
    c->layout = saved_layout;
    return rc;
}

int jw_mv_init(jw_ctx* c, jw_field* U, int Ny, double g, double mu, double mass, const double* host) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (Ny < 1) return fail(JW_ERR_ARG, "Ny must be >= 1");
    if (int e = ensure_ws(c, 24)) return e;
    cplx* R = c->ws;
    HIP_CHECK(jw::set_unit(c, U->d));
    const double sigma = g * mu / sqrt((double)Ny);
    const double m = mass * pow(g, 2.0) * mu;
    for (int s = 0; s < Ny; s++) {
        jw::NoiseSrc ns{nullptr, c->seed, c->traj, (uint32_t)s, JW_TAG_MV};
        if (host) {
            if (int e = jw_noise_inject(c, host + (size_t)s * 8 * c->V, 8)) return e;
            ns.inj = c->inj;
            c->inj_per_site = 0;
        }
        HIP_CHECK(jw::mv_noise_packed(c, R, ns, sigma));
        HIP_CHECK(jw::fft2d(c, R, 4, 1, 1.0 / c->N));
        HIP_CHECK(jw::poisson_packed(c, R, m, g));
        HIP_CHECK(jw::fft2d(c, R, 4, 0, 1.0 / c->N));
        HIP_CHECK(jw::mul_exp_packed(c, U->d, R));
    }
    return JW_OK;
}

// ---------------------------------------------------------------- observables
int jw_measure_corr(jw_ctx* c, const jw_field* U, double* host_corr) {
    if (int e = check_field(U, c, 9, "U")) return e;
    if (!host_corr) return fail(JW_ERR_ARG, "null output");
    if (int e = ensure_ws(c, 24)) return e;
    const size_t V = c->V;
    cplx* Uk = c->ws;
    cplx* tr = Uk + 9 * V;
    cplx* sym = tr + V;
    HIP_CHECK(hipMemcpyAsync(Uk, U->d, 9 * V * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(jw::fft2d(c, Uk, 9, 1, 1.0 / c->N));
    HIP_CHECK(jw::trace(c, Uk, tr));
    HIP_CHECK(jw::average_and_symmetrize(c, tr, sym));
    std::vector<cplx> h(V);
    HIP_CHECK(hipMemcpyAsync(h.data(), sym, V * sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < V; i++) host_corr[i] = h[i].x;
    return JW_OK;
}

int jw_dipole_from_corr(jw_ctx* c, const double* host_corr, double* host_S) {
    if (!c || !host_corr || !host_S) return fail(JW_ERR_ARG, "null argument");
    if (int e = ensure_ws(c, 24)) return e;
    const size_t V = c->V;
    std::vector<cplx> h(V);
    for (size_t i = 0; i < V; i++) h[i] = make_double2(host_corr[i], 0.0);
    HIP_CHECK(hipMemcpyAsync(c->ws, h.data(), V * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(jw::fft2d(c, c->ws, 1, 0, 1.0));
    HIP_CHECK(hipMemcpyAsync(h.data(), c->ws, V * sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const double f = (double)V / 3.0;
    for (size_t i = 0; i < V; i++) host_S[i] = f * h[i].x;
    return JW_OK;
}

int jw_radial_bins(int N, const double* S, int nbins, double* rbar, double* Sb, double* cnt) {
    if (N < 1 || !S || nbins < 1 || !rbar || !Sb || !cnt) return fail(JW_ERR_ARG, "bad argument");
    std::vector<long double> sr(nbins, 0), ss(nbins, 0), sc(nbins, 0);
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++) {
            const int dx = x < N / 2 ? x : x - N, dy = y < N / 2 ? y : y - N;
            const double r = sqrt((double)(dx * dx + dy * dy));
            const int b = (int)floor(r);
            if (b >= nbins) continue;
            sr[b] += r;
            ss[b] += S[(size_t)x * N + y];
            sc[b] += 1;
        }
    for (int b = 0; b < nbins; b++) {
        cnt[b] = (double)sc[b];
        rbar[b] = sc[b] > 0 ? (double)(sr[b] / sc[b]) : 0.0;
        Sb[b] = sc[b] > 0 ? (double)(ss[b] / sc[b]) : 0.0;
    }
    return JW_OK;
}

int jw_saturation_scale(int nbins, const double* rbar, const double* Sb, const double* cnt, double* Qs) {
    if (nbins < 1 || !rbar || !Sb || !cnt || !Qs) return fail(JW_ERR_ARG, "bad argument");
    const double target = exp(-0.5);
    *Qs = 0.0;
    int prev = -1;
    for (int b = 0; b < nbins; b++) {
        if (cnt[b] <= 0) continue;
        if (Sb[b] <= target && prev >= 0) {
            const double t = (Sb[prev] - target) / (Sb[prev] - Sb[b]);
            const double Rs = rbar[prev] + t * (rbar[b] - rbar[prev]);
            *Qs = Rs > 0 ? sqrt(2.0) / Rs : 0.0;
            return JW_OK;
        }
        prev = b;
    }
    return JW_OK;
}

// ---------------------------------------------------------------- instrumentation
static const char* kNames[] = {"fft_rows", "fft_cols", "gen_noise", "kernel_mult", "uxiu",
                               "update", "mv", "observables", "reference_functions",
                               "row_noise_fwd", "col_kmult", "row_uxiu_fwd", "row_update"};

int jw_set_fused(jw_ctx* c, int mode) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    static const int layouts[9] = {0, JW_DEFAULT_LAYOUT, 0, 4, 8, 1, 2, 16, 32};
    if (mode < 0 || mode > 8) return fail(JW_ERR_ARG, "fused mode %d: expected 0..8", mode);
    // N = 4096 runs column strips of layouts 0 and 32, N = 8192 of layout 32;
    // layout 32 exists for those two sizes only
    const bool large = jw::large_supported(c->N);
    if (mode > 1 && ((large && !(mode == 8 || (c->N == 4096 && mode == 2))) || (!large && mode == 8)))
        return fail(JW_ERR_UNSUPPORTED, "fused mode %d is not available at N = %d", mode, c->N);
    c->unfused = mode ? 0 : 1;
    c->layout = layouts[mode];
    return JW_OK;
}

int jw_set_mixed_fft(jw_ctx* c, int mode) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    if (mode != 0 && mode != 1) return fail(JW_ERR_ARG, "mixed FFT mode %d: 0 (FP32 butterflies) or 1 (matrix cores)", mode);
    c->fft_mfma = mode;
    return JW_OK;
}

int jw_timing(jw_ctx* c, int enable) {
    if (!c) return fail(JW_ERR_ARG, "null context");
    c->timing = enable ? 1 : 0;
    for (int k = 0; k < 16; k++) { c->kernel_ms[k] = 0; c->kernel_calls[k] = 0; }
    return JW_OK;
}

int jw_kernel_time(const jw_ctx* c, int which, double* avg_ms, long* calls) {
    if (!c || which < 0 || which >= (int)(sizeof kNames / sizeof kNames[0]) || !avg_ms || !calls)
        return fail(JW_ERR_ARG, "bad argument");
    *calls = c->kernel_calls[which];
    *avg_ms = *calls ? c->kernel_ms[which] / *calls : 0.0;
    return JW_OK;
}

const char* jw_kernel_name(int which) {
    if (which < 0 || which >= (int)(sizeof kNames / sizeof kNames[0])) return "";
    return kNames[which];
}

int jw_calibration_copy(jw_ctx* c, int nplanes) {
    if (!c || nplanes < 1) return fail(JW_ERR_ARG, "bad argument");
    if (int e = ensure_ws(c, (size_t)2 * nplanes)) return e;
    const size_t n = (size_t)nplanes * c->V;
    HIP_CHECK(jw::calibration_copy(c, c->ws, c->ws + n, n));
    return JW_OK;
}

int jw_su3_exp_gen(jw_ctx* c, const double* h8, double s, double* out18, int n) {
    if (!c || !h8 || !out18 || n < 0) return fail(JW_ERR_ARG, "bad argument");
    if (n == 0) return JW_OK;
    double *dh = nullptr, *dout = nullptr;
    HIP_CHECK(hipMalloc(&dh, 8 * sizeof(double) * n));
    if (hipMalloc(&dout, 18 * sizeof(double) * n) != hipSuccess) {
        hipFree(dh);
        return fail(JW_ERR_ALLOC, "allocation failed");
    }
    hipMemcpyAsync(dh, h8, 8 * sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    hipError_t e = jw::exp_batch(c, dh, s, dout, n);
    hipMemcpyAsync(out18, dout, 18 * sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    hipFree(dh);
    hipFree(dout);
    if (e != hipSuccess) return fail(JW_ERR_HIP, "exp kernel: %s", hipGetErrorString(e));
    return JW_OK;
}


// ---------------------------------------------------------------- batches
int jw_batch_create(jw_batch** out, int device, int ngroups, const int* N, const int* count, int fp32) {
    if (!out || ngroups < 1 || ngroups > 64 || !N || !count) return fail(JW_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < ngroups; i++) {
        if (!valid_N(N[i]) || !jw::fused_supported(N[i]))
            return fail(JW_ERR_UNSUPPORTED, "batch group %d: N = %d (batches need 64 <= N <= 2048)", i, N[i]);
        if (count[i] < 1) return fail(JW_ERR_ARG, "batch group %d: count %d", i, count[i]);
    }
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(JW_ERR_ARG, "device %d of %d", device, ndev);
    HIP_CHECK(hipSetDevice(device));
    jw_batch* b = new jw_batch();
    b->device = device;
    b->fp32 = fp32 ? 1 : 0;
    b->g.resize(ngroups);
    int rc = JW_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&b->d_step, sizeof(uint32_t)) != hipSuccess)
        rc = fail(JW_ERR_HIP, "batch stream/event/counter creation failed");
    const size_t es = elem_size(b);
    for (int i = 0; rc == JW_OK && i < ngroups; i++) {
        BatchGroup& g = b->g[i];
        g.N = N[i];
        g.count = count[i];
        g.first = b->ntraj;
        b->ntraj += count[i];
        if ((rc = jw_create(&g.ctx, N[i], device))) break;
        g.ctx->layout = -1;
        const size_t V = g.ctx->V, n = (size_t)g.count * V * es;
        if (hipMalloc(&g.U, u_planes(b) * n) != hipSuccess || hipMalloc(&g.Xi, 8 * n) != hipSuccess ||
            hipMalloc(&g.Vv, 8 * n) != hipSuccess || hipMalloc(&g.ZA, 4 * n) != hipSuccess ||
            hipMalloc(&g.ZB, 4 * n) != hipSuccess) {
            rc = fail(JW_ERR_ALLOC, "batch group %d: %zu bytes of state", i, (24 + u_planes(b)) * n);
            break;
        }
        if (hipEventCreateWithFlags(&g.join, hipEventDisableTiming) != hipSuccess) rc = fail(JW_ERR_HIP, "event");
    }
    if (rc == JW_OK) rc = jw_batch_philox(b, 0, 0, 0);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return JW_OK;
}

void jw_batch_destroy(jw_batch* b) { batch_free(b); }

int jw_batch_size(const jw_batch* b) { return b ? b->ntraj : 0; }

int jw_batch_set(jw_batch* b, int t, const jw_field* U) {
    if (!b || !U) return fail(JW_ERR_ARG, "null argument");
    int k;
    BatchGroup* g = batch_find(b, t, &k);
    if (!g) return fail(JW_ERR_ARG, "trajectory %d of %d", t, b->ntraj);
    if (U->t != 9 || U->ctx->N != g->N) return fail(JW_ERR_ARG, "field is not a 9-component %d^2 lattice", g->N);
    const size_t V = g->ctx->V;
    HIP_CHECK(hipStreamSynchronize(U->ctx->stream));     // pending work on the source field
    HIP_CHECK(hipStreamSynchronize(b->stream));           // and on the batch
    if (b->fp32) {
        HIP_CHECK(jw::u_convert(g->ctx, (jw::cplxf*)g->U + (size_t)k * 6 * V, U->d, 1, 6));
    } else {
        HIP_CHECK(hipMemcpyAsync((cplx*)g->U + (size_t)k * 6 * V, U->d, 6 * V * sizeof(cplx),
                                 hipMemcpyDeviceToDevice, g->ctx->stream));
    }
    HIP_CHECK(hipStreamSynchronize(g->ctx->stream));
    return JW_OK;
}

int jw_batch_get(jw_batch* b, int t, jw_field* U) {
    if (!b || !U) return fail(JW_ERR_ARG, "null argument");
    int k;
    BatchGroup* g = batch_find(b, t, &k);
    if (!g) return fail(JW_ERR_ARG, "trajectory %d of %d", t, b->ntraj);
    if (U->t != 9 || U->ctx->N != g->N) return fail(JW_ERR_ARG, "field is not a 9-component %d^2 lattice", g->N);
    const size_t V = g->ctx->V;
    HIP_CHECK(hipStreamSynchronize(b->stream));
    HIP_CHECK(hipStreamSynchronize(U->ctx->stream));     // pending work on the destination field
    if (b->fp32) {
        HIP_CHECK(jw::u_convert(g->ctx, (jw::cplxf*)g->U + (size_t)k * 6 * V, U->d, 0, 6));
    } else {
        HIP_CHECK(hipMemcpyAsync(U->d, (cplx*)g->U + (size_t)k * 6 * V, 6 * V * sizeof(cplx),
                                 hipMemcpyDeviceToDevice, g->ctx->stream));
        HIP_CHECK(jw::u_rebuild(g->ctx, U->d, g->ctx->stream));
    }
    HIP_CHECK(hipStreamSynchronize(g->ctx->stream));
    return JW_OK;
}

int jw_batch_philox(jw_batch* b, uint32_t seed, uint32_t traj0, uint32_t step) {
    if (!b) return fail(JW_ERR_ARG, "null batch");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->seed = seed;
    b->traj0 = traj0;
    b->step = step;
    HIP_CHECK(hipMemcpy(b->d_step, &step, sizeof step, hipMemcpyHostToDevice));
    if (b->graph) {      // the key of the noise is baked into the graph
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        b->have_key = false;
    }
    return JW_OK;
}

int jw_batch_set_noise_space(jw_batch* b, int space) {
    if (!b) return fail(JW_ERR_ARG, "null batch");
    if (space != JW_NOISE_POSITION && space != JW_NOISE_MOMENTUM) return fail(JW_ERR_ARG, "noise space %d", space);
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->kspace = space == JW_NOISE_MOMENTUM;
    if (b->graph) {      // baked into the graph
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        b->have_key = false;
    }
    return JW_OK;
}

int jw_batch_set_layout(jw_batch* b, int group, int layout) {
    if (!b || group < 0 || group >= (int)b->g.size()) return fail(JW_ERR_ARG, "bad batch group");
    if (layout != -1 && layout != 0 && layout != 2 && layout != 4 && layout != 8 && layout != 16)
        return fail(JW_ERR_ARG, "layout %d", layout);
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (layout > 0) b->launch = 0;      // one launch per kernel class runs row-major planes
    b->g[group].layout = layout;
    b->g[group].ctx->layout = layout;
    b->g[group].fft_mfma = 0;
    b->g[group].ctx->fft_mfma = 0;
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        b->have_key = false;
    }
    return JW_OK;
}

int jw_batch_layout(const jw_batch* b, int group) {
    if (!b || group < 0 || group >= (int)b->g.size()) return -2;
    return b->g[group].layout;
}

int jw_batch_set_launch(jw_batch* b, int mode) {
    if (!b) return fail(JW_ERR_ARG, "null batch");
    if (mode < -1 || mode > 1) return fail(JW_ERR_ARG, "launch mode %d: expected -1, 0 or 1", mode);
    if (mode == 1 && (int)b->g.size() > JW_MAX_GROUPS)
        return fail(JW_ERR_UNSUPPORTED, "one launch per kernel class holds %d groups", JW_MAX_GROUPS);
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->launch = mode;
    if (mode == 1) batch_rowmajor(b);
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        b->have_key = false;
    }
    return JW_OK;
}

int jw_batch_launch(const jw_batch* b) { return b ? b->launch : -2; }

int jw_batch_fft_mode(const jw_batch* b, int group) {
    if (!b || group < 0 || group >= (int)b->g.size()) return -2;
    return b->g[group].fft_mfma;
}

int jw_batch_evolve(jw_batch* b, int coupling, int kernel, double step, const jw_alpha* a, int nsteps,
                    int reunit_every) {
    if (!b) return fail(JW_ERR_ARG, "null batch");
    if (coupling == JW_HATTA_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "HATTA coupling: the reference leaves the coupling uninitialised here");
    if (coupling < 0 || coupling > 4 || kernel != JW_SIN_KERNEL)
        return fail(kernel == JW_LINEAR_KERNEL ? JW_ERR_UNSUPPORTED : JW_ERR_ARG, "batches run the SIN kernel");
    if (!(step >= 0) || nsteps < 0 || reunit_every < 0) return fail(JW_ERR_ARG, "bad step/nsteps/reunit_every");
    if (b->kspace && coupling == JW_NOISE_COUPLING_CONSTANT)
        return fail(JW_ERR_UNSUPPORTED, "momentum-space noise: NOISE coupling scales the noise per site");
    if (nsteps == 0) return JW_OK;
    if (!b->fp32) reunit_every = 0;      // the FP64 form is not projected (as jw_evolve)
    HIP_CHECK(hipSetDevice(b->device));
    const jw::Alpha al = to_alpha(a);
    // steps per graph: one projection period (10 without projection)
    const int K = reunit_every > 0 ? reunit_every : 10;
    const BatchKey key{coupling, reunit_every, step, al};
    const bool same_key = b->have_key && key == b->key;
    // A new coupling, step or alpha may rebuild the coupling tables that the
    // previous call's graph (still running on b->stream) reads, and replaces
    // that graph: wait for the previous call's work first.
    if (!same_key) HIP_CHECK(hipStreamSynchronize(b->stream));
    if (int e = batch_tune(b, coupling, step, al)) return e;
    if (b->launch < 0)
        if (int e = batch_tune_launch(b, coupling, step, al)) return e;
    if (b->graph && !same_key) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    // from here on the tables (and the graph, once built) are this key's
    b->key = key;
    b->have_key = true;
    // every group's stream must be idle before the origin stream forks to it
    for (auto& g : b->g) HIP_CHECK(hipStreamSynchronize(g.ctx->stream));
    if (!b->graph && nsteps >= K) {
        // coupling tables first: their builds are not captured
        for (auto& g : b->g) {
            const double *p, *k;
            if (int e = batch_tables(g, coupling, al, &p, &k)) return e;
            HIP_CHECK(hipStreamSynchronize(g.ctx->stream));
        }
        hipGraph_t graph = nullptr;
        HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = JW_OK;
        for (int it = 0; it < K && rc == JW_OK; it++)
            rc = batch_step_all(b, coupling, step, al, reunit_every > 0 && it == K - 1, it == 0, it < K - 1);
        const hipError_t ec = hipStreamEndCapture(b->stream, &graph);
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess) return fail(JW_ERR_HIP, "step capture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            b->graph = nullptr;
            return fail(JW_ERR_HIP, "step graph: %s", hipGetErrorString(ei));
        }
        b->graph_steps = K;
    }
    int done = 0;
    if (b->graph)
        for (; done + K <= nsteps; done += K) HIP_CHECK(hipGraphLaunch(b->graph, b->stream));
    // the rest, launched directly; projected after the last step
    for (int it = done; it < nsteps; it++)
        if (int e = batch_step_all(b, coupling, step, al, reunit_every > 0 && it == nsteps - 1, it == done,
                                   it < nsteps - 1))
            return e;
    b->step += (uint32_t)nsteps;
    return JW_OK;
}

int jw_batch_sync(jw_batch* b) {
    if (!b) return fail(JW_ERR_ARG, "null batch");
    HIP_CHECK(hipStreamSynchronize(b->stream));
    return JW_OK;
}

}  // extern "C"
