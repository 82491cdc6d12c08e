// jw_step_f32.hip -- the step with FP32 storage and FFTs (the mixed form):
// the fused passes (jw_engine.hpp) and, outside their sizes, the unfused ones
// (jw_packed.hpp); site algebra in FP64.
#include "jw_engine.hpp"
#include "jw_packed.hpp"

namespace jw {

hipError_t step_fused_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const StepArgs& a) {
    return step_fused_t(c, U, Xi, Vv, ZA, ZB, a);
}

hipError_t step_multi_f32(hipStream_t st, const GroupArgs* g, int ng, const StepArgs& a) {
    return step_multi_from<cplxf>(st, g, ng, a);
}

hipError_t step_large_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const StepArgs& a) {
    return step_large_t(c, U, Xi, Vv, ZA, ZB, a);
}

hipError_t step_unfused_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const NoiseSrc& ns,
                            const double* sa_pos, const double* sa_k, int kc, int corr, double ss, int reunit) {
    return step_unfused_t(c, U, Xi, Vv, ZA, ZB, ns, sa_pos, sa_k, kc, corr, ss, reunit);
}

}  // namespace jw
