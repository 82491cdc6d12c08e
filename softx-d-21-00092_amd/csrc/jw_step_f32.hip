// jw_step_f32.hip -- the fused step with FP32 storage and FFTs (the mixed
// form); site algebra in FP64 (jw_engine.hpp).
#include "jw_engine.hpp"

namespace jw {

hipError_t step_fused_f32(jw_ctx* c, cplxf* U, cplxf* Xi, cplxf* Vv, cplxf* ZA, cplxf* ZB, const StepArgs& a) {
    return step_fused_t(c, U, Xi, Vv, ZA, ZB, a);
}

}  // namespace jw
