// k_f32_q37.hip -- fused / collide / stream kernels, storage f32, D2Q{37}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f32, StF32, 37)
}  // namespace lbm
