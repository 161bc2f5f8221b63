// k_f32_q4_5_9.hip -- fused / collide / stream kernels, storage f32, D2Q{4,5,9}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f32, StF32, 4)
LBM_DEFINE_RUN(f32, StF32, 5)
LBM_DEFINE_RUN(f32, StF32, 9)
}  // namespace lbm
