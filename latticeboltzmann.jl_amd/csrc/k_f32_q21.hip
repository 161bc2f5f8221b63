// k_f32_q21.hip -- fused / collide / stream kernels, storage f32, D2Q{21}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f32, StF32, 21)
}  // namespace lbm
