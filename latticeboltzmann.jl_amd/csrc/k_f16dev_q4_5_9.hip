// k_f16dev_q4_5_9.hip -- fused / collide / stream kernels, storage f16dev, D2Q{4,5,9}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f16dev, StF16Dev, 4)
LBM_DEFINE_RUN(f16dev, StF16Dev, 5)
LBM_DEFINE_RUN(f16dev, StF16Dev, 9)
}  // namespace lbm
