// k_f16dev_q21.hip -- fused / collide / stream kernels, storage f16dev, D2Q{21}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f16dev, StF16Dev, 21)
}  // namespace lbm
