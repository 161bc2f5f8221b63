// k_bf16dev_q37.hip -- fused / collide / stream kernels, storage bf16dev, D2Q{37}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(bf16dev, StBF16Dev, 37)
}  // namespace lbm
