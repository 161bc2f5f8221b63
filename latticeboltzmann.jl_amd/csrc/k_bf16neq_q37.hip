// k_bf16neq_q37.hip -- fused / collide / stream kernels, storage bf16neq, D2Q{37}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(bf16neq, StBF16Neq, 37)
}  // namespace lbm
