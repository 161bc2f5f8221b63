// k_bf16neq_q4_5_9.hip -- fused / collide / stream kernels, storage bf16neq, D2Q{4,5,9}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(bf16neq, StBF16Neq, 4)
LBM_DEFINE_RUN(bf16neq, StBF16Neq, 5)
LBM_DEFINE_RUN(bf16neq, StBF16Neq, 9)
}  // namespace lbm
