// k_f16neq_q37.hip -- fused / collide / stream kernels, storage f16neq, D2Q{37}.
// FP32 arithmetic on reduced storage: compiled with -ffp-contract=fast.
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f16neq, StF16Neq, 37)
}  // namespace lbm
