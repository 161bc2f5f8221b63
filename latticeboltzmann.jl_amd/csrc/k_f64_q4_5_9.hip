// k_f64_q4_5_9.hip -- fused / collide / stream kernels, storage f64, D2Q{4,5,9}.
// Plain FP64: compiled with -ffp-contract=off (bit-exact with the CPU restatement).
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f64, StF64, 4)
LBM_DEFINE_RUN(f64, StF64, 5)
LBM_DEFINE_RUN(f64, StF64, 9)
}  // namespace lbm
