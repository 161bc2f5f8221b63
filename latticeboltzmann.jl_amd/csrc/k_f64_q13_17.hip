// k_f64_q13_17.hip -- fused / collide / stream kernels, storage f64, D2Q{13,17}.
// Plain FP64: compiled with -ffp-contract=off (bit-exact with the CPU restatement).
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_RUN(f64, StF64, 13)
LBM_DEFINE_RUN(f64, StF64, 17)
}  // namespace lbm
