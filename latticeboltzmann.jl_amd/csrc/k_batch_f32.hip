// k_batch_f32.hip -- batched periodic D2Q9 lattices, FP32 shifted populations
// (-ffp-contract=fast).
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_BATCH(f32, StF32, 9)
}  // namespace lbm
