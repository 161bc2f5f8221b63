// k_batch_f64.hip -- batched periodic D2Q9 lattices, plain FP64
// (-ffp-contract=off: bit-exact with the CPU restatement per lattice).
#include "lbm_kernels.h"

namespace lbm {
LBM_DEFINE_BATCH(f64, StF64, 9)
}  // namespace lbm
