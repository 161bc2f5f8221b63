// k_adjoint_f32.hip -- reverse-mode step kernels, FP32 shifted populations
// (-ffp-contract=fast).
#include "lbm_adjoint.h"

namespace lbm {
LBM_DEFINE_ADJOINT(f32, StF32)
}  // namespace lbm
