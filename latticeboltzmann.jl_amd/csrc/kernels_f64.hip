// kernels_f64.hip -- the plain FP64 path.  Compiled with -ffp-contract=off:
// every multiply and add rounds separately, as in the reference's Julia code,
// so these kernels reproduce the CPU restatement bit for bit.
#include "lbm_kernels.h"

namespace lbm {

int dispatch_f64(const Launch &L, const StepParams<double> &p, hipStream_t s) {
    return dispatch_q<StF64>(L, p, s);
}

}  // namespace lbm
