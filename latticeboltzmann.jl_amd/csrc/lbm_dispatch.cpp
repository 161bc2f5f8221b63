// lbm_dispatch.cpp -- routes a launch to the (storage, quadrature) unit that
// instantiates its kernels; converts FP64 parameters for the FP32-arithmetic
// storages.
#include "lbm_internal.h"

namespace lbm {

#define LBM_SWITCH_Q(TAG, L, p, s)                          \
    switch ((L).qid) {                                      \
    case 4: return run_##TAG##_q4(L, p, s);                 \
    case 5: return run_##TAG##_q5(L, p, s);                 \
    case 9: return run_##TAG##_q9(L, p, s);                 \
    case 13: return run_##TAG##_q13(L, p, s);               \
    case 17: return run_##TAG##_q17(L, p, s);               \
    case 21: return run_##TAG##_q21(L, p, s);               \
    case 37: return run_##TAG##_q37(L, p, s);               \
    default: return -1;                                     \
    }

int dispatch_f64(const Launch &L, const StepParams<double> &p, hipStream_t s) {
    LBM_SWITCH_Q(f64, L, p, s)
}

int dispatch_lp(const Launch &L, const StepParams<double> &pd, hipStream_t s) {
    const StepParams<float> p = to_float_params(pd);
    switch (L.storage) {
    case 1: LBM_SWITCH_Q(f32, L, p, s)
    case 2: LBM_SWITCH_Q(f16dev, L, p, s)
    case 3: LBM_SWITCH_Q(bf16dev, L, p, s)
    case 4: LBM_SWITCH_Q(f16neq, L, p, s)
    case 5: LBM_SWITCH_Q(bf16neq, L, p, s)
    default: return -1;
    }
}

}  // namespace lbm
