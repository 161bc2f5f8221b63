// kernels_lp.hip -- reduced-precision storages (FP32, FP16/BF16 deviations)
// with FP32 collision arithmetic.  Compiled with FMA contraction allowed.
#include "lbm_kernels.h"

namespace lbm {

static StepParams<float> to_float(const StepParams<double> &d) {
    StepParams<float> f{};
    f.nx = d.nx;
    f.ny = d.ny;
    f.plane = d.plane;
    for (int i = 0; i < 37; ++i) f.w[i] = static_cast<float>(d.w[i]);
    f.s = static_cast<float>(d.s);
    f.neg_s = static_cast<float>(d.neg_s);
    f.s2 = static_cast<float>(d.s2);
    f.three_s = static_cast<float>(d.three_s);
    f.six_s3 = static_cast<float>(d.six_s3);
    f.three_s2 = static_cast<float>(d.three_s2);
    f.s4 = static_cast<float>(d.s4);
    f.omo = static_cast<float>(d.omo);
    f.om = static_cast<float>(d.om);
    f.oms = static_cast<float>(d.oms);
    f.oma = static_cast<float>(d.oma);
    f.tau_force = static_cast<float>(d.tau_force);
    f.force_kind = d.force_kind;
    f.Fx = static_cast<float>(d.Fx);
    f.Fy = static_cast<float>(d.Fy);
    f.Ffield = d.Ffield;
    f.nbc = d.nbc;
    f.bcs = d.bcs;
    f.mw = d.mw;
    f.rho0 = static_cast<float>(d.rho0);
    f.dscale = static_cast<float>(d.dscale);
    f.dscale_inv = static_cast<float>(d.dscale_inv);
    f.mrt_n = 0;
    f.mrt_H = nullptr;
    f.cs_inv = static_cast<float>(d.cs_inv);
    return f;
}

int dispatch_lp(const Launch &L, const StepParams<double> &pd, hipStream_t s) {
    const StepParams<float> p = to_float(pd);
    switch (L.storage) {
    case 1: return dispatch_q<StF32>(L, p, s);
    case 2: return dispatch_q<StF16Dev>(L, p, s);
    case 3: return dispatch_q<StBF16Dev>(L, p, s);
    default: return -1;
    }
}

}  // namespace lbm
