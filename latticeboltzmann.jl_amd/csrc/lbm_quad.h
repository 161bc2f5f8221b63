// lbm_quad.h -- compile-time velocity sets for the gfx950 kernels.
//
// Abscissae and opposite maps of src/quadratures/D2Q{4,5,9,13,17,21,37}.jl and
// the generic opposite rule of src/quadratures.jl:11-19.  Integer tables are
// constexpr so that, after full unrolling over the velocity index, every
// c_i, opp(i) and stencil offset is an immediate.  Weights and 1/c_s^2 are
// doubles computed on the host (lbm_host_quadrature) with the reference's own
// expressions (D2Q17 needs sqrt(193)) and travel in the kernel arguments.
#pragma once

#include <hip/hip_runtime.h>

#define LBM_HD __host__ __device__

namespace lbm {

constexpr int kQMax = 37;

template <int QID>
struct QT;

// D2Q4.jl:16-31
template <>
struct QT<4> {
    static constexpr int Q = 4, FEQ = 1, MAXC = 1, ORDER = 3;
    LBM_HD static constexpr int cx(int i) { constexpr int t[4] = {1, 0, -1, 0}; return t[i]; }
    LBM_HD static constexpr int cy(int i) { constexpr int t[4] = {0, 1, 0, -1}; return t[i]; }
    LBM_HD static constexpr int opp(int i) { return (i + 1 <= 2) ? i + 2 : i - 2; }
};

// D2Q5.jl:17-42
template <>
struct QT<5> {
    static constexpr int Q = 5, FEQ = 1, MAXC = 1, ORDER = 3;
    LBM_HD static constexpr int cx(int i) { constexpr int t[5] = {0, 1, 0, -1, 0}; return t[i]; }
    LBM_HD static constexpr int cy(int i) { constexpr int t[5] = {0, 0, 1, 0, -1}; return t[i]; }
    LBM_HD static constexpr int opp(int i) { return i == 0 ? 0 : (i + 1 <= 3 ? i + 2 : i - 2); }
};

// D2Q9.jl:20-38
template <>
struct QT<9> {
    static constexpr int Q = 9, FEQ = 2, MAXC = 1, ORDER = 5;
    LBM_HD static constexpr int cx(int i) {
        constexpr int t[9] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
        return t[i];
    }
    LBM_HD static constexpr int cy(int i) {
        constexpr int t[9] = {0, 1, 0, -1, -1, -1, 0, 1, 1};
        return t[i];
    }
    LBM_HD static constexpr int opp(int i) { return i == 0 ? 0 : (i + 1 <= 5 ? i + 4 : i - 4); }
};

// generic rule of quadratures.jl:11-19 written 0-based: 0 -> 0, odd -> +1, even -> -1
LBM_HD constexpr int generic_opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

// D2Q13.jl:10-29
template <>
struct QT<13> {
    static constexpr int Q = 13, FEQ = 2, MAXC = 2, ORDER = 5;
    LBM_HD static constexpr int cx(int i) {
        constexpr int t[13] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0};
        return t[i];
    }
    LBM_HD static constexpr int cy(int i) {
        constexpr int t[13] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -2, 2};
        return t[i];
    }
    LBM_HD static constexpr int opp(int i) { return generic_opp(i); }
};

// D2Q17.jl:21-57
template <>
struct QT<17> {
    static constexpr int Q = 17, FEQ = 3, MAXC = 3, ORDER = 7;
    LBM_HD static constexpr int cx(int i) {
        constexpr int t[17] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 2, -2, 3, -3, 0, 0};
        return t[i];
    }
    LBM_HD static constexpr int cy(int i) {
        constexpr int t[17] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 2, -2, -2, 2, 0, 0, 3, -3};
        return t[i];
    }
    LBM_HD static constexpr int opp(int i) { return generic_opp(i); }
};

// D2Q21.jl:15-71 (25 stored populations; the last four have weight 0)
template <>
struct QT<21> {
    static constexpr int Q = 25, FEQ = 3, MAXC = 3, ORDER = 7;
    LBM_HD static constexpr int cx(int i) {
        constexpr int t[25] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0,
                               2, -2, 2, -2, 3, -3, 0, 0, 3, -3, -3, 3};
        return t[i];
    }
    LBM_HD static constexpr int cy(int i) {
        constexpr int t[25] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, 2, -2,
                               2, -2, -2, 2, 0, 0, 3, -3, 3, -3, 3, -3};
        return t[i];
    }
    LBM_HD static constexpr int opp(int i) { return generic_opp(i); }
};

// D2Q37.jl:11-74
template <>
struct QT<37> {
    static constexpr int Q = 37, FEQ = 4, MAXC = 3, ORDER = 9;
    LBM_HD static constexpr int cx(int i) {
        constexpr int t[37] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0, 2, -2, -2, 2, 1, -1,
                               1, -1, 2, -2, 2, -2, 3, -3, 0, 0, 3, -3, 3, -3, 1, -1, -1, 1};
        return t[i];
    }
    LBM_HD static constexpr int cy(int i) {
        constexpr int t[37] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, 2, -2, 1, -1, 1, -1, 2, -2,
                               -2, 2, 2, -2, -2, 2, 0, 0, 3, -3, 1, -1, -1, 1, 3, -3, 3, -3};
        return t[i];
    }
    LBM_HD static constexpr int opp(int i) { return generic_opp(i); }
};

}  // namespace lbm
