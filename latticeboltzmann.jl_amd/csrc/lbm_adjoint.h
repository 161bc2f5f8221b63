// lbm_adjoint.h -- reverse-mode (adjoint) kernels of the periodic SRT/TRT
// collide-stream step, for gradients of a velocity-field loss.
//
// Forward step t (the reference's collide! then stream!, periodic lattice):
//   C_t[x, i]     = Omega_i(S_t[x, :])                       (per node)
//   S_{t+1}[x, i] = C_t[x - c_i, i]                           (pull stream)
// with S_0 = feq(rho0, u0) and loss L = 1/2 sum_x |u(S_T)[x] - u_target[x]|^2.
// Reverse step, lambda_t = dL/dS_t:
//   mu_t[x, i]     = lambda_{t+1}[x + c_i, i]                 (transposed stream: push)
//   lambda_t[x, :] = J_Omega(S_t[x, :])^T mu_t[x, :]          (per node)
// The collision Jacobian.  With nu_i = om mu_i (SRT) or
// nu_i = om_s mu+_i + om_a mu-_i (TRT, mu+- the (anti)symmetric parts under
// i <-> opp(i)), both operators give
//   (J^T mu)_j = mu_j - nu_j + sum_i nu_i dfeq_i/df_j,
//   sum_i nu_i dfeq_i/df_j = A + B . (c_j - u) / rho,
//   A = sum_i nu_i feq_i / rho,   B_d = sum_i nu_i dfeq_i/du_d,
// since rho = sum f, u = sum c f / rho (so du_d/df_j = (c_jd - u_d)/rho) and
// feq_i = w_i rho P_i(u).  Relaxation-time sensitivities accumulate per node:
//   dL/dtau   += sum_i mu_i (f_i - feq_i) / tau^2                 (SRT)
//   dL/dtau_s += sum_i mu+_i (f_i - feq_i) / tau_s^2, likewise tau_a with mu-.
// Adjoint fields are plain (unshifted) arrays of the arithmetic type in the
// structure-of-arrays layout lam[x + nx*(y + ny*i)].
#pragma once

#include "lbm_kernels.h"

namespace lbm {

// d feq_I / d u, for the closed-form equilibria of order FEQ = 1..4 of feq_i
// (the polynomial of maxwell_boltzmann_equilibrium.jl:43-66)
template <int QID, int I, class T>
__device__ __forceinline__ void dfeq_du(T rho, T ux, T uy, T usq, const StepParams<T> &p, T &gx,
                                        T &gy) {
    constexpr int FEQ = QT<QID>::FEQ;
    const T cx = static_cast<T>(QT<QID>::cx(I)), cy = static_cast<T>(QT<QID>::cy(I));
    const T ud = cx * ux + cy * uy;
    // d/du_d of the polynomial = kc * c_d + ku * u_d
    T kc = p.s, ku = T(0);
    if constexpr (FEQ >= 2) {
        kc = kc + p.s2 * ud;
        ku = ku - p.s;
    }
    if constexpr (FEQ >= 3) {
        const T s3 = p.s2 * p.s;
        kc = kc + T(0.5) * (s3 * (ud * ud) - p.s2 * usq);
        ku = ku - p.s2 * ud;
        if constexpr (FEQ >= 4) {
            kc = kc + T(1.0 / 6.0) * (p.s4 * (ud * ud * ud)) - T(0.5) * (s3 * usq * ud);
            ku = ku - T(0.5) * (s3 * (ud * ud)) + T(0.5) * (p.s2 * usq);
        }
    }
    const T wr = p.w[I] * rho;
    gx = wr * (kc * cx + ku * ux);
    gy = wr * (kc * cy + ku * uy);
}

template <int QID, int... Is, class F>
__device__ __forceinline__ void adj_each(F &&fn, std::integer_sequence<int, Is...>) {
    (fn(std::integral_constant<int, Is>{}), ...);
}
#define LBM_ADJ_EACH(...) adj_each<QID>(__VA_ARGS__, LBM_SEQ)

// state S_t of node (x, y): stored directly (DIRECT) or pulled from C_{t-1}
template <int QID, class ST, bool DIRECT>
__device__ __forceinline__ void adj_state(const typename ST::store_t *__restrict__ src,
                                          const StepParams<typename ST::T> &p, int x, int y,
                                          long long n, typename ST::T (&f)[QT<QID>::Q]) {
    if constexpr (DIRECT) load_local<QID, ST>(src, p, n, f, LBM_SEQ);
    else pull_seq<QID, ST>(src, p, x, y, f, LBM_SEQ);
}

// rho, u of a state (the FP32 storages hold g = f - w rho0)
template <int QID, class ST>
__device__ __forceinline__ void adj_moments(const typename ST::T (&f)[QT<QID>::Q],
                                            const StepParams<typename ST::T> &p,
                                            typename ST::T &drho, typename ST::T &rho,
                                            typename ST::T &ux, typename ST::T &uy) {
    if constexpr (ST::kShift) {
        node_moments_dev<QID>(f, p, drho, ux, uy);
        rho = p.rho0 + drho;
    } else {
        node_moments<QID>(f, rho, ux, uy);
        drho = rho - p.rho0;
    }
}

// one reverse step: lam = J^T(S_t) mu, mu pushed from lam_next
template <int QID, class ST, int COLL, bool DIRECT>
__global__ void __launch_bounds__(256)
k_adj_step(const typename ST::T *__restrict__ lam_next, const typename ST::store_t *__restrict__ src,
           typename ST::T *__restrict__ lam, double *__restrict__ gtau,
           const StepParams<typename ST::T> p) {
    using T = typename ST::T;
    constexpr int Q = QT<QID>::Q;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    const long long n = (long long)x + (long long)p.nx * y;
    T f[Q];
    adj_state<QID, ST, DIRECT>(src, p, x, y, n, f);
    T mu[Q];
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        const int sx = wrap(x + QT<QID>::cx(i), p.nx), sy = wrap(y + QT<QID>::cy(i), p.ny);
        mu[i] = lam_next[p.plane * i + (long long)sx + (long long)p.nx * sy];
    });
    T drho, rho, ux, uy;
    adj_moments<QID, ST>(f, p, drho, rho, ux, uy);
    const T usq = ux * ux + uy * uy;
    T feq[Q], neq[Q];
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        if constexpr (ST::kShift) {
            const T g = feq_dev_i<QID, i, T>(drho, rho, ux, uy, usq, p);
            neq[i] = f[i] - g;
            feq[i] = g + p.w[i] * p.rho0;
        } else {
            feq[i] = feq_i<QID, i, T>(rho, ux, uy, usq, p);
            neq[i] = f[i] - feq[i];
        }
    });
    T nu[Q];
    T gs = T(0), ga = T(0);
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        if constexpr (COLL == 0) {
            nu[i] = p.om * mu[i];
            gs = gs + mu[i] * neq[i];
        } else {
            constexpr int o = QT<QID>::opp(i);
            const T mp = T(0.5) * (mu[i] + mu[o]), mm = T(0.5) * (mu[i] - mu[o]);
            nu[i] = p.oms * mp + p.oma * mm;
            gs = gs + mp * neq[i];
            ga = ga + mm * neq[i];
        }
    });
    T A = T(0), Bx = T(0), By = T(0);
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        T gx, gy;
        dfeq_du<QID, i, T>(rho, ux, uy, usq, p, gx, gy);
        A = A + nu[i] * feq[i];
        Bx = Bx + nu[i] * gx;
        By = By + nu[i] * gy;
    });
    const T ir = T(1) / rho;
    A = A * ir;
    Bx = Bx * ir;
    By = By * ir;
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int j = decltype(Ic)::value;
        const T cx = static_cast<T>(QT<QID>::cx(j)), cy = static_cast<T>(QT<QID>::cy(j));
        lam[p.plane * j + n] = (mu[j] - nu[j]) + (A + (Bx * (cx - ux) + By * (cy - uy)));
    });
    if constexpr (COLL == 0) {
        gtau[n] += static_cast<double>(gs) * (static_cast<double>(p.om) * static_cast<double>(p.om));
    } else {
        gtau[n] += static_cast<double>(gs) * (static_cast<double>(p.oms) * static_cast<double>(p.oms));
        gtau[p.plane + n] +=
            static_cast<double>(ga) * (static_cast<double>(p.oma) * static_cast<double>(p.oma));
    }
}

// lambda_T = dL/dS_T and the per-node loss 1/2 |u - u_target|^2
template <int QID, class ST, bool DIRECT>
__global__ void __launch_bounds__(256)
k_adj_seed(const typename ST::store_t *__restrict__ src, const double *__restrict__ ut,
           typename ST::T *__restrict__ lam, double *__restrict__ lossn, double *__restrict__ uT,
           const StepParams<typename ST::T> p) {
    using T = typename ST::T;
    constexpr int Q = QT<QID>::Q;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    const long long n = (long long)x + (long long)p.nx * y;
    T f[Q];
    adj_state<QID, ST, DIRECT>(src, p, x, y, n, f);
    T drho, rho, ux, uy;
    adj_moments<QID, ST>(f, p, drho, rho, ux, uy);
    const double dx = static_cast<double>(ux) - ut[n];
    const double dy = static_cast<double>(uy) - ut[p.plane + n];
    lossn[n] = 0.5 * (dx * dx + dy * dy);
    if (uT) {
        uT[n] = static_cast<double>(ux);
        uT[p.plane + n] = static_cast<double>(uy);
    }
    const double ir = 1.0 / static_cast<double>(rho);
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int j = decltype(Ic)::value;
        const double cx = QT<QID>::cx(j), cy = QT<QID>::cy(j);
        lam[p.plane * j + n] = static_cast<T>(
            (dx * (cx - static_cast<double>(ux)) + dy * (cy - static_cast<double>(uy))) * ir);
    });
}

// S_0 = feq(rho0, u0) (the reference's equilibrium initial condition,
// initial_conditions.jl:7-22, at uniform density)
template <int QID, class ST>
__global__ void __launch_bounds__(256)
k_adj_init(const double *__restrict__ u0, typename ST::store_t *__restrict__ S,
           const StepParams<typename ST::T> p) {
    using T = typename ST::T;
    constexpr int Q = QT<QID>::Q;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    const long long n = (long long)x + (long long)p.nx * y;
    const T ux = static_cast<T>(u0[n]), uy = static_cast<T>(u0[p.plane + n]);
    const T usq = ux * ux + uy * uy;
    T f[Q];
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        if constexpr (ST::kShift) f[i] = feq_dev_i<QID, i, T>(T(0), p.rho0, ux, uy, usq, p);
        else f[i] = feq_i<QID, i, T>(p.rho0, ux, uy, usq, p);
    });
    store_local<QID, ST>(S, p, n, f, LBM_SEQ);
}

// dL/du0 = sum_i lambda_0,i dfeq_i(rho0, u0)/du0
template <int QID, class ST>
__global__ void __launch_bounds__(256)
k_adj_grad_u0(const double *__restrict__ u0, const typename ST::T *__restrict__ lam0,
              double *__restrict__ grad, const StepParams<typename ST::T> p) {
    using T = typename ST::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    const long long n = (long long)x + (long long)p.nx * y;
    const T ux = static_cast<T>(u0[n]), uy = static_cast<T>(u0[p.plane + n]);
    const T usq = ux * ux + uy * uy;
    T gxs = T(0), gys = T(0);
    LBM_ADJ_EACH([&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        T gx, gy;
        dfeq_du<QID, i, T>(p.rho0, ux, uy, usq, p, gx, gy);
        const T l = lam0[p.plane * i + n];
        gxs = gxs + l * gx;
        gys = gys + l * gy;
    });
    grad[n] = static_cast<double>(gxs);
    grad[p.plane + n] = static_cast<double>(gys);
}

template <int QID, class ST>
int launch_adjoint(const AdjLaunch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    using S = typename ST::store_t;
    using T = typename ST::T;
    const int bx = default_block(p.nx);
    const dim3 grid = node_grid(p.nx, p.ny, bx), block(bx, 1, 1);
    const S *st = static_cast<const S *>(L.state);
    const T *li = static_cast<const T *>(L.lam_in);
    T *lo = static_cast<T *>(L.lam_out);
    switch (L.what) {
    case 0:
        k_adj_init<QID, ST><<<grid, block, 0, s>>>(L.u, static_cast<S *>(L.pops), p);
        break;
    case 1:
        if (L.direct) k_adj_seed<QID, ST, true><<<grid, block, 0, s>>>(st, L.u, lo, L.out, L.aux, p);
        else k_adj_seed<QID, ST, false><<<grid, block, 0, s>>>(st, L.u, lo, L.out, L.aux, p);
        break;
    case 2:
        if (L.coll == 0) {
            if (L.direct) k_adj_step<QID, ST, 0, true><<<grid, block, 0, s>>>(li, st, lo, L.out, p);
            else k_adj_step<QID, ST, 0, false><<<grid, block, 0, s>>>(li, st, lo, L.out, p);
        } else if (L.coll == 1) {
            if (L.direct) k_adj_step<QID, ST, 1, true><<<grid, block, 0, s>>>(li, st, lo, L.out, p);
            else k_adj_step<QID, ST, 1, false><<<grid, block, 0, s>>>(li, st, lo, L.out, p);
        } else {
            return -2;
        }
        break;
    case 3:
        k_adj_grad_u0<QID, ST><<<grid, block, 0, s>>>(L.u, li, L.out, p);
        break;
    default:
        return -3;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

#define LBM_DEFINE_ADJOINT(TAG, ST)                                                          \
    int run_adjoint_##TAG(const AdjLaunch &L, const StepParams<typename ST::T> &p,           \
                          hipStream_t s) {                                                   \
        switch (L.qid) {                                                                     \
        case 4: return launch_adjoint<4, ST>(L, p, s);                                       \
        case 5: return launch_adjoint<5, ST>(L, p, s);                                       \
        case 9: return launch_adjoint<9, ST>(L, p, s);                                       \
        case 13: return launch_adjoint<13, ST>(L, p, s);                                     \
        case 17: return launch_adjoint<17, ST>(L, p, s);                                     \
        case 21: return launch_adjoint<21, ST>(L, p, s);                                     \
        case 37: return launch_adjoint<37, ST>(L, p, s);                                     \
        default: return -1;                                                                  \
        }                                                                                    \
    }

}  // namespace lbm
