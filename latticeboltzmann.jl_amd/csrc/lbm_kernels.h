// lbm_kernels.h -- gfx950 collide / stream / boundary kernels (templates).
//
// One thread owns one lattice node (x, y); a 64-lane wavefront covers 64
// consecutive x of one row, so every population plane is read and written as
// contiguous 512 B (FP64) segments.  Populations live in structure-of-arrays
// planes f[x + nx*(y + ny*i)] -- the reference's Array{Float64,3} layout.
//
// Evaluation order.  The FP64 translation unit is compiled with
// -ffp-contract=off and every expression below keeps the operand order of the
// reference's Julia code, so the FP64 kernels reproduce the CPU restatement bit
// for bit.  Where a term is skipped (a product with an abscissa component 0,
// an added +0.0) the skipped operation is an exact identity on the running
// value: the running sums start at +0 and therefore never hold -0.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <utility>

#include "lbm_internal.h"
#include "lbm_quad.h"

namespace lbm {

// ---------------------------------------------------------------------------
// storage traits
// ---------------------------------------------------------------------------
struct StF64 {
    using store_t = double;
    using T = double;
    static constexpr int id = 0;
    __device__ static T load(const store_t *p, long long idx, int, const StepParams<T> &) {
        return p[idx];
    }
    __device__ static void store(store_t *p, long long idx, int, T v, const StepParams<T> &) {
        p[idx] = v;
    }
};

struct StF32 {
    using store_t = float;
    using T = float;
    static constexpr int id = 1;
    __device__ static T load(const store_t *p, long long idx, int, const StepParams<T> &) {
        return p[idx];
    }
    __device__ static void store(store_t *p, long long idx, int, T v, const StepParams<T> &) {
        p[idx] = v;
    }
};

// FP16 deviation storage: g_i = (f_i - w_i rho0) * dscale, stored as IEEE
// binary16 (round to nearest even); f_i = w_i rho0 + g_i / dscale.
struct StF16Dev {
    using store_t = __half;
    using T = float;
    static constexpr int id = 2;
    __device__ static T load(const store_t *p, long long idx, int i, const StepParams<T> &q) {
        return q.w[i] * q.rho0 + __half2float(p[idx]) * q.dscale_inv;
    }
    __device__ static void store(store_t *p, long long idx, int i, T v, const StepParams<T> &q) {
        p[idx] = __float2half_rn((v - q.w[i] * q.rho0) * q.dscale);
    }
};

// BF16 deviation storage, same scheme with bfloat16 (8-bit significand).
struct StBF16Dev {
    using store_t = __hip_bfloat16;
    using T = float;
    static constexpr int id = 3;
    __device__ static T load(const store_t *p, long long idx, int i, const StepParams<T> &q) {
        return q.w[i] * q.rho0 + __bfloat162float(p[idx]) * q.dscale_inv;
    }
    __device__ static void store(store_t *p, long long idx, int i, T v, const StepParams<T> &q) {
        p[idx] = __float2bfloat16((v - q.w[i] * q.rho0) * q.dscale);
    }
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// c * v for an integer abscissa component, exact shortcuts for |c| <= 2
template <int C, class T>
__device__ __forceinline__ T cmul(T v) {
    if constexpr (C == 1) return v;
    else if constexpr (C == -1) return -v;
    else if constexpr (C == 2) return v + v;
    else if constexpr (C == -2) return -(v + v);
    else return static_cast<T>(C) * v;
}

// mod1-style periodic wrap (stream.jl:69-74) of v into [0, n), valid for any |v|
__device__ __forceinline__ int wrap(int v, int n) {
    if (v < 0) {
        v += n;
        if (v < 0) v = ((v % n) + n) % n;
    } else if (v >= n) {
        v -= n;
        if (v >= n) v = v % n;
    }
    return v;
}

// x^4 rounded once from the exact square (two-product), used for the D2Q37
// u_dot_xi^4 term; the CPU restatement computes it the same way.
template <class T>
__device__ __forceinline__ T pow4(T x) {
    T p = x * x;
    T e = __builtin_fma(x, x, -p);
    T hi = p * p;
    T lo = __builtin_fma(p, p, -hi);
    return hi + (lo + (T(2) * p) * e);
}

template <int QID, int I, class T>
__device__ __forceinline__ T udot(T ux, T uy) {
    constexpr int cx = QT<QID>::cx(I), cy = QT<QID>::cy(I);
    if constexpr (cx == 0 && cy == 0) return T(0);
    else if constexpr (cx == 0) return cmul<cy>(uy);
    else if constexpr (cy == 0) return cmul<cx>(ux);
    else return cmul<cx>(ux) + cmul<cy>(uy);
}

// Closed-form equilibrium of population I (maxwell_boltzmann_equilibrium.jl:43-66;
// velocity_distribution_function/quadratures.jl:69-85, 111-127, 161-225).
template <int QID, int I, class T>
__device__ __forceinline__ T feq_i(T rho, T ux, T uy, T usq, const StepParams<T> &p) {
    constexpr int FEQ = QT<QID>::FEQ;
    const T ud = udot<QID, I>(ux, uy);
    const T rw = rho * p.w[I];
    const T a1 = p.s * ud;
    if constexpr (FEQ == 1) {
        return rw * (T(1) + a1);
    } else {
        const T ud2 = ud * ud;
        const T a2 = p.s2 * ud2 + p.neg_s * usq;
        T acc = (T(1) + a1) + T(0.5) * a2;
        if constexpr (FEQ >= 3) {
            const T a3 = (p.s * ud) * (p.s2 * ud2 - p.three_s * usq);
            acc = acc + T(1.0 / 6.0) * a3;
        }
        if constexpr (FEQ >= 4) {
            const T a4 = (p.s4 * pow4(ud) - (p.six_s3 * usq) * ud2) + p.three_s2 * (usq * usq);
            acc = acc + T(1.0 / 24.0) * a4;
        }
        return rw * acc;
    }
}

// ---------------------------------------------------------------------------
// per-node moments: density = sum(f) (moments.jl:3), velocity! (moments.jl:5-19)
// ---------------------------------------------------------------------------
template <int QID, class T, int... Is>
__device__ __forceinline__ void moments_seq(const T (&f)[QT<QID>::Q], T &rho, T &ux, T &uy,
                                            std::integer_sequence<int, Is...>) {
    T r = f[0] + f[1];
    ((r = (Is >= 2 ? r + f[Is] : r)), ...);
    T mx = T(0), my = T(0);
    ((mx = (QT<QID>::cx(Is) != 0 ? mx + cmul<QT<QID>::cx(Is)>(f[Is]) : mx)), ...);
    ((my = (QT<QID>::cy(Is) != 0 ? my + cmul<QT<QID>::cy(Is)>(f[Is]) : my)), ...);
    rho = r;
    ux = mx / r;
    uy = my / r;
}

template <int QID, class T>
__device__ __forceinline__ void node_moments(const T (&f)[QT<QID>::Q], T &rho, T &ux, T &uy) {
    moments_seq<QID, T>(f, rho, ux, uy, std::make_integer_sequence<int, QT<QID>::Q>{});
}

template <class T>
__device__ __forceinline__ void node_force(const StepParams<T> &p, int x, int y, T &Fx, T &Fy) {
    if (p.force_kind == 1) {
        Fx = p.Fx;
        Fy = p.Fy;
    } else {
        const long long n = (long long)x + (long long)p.nx * y;
        Fx = static_cast<T>(p.Ffield[n]);
        Fy = static_cast<T>(p.Ffield[p.plane + n]);
    }
}

// ---------------------------------------------------------------------------
// collision of one node, in place (f -> post-collision f)
// ---------------------------------------------------------------------------
template <int QID, int COLL, class T, int... Is>
__device__ __forceinline__ void collide_seq(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                            int y, std::integer_sequence<int, Is...>) {
    constexpr int Q = QT<QID>::Q;
    T rho, ux, uy;
    node_moments<QID, T>(f, rho, ux, uy);
    if (p.force_kind != 0) {
        T Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;  // srt.jl:73 / trt.jl:160
        uy = uy + p.tau_force * Fy;
    }
    const T usq = ux * ux + uy * uy;
    T feq[Q];
    ((feq[Is] = feq_i<QID, Is, T>(rho, ux, uy, usq, p)), ...);
    if constexpr (COLL == 0) {
        // srt.jl:76-78
        ((f[Is] = p.omo * f[Is] + p.om * feq[Is]), ...);
    } else {
        // trt.jl:163-175 (each population reads the pre-collision pair)
        T g[Q];
        ((g[Is] = f[Is]), ...);
        auto trt = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            constexpr int j = QT<QID>::opp(i);
            const T feq_s = T(0.5) * (feq[i] + feq[j]);
            const T feq_a = T(0.5) * (feq[i] - feq[j]);
            const T f_s = T(0.5) * (g[i] + g[j]);
            const T f_a = T(0.5) * (g[i] - g[j]);
            f[i] = g[i] + ((-p.oms) * (f_s - feq_s) - p.oma * (f_a - feq_a));
        };
        (trt(std::integral_constant<int, Is>{}), ...);
    }
}

template <int QID, int COLL, class T>
__device__ __forceinline__ void collide_node(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                             int y) {
    collide_seq<QID, COLL, T>(f, p, x, y, std::make_integer_sequence<int, QT<QID>::Q>{});
}

// ---------------------------------------------------------------------------
// regularised MRT (mrt.jl:56-118), FP64.  Hermite tables mrt_H: for order n
// (2..4) a block of Q x 2^n values at offset Q*(0, 4, 12)[n-2], row i,
// column k = column-major flattened tensor index.
// ---------------------------------------------------------------------------
template <int QID, class T>
__device__ __forceinline__ void collide_mrt_node(T (&f)[QT<QID>::Q], const StepParams<T> &p,
                                                 int x, int y) {
    constexpr int Q = QT<QID>::Q;
    constexpr int N = QT<QID>::ORDER / 2;
    T rho, ux, uy;
    node_moments<QID, T>(f, rho, ux, uy);
    if (p.force_kind != 0) {
        T Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;  // mrt.jl:94
        uy = uy + p.tau_force * Fy;
    }
    const T u[2] = {ux, uy};
    T acoll[3][16];
#pragma unroll
    for (int n = 2; n <= N; ++n) {
        const int m = 1 << n;
        const double *H = p.mrt_H + Q * (n == 2 ? 0 : (n == 3 ? 4 : 12));
#pragma unroll
        for (int k = 0; k < m; ++k) {
            // a_eq (hermite.jl:45-77 at T = 1)
            T pr = u[k & 1] * u[(k >> 1) & 1];
            if (n >= 3) pr = pr * u[(k >> 2) & 1];
            if (n >= 4) pr = pr * u[(k >> 3) & 1];
            const T aeq = rho * pr;
            // a_f = sum_i f_i H_n(c_i) (mrt.jl:102)
            T a = f[0] * H[k] + f[1] * H[m + k];
#pragma unroll
            for (int i = 2; i < Q; ++i) a = a + f[i] * H[i * m + k];
            acoll[n - 2][k] = p.mrt_omo[n] * a + p.mrt_om[n] * aeq;  // mrt.jl:104
        }
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const T ru = ux * static_cast<T>(QT<QID>::cx(i)) + uy * static_cast<T>(QT<QID>::cy(i));
        T hsum = T(0);
#pragma unroll
        for (int n = 2; n <= N; ++n) {
            const int m = 1 << n;
            const double *H = p.mrt_H + Q * (n == 2 ? 0 : (n == 3 ? 4 : 12));
            T d = T(0);
#pragma unroll
            for (int k = 0; k < m; ++k) d = d + acoll[n - 2][k] * H[i * m + k];
            const T fact = n == 2 ? T(2) : (n == 3 ? T(6) : T(24));
            const T e = (p.mrt_sn[n] * d) / fact;
            hsum = (n == 2) ? e : hsum + e;
        }
        f[i] = p.w[i] * ((rho + (p.s * rho) * ru) + hsum);  // mrt.jl:106-115
    }
}

// ---------------------------------------------------------------------------
// boundary conditions of the ordered list for node (x, y), applied to the
// pulled populations f (bounce_back.jl:117-181, moving_wall.jl:198-219):
// population i is replaced by the same node's post-collision opp(i) (+ the
// moving-wall term) when its pull source lies beyond the wall.
// ---------------------------------------------------------------------------
template <int QID, class ST, int... Is>
__device__ __forceinline__ void apply_bc_seq(const typename ST::store_t *__restrict__ C,
                                             const StepParams<typename ST::T> &p, int x, int y,
                                             typename ST::T (&f)[QT<QID>::Q],
                                             std::integer_sequence<int, Is...>) {
    using T = typename ST::T;
    constexpr int Q = QT<QID>::Q;
    const int nx = p.nx, ny = p.ny;
    const long long node = (long long)x + (long long)nx * y;
    for (int b = 0; b < p.nbc; ++b) {
        const BCDev bc = p.bcs[b];
        const bool inr = bc.kind == 1 || (x >= bc.x0 && x <= bc.x1 && y >= bc.y0 && y <= bc.y1);
        if (!inr) continue;
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i), j = QT<QID>::opp(i);
            bool hit;
            switch (bc.dir) {
            case 0: hit = (y - cy) >= ny; break;
            case 2: hit = (y - cy) < 0; break;
            case 1: hit = (x - cx) >= nx; break;
            default: hit = (x - cx) < 0; break;
            }
            if (hit) {
                T v = ST::load(C + p.plane * j, node, j, p);
                if (bc.kind == 1) v = v + static_cast<T>(p.mw[b * Q + i]);
                f[i] = v;
            }
        };
        (one(std::integral_constant<int, Is>{}), ...);
    }
}

// Pull the post-stream, post-BC populations of node (x, y) from the
// post-collision field C (stream.jl:19-30 + apply!).
template <int QID, class ST, int... Is>
__device__ __forceinline__ void pull_seq(const typename ST::store_t *__restrict__ C,
                                         const StepParams<typename ST::T> &p, int x, int y,
                                         typename ST::T (&f)[QT<QID>::Q],
                                         std::integer_sequence<int, Is...> seq) {
    constexpr int M = QT<QID>::MAXC;
    const int nx = p.nx, ny = p.ny;
    const bool edge = (x < M) | (x >= nx - M) | (y < M) | (y >= ny - M);
    if (!edge) {
        const long long node = (long long)x + (long long)nx * y;
        ((f[Is] = ST::load(C + p.plane * Is,
                           node - QT<QID>::cx(Is) - (long long)nx * QT<QID>::cy(Is), Is, p)),
         ...);
    } else {
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            const int sx = wrap(x - QT<QID>::cx(i), nx);
            const int sy = wrap(y - QT<QID>::cy(i), ny);
            f[i] = ST::load(C + p.plane * i, (long long)sx + (long long)nx * sy, i, p);
        };
        (one(std::integral_constant<int, Is>{}), ...);
        if (p.nbc > 0) apply_bc_seq<QID, ST>(C, p, x, y, f, seq);
    }
}

template <int QID, class ST, int... Is>
__device__ __forceinline__ void load_local(const typename ST::store_t *__restrict__ F,
                                           const StepParams<typename ST::T> &p, long long node,
                                           typename ST::T (&f)[QT<QID>::Q],
                                           std::integer_sequence<int, Is...>) {
    ((f[Is] = ST::load(F + p.plane * Is, node, Is, p)), ...);
}

template <int QID, class ST, int... Is>
__device__ __forceinline__ void store_local(typename ST::store_t *__restrict__ F,
                                            const StepParams<typename ST::T> &p, long long node,
                                            const typename ST::T (&f)[QT<QID>::Q],
                                            std::integer_sequence<int, Is...>) {
    ((ST::store(F + p.plane * Is, node, Is, f[Is], p)), ...);
}

template <int QID, int COLL, class T>
__device__ __forceinline__ void collide_any(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                            int y) {
    if constexpr (COLL == 2) collide_mrt_node<QID, T>(f, p, x, y);
    else collide_node<QID, COLL, T>(f, p, x, y);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
#define LBM_SEQ std::make_integer_sequence<int, QT<QID>::Q>{}

// collide!: f_out = collide(f_in), node-local
template <int QID, int COLL, class ST>
__global__ void __launch_bounds__(256)
k_collide(const typename ST::store_t *__restrict__ fin, typename ST::store_t *__restrict__ fout,
          const StepParams<typename ST::T> p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    typename ST::T f[QT<QID>::Q];
    const long long node = (long long)x + (long long)p.nx * y;
    load_local<QID, ST>(fin, p, node, f, LBM_SEQ);
    collide_any<QID, COLL>(f, p, x, y);
    store_local<QID, ST>(fout, p, node, f, LBM_SEQ);
}

// stream! + apply!: f_new = BCs(stream(f), f)
template <int QID, class ST>
__global__ void __launch_bounds__(256)
k_stream_bc(const typename ST::store_t *__restrict__ C, typename ST::store_t *__restrict__ S,
            const StepParams<typename ST::T> p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    typename ST::T f[QT<QID>::Q];
    pull_seq<QID, ST>(C, p, x, y, f, LBM_SEQ);
    store_local<QID, ST>(S, p, (long long)x + (long long)p.nx * y, f, LBM_SEQ);
}

// one full time step on post-collision state: C' = collide(BCs(stream(C), C))
template <int QID, int COLL, class ST>
__global__ void __launch_bounds__(256)
k_fused(const typename ST::store_t *__restrict__ C, typename ST::store_t *__restrict__ Cn,
        const StepParams<typename ST::T> p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    typename ST::T f[QT<QID>::Q];
    pull_seq<QID, ST>(C, p, x, y, f, LBM_SEQ);
    collide_any<QID, COLL>(f, p, x, y);
    store_local<QID, ST>(Cn, p, (long long)x + (long long)p.nx * y, f, LBM_SEQ);
}

// apply!(bcs, q, f_new, f_old) in place on f_new (only nodes within MAXC of an edge)
template <int QID, class ST, int... Is>
__device__ __forceinline__ void apply_only(const typename ST::store_t *__restrict__ fold,
                                           typename ST::store_t *__restrict__ fnew,
                                           const StepParams<typename ST::T> &p, int x, int y,
                                           std::integer_sequence<int, Is...> seq) {
    using T = typename ST::T;
    constexpr int Q = QT<QID>::Q;
    const long long node = (long long)x + (long long)p.nx * y;
    T f[Q];
    bool touched[Q];
    ((touched[Is] = false), ...);
    ((f[Is] = T(0)), ...);
    for (int b = 0; b < p.nbc; ++b) {
        const BCDev bc = p.bcs[b];
        const bool inr = bc.kind == 1 || (x >= bc.x0 && x <= bc.x1 && y >= bc.y0 && y <= bc.y1);
        if (!inr) continue;
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i), j = QT<QID>::opp(i);
            bool hit;
            switch (bc.dir) {
            case 0: hit = (y - cy) >= p.ny; break;
            case 2: hit = (y - cy) < 0; break;
            case 1: hit = (x - cx) >= p.nx; break;
            default: hit = (x - cx) < 0; break;
            }
            if (hit) {
                T v = ST::load(fold + p.plane * j, node, j, p);
                if (bc.kind == 1) v = v + static_cast<T>(p.mw[b * Q + i]);
                f[i] = v;
                touched[i] = true;
            }
        };
        (one(std::integral_constant<int, Is>{}), ...);
    }
    auto st = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        if (touched[i]) ST::store(fnew + p.plane * i, node, i, f[i], p);
    };
    (st(std::integral_constant<int, Is>{}), ...);
}

template <int QID, class ST>
__global__ void __launch_bounds__(256)
k_apply_bc(const typename ST::store_t *__restrict__ fold, typename ST::store_t *__restrict__ fnew,
           const StepParams<typename ST::T> p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    constexpr int M = QT<QID>::MAXC;
    const bool edge = (x < M) | (x >= p.nx - M) | (y < M) | (y >= p.ny - M);
    if (!edge) return;
    apply_only<QID, ST>(fold, fnew, p, x, y, LBM_SEQ);
}

// plain copy of n elements of U (field copy; also the HBM calibration kernel)
template <class U>
__global__ void __launch_bounds__(256) k_copy(const U *__restrict__ src, U *__restrict__ dst,
                                              long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = i; k < n; k += stride) dst[k] = src[k];
}

// ---------------------------------------------------------------------------
// host-side launch helper
// ---------------------------------------------------------------------------
inline dim3 node_grid(int nx, int ny, int bx) { return dim3((nx + bx - 1) / bx, ny, 1); }

inline int default_block(int nx) {
    if (nx >= 256) return 256;
    if (nx >= 128) return 128;
    return 64;
}

template <int QID, class ST>
int launch_op(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    using S = typename ST::store_t;
    const int bx = L.block > 0 ? L.block : default_block(p.nx);
    const dim3 grid = node_grid(p.nx, p.ny, bx), block(bx, 1, 1);
    const S *src = static_cast<const S *>(L.src);
    S *dst = static_cast<S *>(L.dst);
    switch (L.op) {
    case Op::Collide:
        if (L.coll == 0) k_collide<QID, 0, ST><<<grid, block, 0, s>>>(src, dst, p);
        else if (L.coll == 1) k_collide<QID, 1, ST><<<grid, block, 0, s>>>(src, dst, p);
        else if constexpr (ST::id == 0) k_collide<QID, 2, ST><<<grid, block, 0, s>>>(src, dst, p);
        else return -2;
        break;
    case Op::Fused:
        if (L.coll == 0) k_fused<QID, 0, ST><<<grid, block, 0, s>>>(src, dst, p);
        else if (L.coll == 1) k_fused<QID, 1, ST><<<grid, block, 0, s>>>(src, dst, p);
        else if constexpr (ST::id == 0) k_fused<QID, 2, ST><<<grid, block, 0, s>>>(src, dst, p);
        else return -2;
        break;
    case Op::StreamBC:
        k_stream_bc<QID, ST><<<grid, block, 0, s>>>(src, dst, p);
        break;
    case Op::ApplyBC:
        k_apply_bc<QID, ST><<<grid, block, 0, s>>>(src, dst, p);
        break;
    default:
        return -3;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <class ST>
int dispatch_q(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    switch (L.qid) {
    case 4: return launch_op<4, ST>(L, p, s);
    case 5: return launch_op<5, ST>(L, p, s);
    case 9: return launch_op<9, ST>(L, p, s);
    case 13: return launch_op<13, ST>(L, p, s);
    case 17: return launch_op<17, ST>(L, p, s);
    case 21: return launch_op<21, ST>(L, p, s);
    case 37: return launch_op<37, ST>(L, p, s);
    default: return -1;
    }
}

}  // namespace lbm
