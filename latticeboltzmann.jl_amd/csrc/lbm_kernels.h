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

#include <type_traits>
#include <utility>

#include "lbm_internal.h"
#include "lbm_quad.h"

namespace lbm {

// ---------------------------------------------------------------------------
// storage traits
// ---------------------------------------------------------------------------
// decode(stored, i) -> arithmetic value of population i; encode is its inverse.
template <class Derived, class S, class TT>
struct StBase {
    using store_t = S;
    using T = TT;
    static constexpr bool kNeq = false;
    // kShift: arithmetic runs on shifted populations g_i = f_i - w_i rho0
    static constexpr bool kShift = !std::is_same<TT, double>::value;
    __device__ static T load(const S *p, long long idx, int i, const StepParams<T> &q) {
        return Derived::decode(p[idx], i, q);
    }
    __device__ static void store(S *p, long long idx, int i, T v, const StepParams<T> &q) {
        p[idx] = Derived::encode(v, i, q);
    }
};

struct StF64 : StBase<StF64, double, double> {
    static constexpr int id = 0;
    __device__ static double decode(double s, int, const StepParams<double> &) { return s; }
    __device__ static double encode(double v, int, const StepParams<double> &) { return v; }
};

// The FP32-arithmetic storages all work on shifted populations
// g_i = f_i - w_i rho0: the collision never forms f_i ~ w_i, so FP32 keeps its
// relative precision on the deviations that carry rho - rho0 and u.
// FP32 storage: g_i in binary32.
struct StF32 : StBase<StF32, float, float> {
    static constexpr int id = 1;
    __device__ static float decode(float s, int, const StepParams<float> &) { return s; }
    __device__ static float encode(float v, int, const StepParams<float> &) { return v; }
};

// FP16 deviation storage: g_i * dscale in IEEE binary16 (round to nearest even).
struct StF16Dev : StBase<StF16Dev, __half, float> {
    static constexpr int id = 2;
    __device__ static float decode(__half s, int, const StepParams<float> &q) {
        return __half2float(s) * q.dscale_inv;
    }
    __device__ static __half encode(float v, int, const StepParams<float> &q) {
        return __float2half_rn(v * q.dscale);
    }
};

// BF16 deviation storage, same scheme with bfloat16 (8-bit significand).
struct StBF16Dev : StBase<StBF16Dev, __hip_bfloat16, float> {
    static constexpr int id = 3;
    __device__ static float decode(__hip_bfloat16 s, int, const StepParams<float> &q) {
        return __bfloat162float(s) * q.dscale_inv;
    }
    __device__ static __hip_bfloat16 encode(float v, int, const StepParams<float> &q) {
        return __float2bfloat16(v * q.dscale);
    }
};

// Non-equilibrium storage: per node one 8-byte equilibrium record (below)
// and, per population, the 16-bit non-equilibrium part
// (f_i - feq_i) * dscale; the arithmetic value is
// g_i = f_i - w_i rho0 = (feq_i - w_i rho0) + (f_i - feq_i).
// The field holds the node records first (one dwordx2 per neighbour read),
// then Q 16-bit planes.
//
// Equilibrium record.  For first- and second-order equilibria it holds
//   a = (drho - s rho |u|^2 / 2,  s rho ux,  s rho uy,  1 / (2 rho))
// so that feq_i - w_i rho0 = w_i (a0 + c_i.A + a3 (c_i.A)^2), A = (a1, a2):
// six operations per population instead of the full polynomial.  For the
// third- and fourth-order sets it holds (drho, ux, uy, 0).  In memory the
// record is four binary16 values (a0 scaled by 2^10).  The stored
// non-equilibrium parts are taken against the rounded record, so the record's
// rounding is carried exactly by them and costs no accuracy beyond the
// rounding of slightly larger non-equilibrium parts.
struct alignas(16) Mom4 {
    float a0, a1, a2, a3;
};

struct alignas(8) Rec16 {
    __half a0, a1, a2, a3;
};

template <class S, int ID>
struct StNeq {
    using store_t = S;
    using T = float;
    static constexpr int id = ID;
    static constexpr bool kNeq = true;
    static constexpr bool kShift = true;
    static constexpr float kA0 = 1024.0f;
    __device__ static float to_f(S v, const StepParams<float> &q) {
        if constexpr (ID == 4) return __half2float(v) * q.dscale_inv;
        else return __bfloat162float(v) * q.dscale_inv;
    }
    __device__ static S from_f(float v, const StepParams<float> &q) {
        if constexpr (ID == 4) return __float2half_rn(v * q.dscale);
        else return __float2bfloat16(v * q.dscale);
    }
    // an already scaled value to storage
    __device__ static S raw(float v) {
        if constexpr (ID == 4) return __float2half_rn(v);
        else return __float2bfloat16(v);
    }
    __device__ static Rec16 pack(const Mom4 &m) {
        return Rec16{__float2half_rn(m.a0 * kA0), __float2half_rn(m.a1), __float2half_rn(m.a2),
                     __float2half_rn(m.a3)};
    }
    __device__ static Mom4 unpack(const Rec16 &r) {
        return Mom4{__half2float(r.a0) * (1.0f / kA0), __half2float(r.a1), __half2float(r.a2),
                    __half2float(r.a3)};
    }
    __device__ static const Rec16 *recs(const S *base) { return reinterpret_cast<const Rec16 *>(base); }
    __device__ static Rec16 *recs(S *base) { return reinterpret_cast<Rec16 *>(base); }
    // population plane 0 starts after the 8-byte records = 4 * plane 16-bit words
    __device__ static const S *pops(const S *base, long long plane) { return base + 4 * plane; }
    __device__ static S *pops(S *base, long long plane) { return base + 4 * plane; }
};
using StF16Neq = StNeq<__half, 4>;
using StBF16Neq = StNeq<__hip_bfloat16, 5>;

// VEC consecutive stored values of one plane (one thread's x-run)
template <class S, int VEC>
struct alignas(sizeof(S)) Pack {
    S v[VEC];
};

// raw bit container of a Pack for non-temporal stores
template <int BYTES>
struct RawOf;
template <> struct RawOf<2> { using type = unsigned short; };
template <> struct RawOf<4> { using type = unsigned int; };
template <> struct RawOf<8> { using type = unsigned int __attribute__((ext_vector_type(2))); };
template <> struct RawOf<16> { using type = unsigned int __attribute__((ext_vector_type(4))); };
template <> struct RawOf<32> { using type = unsigned int __attribute__((ext_vector_type(8))); };

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
// velocity_distribution_function/quadratures.jl:3-19, 45-61, 95-124, 129-159).
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

template <int QID, class T, int... Is>
__device__ __forceinline__ void feq_all_seq(T rho, T ux, T uy, const StepParams<T> &p,
                                            T (&feq)[QT<QID>::Q],
                                            std::integer_sequence<int, Is...>) {
    const T usq = ux * ux + uy * uy;
    ((feq[Is] = feq_i<QID, Is, T>(rho, ux, uy, usq, p)), ...);
}

template <int QID, class T>
__device__ __forceinline__ void feq_all(T rho, T ux, T uy, const StepParams<T> &p,
                                        T (&feq)[QT<QID>::Q]) {
    feq_all_seq<QID, T>(rho, ux, uy, p, feq, std::make_integer_sequence<int, QT<QID>::Q>{});
}

// population I of node n of field F, in arithmetic precision
template <int QID, int I, class ST>
__device__ __forceinline__ typename ST::T ld_pop(const typename ST::store_t *__restrict__ F,
                                                 const StepParams<typename ST::T> &p,
                                                 long long n) {
    if constexpr (ST::kNeq) {
        return geq_rec<QID, I>(ST::unpack(ST::recs(F)[n]), p) +
               ST::to_f(ST::pops(F, p.plane)[p.plane * I + n], p);
    } else {
        return ST::load(F + p.plane * I, n, I, p);
    }
}

// feq_i - w_i rho0 = w_i (drho + rho (a1 + a2/2 + a3/6 + a4/24)), the shifted
// equilibrium of the FP32 paths
template <int QID, int I, class T>
__device__ __forceinline__ T feq_dev_i(T drho, T rho, T ux, T uy, T usq, const StepParams<T> &p) {
    constexpr int FEQ = QT<QID>::FEQ;
    const T ud = udot<QID, I>(ux, uy);
    const T a1 = p.s * ud;
    T poly = a1;
    if constexpr (FEQ >= 2) {
        const T ud2 = ud * ud;
        poly = poly + T(0.5) * (p.s2 * ud2 + p.neg_s * usq);
        if constexpr (FEQ >= 3) poly = poly + T(1.0 / 6.0) * ((p.s * ud) * (p.s2 * ud2 - p.three_s * usq));
        if constexpr (FEQ >= 4)
            poly = poly + T(1.0 / 24.0) * ((p.s4 * (ud2 * ud2) - (p.six_s3 * usq) * ud2) +
                                           p.three_s2 * (usq * usq));
    }
    return p.w[I] * (drho + rho * poly);
}

template <int QID, class T, int... Is>
__device__ __forceinline__ void feq_dev_all_seq(T drho, T ux, T uy, const StepParams<T> &p,
                                                T (&geq)[QT<QID>::Q],
                                                std::integer_sequence<int, Is...>) {
    const T usq = ux * ux + uy * uy;
    const T rho = p.rho0 + drho;
    ((geq[Is] = feq_dev_i<QID, Is, T>(drho, rho, ux, uy, usq, p)), ...);
}

template <int QID, class T>
__device__ __forceinline__ void feq_dev_all(T drho, T ux, T uy, const StepParams<T> &p,
                                            T (&geq)[QT<QID>::Q]) {
    feq_dev_all_seq<QID, T>(drho, ux, uy, p, geq, std::make_integer_sequence<int, QT<QID>::Q>{});
}

// 1/x: the hardware reciprocal (1 ulp) for the FP32-arithmetic storages, the
// correctly rounded quotient in FP64
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_fast(double x) { return 1.0 / x; }

// equilibrium record of (drho, u) for quadrature QID (see Mom4)
template <int QID>
__device__ __forceinline__ Mom4 make_rec(float drho, float ux, float uy, const StepParams<float> &p) {
    constexpr int FEQ = QT<QID>::FEQ;
    if constexpr (FEQ <= 2) {
        const float rho = p.rho0 + drho;
        const float sr = p.s * rho;
        const float a0 = FEQ == 2 ? drho - 0.5f * sr * (ux * ux + uy * uy) : drho;
        return Mom4{a0, sr * ux, sr * uy, FEQ == 2 ? 0.5f * rcp_fast(rho) : 0.0f};
    } else {
        return Mom4{drho, ux, uy, 0.0f};
    }
}

// A node's own shifted equilibrium, shared by all of its populations.  For
// FEQ >= 3 the expansion of maxwell_boltzmann_equilibrium.jl in t = c_i.u,
//   s t + (s^2 t^2 - s u^2)/2 + s t (s^2 t^2 - 3 s u^2)/6
//       + (s^4 t^4 - 6 s^3 u^2 t^2 + 3 s^2 u^4)/24,
// is a polynomial in t with per-node coefficients; with rho folded in,
// feq_i - w_i rho0 = w_i (k0 + t (k1 + t (k2 + t (k3 + t k4)))): four FMAs per
// population by Horner's rule.  FEQ <= 2 keeps the record form.
struct NodeEq {
    Mom4 rec;
    float k0, k1, k2, k3, k4, ux, uy;
};

template <int QID>
__device__ __forceinline__ NodeEq node_eq(float drho, float ux, float uy,
                                          const StepParams<float> &p) {
    constexpr int FEQ = QT<QID>::FEQ;
    NodeEq e{};
    if constexpr (FEQ <= 2) {
        e.rec = make_rec<QID>(drho, ux, uy, p);
    } else {
        const float rho = p.rho0 + drho;
        const float usq = ux * ux + uy * uy;
        const float s3 = p.s * p.s2;
        float c0 = -0.5f * p.s * usq;
        const float c1 = p.s - 0.5f * p.s2 * usq;
        float c2 = 0.5f * p.s2;
        const float c3 = (1.0f / 6.0f) * s3;
        float c4 = 0.0f;
        if constexpr (FEQ >= 4) {
            c0 = c0 + 0.125f * p.s2 * (usq * usq);
            c2 = c2 - 0.25f * s3 * usq;
            c4 = (1.0f / 24.0f) * p.s4;
        }
        e.k0 = drho + rho * c0;
        e.k1 = rho * c1;
        e.k2 = rho * c2;
        e.k3 = rho * c3;
        e.k4 = rho * c4;
        e.ux = ux;
        e.uy = uy;
    }
    return e;
}

// w_I (k0 + t (k1 + t (k2 + t (k3 + t k4)))) with t = c_I.u (FEQ >= 3)
template <int QID, int I>
__device__ __forceinline__ float horner_geq(const NodeEq &e, const StepParams<float> &p) {
    const float t = udot<QID, I>(e.ux, e.uy);
    float h = e.k3;
    if constexpr (QT<QID>::FEQ >= 4) h = h + t * e.k4;
    h = e.k2 + t * h;
    h = e.k1 + t * h;
    h = e.k0 + t * h;
    return p.w[I] * h;
}

// feq_I - w_I rho0 from an equilibrium record.  For FEQ >= 3 the record's
// (drho, u) go through the same per-node coefficients as a node's own
// equilibrium: the writer of a record (the non-equilibrium parts are taken
// against it) forms them once for all Q populations, and every reader
// evaluates the identical expression.
template <int QID, int I>
__device__ __forceinline__ float geq_rec(const Mom4 &m, const StepParams<float> &p) {
    constexpr int FEQ = QT<QID>::FEQ;
    if constexpr (FEQ <= 2) {
        const float cA = udot<QID, I>(m.a1, m.a2);
        if constexpr (FEQ == 2) return p.w[I] * ((m.a0 + cA) + m.a3 * (cA * cA));
        else return p.w[I] * (m.a0 + cA);
    } else if constexpr (FEQ == 3) {
        // node_eq's k0 + t (k1 + t (k2 + t k3)) with k0 = drho + rho c0 and
        // k_n = rho c_n, factored as drho + rho (c0 + t (c1 + t (c2 + t c3))):
        // a reader forms two per-record coefficients (c2, c3 are constants)
        // instead of four.  Every reader and the writer of a record use this
        // same expression.
        const float drho = m.a0, ux = m.a1, uy = m.a2;
        const float rho = p.rho0 + drho;
        const float usq = ux * ux + uy * uy;
        const float c0 = -0.5f * p.s * usq;
        const float c1 = p.s - 0.5f * p.s2 * usq;
        const float c2 = 0.5f * p.s2;
        const float c3 = (1.0f / 6.0f) * (p.s * p.s2);
        const float t = udot<QID, I>(ux, uy);
        return p.w[I] * (drho + rho * (c0 + t * (c1 + t * (c2 + t * c3))));
    } else {
        return horner_geq<QID, I>(node_eq<QID>(m.a0, m.a1, m.a2, p), p);
    }
}

template <int QID, int I>
__device__ __forceinline__ float geq_node(const NodeEq &e, const StepParams<float> &p) {
    if constexpr (QT<QID>::FEQ <= 2) return geq_rec<QID, I>(e.rec, p);
    else return horner_geq<QID, I>(e, p);
}

// shifted moments: drho = sum g, rho u = sum c g (sum_i c_i w_i = 0)
template <int QID, class T, int... Is>
__device__ __forceinline__ void moments_dev_seq(const T (&g)[QT<QID>::Q], const StepParams<T> &p,
                                                T &drho, T &ux, T &uy,
                                                std::integer_sequence<int, Is...>) {
    T r = T(0), mx = T(0), my = T(0);
    ((r = r + g[Is]), ...);
    ((mx = (QT<QID>::cx(Is) != 0 ? mx + cmul<QT<QID>::cx(Is)>(g[Is]) : mx)), ...);
    ((my = (QT<QID>::cy(Is) != 0 ? my + cmul<QT<QID>::cy(Is)>(g[Is]) : my)), ...);
    const T rho = p.rho0 + r;
    drho = r;
    const T ir = rcp_fast(rho);
    ux = mx * ir;
    uy = my * ir;
}

template <int QID, class T>
__device__ __forceinline__ void node_moments_dev(const T (&g)[QT<QID>::Q], const StepParams<T> &p,
                                                 T &drho, T &ux, T &uy) {
    moments_dev_seq<QID, T>(g, p, drho, ux, uy, std::make_integer_sequence<int, QT<QID>::Q>{});
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
template <int QID, int COLL, bool SHIFT, class T, int... Is>
__device__ __forceinline__ void collide_seq(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                            int y, std::integer_sequence<int, Is...>) {
    constexpr int Q = QT<QID>::Q;
    T rho, ux, uy, drho = T(0);
    if constexpr (SHIFT) {
        node_moments_dev<QID, T>(f, p, drho, ux, uy);
        rho = p.rho0 + drho;
    } else {
        node_moments<QID, T>(f, rho, ux, uy);
    }
    if (p.force_kind != 0) {
        T Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;  // srt.jl:54 / trt.jl:79
        uy = uy + p.tau_force * Fy;
    }
    T feq[Q];
    if constexpr (SHIFT) {
        // FP32 storages: the shifted equilibrium of the node (see NodeEq)
        const NodeEq e = node_eq<QID>(drho, ux, uy, p);
        ((feq[Is] = geq_node<QID, Is>(e, p)), ...);
    } else {
        const T usq = ux * ux + uy * uy;
        ((feq[Is] = feq_i<QID, Is, T>(rho, ux, uy, usq, p)), ...);
    }
    if constexpr (COLL == 0) {
        // srt.jl:57-59
        ((f[Is] = p.omo * f[Is] + p.om * feq[Is]), ...);
    } else if constexpr (SHIFT) {
        // trt.jl:82-94 rewritten on n = f - feq (FP32 storages): with
        // n+- = (n_i +- n_opp)/2, f_i - (n+/tau_s + n-/tau_a) = feq_i + ka n_i + kb n_opp
        const T ka = T(1) - T(0.5) * (p.oms + p.oma), kb = T(0.5) * (p.oma - p.oms);
        T n[Q];
        ((n[Is] = f[Is] - feq[Is]), ...);
        ((f[Is] = feq[Is] + (ka * n[Is] + kb * n[QT<QID>::opp(Is)])), ...);
    } else {
        // trt.jl:82-94 (each population reads the pre-collision pair)
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

// Collision for the non-equilibrium storages, done on n = f - feq: the
// relaxed part is (1 - 1/tau) n_i (SRT) or ka n_i + kb n_opp(i) (TRT, as
// above); the stored part adds the difference between the exact equilibrium
// and that of the rounded record.  Returned multiplied by the storage scale,
// with the rounded record of the node.
//
// For the first- and second-order sets it runs one opposite pair (i, j =
// opp(i), c_j = -c_i) at a time.  A node's shifted equilibrium splits into an
// even and an odd part: feq_i = w (Ev + Od), feq_j = w (Ev - Od) with the
// record's Ev = a0 + a3 (c.a)^2, Od = c.a.  The stored part
//   nq_i = k_a n_i + k_b n_j + (feq_i(exact) - feq_i(record)),
//   n_i = f_i - w (Ev + Od),  n_j = f_j - w (Ev - Od),
//   feq_i(exact) - feq_i(record) = w ((Ev - Ev_q) + (Od - Od_q))
// (subscript q: the rounded record; SRT: k_b = 0).  A pair costs two even and
// two odd parts instead of four full equilibria.  The non-equilibrium parts
// stay differences of nearly equal values (exact in FP32), as in the
// per-population form: relaxing the full populations instead (k_a f_i +
// k_b f_j + ...) rounds at their size and doubled the FP16-NEQ velocity
// error of the 4096^2 vortex after 5000 steps (5e-5 -> 1.4e-4).
template <int QID, int COLL, class ST, int... Is>
__device__ __forceinline__ void collide_neq_pairs(const float (&f)[QT<QID>::Q],
                                                  const StepParams<float> &p, int x, int y,
                                                  Rec16 &rec, float (&nq)[QT<QID>::Q],
                                                  std::integer_sequence<int, Is...>) {
    constexpr int FEQ = QT<QID>::FEQ;
    static_assert(FEQ <= 2, "the pair form uses the record form of the equilibrium");
    float drho, ux, uy;
    node_moments_dev<QID, float>(f, p, drho, ux, uy);
    if (p.force_kind != 0) {
        float Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;
        uy = uy + p.tau_force * Fy;
    }
    const Mom4 mt = make_rec<QID>(drho, ux, uy, p);
    rec = ST::pack(mt);
    const Mom4 mq = ST::unpack(rec);
    // even and odd parts of the node's (exact or rounded) equilibrium along c_I
    auto parts = [&](auto Ic, const Mom4 &m, float &ev, float &od) {
        constexpr int I = decltype(Ic)::value;
        od = udot<QID, I>(m.a1, m.a2);
        ev = FEQ == 2 ? m.a0 + m.a3 * (od * od) : m.a0;
    };
    float ka, kb;
    if constexpr (COLL == 0) {
        ka = p.omo * p.dscale;
        kb = 0.0f;
    } else {
        ka = (1.0f - 0.5f * (p.oms + p.oma)) * p.dscale;
        kb = (0.5f * (p.oma - p.oms)) * p.dscale;
    }
    auto pair = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        constexpr int j = QT<QID>::opp(i);
        if constexpr (i <= j) {
            const float w = p.w[i], wd = w * p.dscale;
            float ev, od, evq, odq;
            parts(Ic, mt, ev, od);
            parts(Ic, mq, evq, odq);
            const float dE = ev - evq;
            if constexpr (i == j) {
                nq[i] = (ka + kb) * (f[i] - w * ev) + wd * dE;
            } else {
                const float dO = od - odq;
                const float ni = f[i] - w * (ev + od), nj = f[j] - w * (ev - od);
                nq[i] = ka * ni + (kb * nj + wd * (dE + dO));
                nq[j] = ka * nj + (kb * ni + wd * (dE - dO));
            }
        }
    };
    (pair(std::integral_constant<int, Is>{}), ...);
}

// The per-population form of the same collision, for the third- and
// fourth-order sets: there the opposite-pair split of the Horner polynomial
// costs registers and scheduling (D2Q37 FP16-NEQ 8192 x 2048: 1.09 ms with
// the pair form, 0.89 ms with this one).
template <int QID, int COLL, class ST, int... Is>
__device__ __forceinline__ void collide_neq_seq(const float (&f)[QT<QID>::Q],
                                                const StepParams<float> &p, int x, int y,
                                                Rec16 &rec, float (&nq)[QT<QID>::Q],
                                                std::integer_sequence<int, Is...>) {
    constexpr int Q = QT<QID>::Q;
    float drho, ux, uy;
    node_moments_dev<QID, float>(f, p, drho, ux, uy);
    if (p.force_kind != 0) {
        float Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;
        uy = uy + p.tau_force * Fy;
    }
    // relax towards the exact equilibrium; store against the rounded record:
    // C_i - feq_i(record) = relax(n)_i + (feq_i(exact) - feq_i(record))
    const Mom4 mt = make_rec<QID>(drho, ux, uy, p);
    rec = ST::pack(mt);
    const Mom4 mq = ST::unpack(rec);
    const NodeEq e = node_eq<QID>(drho, ux, uy, p);
    float n[Q], d[Q];
    auto split = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        const float ge = geq_node<QID, i>(e, p);
        n[i] = f[i] - ge;
        d[i] = (ge - geq_rec<QID, i>(mq, p)) * p.dscale;
    };
    (split(std::integral_constant<int, Is>{}), ...);
    if constexpr (COLL == 0) {
        const float k = p.omo * p.dscale;
        ((nq[Is] = k * n[Is] + d[Is]), ...);
    } else {
        const float ka = (1.0f - 0.5f * (p.oms + p.oma)) * p.dscale;
        const float kb = (0.5f * (p.oma - p.oms)) * p.dscale;
        ((nq[Is] = (ka * n[Is] + kb * n[QT<QID>::opp(Is)]) + d[Is]), ...);
    }
}

// the NEQ collision of a node: the opposite-pair form for the first- and
// second-order sets, the per-population form for the third- and fourth-order
template <int QID, int COLL, class ST, int... Is>
__device__ __forceinline__ void collide_neq(const float (&f)[QT<QID>::Q], const StepParams<float> &p,
                                            int x, int y, Rec16 &rec, float (&nq)[QT<QID>::Q],
                                            std::integer_sequence<int, Is...> seq) {
    if constexpr (QT<QID>::FEQ <= 2) collide_neq_pairs<QID, COLL, ST>(f, p, x, y, rec, nq, seq);
    else collide_neq_seq<QID, COLL, ST>(f, p, x, y, rec, nq, seq);
}

template <int QID, int COLL, bool SHIFT, class T>
__device__ __forceinline__ void collide_node(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                             int y) {
    collide_seq<QID, COLL, SHIFT, T>(f, p, x, y, std::make_integer_sequence<int, QT<QID>::Q>{});
}

// ---------------------------------------------------------------------------
// regularised MRT (mrt.jl:56-118), FP64.
//
// The Hermite tensor H_n(c_i) (hermite_polynomials.jl:47-82) has entries
// A - cs B (+ cs^2 C for n = 4) with cs = 1/s and integers A, B, C that only
// depend on the abscissa and the flattened (column-major) tensor index k.
// They are formed here at compile time, so every table value is either an
// integer constant or one or two FP64 operations on cs: no table is read,
// and products with an all-zero entry are left out.  Value by value this is
// the host table's arithmetic (an integer term of zero adds an exact +0
// there), and a left-out product only adds a signed zero to the sums, so
// the populations equal the table-driven sums (mrt.jl:102, 106-115) bit for
// bit, up to the sign of an exactly zero result.
// ---------------------------------------------------------------------------
struct HermInt {
    int A, B, C;
};
LBM_HD constexpr int kdelta(int a, int b) { return a == b ? 1 : 0; }
template <int QID>
LBM_HD constexpr HermInt herm_int(int n, int i, int k) {
    const int xi[2] = {QT<QID>::cx(i), QT<QID>::cy(i)};
    const int a = k & 1, b = (k >> 1) & 1, g = (k >> 2) & 1, e = (k >> 3) & 1;
    if (n == 2) return HermInt{xi[b] * xi[a], kdelta(a, b), 0};
    if (n == 3)
        return HermInt{xi[g] * xi[b] * xi[a],
                       xi[a] * kdelta(b, g) + xi[b] * kdelta(a, g) + xi[g] * kdelta(a, b), 0};
    return HermInt{xi[e] * xi[g] * xi[b] * xi[a],
                   xi[a] * xi[b] * kdelta(g, e) + xi[a] * xi[g] * kdelta(b, e) +
                       xi[a] * xi[e] * kdelta(b, g) + xi[b] * xi[g] * kdelta(a, e) +
                       xi[b] * xi[e] * kdelta(a, g) + xi[g] * xi[e] * kdelta(a, b),
                   kdelta(a, b) * kdelta(g, e) + kdelta(a, g) * kdelta(b, e) +
                       kdelta(a, e) * kdelta(b, g)};
}
// the table value ((A - cs B) + cs^2 C), with the zero-integer terms dropped
template <int A, int B, int C, class T>
__device__ __forceinline__ T herm_value(T cs, T cs2) {
    if constexpr (B == 0 && C == 0) return T(A);
    else if constexpr (C == 0) return T(A) - cs * T(B);
    else if constexpr (B == 0) return T(A) + cs2 * T(C);
    else return (T(A) - cs * T(B)) + cs2 * T(C);
}
template <int QID, int N, int I, int K>
LBM_HD constexpr bool herm_nonzero() {
    constexpr HermInt h = herm_int<QID>(N, I, K);
    return h.A != 0 || h.B != 0 || h.C != 0;
}

// the first population with a nonzero entry k of H_n (Q if there is none)
template <int QID, int N, int K>
LBM_HD constexpr int herm_first() {
    for (int i = 0; i < QT<QID>::Q; ++i) {
        const HermInt h = herm_int<QID>(N, i, K);
        if (h.A != 0 || h.B != 0 || h.C != 0) return i;
    }
    return QT<QID>::Q;
}

// population i's terms of a_f[k] = sum_i f_i H_n(c_i)[k] (mrt.jl:102): the
// sums run over i in the outer loop, so each a_f[k] is the same left fold
// over i while f_i is consumed as it goes
template <int QID, int N, int I, class T, int... Ks>
__device__ __forceinline__ void mrt_project_pop(const T fi, T cs, T cs2, T (&a)[16],
                                                std::integer_sequence<int, Ks...>) {
    auto term = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (herm_nonzero<QID, N, I, k>()) {
            constexpr HermInt h = herm_int<QID>(N, I, k);
            const T v = fi * herm_value<h.A, h.B, h.C>(cs, cs2);
            if constexpr (herm_first<QID, N, k>() == I) a[k] = v;
            else a[k] = a[k] + v;
        }
    };
    (term(std::integral_constant<int, Ks>{}), ...);
}

// sum_k a_coll[k] H_n(c_i)[k] as a left fold over k (mrt.jl:106-115)
template <int QID, int N, int I, class T, int... Ks>
__device__ __forceinline__ T mrt_expand(const T (&ac)[16], T cs, T cs2,
                                        std::integer_sequence<int, Ks...>) {
    T d = T(0);
    bool first = true;
    auto term = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (herm_nonzero<QID, N, I, k>()) {
            constexpr HermInt h = herm_int<QID>(N, I, k);
            const T v = ac[k] * herm_value<h.A, h.B, h.C>(cs, cs2);
            d = first ? v : d + v;
            first = false;
        }
    };
    (term(std::integral_constant<int, Ks>{}), ...);
    return d;
}

// a_coll[k] = (1 - 1/tau_n) a_f[k] + (1/tau_n) a_eq[k] (mrt.jl:99-104)
template <int N, class T, int... Ks>
__device__ __forceinline__ void mrt_relax(const StepParams<T> &p, T rho, const T (&u)[2],
                                          T (&ac)[16], std::integer_sequence<int, Ks...>) {
    auto one = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        // a_eq (hermite.jl:45-77 at T = 1)
        T pr = u[k & 1] * u[(k >> 1) & 1];
        if constexpr (N >= 3) pr = pr * u[(k >> 2) & 1];
        if constexpr (N >= 4) pr = pr * u[(k >> 3) & 1];
        const T aeq = rho * pr;
        ac[k] = p.mrt_omo[N] * ac[k] + p.mrt_om[N] * aeq;  // mrt.jl:104
    };
    (one(std::integral_constant<int, Ks>{}), ...);
}

// RN(x / 3) for finite x without a division: q = RN(x RN(1/3)) is within
// one ulp of x / 3, its remainder x - 3q is exact in one fused operation, and
// RN(q + rem RN(1/3)) is then the correctly rounded quotient (Markstein's
// correction theorem; see Muller et al., Handbook of Floating-Point
// Arithmetic, 2nd ed., sec. 4.9).  Dividing by 3! = 2 * 3 and 4! = 8 * 3
// scales by a power of two first, which is exact away from the subnormals.
__device__ __forceinline__ double div3_rn(double x) {
    constexpr double r = 1.0 / 3.0;
    const double q = x * r;
    const double rem = __builtin_fma(-q, 3.0, x);
    return __builtin_fma(rem, r, q);
}

template <class F, int... Is>
__device__ __forceinline__ void mrt_each(F &&fn, std::integer_sequence<int, Is...>) {
    (fn(std::integral_constant<int, Is>{}), ...);
}

template <int QID, class T>
__device__ __forceinline__ void collide_mrt_node(T (&f)[QT<QID>::Q], const StepParams<T> &p,
                                                 int x, int y) {
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
    const T cs = p.cs_inv, cs2 = cs * cs;
    T ac2[16], ac3[16], ac4[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) ac2[k] = ac3[k] = ac4[k] = T(0);
    auto project = [&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (N >= 2)
            mrt_project_pop<QID, 2, i>(f[i], cs, cs2, ac2, std::make_integer_sequence<int, 4>{});
        if constexpr (N >= 3)
            mrt_project_pop<QID, 3, i>(f[i], cs, cs2, ac3, std::make_integer_sequence<int, 8>{});
        if constexpr (N >= 4)
            mrt_project_pop<QID, 4, i>(f[i], cs, cs2, ac4, std::make_integer_sequence<int, 16>{});
    };
    mrt_each(project, std::make_integer_sequence<int, QT<QID>::Q>{});
    if constexpr (N >= 2) mrt_relax<2>(p, rho, u, ac2, std::make_integer_sequence<int, 4>{});
    if constexpr (N >= 3) mrt_relax<3>(p, rho, u, ac3, std::make_integer_sequence<int, 8>{});
    if constexpr (N >= 4) mrt_relax<4>(p, rho, u, ac4, std::make_integer_sequence<int, 16>{});
    auto out = [&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const T ru = ux * static_cast<T>(QT<QID>::cx(i)) + uy * static_cast<T>(QT<QID>::cy(i));
        T hsum = T(0);
        if constexpr (N >= 2) {
            const T d = mrt_expand<QID, 2, i>(ac2, cs, cs2, std::make_integer_sequence<int, 4>{});
            hsum = (p.mrt_sn[2] * d) / T(2);
        }
        if constexpr (N >= 3) {
            const T d = mrt_expand<QID, 3, i>(ac3, cs, cs2, std::make_integer_sequence<int, 8>{});
            hsum = hsum + div3_rn(T(0.5) * (p.mrt_sn[3] * d));       // (s^3 d) / 3!
        }
        if constexpr (N >= 4) {
            const T d = mrt_expand<QID, 4, i>(ac4, cs, cs2, std::make_integer_sequence<int, 16>{});
            hsum = hsum + div3_rn(T(0.125) * (p.mrt_sn[4] * d));     // (s^4 d) / 4!
        }
        f[i] = p.w[i] * ((rho + (p.s * rho) * ru) + hsum);  // mrt.jl:106-115
    };
    mrt_each(out, std::make_integer_sequence<int, QT<QID>::Q>{});
}


// ---------------------------------------------------------------------------
// boundary conditions of the ordered list for node (x, y), applied to the
// pulled populations f (bounce_back.jl:8-72, moving_wall.jl:17-38):
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
                T v = ld_pop<QID, j, ST>(C, p, node);
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
        ((f[Is] = ld_pop<QID, Is, ST>(
              C, p, node - QT<QID>::cx(Is) - (long long)nx * QT<QID>::cy(Is))),
         ...);
    } else {
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            const int sx = wrap(x - QT<QID>::cx(i), nx);
            const int sy = wrap(y - QT<QID>::cy(i), ny);
            f[i] = ld_pop<QID, i, ST>(C, p, (long long)sx + (long long)nx * sy);
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
    ((f[Is] = ld_pop<QID, Is, ST>(F, p, node)), ...);
}

// the rounded equilibrium record of a shifted population vector and its
// non-equilibrium parts f_i - feq_i(record)
template <int QID, class ST, int... Is>
__device__ __forceinline__ void neq_split_seq(const float (&f)[QT<QID>::Q],
                                              const StepParams<float> &p, Rec16 &rec,
                                              float (&g)[QT<QID>::Q],
                                              std::integer_sequence<int, Is...>) {
    float drho, ux, uy;
    node_moments_dev<QID, float>(f, p, drho, ux, uy);
    rec = ST::pack(make_rec<QID>(drho, ux, uy, p));
    const Mom4 m = ST::unpack(rec);
    ((g[Is] = f[Is] - geq_rec<QID, Is>(m, p)), ...);
}

template <int QID, class ST>
__device__ __forceinline__ void neq_split(const float (&f)[QT<QID>::Q], const StepParams<float> &p,
                                          Rec16 &rec, float (&g)[QT<QID>::Q]) {
    neq_split_seq<QID, ST>(f, p, rec, g, std::make_integer_sequence<int, QT<QID>::Q>{});
}

template <int QID, class ST, int... Is>
__device__ __forceinline__ void store_local(typename ST::store_t *__restrict__ F,
                                            const StepParams<typename ST::T> &p, long long node,
                                            const typename ST::T (&f)[QT<QID>::Q],
                                            std::integer_sequence<int, Is...>) {
    if constexpr (ST::kNeq) {
        Rec16 m;
        float g[QT<QID>::Q];
        neq_split<QID, ST>(f, p, m, g);
        ST::recs(F)[node] = m;
        auto *P = ST::pops(F, p.plane);
        ((P[p.plane * Is + node] = ST::from_f(g[Is], p)), ...);
    } else {
        ((ST::store(F + p.plane * Is, node, Is, f[Is], p)), ...);
    }
}

// collide!(::IterativeInitializationCollisionModel) (iterative_initialization.jl:41-58):
// f_i = (1 - 1/tau) f_i + (1/tau) (w_i rho + N_i), N the precomputed nonlinear
// part of the equilibrium of the prescribed velocity; FP64
template <int QID>
__device__ __forceinline__ void collide_mei_node(double (&f)[QT<QID>::Q],
                                                 const StepParams<double> &p, int x, int y) {
    constexpr int Q = QT<QID>::Q;
    double rho = f[0] + f[1];
#pragma unroll
    for (int i = 2; i < Q; ++i) rho = rho + f[i];
    const long long n = (long long)x + (long long)p.nx * y;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double feq = p.w[i] * rho + p.mei[p.plane * i + n];
        f[i] = p.omo * f[i] + p.om * feq;
    }
}

// FP64 runs on the populations themselves (the reference's arithmetic); the
// FP32 storages on shifted populations
template <int QID, int COLL, class T>
__device__ __forceinline__ void collide_any(T (&f)[QT<QID>::Q], const StepParams<T> &p, int x,
                                            int y) {
    constexpr bool SHIFT = !std::is_same<T, double>::value;
    if constexpr (COLL == 3) collide_mei_node<QID>(f, p, x, y);
    else if constexpr (COLL == 2) collide_mrt_node<QID, T>(f, p, x, y);
    else collide_node<QID, COLL, SHIFT, T>(f, p, x, y);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
#define LBM_SEQ std::make_integer_sequence<int, QT<QID>::Q>{}

// fn(std::integral_constant<int, i>) for every population i of QID
template <int QID, class F, int... Is>
__device__ __forceinline__ void each_node_seq(F &&fn, std::integer_sequence<int, Is...>) {
    (fn(std::integral_constant<int, Is>{}), ...);
}
template <int QID, class F>
__device__ __forceinline__ void adj_each_node(F &&fn) {
    each_node_seq<QID>(fn, std::make_integer_sequence<int, QT<QID>::Q>{});
}

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

// ---------------------------------------------------------------------------
// fused step, VEC consecutive nodes per thread.  In the interior every plane
// is read as one VEC-wide (unaligned for shifted planes) vector load and
// written as one aligned vector store; requires nx % VEC == 0 (host-checked).
// NT: non-temporal stores (the written lattice is not re-read this step).
// ---------------------------------------------------------------------------
template <int QID, class ST, int VEC, int... Is>
__device__ __forceinline__ void pull_interior_vec(const typename ST::store_t *__restrict__ C,
                                                  const StepParams<typename ST::T> &p,
                                                  long long node0,
                                                  typename ST::T (&f)[VEC][QT<QID>::Q],
                                                  std::integer_sequence<int, Is...>) {
    using S = typename ST::store_t;
    auto one = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        const long long src = node0 - QT<QID>::cx(i) - (long long)p.nx * QT<QID>::cy(i);
        Pack<S, VEC> pk;
        if constexpr (ST::kNeq) {
            __builtin_memcpy(&pk, ST::pops(C, p.plane) + p.plane * i + src, sizeof(pk));
            const Rec16 *m = ST::recs(C) + src;
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                f[v][i] = geq_rec<QID, i>(ST::unpack(m[v]), p) + ST::to_f(pk.v[v], p);
        } else {
            __builtin_memcpy(&pk, C + p.plane * i + src, sizeof(pk));
#pragma unroll
            for (int v = 0; v < VEC; ++v) f[v][i] = ST::decode(pk.v[v], i, p);
        }
    };
    (one(std::integral_constant<int, Is>{}), ...);
}

// Interior pull with wave-uniform row bases.  Every wavefront lies in one
// lattice row (blockDim.x is a multiple of 64), so with yw the row made
// uniform the plane and row offsets are scalar and a lane adds only a 32-bit
// byte offset (host guarantees nx < 2^27).
template <class U>
__device__ __forceinline__ const U *at32(const U *base, int idx) {
    return reinterpret_cast<const U *>(reinterpret_cast<const char *>(base) +
                                       static_cast<unsigned>(idx) * static_cast<unsigned>(sizeof(U)));
}
template <class U>
__device__ __forceinline__ U *at32(U *base, int idx) {
    return reinterpret_cast<U *>(reinterpret_cast<char *>(base) +
                                 static_cast<unsigned>(idx) * static_cast<unsigned>(sizeof(U)));
}

// Wave-uniform row addressing inside a loop.  at_s(base, idx) gives
// base + idx with the base made an opaque, wave-uniform value (see
// opaque_uniform) and the 32-bit lane offset formed at the access.  Without
// this the compiler hoists the zero-extended lane offset out of the loop,
// shares one 64-bit VGPR address between planes and adds each plane's
// offset with 64-bit VALU arithmetic, instead of using the scalar-base
// ("saddr") form of the global load or store.
template <class U>
using gptr = __attribute__((address_space(1))) U *;

// A wave-uniform 64-bit integer the compiler can neither hoist out of a loop
// nor treat as divergent: an asm statement hides its value (its result counts
// as divergent), readfirstlane restores the uniformity (a copy on an SGPR).
__device__ __forceinline__ long long opaque_uniform(long long v) {
    asm volatile("" : "+s"(v));
    const unsigned long long u = static_cast<unsigned long long>(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

__device__ __forceinline__ long long first_lane(long long v) {
    const unsigned long long u = static_cast<unsigned long long>(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// base + a byte offset: base wave-uniform (SGPR pair), off a lane's offset
template <class U>
__device__ __forceinline__ gptr<U> at_b(U *base, unsigned off) {
    const long long b = first_lane(reinterpret_cast<long long>(base));
    return reinterpret_cast<gptr<U>>(b + off);
}
// idx * size as a byte offset the compiler cannot hoist: formed once per loop
// iteration and shared by every access of that iteration (each stencil
// shift goes into the uniform base instead)
__device__ __forceinline__ unsigned lane_bytes(int idx, unsigned size) {
    unsigned off = static_cast<unsigned>(idx) * size;
    asm volatile("" : "+v"(off));
    return off;
}

// a load through a global pointer (copied as raw bits: class types cannot
// bind a reference in the global address space)
template <class U>
__device__ __forceinline__ typename std::remove_const<U>::type ld_g(gptr<U> ptr) {
    using V = typename std::remove_const<U>::type;
    using R = typename RawOf<sizeof(V)>::type;
    const R r = *reinterpret_cast<gptr<const R>>(ptr);
    V v;
    __builtin_memcpy(&v, &r, sizeof(V));
    return v;
}

template <class U>
__device__ __forceinline__ gptr<U> at_s(U *base, int idx) {
    const long long b = first_lane(reinterpret_cast<long long>(base));
    const unsigned off = static_cast<unsigned>(idx) * static_cast<unsigned>(sizeof(U));
    return reinterpret_cast<gptr<U>>(b + off);
}

template <int QID, class ST, int VEC, int... Is>
__device__ __forceinline__ void pull_interior_row(const typename ST::store_t *__restrict__ C,
                                                  const StepParams<typename ST::T> &p, int yw,
                                                  int x0, typename ST::T (&f)[VEC][QT<QID>::Q],
                                                  std::integer_sequence<int, Is...>) {
    using S = typename ST::store_t;
    auto one = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
        const long long row = (long long)p.nx * (yw - cy);
        Pack<S, VEC> pk;
        if constexpr (ST::kNeq) {
            __builtin_memcpy(&pk, at32(ST::pops(C, p.plane) + p.plane * i + row, x0 - cx),
                             sizeof(pk));
            const Rec16 *m = at32(ST::recs(C) + row, x0 - cx);
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                f[v][i] = geq_rec<QID, i>(ST::unpack(m[v]), p) + ST::to_f(pk.v[v], p);
        } else {
            __builtin_memcpy(&pk, at32(C + p.plane * i + row, x0 - cx), sizeof(pk));
#pragma unroll
            for (int v = 0; v < VEC; ++v) f[v][i] = ST::decode(pk.v[v], i, p);
        }
    };
    (one(std::integral_constant<int, Is>{}), ...);
}

template <bool NT, class S, int VEC>
__device__ __forceinline__ void put_pack(S *dst, const Pack<S, VEC> &pk) {
    if constexpr (NT) {
        using R = typename RawOf<sizeof(pk)>::type;
        R r;
        __builtin_memcpy(&r, &pk, sizeof(pk));
        __builtin_nontemporal_store(r, reinterpret_cast<R *>(dst));
    } else {
        __builtin_memcpy(dst, &pk, sizeof(pk));
    }
}
template <bool NT, class S, int VEC>
__device__ __forceinline__ void put_pack(gptr<S> dst, const Pack<S, VEC> &pk) {
    using R = typename RawOf<sizeof(pk)>::type;
    R r;
    __builtin_memcpy(&r, &pk, sizeof(pk));
    if constexpr (NT) __builtin_nontemporal_store(r, reinterpret_cast<gptr<R>>(dst));
    else *reinterpret_cast<gptr<R>>(dst) = r;
}

// Stores of VEC consecutive nodes starting at (x0, y).  ROW: wave-uniform row
// addressing (see pull_interior_row); otherwise 64-bit node indices.  NEQ
// storages store the collision's equilibrium records and the scaled
// non-equilibrium parts nq.
template <bool ROW, class U>
__device__ __forceinline__ U *node_ptr(U *plane_base, int nx, int y, int x0) {
    if constexpr (ROW) return at32(plane_base + (long long)nx * y, x0);
    else return plane_base + ((long long)nx * y + x0);
}

template <int QID, class ST, int VEC, bool NT, bool ROW, int... Is>
__device__ __forceinline__ void store_vec(typename ST::store_t *__restrict__ F,
                                          const StepParams<typename ST::T> &p, int y, int x0,
                                          const typename ST::T (&f)[VEC][QT<QID>::Q],
                                          std::integer_sequence<int, Is...>,
                                          const Rec16 *rec = nullptr,
                                          const float (*nq)[QT<QID>::Q] = nullptr) {
    using S = typename ST::store_t;
    if constexpr (ST::kNeq) {
        Rec16 *m = node_ptr<ROW>(ST::recs(F), p.nx, y, x0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            Pack<Rec16, 1> pm;
            pm.v[0] = rec[v];
            put_pack<NT>(m + v, pm);
        }
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            Pack<S, VEC> pk;
#pragma unroll
            for (int v = 0; v < VEC; ++v) pk.v[v] = ST::raw(nq[v][i]);
            put_pack<NT>(node_ptr<ROW>(ST::pops(F, p.plane) + p.plane * i, p.nx, y, x0), pk);
        };
        (one(std::integral_constant<int, Is>{}), ...);
    } else {
        auto one = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            Pack<S, VEC> pk;
#pragma unroll
            for (int v = 0; v < VEC; ++v) pk.v[v] = ST::encode(f[v][i], i, p);
            put_pack<NT>(node_ptr<ROW>(F + p.plane * i, p.nx, y, x0), pk);
        };
        (one(std::integral_constant<int, Is>{}), ...);
    }
}

// Collision and store of one node, one population (SRT) or one opposite pair
// (TRT) at a time, so that only the pulled populations stay live (not a
// second array of equilibria): wide stencils (D2Q17, D2Q37) otherwise hold
// 2 Q values per lane.  Same expressions as collide_seq (bit for bit).
template <int QID, int COLL, class ST, bool NT, bool ROW, int... Is>
__device__ __forceinline__ void collide_store_seq(const typename ST::T (&f)[QT<QID>::Q],
                                                  const StepParams<typename ST::T> &p, int x,
                                                  int y, typename ST::store_t *__restrict__ F,
                                                  std::integer_sequence<int, Is...>) {
    using T = typename ST::T;
    using S = typename ST::store_t;
    constexpr bool SHIFT = !std::is_same<T, double>::value;
    T rho, ux, uy, drho = T(0);
    if constexpr (SHIFT) {
        node_moments_dev<QID, T>(f, p, drho, ux, uy);
        rho = p.rho0 + drho;
    } else {
        node_moments<QID, T>(f, rho, ux, uy);
    }
    if (p.force_kind != 0) {
        T Fx, Fy;
        node_force(p, x, y, Fx, Fy);
        ux = ux + p.tau_force * Fx;
        uy = uy + p.tau_force * Fy;
    }
    const T usq = ux * ux + uy * uy;
    NodeEq e{};
    if constexpr (SHIFT) e = node_eq<QID>(drho, ux, uy, p);
    auto feq = [&](auto Ic) -> T {
        constexpr int i = decltype(Ic)::value;
        if constexpr (SHIFT) return geq_node<QID, i>(e, p);
        else return feq_i<QID, i, T>(rho, ux, uy, usq, p);
    };
    auto put = [&](auto Ic, T v) {
        constexpr int i = decltype(Ic)::value;
        Pack<S, 1> pk;
        pk.v[0] = ST::encode(v, i, p);
        put_pack<NT>(node_ptr<ROW>(F + p.plane * i, p.nx, y, x), pk);
    };
    auto one = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        constexpr int j = QT<QID>::opp(i);
        if constexpr (COLL == 0) {
            put(Ic, p.omo * f[i] + p.om * feq(Ic));  // srt.jl:57-59
        } else if constexpr (i <= j) {
            const T ei = feq(Ic), ej = feq(std::integral_constant<int, j>{});
            if constexpr (SHIFT) {
                const T ka = T(1) - T(0.5) * (p.oms + p.oma), kb = T(0.5) * (p.oma - p.oms);
                const T ni = f[i] - ei, nj = f[j] - ej;
                put(Ic, ei + (ka * ni + kb * nj));
                if constexpr (j != i) put(std::integral_constant<int, j>{}, ej + (ka * nj + kb * ni));
            } else {
                // trt.jl:82-94 for i and for opp(i)
                auto trt = [&](T gi, T gj, T qi, T qj) {
                    const T feq_s = T(0.5) * (qi + qj);
                    const T feq_a = T(0.5) * (qi - qj);
                    const T f_s = T(0.5) * (gi + gj);
                    const T f_a = T(0.5) * (gi - gj);
                    return gi + ((-p.oms) * (f_s - feq_s) - p.oma * (f_a - feq_a));
                };
                put(Ic, trt(f[i], f[j], ei, ej));
                if constexpr (j != i) put(std::integral_constant<int, j>{}, trt(f[j], f[i], ej, ei));
            }
        }
    };
    (one(std::integral_constant<int, Is>{}), ...);
}

template <int QID, int COLL, class ST, int VEC, bool NT>
__global__ void __launch_bounds__(256)
k_fused_v(const typename ST::store_t *__restrict__ C, typename ST::store_t *__restrict__ Cn,
          const StepParams<typename ST::T> p) {
    constexpr int M = QT<QID>::MAXC;
    const int x0 = (blockIdx.x * blockDim.x + threadIdx.x) * VEC;
    // blockDim.y rows per block; a wavefront never spans two rows (blockDim.x
    // is a multiple of 64), so the row is wave-uniform
    const int y = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + threadIdx.y);
    if (x0 >= p.nx || y >= p.ny) return;
    typename ST::T f[VEC][QT<QID>::Q];
    const bool row = p.nx < (1 << 27);  // 32-bit byte offsets within a row
    const bool interior = (x0 >= M) && (x0 + VEC <= p.nx - M) && (y >= M) && (y < p.ny - M);
    if (interior && row) {
        pull_interior_row<QID, ST, VEC>(C, p, y, x0, f, LBM_SEQ);
    } else if (interior) {
        pull_interior_vec<QID, ST, VEC>(C, p, (long long)x0 + (long long)p.nx * y, f, LBM_SEQ);
    } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) pull_seq<QID, ST>(C, p, x0 + v, y, f[v], LBM_SEQ);
    }
    if constexpr (ST::kNeq) {
        Rec16 rec[VEC];
        float nq[VEC][QT<QID>::Q];
#pragma unroll
        for (int v = 0; v < VEC; ++v)
            collide_neq<QID, COLL, ST>(f[v], p, x0 + v, y, rec[v], nq[v], LBM_SEQ);
        if (row) store_vec<QID, ST, VEC, NT, true>(Cn, p, y, x0, f, LBM_SEQ, rec, nq);
        else store_vec<QID, ST, VEC, NT, false>(Cn, p, y, x0, f, LBM_SEQ, rec, nq);
    } else if constexpr (VEC == 1 && COLL != 2) {
        if (row) collide_store_seq<QID, COLL, ST, NT, true>(f[0], p, x0, y, Cn, LBM_SEQ);
        else collide_store_seq<QID, COLL, ST, NT, false>(f[0], p, x0, y, Cn, LBM_SEQ);
    } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) collide_any<QID, COLL>(f[v], p, x0 + v, y);
        if (row) store_vec<QID, ST, VEC, NT, true>(Cn, p, y, x0, f, LBM_SEQ);
        else store_vec<QID, ST, VEC, NT, false>(Cn, p, y, x0, f, LBM_SEQ);
    }
}

// Fused step over the flat node index n = x + nx y, one node per thread, for
// lattices whose rows are not a whole number of 128-byte lines.  A wave then
// holds 64 consecutive nodes from a multiple of 64, so its stores (and the
// loads of the populations that do not move along x) cover whole lines of a
// plane instead of straddling one at every row start.  Same pull, boundary
// conditions and collision as the row-shaped kernel, node by node.  D2Q9
// TRT on a 4100^2 lattice: 5.4 TB/s of algorithmic traffic in FP32 and
// 5.5 TB/s in FP64, against 4.3 and 5.0 TB/s with row-shaped blocks.
template <int QID, int COLL, class ST, bool NT>
__global__ void __launch_bounds__(256)
k_fused_flat(const typename ST::store_t *__restrict__ C, typename ST::store_t *__restrict__ Cn,
             const StepParams<typename ST::T> p) {
    using S = typename ST::store_t;
    const long long nodes = (long long)p.nx * p.ny;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    // row of n from the correctly rounded 1/nx, off by at most one
    long long y = (long long)((double)n * (1.0 / (double)p.nx));
    long long x = n - y * p.nx;
    if (x < 0) {
        --y;
        x += p.nx;
    } else if (x >= p.nx) {
        ++y;
        x -= p.nx;
    }
    typename ST::T f[QT<QID>::Q];
    pull_seq<QID, ST>(C, p, (int)x, (int)y, f, LBM_SEQ);
    if constexpr (COLL == 2) {
        collide_any<QID, 2>(f, p, (int)x, (int)y);
        auto st = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            Pack<S, 1> pk;
            pk.v[0] = ST::encode(f[i], i, p);
            put_pack<NT>(Cn + (p.plane * i + n), pk);
        };
        adj_each_node<QID>(st);
    } else {
        collide_store_seq<QID, COLL, ST, NT, false>(f, p, (int)x, (int)y, Cn, LBM_SEQ);
    }
}

// The nodes within M of the lattice edge (x < M, x >= nx - M, y < M or
// y >= ny - M), numbered: first the M bottom and M top rows, then the side
// columns of the rows between them.  Every other node has all its pull
// sources inside the lattice.
// The edge band: the MY rows at the top and bottom and the MX columns at the
// left and right of the lattice (every node when there is no interior),
// enumerated rows first.
__host__ __device__ inline long long edge_count(int nx, int ny, int MX, int MY) {
    if (ny <= 2 * MY || nx <= 2 * MX) return (long long)nx * ny;
    return 2LL * MY * nx + (long long)(ny - 2 * MY) * (2 * MX);
}
__host__ __device__ inline long long edge_count(int nx, int ny, int M) {
    return edge_count(nx, ny, M, M);
}
__device__ inline bool edge_node(long long k, int nx, int ny, int MX, int MY, int &x, int &y) {
    if (ny <= 2 * MY || nx <= 2 * MX) {
        if (k >= (long long)nx * ny) return false;
        y = (int)(k / nx);
        x = (int)(k - (long long)y * nx);
        return true;
    }
    const long long band = 2LL * MY * nx;
    if (k < band) {
        const int r = (int)(k / nx);
        x = (int)(k - (long long)r * nx);
        y = r < MY ? r : ny - 2 * MY + r;
        return true;
    }
    const long long j = k - band;
    if (j >= (long long)(ny - 2 * MY) * (2 * MX)) return false;
    const int r = (int)(j / (2 * MX)), c = (int)(j - (long long)r * (2 * MX));
    y = MY + r;
    x = c < MX ? c : nx - 2 * MX + c;
    return true;
}
__device__ inline bool edge_node(long long k, int nx, int ny, int M, int &x, int &y) {
    return edge_node(k, nx, ny, M, M, x, y);
}

// one edge node: general pull (wrap + boundary conditions), collide, store
template <int QID, int COLL, class ST>
__device__ __forceinline__ void neq_edge_update(const typename ST::store_t *__restrict__ C,
                                                typename ST::store_t *__restrict__ Cn,
                                                const StepParams<float> &p, int x, int y) {
    constexpr int Q = QT<QID>::Q;
    float f[Q];
    pull_seq<QID, ST>(C, p, x, y, f, LBM_SEQ);
    Rec16 rec;
    float nq[Q];
    collide_neq<QID, COLL, ST>(f, p, x, y, rec, nq, LBM_SEQ);
    const long long n = (long long)x + (long long)p.nx * y;
    ST::recs(Cn)[n] = rec;
    auto st = [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        ST::pops(Cn, p.plane)[p.plane * i + n] = ST::raw(nq[i]);
    };
    adj_each_node<QID>(st);
}

// Non-equilibrium storages: every thread advances R consecutive rows of one
// lattice column.  The equilibrium records of the 2M + 1 rows around the
// current one (M = the stencil's largest |c|) stay in LDS, packed (8 bytes
// each): 2M + 2 rotating row slots (4 for M = 1, 8 for M = 2, 3) of
// blockDim.x + 2M records (the block's columns and M halo columns on each
// side).  Every thread loads the record of its own column for the row ahead
// and one barrier per row publishes it, so a node loads one record instead of
// one per pulled population.  The blocks of the last grid column update the
// nodes within M of the lattice edge instead (general pull: wrap + boundary
// conditions), a share per row of blocks, so the edge runs alongside the
// interior rather than as a separate latency-bound launch.  Measured on
// MI355X (D2Q9 4096^2, FP16-NEQ): 8 rows, 204 us per step, against 305 us
// with one node per thread gathering every neighbour's record.
template <int QID, int COLL, class ST, int R, bool NT>
__global__ void __launch_bounds__(256)
k_fused_neq_rows(const typename ST::store_t *__restrict__ C,
                 typename ST::store_t *__restrict__ Cn, const StepParams<float> p) {
    constexpr int Q = QT<QID>::Q;
    constexpr int M = QT<QID>::MAXC;
    constexpr int SL = M == 1 ? 4 : 8;  // row slots: at least 2M + 2, a power of two
    static_assert(2 * M + 2 <= SL, "record window too small for the stencil");
    using S = typename ST::store_t;
    __shared__ Rec16 win[SL][256 + 2 * M];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x == (int)gridDim.x - 1) {
        // the last block column: this row of blocks' share of the edge nodes
        const long long ne = edge_count(p.nx, p.ny, M);
        const long long chunk = (ne + gridDim.y - 1) / gridDim.y;
        const long long k1 = min(ne, (long long)(blockIdx.y + 1) * chunk);
        for (long long k = (long long)blockIdx.y * chunk + tid; k < k1; k += blockDim.x) {
            int ex, ey;
            if (edge_node(k, p.nx, p.ny, M, ex, ey)) neq_edge_update<QID, COLL, ST>(C, Cn, p, ex, ey);
        }
        return;
    }
    const int xb = blockIdx.x * blockDim.x;  // first column of the block
    const int x = xb + tid;
    const int y0 = __builtin_amdgcn_readfirstlane((int)blockIdx.y * R);
    const int last = min((int)blockDim.x, p.nx - xb) - 1;  // last active thread
    const bool active = tid <= last;
    const Rec16 *rb = ST::recs(C);
    // halo columns, wrapped once (the launch needs nx > 2M)
    const int xl = xb - M + tid, xr = xb + last + 1 + tid;
    const int xlw = xl < 0 ? xl + p.nx : xl, xrw = xr >= p.nx ? xr - p.nx : xr;
    auto row_of = [&](int yy) {
        const int yw = yy < 0 ? yy + p.ny : (yy >= p.ny ? yy - p.ny : yy);
        return rb + (long long)p.nx * yw;
    };
    auto load_row = [&](int yy) {
        const Rec16 *r = row_of(yy);
        Rec16 *w = win[yy & (SL - 1)];
        if (active) w[M + tid] = ld_g(at_s(r, x));
        if (tid < M) {
            w[tid] = ld_g(at_s(r, xlw));
            w[M + last + 1 + tid] = ld_g(at_s(r, xrw));
        }
    };
#pragma unroll
    for (int k = -M; k < M; ++k) load_row(y0 + k);
    const bool xin = active && (x >= M) && (x < p.nx - M);
    // Software pipeline (one-speed stencils): the record of row y + 1 + M and
    // the populations of row y + 1 are loaded into registers while row y is
    // collided, and published / used one iteration later, so neither load's
    // latency sits between a row's barrier and its arithmetic.  Two register
    // sets alternate between even and odd rows (no copies).  D2Q9 FP16-NEQ
    // 4096^2: 194 us against 197 us; on the wide stencils the extra Q
    // registers cost more occupancy than the overlap gains (D2Q17 385 -> 423
    // us), so they load in the row itself.
    //
    // Addresses (up to nine planes): every plane row's base is formed in
    // scalar registers (the stencil shift included) from an opaque, per-row
    // uniform plane size and row index, and every access adds one per-row lane
    // byte offset, so the loads and stores take the scalar-base form with no
    // 64-bit VALU address arithmetic (D2Q9 row loop: 316 -> 238 VALU
    // instructions per node).  With 13 to 37 planes the 2Q live bases
    // overflow the scalar registers (D2Q37: 423 spilled SGPRs, more VALU
    // than before), so those keep 64-bit VGPR addresses.
    constexpr bool kPipe = M == 1;
    constexpr bool kScalarBase = Q <= 9;
    auto pop_in = [&](long long pl, int yy, unsigned o2, auto Ic) -> S {
        constexpr int i = decltype(Ic)::value;
        constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
        const S *row = ST::pops(C, pl) + pl * i + (long long)p.nx * (yy - cy);
        if constexpr (kScalarBase) return ld_g(at_b(row - cx, o2));
        else return *at32(row, x - cx);
    };
    Rec16 rc{}, rl{}, rr{};
    S hA[kPipe ? Q : 1], hB[kPipe ? Q : 1];
    auto fetch = [&](int yy, long long pl, unsigned o2, unsigned o8, S (&h)[kPipe ? Q : 1]) {
        const Rec16 *r = row_of(yy + M);
        if (active) {
            if constexpr (kScalarBase) rc = ld_g(at_b(r, o8));
            else rc = *at32(r, x);
        }
        if (tid < M) {
            rl = ld_g(at_s(r, xlw));
            rr = ld_g(at_s(r, xrw));
        }
        if constexpr (kPipe) {
            if (xin && yy >= M && yy < p.ny - M) {
                auto one = [&](auto Ic) {
                    constexpr int i = decltype(Ic)::value;
                    h[i] = pop_in(pl, yy, o2, Ic);
                };
                adj_each_node<QID>(one);
            }
        }
    };
    // one row; returns false past the last lattice row
    auto row = [&](int r, S (&hc)[kPipe ? Q : 1], S (&hn)[kPipe ? Q : 1]) -> bool {
        // scalar bases: an opaque uniform row, so no loop-carried (and
        // VGPR-held) row offsets
        const int y = kScalarBase ? __builtin_amdgcn_readfirstlane(y0 + r) : y0 + r;
        if (y >= p.ny) return false;
        const unsigned o2 = kScalarBase ? lane_bytes(x, sizeof(S)) : 0u;
        const unsigned o8 = kScalarBase ? lane_bytes(x, sizeof(Rec16)) : 0u;
        if constexpr (kPipe) {
            Rec16 *w = win[(y + M) & (SL - 1)];
            if (active) w[M + tid] = rc;
            if (tid < M) {
                w[tid] = rl;
                w[M + last + 1 + tid] = rr;
            }
        } else {
            load_row(y + M);
        }
        __syncthreads();
        // the plane size made opaque per row: the Q plane-row addresses are
        // formed when used instead of being kept live across the loop (2Q
        // registers for the wide stencils)
        long long pl;
        if constexpr (kScalarBase) {
            pl = opaque_uniform(p.plane);
        } else {
            pl = p.plane;
            asm volatile("" : "+s"(pl));
        }
        if constexpr (kPipe) {
            if (r + 1 < R && y + 1 < p.ny) fetch(y + 1, pl, o2, o8, hn);
        }
        if (xin && y >= M && y < p.ny - M) {
            float f[Q];
            auto one = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
                S hv;
                if constexpr (kPipe) hv = hc[i];
                else hv = pop_in(pl, y, o2, Ic);
                f[i] = geq_rec<QID, i>(ST::unpack(win[(y - cy) & (SL - 1)][M + tid - cx]), p) +
                       ST::to_f(hv, p);
            };
            adj_each_node<QID>(one);
            Rec16 rec;
            float nq[Q];
            collide_neq<QID, COLL, ST>(f, p, x, y, rec, nq, LBM_SEQ);
            Pack<Rec16, 1> pm;
            pm.v[0] = rec;
            Rec16 *rrow = ST::recs(Cn) + (long long)p.nx * y;
            if constexpr (kScalarBase) put_pack<NT>(at_b(rrow, o8), pm);
            else put_pack<NT>(at32(rrow, x), pm);
            auto st = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                Pack<S, 1> pk;
                pk.v[0] = ST::raw(nq[i]);
                S *row = ST::pops(Cn, pl) + pl * i + (long long)p.nx * y;
                if constexpr (kScalarBase) put_pack<NT>(at_b(row, o2), pk);
                else put_pack<NT>(at32(row, x), pk);
            };
            adj_each_node<QID>(st);
        }
        return true;
    };
    if constexpr (kPipe)
        fetch(y0, p.plane, kScalarBase ? lane_bytes(x, sizeof(S)) : 0u,
              kScalarBase ? lane_bytes(x, sizeof(Rec16)) : 0u, hA);
    if constexpr (kPipe) {
        // two rows per iteration: the register sets swap roles without copies
        static_assert(R % 2 == 0, "rows per thread come in pairs");
#pragma unroll 1
        for (int r = 0; r < R; r += 2) {
            if (!row(r, hA, hB)) break;
            if (!row(r + 1, hB, hA)) break;
        }
    } else {
        // one copy of the (large) row body: D2Q37 FP16-NEQ 8192 x 2048 ran
        // 1.14 ms with two copies against 0.91 ms with one
#pragma unroll 1
        for (int r = 0; r < R; ++r)
            if (!row(r, hA, hB)) break;
    }
}

// Equilibrium coefficients of one record (the third- and fourth-order sets),
// 20 bytes: k0, k1, k2 of node_eq, the density rho (k3 = rho c3 and k4 = rho c4
// are formed from it on reading, as node_eq forms them) and the record's two
// binary16 velocity components.
struct alignas(4) EqCoef {
    float k0, k1, k2, rho;
    __half ux, uy;
};
template <int QID, class ST>
__device__ __forceinline__ EqCoef eq_coef(const Rec16 &r, const StepParams<float> &p) {
    const Mom4 m = ST::unpack(r);
    const NodeEq e = node_eq<QID>(m.a0, m.a1, m.a2, p);
    return EqCoef{e.k0, e.k1, e.k2, p.rho0 + m.a0, r.a1, r.a2};
}
// geq_rec of the record the coefficients came from: the same Horner evaluation
template <int QID, int I>
__device__ __forceinline__ float horner_coef(const EqCoef &c, const StepParams<float> &p) {
    const float ux = __half2float(c.ux), uy = __half2float(c.uy);
    const float t = udot<QID, I>(ux, uy);
    const float s3 = p.s * p.s2;
    float h = c.rho * ((1.0f / 6.0f) * s3);                                  // k3
    if constexpr (QT<QID>::FEQ >= 4) h = h + t * (c.rho * ((1.0f / 24.0f) * p.s4));  // k4
    h = c.k2 + t * h;
    h = c.k1 + t * h;
    h = c.k0 + t * h;
    return p.w[I] * h;
}

// The row kernel for the fourth-order set (D2Q37) with a coefficient window:
// as a record enters the LDS window it is turned into its equilibrium
// coefficients (EqCoef, 20 bytes) once, and each of the Q nodes that pull
// from it evaluates one Horner polynomial instead of unpacking the record and
// forming the coefficients again.  Everything else (row marching, 2M + 2 row
// slots, halo columns, the edge nodes in the last block column) is as in
// k_fused_neq_rows.  D2Q37 FP16-NEQ 8192 x 2048 on MI355X, 256 threads: 747
// us per step with 8 rows per thread against 886 us for k_fused_neq_rows at
// its best shape.  For D2Q17, whose third-order coefficients are cheap to
// rebuild, the window's LDS (occupancy 3 instead of 4-5 waves per SIMD) cost
// more than it saved (438 against 382 us), so the third-order sets keep
// k_fused_neq_rows.  The window is sized for 256 threads: smaller blocks
// hold the same LDS and run fewer waves.
constexpr int kCwGroup = 8;
// (three waves per SIMD, what the window's LDS allows: the BF16 instance
// would otherwise take 191 VGPRs and two waves)
template <int QID, int COLL, class ST, int R, bool NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_fused_neq_rows_cw(const typename ST::store_t *__restrict__ C,
                    typename ST::store_t *__restrict__ Cn, const StepParams<float> p) {
    constexpr int Q = QT<QID>::Q;
    constexpr int M = QT<QID>::MAXC;
    static_assert(QT<QID>::FEQ >= 4 && M <= 3, "coefficient window: the fourth-order set");
    constexpr int SL = 8;  // row slots: at least 2M + 2, a power of two
    using S = typename ST::store_t;
    __shared__ EqCoef win[SL][256 + 2 * M];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x == (int)gridDim.x - 1) {
        const long long ne = edge_count(p.nx, p.ny, M);
        const long long chunk = (ne + gridDim.y - 1) / gridDim.y;
        const long long k1 = min(ne, (long long)(blockIdx.y + 1) * chunk);
        for (long long k = (long long)blockIdx.y * chunk + tid; k < k1; k += blockDim.x) {
            int ex, ey;
            if (edge_node(k, p.nx, p.ny, M, ex, ey)) neq_edge_update<QID, COLL, ST>(C, Cn, p, ex, ey);
        }
        return;
    }
    const int xb = blockIdx.x * blockDim.x;
    const int x = xb + tid;
    const int y0 = __builtin_amdgcn_readfirstlane((int)blockIdx.y * R);
    const int last = min((int)blockDim.x, p.nx - xb) - 1;
    const bool active = tid <= last;
    const Rec16 *rb = ST::recs(C);
    const int xl = xb - M + tid, xr = xb + last + 1 + tid;
    const int xlw = xl < 0 ? xl + p.nx : xl, xrw = xr >= p.nx ? xr - p.nx : xr;
    auto row_of = [&](int yy) {
        const int yw = yy < 0 ? yy + p.ny : (yy >= p.ny ? yy - p.ny : yy);
        return rb + (long long)p.nx * yw;
    };
    auto load_row = [&](int yy) {
        const Rec16 *r = row_of(yy);
        EqCoef *w = win[yy & (SL - 1)];
        if (active) w[M + tid] = eq_coef<QID, ST>(ld_g(at_s(r, x)), p);
        if (tid < M) {
            w[tid] = eq_coef<QID, ST>(ld_g(at_s(r, xlw)), p);
            w[M + last + 1 + tid] = eq_coef<QID, ST>(ld_g(at_s(r, xrw)), p);
        }
    };
#pragma unroll
    for (int k = -M; k < M; ++k) load_row(y0 + k);
    const bool xin = active && (x >= M) && (x < p.nx - M);
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        if (y >= p.ny) break;
        load_row(y + M);
        __syncthreads();
        long long pl = p.plane;
        asm volatile("" : "+s"(pl));
        if (xin && y >= M && y < p.ny - M) {
            // the Q stored parts first, all in flight together
            S hv[Q];
            auto ld = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
                const S *row = ST::pops(C, pl) + pl * i + (long long)p.nx * (y - cy);
                hv[i] = *at32(row, x - cx);
            };
            adj_each_node<QID>(ld);
            float f[Q];
            // groups of kCwGroup populations: the group's window entries are read
            // together, then evaluated, and each population is pinned where it
            // is formed (left to itself the scheduler sinks every Horner
            // evaluation to the collision and keeps all Q coefficient sets
            // live: D2Q37 326 VGPRs)
            auto group = [&](auto Gc) {
                constexpr int g0 = decltype(Gc)::value * kCwGroup;
                EqCoef e[kCwGroup];
                auto rd = [&](auto Kc) {
                    constexpr int i = g0 + decltype(Kc)::value;
                    if constexpr (i < Q) {
                        constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
                        e[i - g0] = win[(y - cy) & (SL - 1)][M + tid - cx];
                    }
                };
                mrt_each(rd, std::make_integer_sequence<int, kCwGroup>{});
                auto ev = [&](auto Kc) {
                    constexpr int i = g0 + decltype(Kc)::value;
                    if constexpr (i < Q) {
                        f[i] = horner_coef<QID, i>(e[i - g0], p) + ST::to_f(hv[i], p);
                        asm volatile("" : "+v"(f[i]));
                    }
                };
                mrt_each(ev, std::make_integer_sequence<int, kCwGroup>{});
                asm volatile("" ::: "memory");
            };
            mrt_each(group, std::make_integer_sequence<int, (Q + kCwGroup - 1) / kCwGroup>{});
            Rec16 rec;
            float nq[Q];
            collide_neq<QID, COLL, ST>(f, p, x, y, rec, nq, LBM_SEQ);
            Pack<Rec16, 1> pm;
            pm.v[0] = rec;
            put_pack<NT>(at32(ST::recs(Cn) + (long long)p.nx * y, x), pm);
            auto st = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                Pack<S, 1> pk;
                pk.v[0] = ST::raw(nq[i]);
                put_pack<NT>(at32(ST::pops(Cn, pl) + pl * i + (long long)p.nx * y, x), pk);
            };
            adj_each_node<QID>(st);
        }
    }
}

// The row kernel with two adjacent columns per lane, for the one-speed
// stencils (D2Q4/5/9) on lattices of even width.  Everything else is as in
// k_fused_neq_rows (LDS record window, software-pipelined rows, scalar row
// bases): a lane loads and stores each plane's two populations as one
// 32-bit access (unaligned for the planes shifted in x) and its two records
// as one 16-byte access, unpacks each of the four records of a window row
// it reads once for both nodes, and shares the per-row scalar address work
// between them.  The lane pairs are (x, x + 1) with x even, so the edge band
// the last block column updates is two columns wide in x.  D2Q9 FP16-NEQ
// 4096^2 on MI355X, 256 threads: 166-169 us per step with 2 rows per thread
// (13 global memory instructions per node instead of 24), against 179-186 us
// for one column per lane at its best row count (4); 1 row per thread, 182
// us.  Without the register double-buffer of the populations the kernel
// needs 80 VGPRs instead of 101 and runs the same at 2 rows, 3% slower at 4.
template <int QID, int COLL, class ST, int R, bool NT>
__global__ void __launch_bounds__(256)
k_fused_neq_rows2(const typename ST::store_t *__restrict__ C,
                  typename ST::store_t *__restrict__ Cn, const StepParams<float> p) {
    constexpr int Q = QT<QID>::Q;
    static_assert(QT<QID>::MAXC == 1 && Q <= 9, "two columns per lane: one-speed stencils");
    static_assert(R % 2 == 0, "rows per thread come in pairs");
    constexpr int SL = 4;  // row slots (2M + 2)
    constexpr int MX = 2;  // edge band width in x
    using S = typename ST::store_t;
    using P2 = Pack<S, 2>;
    using R2 = Pack<Rec16, 2>;
    // slot j + 2 holds column xb + j: the left halo at 1, a lane's pair
    // 16-byte aligned at 2 + 2 tid, the right halo after the last pair
    __shared__ Rec16 win[SL][2 * 256 + 4];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x == (int)gridDim.x - 1) {
        const long long ne = edge_count(p.nx, p.ny, MX, 1);
        const long long chunk = (ne + gridDim.y - 1) / gridDim.y;
        const long long k1 = min(ne, (long long)(blockIdx.y + 1) * chunk);
        for (long long k = (long long)blockIdx.y * chunk + tid; k < k1; k += blockDim.x) {
            int ex, ey;
            if (edge_node(k, p.nx, p.ny, MX, 1, ex, ey))
                neq_edge_update<QID, COLL, ST>(C, Cn, p, ex, ey);
        }
        return;
    }
    const int xb = blockIdx.x * (2 * blockDim.x);  // first column of the block
    const int x = xb + 2 * tid;                    // the lane's first column
    const int y0 = __builtin_amdgcn_readfirstlane((int)blockIdx.y * R);
    const int ncol = min(2 * (int)blockDim.x, p.nx - xb);  // even (nx even)
    const bool active = 2 * tid < ncol;
    const Rec16 *rb = ST::recs(C);
    const int xl = xb - 1, xr = xb + ncol;  // halo columns, wrapped once
    const int xlw = xl < 0 ? xl + p.nx : xl, xrw = xr >= p.nx ? xr - p.nx : xr;
    auto row_of = [&](int yy) {
        const int yw = yy < 0 ? yy + p.ny : (yy >= p.ny ? yy - p.ny : yy);
        return rb + (long long)p.nx * yw;
    };
    const bool xin = active && x >= MX && x + 1 < p.nx - MX;
    auto pop_in = [&](long long pl, int yy, unsigned o2, auto Ic) -> P2 {
        constexpr int i = decltype(Ic)::value;
        constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
        const S *row = ST::pops(C, pl) + pl * i + (long long)p.nx * (yy - cy);
        return ld_g(at_b(reinterpret_cast<const P2 *>(row - cx), o2));
    };
    R2 rc{};
    Rec16 rl{}, rr{};
    P2 hA[Q], hB[Q];
    auto fetch = [&](int yy, long long pl, unsigned o2, unsigned o8, P2 (&h)[Q]) {
        const Rec16 *r = row_of(yy + 1);
        if (active) rc = ld_g(at_b(reinterpret_cast<const R2 *>(r), o8));
        if (tid == 0) {
            rl = ld_g(at_s(r, xlw));
            rr = ld_g(at_s(r, xrw));
        }
        if (xin && yy >= 1 && yy < p.ny - 1) {
            auto one = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                h[i] = pop_in(pl, yy, o2, Ic);
            };
            adj_each_node<QID>(one);
        }
    };
    auto row = [&](int r, P2 (&hc)[Q], P2 (&hn)[Q]) -> bool {
        const int y = __builtin_amdgcn_readfirstlane(y0 + r);
        if (y >= p.ny) return false;
        // x * sizeof(S) bytes for the population pair, 4 x that for the records
        const unsigned o2 = lane_bytes(x, sizeof(S));
        const unsigned o8 = o2 * (unsigned)(sizeof(Rec16) / sizeof(S));
        {
            Rec16 *w = win[(y + 1) & (SL - 1)];
            if (active) *reinterpret_cast<R2 *>(w + 2 + 2 * tid) = rc;
            if (tid == 0) {
                w[1] = rl;
                w[2 + ncol] = rr;
            }
        }
        __syncthreads();
        const long long pl = opaque_uniform(p.plane);
        if (r + 1 < R && y + 1 < p.ny) fetch(y + 1, pl, o2, o8, hn);
        if (xin && y >= 1 && y < p.ny - 1) {
            float f[2][Q];
            // window rows y + 1, y, y - 1 (c_y = -1, 0, +1): the four records
            // of columns x - 1 .. x + 2, unpacked once for both nodes
            auto wrow = [&](auto Cy) {
                constexpr int CY = decltype(Cy)::value;
                const Rec16 *w = win[(y - CY) & (SL - 1)] + 1 + 2 * tid;
                Mom4 m[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) m[j] = ST::unpack(w[j]);
                auto one = [&](auto Ic) {
                    constexpr int i = decltype(Ic)::value;
                    constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
                    if constexpr (cy == CY) {
                        f[0][i] = geq_rec<QID, i>(m[1 - cx], p) + ST::to_f(hc[i].v[0], p);
                        f[1][i] = geq_rec<QID, i>(m[2 - cx], p) + ST::to_f(hc[i].v[1], p);
                    }
                };
                adj_each_node<QID>(one);
            };
            wrow(std::integral_constant<int, 1>{});
            wrow(std::integral_constant<int, 0>{});
            wrow(std::integral_constant<int, -1>{});
            R2 recs;
            float nq[2][Q];
            collide_neq<QID, COLL, ST>(f[0], p, x, y, recs.v[0], nq[0], LBM_SEQ);
            collide_neq<QID, COLL, ST>(f[1], p, x + 1, y, recs.v[1], nq[1], LBM_SEQ);
            Pack<R2, 1> pm;
            pm.v[0] = recs;
            Rec16 *rrow = ST::recs(Cn) + (long long)p.nx * y;
            put_pack<NT>(at_b(reinterpret_cast<R2 *>(rrow), o8), pm);
            auto st = [&](auto Ic) {
                constexpr int i = decltype(Ic)::value;
                Pack<P2, 1> pk;
                pk.v[0].v[0] = ST::raw(nq[0][i]);
                pk.v[0].v[1] = ST::raw(nq[1][i]);
                S *prow = ST::pops(Cn, pl) + pl * i + (long long)p.nx * y;
                put_pack<NT>(at_b(reinterpret_cast<P2 *>(prow), o2), pk);
            };
            adj_each_node<QID>(st);
        }
        return true;
    };
#pragma unroll
    for (int k = -1; k < 1; ++k) {
        const Rec16 *r = row_of(y0 + k);
        Rec16 *w = win[(y0 + k) & (SL - 1)];
        if (active) *reinterpret_cast<R2 *>(w + 2 + 2 * tid) = ld_g(at_s(reinterpret_cast<const R2 *>(r), x / 2));
        if (tid == 0) {
            w[1] = ld_g(at_s(r, xlw));
            w[2 + ncol] = ld_g(at_s(r, xrw));
        }
    }
    fetch(y0, p.plane, lane_bytes(x, sizeof(S)), lane_bytes(x, sizeof(Rec16)), hA);
#pragma unroll 1
    for (int r = 0; r < R; r += 2) {
        if (!row(r, hA, hB)) break;
        if (!row(r + 1, hB, hA)) break;
    }
}

// the edge nodes alone (lattices without interior nodes)
template <int QID, int COLL, class ST>
__global__ void __launch_bounds__(256)
k_fused_edge(const typename ST::store_t *__restrict__ C, typename ST::store_t *__restrict__ Cn,
             const StepParams<float> p) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int x, y;
    if (!edge_node(k, p.nx, p.ny, QT<QID>::MAXC, x, y)) return;
    neq_edge_update<QID, COLL, ST>(C, Cn, p, x, y);
}

// largest nodes-per-thread worth instantiating (keeps VEC * Q registers bounded)
template <int QID, class ST>
constexpr int max_vec() {
    constexpr int Q = QT<QID>::Q;
    constexpr int cap = sizeof(typename ST::store_t) >= 8 ? 2 : (sizeof(typename ST::store_t) == 4 ? 4 : 8);
    return (Q <= 9) ? cap : (Q <= 17 ? (cap < 4 ? cap : 4) : 2);
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
                T v = ld_pop<QID, j, ST>(fold, p, node);
                if (bc.kind == 1) v = v + static_cast<T>(p.mw[b * Q + i]);
                f[i] = v;
                touched[i] = true;
            }
        };
        (one(std::integral_constant<int, Is>{}), ...);
    }
    if constexpr (ST::kNeq) {
        // a node's 16-bit parts are relative to its own moments: re-encode the node
        bool any = false;
        ((any = any || touched[Is]), ...);
        if (!any) return;
        T g[Q];
        load_local<QID, ST>(fnew, p, node, g, seq);
        ((g[Is] = touched[Is] ? f[Is] : g[Is]), ...);
        store_local<QID, ST>(fnew, p, node, g, seq);
    } else {
        auto st = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            if (touched[i]) ST::store(fnew + p.plane * i, node, i, f[i], p);
        };
        (st(std::integral_constant<int, Is>{}), ...);
    }
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

template <int QID, int COLL, class ST, int VEC>
int launch_fused_vn(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    using S = typename ST::store_t;
    const int bx = L.block > 0 ? L.block : default_block(p.nx / VEC);
    const int by = (L.block_y > 0 && bx * L.block_y <= 256) ? L.block_y : 1;
    if constexpr (VEC == 1 && !ST::kNeq) {
        // rows that are not whole 128-byte lines: the flat-index kernel
        if (((long long)p.nx * (long long)sizeof(S)) % 128 != 0) {
            const int threads = bx * by;
            const long long nodes = (long long)p.nx * p.ny;
            const unsigned blocks = (unsigned)((nodes + threads - 1) / threads);
            const S *src = static_cast<const S *>(L.src);
            S *dst = static_cast<S *>(L.dst);
            if (L.nt) k_fused_flat<QID, COLL, ST, true><<<blocks, threads, 0, s>>>(src, dst, p);
            else k_fused_flat<QID, COLL, ST, false><<<blocks, threads, 0, s>>>(src, dst, p);
            return hipGetLastError() == hipSuccess ? 0 : -4;
        }
    }
    const dim3 grid((p.nx / VEC + bx - 1) / bx, (p.ny + by - 1) / by, 1), block(bx, by, 1);
    const S *src = static_cast<const S *>(L.src);
    S *dst = static_cast<S *>(L.dst);
    if (L.nt) k_fused_v<QID, COLL, ST, VEC, true><<<grid, block, 0, s>>>(src, dst, p);
    else k_fused_v<QID, COLL, ST, VEC, false><<<grid, block, 0, s>>>(src, dst, p);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// non-equilibrium storages, one-speed stencils: L.vec > 1 = rows per thread
template <int QID, int COLL, class ST, int R>
int launch_neq_rows_r(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    using S = typename ST::store_t;
    const int bx = L.block > 0 ? L.block : default_block(p.nx);
    const dim3 grid((p.nx + bx - 1) / bx, (p.ny + R - 1) / R, 1), block(bx, 1, 1);
    const S *src = static_cast<const S *>(L.src);
    S *dst = static_cast<S *>(L.dst);
    constexpr int M = QT<QID>::MAXC;
    if constexpr (M == 1 && QT<QID>::Q <= 9) {
        // two columns per lane on even widths (edge band two columns wide in x)
        if (p.nx % 2 == 0 && p.nx > 4 && p.ny > 2) {
            const dim3 g2((p.nx / 2 + bx - 1) / bx + 1, grid.y, 1);
            if (L.nt) k_fused_neq_rows2<QID, COLL, ST, R, true><<<g2, block, 0, s>>>(src, dst, p);
            else k_fused_neq_rows2<QID, COLL, ST, R, false><<<g2, block, 0, s>>>(src, dst, p);
            return hipGetLastError() == hipSuccess ? 0 : -4;
        }
    }
    if constexpr (QT<QID>::FEQ >= 4) {
        // the fourth-order set (D2Q37): the coefficient window
        if (p.nx > 2 * M && p.ny > 2 * M) {
            const dim3 g2(grid.x + 1, grid.y, 1);
            if (L.nt) k_fused_neq_rows_cw<QID, COLL, ST, R, true><<<g2, block, 0, s>>>(src, dst, p);
            else k_fused_neq_rows_cw<QID, COLL, ST, R, false><<<g2, block, 0, s>>>(src, dst, p);
            return hipGetLastError() == hipSuccess ? 0 : -4;
        }
    }
    if (p.nx > 2 * M && p.ny > 2 * M) {
        const dim3 g2(grid.x + 1, grid.y, 1);  // + one block column for the edge
        if (L.nt) k_fused_neq_rows<QID, COLL, ST, R, true><<<g2, block, 0, s>>>(src, dst, p);
        else k_fused_neq_rows<QID, COLL, ST, R, false><<<g2, block, 0, s>>>(src, dst, p);
    } else {
        const long long nedge = edge_count(p.nx, p.ny, M);
        k_fused_edge<QID, COLL, ST><<<(unsigned)((nedge + 255) / 256), 256, 0, s>>>(src, dst, p);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---------------------------------------------------------------------------
// Regularised MRT on the matrix cores (FP64 MFMA), periodic lattices.
//
// The Hermite tensors are symmetric, so the projection and reconstruction
// need only their distinct entries: component m = 0 (H0 = 1), 1, 2 (H1 = c)
// and, per order n = 2..N, one entry per (a, b) = (number of x, number of y
// indices), a + b = n -- 15 components for N = 4, 10 for N = 3, 6 for N = 2,
// padded to 16.  Per tile of 16 nodes:
//   a_f[m][node] = sum_i H_m(c_i) f_i[node]       (16 x Qp) x (Qp x 16)
//   f_i[node]    = sum_m W[i][m] a_coll[m][node]  (Qr x 16) x (16 x 16)
// with W[i][m] = w_i {1, s c_x, s c_y, s^n / n! binom(n, a) H_(a,b)(c_i)}
// (the binomial counts the equal entries of the full tensor) and a_coll the
// relaxed coefficients of mrt.jl:99-104 (order 0 and 1: rho and rho u with the
// force-shifted u, as in mrt.jl:106-115).  v_mfma_f64_16x16x4f64 takes A
// (16 x 4) with lane l holding A[l % 16][l / 16], B (4 x 16) with lane l
// holding B[l / 16][l % 16], and returns D (16 x 16) with lane l, register r
// holding D[l / 16 + 4 r][l % 16] (checked on gfx950 with a probe of known
// operands).  So lane l = 16 j + n pulls the
// populations i = j, j + 4, j + 8, ... of node n -- the B operand of the
// projection's K-steps as loaded -- and the A operands (coefficients) sit in
// LDS in lane order.  The projection's D goes through a per-wave LDS buffer
// to become the reconstruction's B.  Sums run in the matrix core's order:
// parity with the oracle is to rounding (1e-12 relative over a run), not bit
// for bit.
// ---------------------------------------------------------------------------
template <int QID>
struct MrtTile {
    static constexpr int Q = QT<QID>::Q;
    static constexpr int N = QT<QID>::ORDER / 2;
    static constexpr int KP = (Q + 3) / 4;    // projection K-steps (4 populations each)
    static constexpr int CR = (Q + 15) / 16;  // reconstruction row blocks (16 populations)
    static constexpr int NC = 3 + (N >= 2 ? 3 : 0) + (N >= 3 ? 4 : 0) + (N >= 4 ? 5 : 0);
    static_assert(NC <= 16, "distinct Hermite components exceed one 16-wide block");
};
// component m -> (order n, number a of x indices); m = 0: (0, 0), 1: (1, 1), 2: (1, 0)
LBM_HD constexpr void mrt_comp(int m, int &n, int &a) {
    if (m == 0) { n = 0; a = 0; return; }
    if (m <= 2) { n = 1; a = m == 1 ? 1 : 0; return; }
    int base = 3;
    for (int o = 2; o <= 4; ++o) {
        if (m < base + o + 1) { n = o; a = o - (m - base); return; }
        base += o + 1;
    }
    n = -1; a = 0;
}
LBM_HD constexpr int binom_small(int n, int k) {
    int r = 1;
    for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
    return r;
}
// H_(a,b)(c_i) of order n through the full-tensor entry whose first a indices
// are x (bit 0) and the rest y (bit 1)
template <int QID>
__device__ double mrt_herm_distinct(int n, int a, int i, double cs, double cs2) {
    if (n == 0) return 1.0;
    if (n == 1) return (double)(a == 1 ? QT<QID>::cx(i) : QT<QID>::cy(i));
    const int k = ((1 << (n - a)) - 1) << a;
    const HermInt h = herm_int<QID>(n, i, k);
    return ((double)h.A - cs * (double)h.B) + cs2 * (double)h.C;
}

template <int QID, bool NT>
__global__ void __launch_bounds__(256)
k_mrt_mfma(const double *__restrict__ C, double *__restrict__ Cn, const StepParams<double> p,
           int iters) {
    using TT = MrtTile<QID>;
    constexpr int Q = TT::Q, KP = TT::KP, CR = TT::CR, NC = TT::NC;
    using d4 = double __attribute__((ext_vector_type(4)));
    __shared__ double tabP[KP][64];       // projection A operands
    __shared__ double tabR[CR][4][64];    // reconstruction A operands
    __shared__ double xch[4][16][17];     // per wave: a_f of its 16 nodes
    const double cs = p.cs_inv, cs2 = cs * cs;
    for (int e = threadIdx.x; e < KP * 64; e += blockDim.x) {
        const int kk = e / 64, l = e % 64, m = l % 16, i = 4 * kk + l / 16;
        int n, a;
        mrt_comp(m, n, a);
        tabP[kk][l] = (m < NC && i < Q) ? mrt_herm_distinct<QID>(n, a, i, cs, cs2) : 0.0;
    }
    for (int e = threadIdx.x; e < CR * 4 * 64; e += blockDim.x) {
        const int c = e / 256, kk = (e / 64) % 4, l = e % 64;
        const int i = 16 * c + l % 16, m = 4 * kk + l / 16;
        double v = 0.0;
        if (m < NC && i < Q) {
            int n, a;
            mrt_comp(m, n, a);
            double fac = 1.0;
            if (n >= 1) fac = p.s;
            if (n >= 2) fac = p.mrt_sn[n] * (double)binom_small(n, a) / (double)(n == 2 ? 2 : (n == 3 ? 6 : 24));
            v = p.w[i] * fac * mrt_herm_distinct<QID>(n, a, i, cs, cs2);
        }
        tabR[c][kk][l] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = lane >> 4, nd = lane & 15;
    // this lane's reconstruction components m = 4 kk + j, packed (order + 1)
    // | (x count << 4) per byte; order 0 = padding
    unsigned tags = 0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int m = 4 * kk + j;
        int n = -1, a = 0;
        if (m < NC) mrt_comp(m, n, a);
        tags |= (unsigned)((n + 1) | (a << 4)) << (8 * kk);
    }
    // the pulled populations i = j + 4 t: byte offset of the source from the
    // node within plane i (interior tiles)
    int soff[KP];
#pragma unroll
    for (int t = 0; t < KP; ++t) {
        const int i = j + 4 * t < Q ? j + 4 * t : 0;
        soff[t] = 8 * (-QT<QID>::cy(i) * p.nx - QT<QID>::cx(i));
    }
    constexpr int M = QT<QID>::MAXC;
    const int tiles_x = p.nx / 16;
    // wave-uniform tile coordinates, stepped by the grid's stride each iteration
    const int stride = (int)gridDim.x * 4;
    const int t0 = (int)blockIdx.x * 4 + wv;
    int ty = t0 / tiles_x, tx = t0 - (t0 / tiles_x) * tiles_x;
    const int sty = stride / tiles_x, stx = stride - sty * tiles_x;
    // lane part of the byte offsets: plane j plus the node (loads and stores)
    const unsigned ld_lane = 8u * (unsigned)(j * p.plane);
    double bf[KP];
    // the populations of tile (tx, ty) into bf
    auto pull = [&](int ptx, int pty) {
        const bool plive = pty < p.ny;
        const int x = ptx * 16 + nd, y = plive ? pty : 0;
        const unsigned node8 = 8u * (unsigned)(p.nx * y + x);
        const bool inner = ptx * 16 >= M && ptx * 16 + 16 <= p.nx - M && y >= M && y < p.ny - M;
        if (inner) {
#pragma unroll
            for (int t = 0; t < KP; ++t) {
                // plane 4 t from a uniform base, plane j and the shift per lane
                const double *base = C + (long long)(4 * t) * p.plane;
                bf[t] = j + 4 * t < Q ? ld_g(at_b(base, ld_lane + node8 + (unsigned)soff[t])) : 0.0;
            }
        } else {
#pragma unroll
            for (int t = 0; t < KP; ++t) {
                double v = 0.0;
                const int i = j + 4 * t;
                if (plive && i < Q) {
                    int xs = x - QT<QID>::cx(i), ys = y - QT<QID>::cy(i);
                    xs = xs < 0 ? xs + p.nx : (xs >= p.nx ? xs - p.nx : xs);
                    ys = ys < 0 ? ys + p.ny : (ys >= p.ny ? ys - p.ny : ys);
                    v = C[p.plane * i + (long long)p.nx * ys + xs];
                }
                bf[t] = v;
            }
        }
    };
    pull(tx, ty);
    for (int it = 0; it < iters; ++it) {
        const bool live = ty < p.ny;
        const int x = tx * 16 + nd, y = live ? ty : 0;
        const unsigned node8 = 8u * (unsigned)(p.nx * y + x);
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < KP; ++t)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(tabP[t][lane], bf[t], acc, 0, 0, 0);
        // the next tile's populations load while this one relaxes and
        // reconstructs (into the registers the projection has consumed)
        tx += stx;
        ty += sty;
        if (tx >= tiles_x) {
            tx -= tiles_x;
            ++ty;
        }
        if (it + 1 < iters) pull(tx, ty);
#pragma unroll
        for (int r = 0; r < 4; ++r) xch[wv][nd][j + 4 * r] = acc[r];
        __builtin_amdgcn_wave_barrier();
        const double rho = xch[wv][nd][0];
        double ux = xch[wv][nd][1] / rho, uy = xch[wv][nd][2] / rho;
        if (p.force_kind != 0 && live) {
            double Fx, Fy;
            node_force(p, x, y, Fx, Fy);
            ux = ux + p.tau_force * Fx;  // mrt.jl:94
            uy = uy + p.tau_force * Fy;
        }
        double bc[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const unsigned tg = (tags >> (8 * kk)) & 0xffu;
            const int n = (int)(tg & 15u) - 1, a = (int)(tg >> 4);
            double v;
            if (n == 0) v = rho;
            else if (n == 1) v = rho * (a == 1 ? ux : uy);
            else if (n < 0) v = 0.0;
            else {
                double eq = rho;
#pragma unroll
                for (int q = 0; q < 4; ++q) eq = eq * (q < a ? ux : (q < n ? uy : 1.0));
                const double omo = n == 2 ? p.mrt_omo[2] : (n == 3 ? p.mrt_omo[3] : p.mrt_omo[4]);
                const double om = n == 2 ? p.mrt_om[2] : (n == 3 ? p.mrt_om[3] : p.mrt_om[4]);
                v = omo * xch[wv][nd][4 * kk + j] + om * eq;  // mrt.jl:99-104
            }
            bc[kk] = v;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < CR; ++c) {
            d4 o = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                o = __builtin_amdgcn_mfma_f64_16x16x4f64(tabR[c][kk][lane], bc[kk], o, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * c + 4 * r + j;
                if (live && i < Q) {
                    // plane 16 c + 4 r from a uniform base, plane j and the node per lane
                    double *base = Cn + (long long)(16 * c + 4 * r) * p.plane;
                    gptr<double> dst = at_b(base, ld_lane + node8);
                    if constexpr (NT) __builtin_nontemporal_store(o[r], dst);
                    else *dst = o[r];
                }
            }
        }
    }
}

// 16 tiles of 16 nodes per wave, four waves per block; periodic lattices
// (no boundary list) with nx a multiple of 16
template <int QID>
int launch_mrt_mfma(const void *src, void *dst, bool nt, const StepParams<double> &p,
                    hipStream_t s) {
    if (p.nbc != 0 || p.nx % 16 != 0) return -5;
    // 32-bit byte offsets: plane j <= 3 plus the node and a shift below 4 GiB
    if (4 * 8 * p.plane >= (1LL << 32)) return -5;
    const long long ntiles = (long long)(p.nx / 16) * p.ny;
    const long long blocks = (ntiles + 63) / 64;
    const int iters = (int)((ntiles + blocks * 4 - 1) / (blocks * 4));
    const double *C = static_cast<const double *>(src);
    double *Cn = static_cast<double *>(dst);
    if (nt) k_mrt_mfma<QID, true><<<(unsigned)blocks, 256, 0, s>>>(C, Cn, p, iters);
    else k_mrt_mfma<QID, false><<<(unsigned)blocks, 256, 0, s>>>(C, Cn, p, iters);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <int QID, int COLL, class ST>
int launch_fused_c(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    constexpr int MV = max_vec<QID, ST>();
    const int vec = L.vec < 1 ? 1 : L.vec;
    if constexpr (ST::kNeq) {
        if (vec > 1) {
            if (p.nx >= (1 << 27)) return -5;
            switch (vec) {
            case 2: return launch_neq_rows_r<QID, COLL, ST, 2>(L, p, s);
            case 4: return launch_neq_rows_r<QID, COLL, ST, 4>(L, p, s);
            case 8: return launch_neq_rows_r<QID, COLL, ST, 8>(L, p, s);
            case 16: return launch_neq_rows_r<QID, COLL, ST, 16>(L, p, s);
            default: return -5;
            }
        }
    }
    if (vec > MV || p.nx % vec != 0) return -5;
    switch (vec) {
    case 1: return launch_fused_vn<QID, COLL, ST, 1>(L, p, s);
    case 2: if constexpr (MV >= 2) return launch_fused_vn<QID, COLL, ST, 2>(L, p, s); break;
    case 4: if constexpr (MV >= 4) return launch_fused_vn<QID, COLL, ST, 4>(L, p, s); break;
    case 8: if constexpr (MV >= 8) return launch_fused_vn<QID, COLL, ST, 8>(L, p, s); break;
    default: break;
    }
    return -5;
}

template <int QID, class ST>
int launch_fused_v(const Launch &L, const StepParams<typename ST::T> &p, hipStream_t s) {
    return L.coll == 0 ? launch_fused_c<QID, 0, ST>(L, p, s) : launch_fused_c<QID, 1, ST>(L, p, s);
}

// The first step of a run, from the canonical state S_0, in one pass: the
// run otherwise starts with collide! over the whole lattice (C_0) and then
// pulls from C_0 in the first fused step, two passes over HBM.  Here a block
// collides the S_0 of its tile and a one-node halo into LDS, kept encoded as
// collide! would store it, then each node of the tile decodes its pulls from
// LDS, applies the boundary list (its own C_0 from LDS), collides and stores
// C_1.  The collision and boundary code is the fused kernel's, so in FP64 C_1
// is bit-identical to the two-pass path (the FP32-arithmetic storages agree to
// the rounding of their FP32 operations, which the two paths order
// differently).  One-speed stencils (halo 1), storages with one value per
// population; tiles of 64 x 8 nodes (the host requires nx % 64 == 0 and
// ny % 8 == 0), two nodes per thread; the halo costs 29% more collisions and
// reads, most of the re-read rows coming from L2.
template <int QID, int COLL, class ST, bool NT>
__global__ void __launch_bounds__(256)
k_fused_first(const typename ST::store_t *__restrict__ S, typename ST::store_t *__restrict__ Cn,
              const StepParams<typename ST::T> p) {
    constexpr int Q = QT<QID>::Q;
    static_assert(QT<QID>::MAXC == 1 && !ST::kNeq, "one-speed stencils, one value per population");
    using T = typename ST::T;
    using St = typename ST::store_t;
    constexpr int TX = 64, TY = 8, HX = TX + 2, HY = TY + 2, HN = HX * HY;
    __shared__ St c0[Q][HN];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    for (int h = tid; h < HN; h += blockDim.x) {
        const int hx = h % HX, hy = h / HX;
        const int gx = wrap(x0 - 1 + hx, p.nx), gy = wrap(y0 - 1 + hy, p.ny);
        T f[Q];
        load_local<QID, ST>(S, p, (long long)gx + (long long)p.nx * gy, f, LBM_SEQ);
        collide_any<QID, COLL>(f, p, gx, gy);
#pragma unroll
        for (int i = 0; i < Q; ++i) c0[i][h] = ST::encode(f[i], i, p);
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < TY / 4; ++r) {
        const int tx = tid % TX, ty = tid / TX + 4 * r;
        const int x = x0 + tx, y = y0 + ty;
        const int h = (ty + 1) * HX + (tx + 1);
        T f[Q];
        auto pull = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            f[i] = ST::decode(c0[i][h - QT<QID>::cx(i) - HX * QT<QID>::cy(i)], i, p);
        };
        adj_each_node<QID>(pull);
        // the boundary list (bounce_back.jl:8-72, moving_wall.jl:17-38) as in
        // apply_bc_seq, with the node's own post-collision populations from LDS
        if (p.nbc > 0 && (x == 0 || x == p.nx - 1 || y == 0 || y == p.ny - 1)) {
            for (int b = 0; b < p.nbc; ++b) {
                const BCDev bc = p.bcs[b];
                const bool inr =
                    bc.kind == 1 || (x >= bc.x0 && x <= bc.x1 && y >= bc.y0 && y <= bc.y1);
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
                        T v = ST::decode(c0[j][h], j, p);
                        if (bc.kind == 1) v = v + static_cast<T>(p.mw[b * Q + i]);
                        f[i] = v;
                    }
                };
                adj_each_node<QID>(one);
            }
        }
        collide_any<QID, COLL>(f, p, x, y);
        const long long node = (long long)x + (long long)p.nx * y;
        auto st = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            Pack<St, 1> pk;
            pk.v[0] = ST::encode(f[i], i, p);
            put_pack<NT>(Cn + p.plane * i + node, pk);
        };
        adj_each_node<QID>(st);
    }
}

// The same one-pass first step for the non-equilibrium storages (one-speed
// stencils): the halo tile's C_0 is what collide! stores -- each node's S_0
// decoded, collided and re-encoded (an 8-byte record and Q 16-bit parts) --
// kept in LDS; each node of the tile then decodes its pulls from it, applies
// the boundary list with its own decoded C_0, and collides and encodes as the
// fused NEQ kernels do.  Its FP32 operations are those of collide! and of the
// fused kernels; like any two NEQ kernels, the two-pass start and this one
// agree to the rounding of single 16-bit values.
template <int QID, int COLL, class ST, bool NT>
__global__ void __launch_bounds__(256)
k_fused_first_neq(const typename ST::store_t *__restrict__ S, typename ST::store_t *__restrict__ Cn,
                  const StepParams<float> p) {
    constexpr int Q = QT<QID>::Q;
    static_assert(QT<QID>::MAXC == 1 && ST::kNeq, "one-speed stencils, non-equilibrium storages");
    using St = typename ST::store_t;
    constexpr int TX = 64, TY = 8, HX = TX + 2, HY = TY + 2, HN = HX * HY;
    __shared__ Rec16 rc0[HN];
    __shared__ St nc0[Q][HN];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    for (int h = tid; h < HN; h += blockDim.x) {
        const int hx = h % HX, hy = h / HX;
        const int gx = wrap(x0 - 1 + hx, p.nx), gy = wrap(y0 - 1 + hy, p.ny);
        float f[Q];
        load_local<QID, ST>(S, p, (long long)gx + (long long)p.nx * gy, f, LBM_SEQ);
        collide_any<QID, COLL>(f, p, gx, gy);
        Rec16 m;
        float g[Q];
        neq_split<QID, ST>(f, p, m, g);
        rc0[h] = m;
#pragma unroll
        for (int i = 0; i < Q; ++i) nc0[i][h] = ST::from_f(g[i], p);
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < TY / 4; ++r) {
        const int tx = tid % TX, ty = tid / TX + 4 * r;
        const int x = x0 + tx, y = y0 + ty;
        const int h = (ty + 1) * HX + (tx + 1);
        float f[Q];
        auto pull = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            const int hs = h - QT<QID>::cx(i) - HX * QT<QID>::cy(i);
            f[i] = geq_rec<QID, i>(ST::unpack(rc0[hs]), p) + ST::to_f(nc0[i][hs], p);
        };
        adj_each_node<QID>(pull);
        // the boundary list as in apply_bc_seq, with the node's own decoded C_0
        if (p.nbc > 0 && (x == 0 || x == p.nx - 1 || y == 0 || y == p.ny - 1)) {
            for (int b = 0; b < p.nbc; ++b) {
                const BCDev bc = p.bcs[b];
                const bool inr =
                    bc.kind == 1 || (x >= bc.x0 && x <= bc.x1 && y >= bc.y0 && y <= bc.y1);
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
                        float v = geq_rec<QID, j>(ST::unpack(rc0[h]), p) + ST::to_f(nc0[j][h], p);
                        if (bc.kind == 1) v = v + static_cast<float>(p.mw[b * Q + i]);
                        f[i] = v;
                    }
                };
                adj_each_node<QID>(one);
            }
        }
        Rec16 rec;
        float nq[Q];
        collide_neq<QID, COLL, ST>(f, p, x, y, rec, nq, LBM_SEQ);
        const long long node = (long long)x + (long long)p.nx * y;
        Pack<Rec16, 1> pm;
        pm.v[0] = rec;
        put_pack<NT>(ST::recs(Cn) + node, pm);
        auto st = [&](auto Ic) {
            constexpr int i = decltype(Ic)::value;
            Pack<St, 1> pk;
            pk.v[0] = ST::raw(nq[i]);
            put_pack<NT>(ST::pops(Cn, p.plane) + p.plane * i + node, pk);
        };
        adj_each_node<QID>(st);
    }
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
        else if (L.coll == 2) {
            if constexpr (ST::id == 0) k_collide<QID, 2, ST><<<grid, block, 0, s>>>(src, dst, p);
            else return -2;
        } else if (L.coll == 3) {
            if constexpr (ST::id == 0) k_collide<QID, 3, ST><<<grid, block, 0, s>>>(src, dst, p);
            else return -2;
        } else {
            return -2;
        }
        break;
    case Op::Fused:
        if (L.coll == 3) return -2;  // collide-only model: the host runs collide + stream
        if (L.coll == 2) {
            // MRT (FP64 only): with non-temporal stores through the row-shaped
            // kernel, one node per thread; 16 nodes per thread asks for the
            // matrix-core tiles (periodic lattices)
            if constexpr (ST::id == 0) {
                if (L.vec == 16) return launch_mrt_mfma<QID>(L.src, L.dst, L.nt != 0, p, s);
                if (L.nt) return launch_fused_vn<QID, 2, ST, 1>(L, p, s);
                k_fused<QID, 2, ST><<<grid, block, 0, s>>>(src, dst, p);
            } else {
                return -2;
            }
            break;
        }
        if (L.vec > 1 || L.nt || ST::kNeq) return launch_fused_v<QID, ST>(L, p, s);
        if (L.coll == 0) k_fused<QID, 0, ST><<<grid, block, 0, s>>>(src, dst, p);
        else k_fused<QID, 1, ST><<<grid, block, 0, s>>>(src, dst, p);
        break;
    case Op::StreamBC:
        k_stream_bc<QID, ST><<<grid, block, 0, s>>>(src, dst, p);
        break;
    case Op::FusedFirst:
        if constexpr (!ST::kNeq && QT<QID>::MAXC == 1) {
            // MRT is FP64 only
            if (p.nx % 64 != 0 || p.ny % 8 != 0 || L.coll > (ST::id == 0 ? 2 : 1)) return -5;
            const dim3 g(p.nx / 64, p.ny / 8, 1);
            const bool nt = L.nt != 0;
            if (L.coll == 0) {
                if (nt) k_fused_first<QID, 0, ST, true><<<g, 256, 0, s>>>(src, dst, p);
                else k_fused_first<QID, 0, ST, false><<<g, 256, 0, s>>>(src, dst, p);
            } else if (L.coll == 1) {
                if (nt) k_fused_first<QID, 1, ST, true><<<g, 256, 0, s>>>(src, dst, p);
                else k_fused_first<QID, 1, ST, false><<<g, 256, 0, s>>>(src, dst, p);
            } else if constexpr (ST::id == 0) {
                if (nt) k_fused_first<QID, 2, ST, true><<<g, 256, 0, s>>>(src, dst, p);
                else k_fused_first<QID, 2, ST, false><<<g, 256, 0, s>>>(src, dst, p);
            }
        } else if constexpr (ST::kNeq && QT<QID>::MAXC == 1) {
            if (p.nx % 64 != 0 || p.ny % 8 != 0 || L.coll > 1) return -5;
            const dim3 g(p.nx / 64, p.ny / 8, 1);
            const bool nt = L.nt != 0;
            if (L.coll == 0) {
                if (nt) k_fused_first_neq<QID, 0, ST, true><<<g, 256, 0, s>>>(src, dst, p);
                else k_fused_first_neq<QID, 0, ST, false><<<g, 256, 0, s>>>(src, dst, p);
            } else {
                if (nt) k_fused_first_neq<QID, 1, ST, true><<<g, 256, 0, s>>>(src, dst, p);
                else k_fused_first_neq<QID, 1, ST, false><<<g, 256, 0, s>>>(src, dst, p);
            }
        } else {
            return -5;
        }
        break;
    case Op::ApplyBC:
        k_apply_bc<QID, ST><<<grid, block, 0, s>>>(src, dst, p);
        break;
    default:
        return -3;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---------------------------------------------------------------------------
// batched independent periodic lattices: a persistent grid strides over a
// table of tiles (VEC * blockDim.x consecutive nodes of one lattice each);
// lattice geometry and relaxation times come from the lattice table (uniform
// per tile, so they live in scalar registers and the plane bases are scalar:
// a lane adds a 32-bit offset).  A lane holds VEC nodes blockDim.x apart, so
// every load instruction stays coalesced, and issues all their loads before
// it collides any of them.  OP: 0 collide, 1 fused step, 2 stream.  All
// lattices share the ping-pong parity; stores are non-temporal.
// ---------------------------------------------------------------------------
template <int QID, class ST, int OP, int VEC>
__global__ void __launch_bounds__(256)
k_batch(const BatchLattice *__restrict__ lats, const BatchTile *__restrict__ tiles, int ntiles,
        typename ST::store_t *__restrict__ pool, int parity,
        const StepParams<typename ST::T> base) {
    using T = typename ST::T;
    using S = typename ST::store_t;
    constexpr int Q = QT<QID>::Q;
    static_assert(QT<QID>::MAXC == 1, "batches pull from the nearest neighbours only");
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const BatchTile tile = tiles[t];
        // the lattice record is read in place, field by field (a local copy
        // indexed by the parity would live in memory): with a wave-uniform
        // index every field is a scalar load and the plane bases stay scalar
        const BatchLattice &L = lats[__builtin_amdgcn_readfirstlane(tile.lat)];
        StepParams<T> p = base;
        p.nx = L.nx;
        p.ny = L.ny;
        p.plane = L.plane;
        p.omo = static_cast<T>(L.omo);
        p.om = static_cast<T>(L.om);
        p.oms = static_cast<T>(L.oms);
        p.oma = static_cast<T>(L.oma);
        const S *src = pool + L.off[parity];
        S *dst = pool + L.off[parity ^ 1];
        const int last = (int)L.plane - 1, pitch = L.pitch;
        T f[VEC][Q];
        int nn[VEC], xs[VEC], ys[VEC];
        bool live[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            // a lane past the lattice's end or on a row's padding loads the
            // row's last node and stores nothing
            const int n0 = tile.n0 + (int)threadIdx.x + v * (int)blockDim.x;
            const int s0 = n0 <= last ? n0 : last;
            int y = (int)((double)s0 * L.inv_pitch);
            int x = s0 - y * pitch;
            if (x < 0) {
                --y;
                x += pitch;
            } else if (x >= pitch) {
                ++y;
                x -= pitch;
            }
            live[v] = n0 <= last && x < L.nx;
            if (x >= L.nx) x = L.nx - 1;
            const int n = x + pitch * y;
            nn[v] = n;
            xs[v] = x;
            ys[v] = y;
            if constexpr (OP == 0) {
                auto ld = [&](auto Ic) {
                    constexpr int i = decltype(Ic)::value;
                    f[v][i] = ST::decode(*at32(src + p.plane * i, n), i, p);
                };
                adj_each_node<QID>(ld);
            } else {
                // periodic pull without a branch: the wrapped neighbour columns
                // and rows are selected per node, so the lanes of a wave that
                // holds a lattice edge (most waves, for lattices narrower than
                // a few hundred nodes) do not split into an interior and an
                // edge path
                const int xm = x == 0 ? L.nx - 1 : x - 1, xp = x == L.nx - 1 ? 0 : x + 1;
                const int ym = y == 0 ? L.ny - 1 : y - 1, yp = y == L.ny - 1 ? 0 : y + 1;
                auto ld = [&](auto Ic) {
                    constexpr int i = decltype(Ic)::value;
                    constexpr int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
                    const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
                    const int sy = cy == 0 ? y : (cy > 0 ? ym : yp);
                    f[v][i] = ST::decode(*at32(src + p.plane * i, sx + pitch * sy), i, p);
                };
                adj_each_node<QID>(ld);
            }
        }
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            if constexpr (OP != 2) {
                if (L.coll == 0) collide_any<QID, 0>(f[v], p, xs[v], ys[v]);
                else collide_any<QID, 1>(f[v], p, xs[v], ys[v]);
            }
            if (live[v]) {
                auto st = [&](auto Ic) {
                    constexpr int i = decltype(Ic)::value;
                    Pack<S, 1> pk;
                    pk.v[0] = ST::encode(f[v][i], i, p);
                    put_pack<true>(at32(dst + p.plane * i, nn[v]), pk);
                };
                adj_each_node<QID>(st);
            }
        }
    }
}

template <int QID, class ST, int VEC>
int launch_batch(int op, const BatchLattice *lats, const BatchTile *tiles, int ntiles,
                 typename ST::store_t *pl, int parity, int block, int grid,
                 const StepParams<typename ST::T> &p, hipStream_t s) {
    if (op == 0) k_batch<QID, ST, 0, VEC><<<grid, block, 0, s>>>(lats, tiles, ntiles, pl, parity, p);
    else if (op == 1) k_batch<QID, ST, 1, VEC><<<grid, block, 0, s>>>(lats, tiles, ntiles, pl, parity, p);
    else k_batch<QID, ST, 2, VEC><<<grid, block, 0, s>>>(lats, tiles, ntiles, pl, parity, p);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

#define LBM_DEFINE_BATCH(TAG, ST, QID)                                                       \
    int run_batch_##TAG##_q##QID(int op, const BatchLattice *lats, const BatchTile *tiles,   \
                                 int ntiles, void *pool, int parity, int block, int grid,    \
                                 int vec, const StepParams<typename ST::T> &p,               \
                                 hipStream_t s) {                                            \
        using S = typename ST::store_t;                                                      \
        S *pl = static_cast<S *>(pool);                                                      \
        if (vec == 1)                                                                        \
            return launch_batch<QID, ST, 1>(op, lats, tiles, ntiles, pl, parity, block, grid, p, s); \
        if (vec == 2)                                                                        \
            return launch_batch<QID, ST, 2>(op, lats, tiles, ntiles, pl, parity, block, grid, p, s); \
        if (vec == 4)                                                                        \
            return launch_batch<QID, ST, 4>(op, lats, tiles, ntiles, pl, parity, block, grid, p, s); \
        return -2;                                                                           \
    }

// One translation unit per (storage, quadrature group) instantiates these
// entry points; lbm_dispatch.cpp routes a launch to them.
#define LBM_DEFINE_RUN(TAG, ST, QID)                                                        \
    int run_##TAG##_q##QID(const Launch &L, const StepParams<typename ST::T> &p,           \
                           hipStream_t s) {                                                 \
        return launch_op<QID, ST>(L, p, s);                                                 \
    }

}  // namespace lbm
