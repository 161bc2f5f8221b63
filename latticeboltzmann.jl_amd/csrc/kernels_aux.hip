// kernels_aux.hip -- moments, FP64 reductions and storage conversion.
// Compiled with -ffp-contract=off (FP64 arithmetic in the reference's order).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "lbm_internal.h"
#include "lbm_kernels.h"
#include "lbm_quad.h"

namespace lbm {

// The FP32-arithmetic storages hold shifted populations g_i = f_i - w_i rho0
// with the weights and rho0 the kernels use (binary32 values).
__device__ __forceinline__ double shift_of(const StepParams<double> &p, int i) {
    return (double)(float)p.w[i] * (double)(float)p.rho0;
}

// population i at linear node n, reconstructed in FP64 from any storage
template <int SID>
__device__ __forceinline__ double load_d(const void *base, long long plane, long long n, int i,
                                         const StepParams<double> &p) {
    const long long k = plane * i + n;
    if constexpr (SID == 0) {
        return static_cast<const double *>(base)[k];
    } else if constexpr (SID == 1) {
        const double g = static_cast<double>(static_cast<const float *>(base)[k]);
        return shift_of(p, i) + g;
    } else if constexpr (SID == 2) {
        const double g = static_cast<double>(__half2float(static_cast<const __half *>(base)[k]));
        return shift_of(p, i) + g * p.dscale_inv;
    } else {
        const double g =
            static_cast<double>(__bfloat162float(static_cast<const __hip_bfloat16 *>(base)[k]));
        return shift_of(p, i) + g * p.dscale_inv;
    }
}

template <int SID>
__device__ __forceinline__ void store_d(void *base, long long plane, long long n, int i, double v,
                                        const StepParams<double> &p) {
    const long long k = plane * i + n;
    if constexpr (SID == 0) {
        static_cast<double *>(base)[k] = v;
    } else if constexpr (SID == 1) {
        static_cast<float *>(base)[k] = static_cast<float>(v - shift_of(p, i));
    } else if constexpr (SID == 2) {
        const double g = (v - shift_of(p, i)) * p.dscale;
        static_cast<__half *>(base)[k] = __float2half_rn(static_cast<float>(g));
    } else {
        const double g = (v - shift_of(p, i)) * p.dscale;
        static_cast<__hip_bfloat16 *>(base)[k] = __float2bfloat16(static_cast<float>(g));
    }
}

// host Float64 (nx, ny, Q) <-> device storage, elementwise
template <int SID>
__global__ void __launch_bounds__(256)
k_convert(const double *__restrict__ hostlike, void *dev, int Q, int to_device,
          double *__restrict__ out, const StepParams<double> p) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.plane) return;
    for (int i = 0; i < Q; ++i) {
        if (to_device) store_d<SID>(dev, p.plane, n, i, hostlike[p.plane * i + n], p);
        else out[p.plane * i + n] = load_d<SID>(dev, p.plane, n, i, p);
    }
}

// density / velocity! / deviatoric_tensor per node, in the CPU restatement's order
// (moments.jl:1-19, 81-96; hermite_polynomials.jl:47-51; hermite.jl:45-48)
template <int QID, int SID>
__global__ void __launch_bounds__(256)
k_moments(const void *f, const StepParams<double> p, double tau, double *__restrict__ rho_o,
          double *__restrict__ ux_o, double *__restrict__ uy_o, double *__restrict__ sxx_o,
          double *__restrict__ sxy_o, double *__restrict__ syy_o) {
    constexpr int Q = QT<QID>::Q;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= p.nx) return;
    const long long n = (long long)x + (long long)p.nx * y;
    double fl[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fl[i] = load_d<SID>(f, p.plane, n, i, p);
    double rho = fl[0] + fl[1];
#pragma unroll
    for (int i = 2; i < Q; ++i) rho = rho + fl[i];
    double ux = 0.0, uy = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) ux = ux + fl[i] * static_cast<double>(QT<QID>::cx(i));
    ux = ux / rho;
#pragma unroll
    for (int i = 0; i < Q; ++i) uy = uy + fl[i] * static_cast<double>(QT<QID>::cy(i));
    uy = uy / rho;
    if (rho_o) rho_o[n] = rho;
    if (ux_o) ux_o[n] = ux;
    if (uy_o) uy_o[n] = uy;
    if (!(sxx_o || sxy_o || syy_o)) return;
    const double cs = p.cs_inv;
    auto h2 = [&](int i, int k) -> double {
        const int a = k & 1, b = k >> 1;
        const int ca = a == 0 ? QT<QID>::cx(i) : QT<QID>::cy(i);
        const int cb = b == 0 ? QT<QID>::cx(i) : QT<QID>::cy(i);
        return static_cast<double>(cb * ca) - cs * static_cast<double>(a == b ? 1 : 0);
    };
    double abar[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double a = fl[0] * h2(0, k) + fl[1] * h2(1, k);
#pragma unroll
        for (int i = 2; i < Q; ++i) a = a + fl[i] * h2(i, k);
        abar[k] = a;
    }
    const double u[2] = {ux, uy};
    double sg[4];
    const double den = 1.0 + 1.0 / (2.0 * tau);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double aeq = rho * (u[k & 1] * u[k >> 1] + 0.0);
        sg[k] = (abar[k] - aeq) / den;
    }
    const double tr2 = (sg[0] + sg[3]) / 2.0;
    if (sxx_o) sxx_o[n] = sg[0] - tr2;
    if (sxy_o) sxy_o[n] = sg[2];
    if (syy_o) syy_o[n] = sg[3] - tr2;
}

// per-block FP64 partial sums: rho, rho ux, rho uy, rho |u|^2, #non-finite
// nodes; a fixed grid strides over the lattice, and the partials are stored
// column by column (partials[k * gridDim.x + block])
template <int QID, int SID>
__global__ void __launch_bounds__(256)
k_reduce_partial(const void *f, const StepParams<double> p, double *__restrict__ partials) {
    constexpr int Q = QT<QID>::Q;
    __shared__ double red[5][256];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < p.plane;
         n += stride) {
        double fl[Q];
        bool finite = true;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            fl[i] = load_d<SID>(f, p.plane, n, i, p);
            finite = finite && isfinite(fl[i]);
        }
        double rho = fl[0] + fl[1];
#pragma unroll
        for (int i = 2; i < Q; ++i) rho = rho + fl[i];
        double ux = 0.0, uy = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) ux = ux + fl[i] * static_cast<double>(QT<QID>::cx(i));
        ux = ux / rho;
#pragma unroll
        for (int i = 0; i < Q; ++i) uy = uy + fl[i] * static_cast<double>(QT<QID>::cy(i));
        uy = uy / rho;
        if (finite) {
            v[0] += rho;
            v[1] += rho * ux;
            v[2] += rho * uy;
            v[3] += rho * (ux * ux + uy * uy);
        } else {
            v[4] += 1.0;
        }
    }
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) partials[(long long)k * gridDim.x + blockIdx.x] = red[k][0];
}

// fixed-order sum of each column of partials (one block per column)
__global__ void __launch_bounds__(256)
k_reduce_final(const double *__restrict__ partials, int npart, double *__restrict__ out) {
    __shared__ double red[256];
    const double *col = partials + (long long)blockIdx.x * npart;
    double v = 0.0;
    for (int b = threadIdx.x; b < npart; b += blockDim.x) v += col[b];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// On-device processing: one pass over f gives the sums behind a row of
// next!(::TrackHydrodynamicErrors) (track_hydrodynamic_errors.jl:115-203) and
// behind the stopping criteria (stopping_criteria.jl:17-99).  Each per-node
// value is formed in the reference's order; the FP64 sums run over a fixed
// grid-stride and tree order, so a result is reproducible bit for bit.
template <int QID, int SID, bool ERR>
__global__ void __launch_bounds__(256)
k_track(const void *f, const StepParams<double> p, const TrackArgs a,
        double *__restrict__ partials) {
    constexpr int Q = QT<QID>::Q;
    constexpr int NC = kTrackCols;
    __shared__ double red[4][NC];
    double acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < p.plane;
         n += stride) {
        double fl[Q];
        bool finite = true;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            fl[i] = load_d<SID>(f, p.plane, n, i, p);
            finite = finite && isfinite(fl[i]);
        }
        // density and velocity! (moments.jl:1-19)
        double rho = fl[0] + fl[1];
#pragma unroll
        for (int i = 2; i < Q; ++i) rho = rho + fl[i];
        double ux = 0.0, uy = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) ux = ux + fl[i] * static_cast<double>(QT<QID>::cx(i));
        ux = ux / rho;
#pragma unroll
        for (int i = 0; i < Q; ++i) uy = uy + fl[i] * static_cast<double>(QT<QID>::cy(i));
        uy = uy / rho;
        if (!finite) acc[kNonFinite] += 1.0;
        // MeanVelocityStoppingCriteria: sum of u_x (stopping_criteria.jl:23-37)
        acc[kSumUx] += ux;
        acc[kSumUy] += uy;
        // VelocityConvergenceStoppingCriteria (stopping_criteria.jl:81-99)
        if (a.u_old) {
            const double o1 = a.u_old[n], o2 = a.u_old[p.plane + n];
            acc[kDu2] += (ux - o1) * (ux - o1) + (uy - o2) * (uy - o2);
            acc[kUold2] += o1 * o1 + o2 * o2;
            a.u_old[n] = ux;
            a.u_old[p.plane + n] = uy;
        }
        if constexpr (ERR) {
            // expected rho, u_x, u_y, p, sigma_xx, sigma_xy, sigma_yx, sigma_yy
            double e[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) e[k] = a.aff[2 * k] + a.aff[2 * k + 1] * a.exp[p.plane * k + n];
            // a_bar_2 = sum_i f_i H2_i, H2 = c c' - cs I (hermite_polynomials.jl:47-51);
            // index k = alpha + 2 beta, column-major 2x2
            const double cs = p.cs_inv;
            double abar[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                auto h2 = [&](int i) -> double {
                    const int al = k & 1, be = k >> 1;
                    const int ca = al == 0 ? QT<QID>::cx(i) : QT<QID>::cy(i);
                    const int cb = be == 0 ? QT<QID>::cx(i) : QT<QID>::cy(i);
                    return static_cast<double>(cb * ca) - cs * static_cast<double>(al == be ? 1 : 0);
                };
                double s = fl[0] * h2(0) + fl[1] * h2(1);
#pragma unroll
                for (int i = 2; i < Q; ++i) s = s + fl[i] * h2(i);
                abar[k] = s;
            }
            const double u[2] = {ux, uy};
            // a_eq_2 = rho (u u' + cs (T - 1) I) at T = 1 (hermite.jl:45-48)
            double aeq[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int al = k & 1, be = k >> 1;
                aeq[k] = al == be ? rho * (u[al] * u[be] + 0.0) : rho * (u[al] * u[be]);
            }
            const double c = 1.0 / (2.0 * a.tau);
            const double den = 1.0 + 1.0 / (2.0 * a.tau);
            // a_2 = (a_bar_2 + a_eq_2 / (2 tau)) / (1 + 1/(2 tau)), P = a_2 - rho (u u' - I),
            // p = tr(P) / D (track_hydrodynamic_errors.jl:153, 173, 181)
            double P[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int al = k & 1, be = k >> 1;
                const double a2 = (abar[k] + c * aeq[k]) / den;
                P[k] = a2 - rho * (al == be ? u[al] * u[be] - 1.0 : u[al] * u[be]);
            }
            const double pr = ((0.0 + P[0]) + P[3]) / 2.0;
            // sigma_lb = deviatoric_tensor(q, tau, f, rho, u) (moments.jl:81-96)
            double sg[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) sg[k] = (abar[k] - aeq[k]) / den;
            const double tr2 = ((0.0 + sg[0]) + sg[3]) / 2.0;
            sg[0] = sg[0] - tr2;
            sg[3] = sg[3] - tr2;
            // dimensionless_velocity / _stress (problems.jl:127-134)
            const double d1 = ux / a.u_max, d2 = uy / a.u_max;
#pragma unroll
            for (int k = 0; k < 4; ++k) sg[k] = sg[k] * a.sfac;
            // squared differences and norms with the cell weight 1
            // (track_hydrodynamic_errors.jl:127-132, 192-202)
            acc[kErrP] += (pr - e[3]) * (pr - e[3]);
            acc[kErrRho] += (rho - e[0]) * (rho - e[0]);
            acc[kErrU] += (d1 - e[1]) * (d1 - e[1]) + (d2 - e[2]) * (d2 - e[2]);
            acc[kErrSxx] += (e[4] - sg[0]) * (e[4] - sg[0]);
            acc[kErrSxy] += (e[5] - sg[2]) * (e[5] - sg[2]);
            acc[kErrSyx] += (e[6] - sg[1]) * (e[6] - sg[1]);
            acc[kErrSyy] += (e[7] - sg[3]) * (e[7] - sg[3]);
            acc[kTotU2] += e[1] * e[1] + e[2] * e[2];
            acc[kTotSxx2] += e[4] * e[4];
            acc[kTotSyy2] += e[7] * e[7];
            acc[kTotSxy2] += e[5] * e[5];
            acc[kTotSyx2] += e[6] * e[6];
            acc[kTotP2] += e[3] * e[3];
            acc[kMass] += rho;
            acc[kMomentum] += rho * (d1 + d2);
            acc[kEnergy] += rho * (d1 * d1 + d2 * d2);
        }
    }
    // wave sums by fixed shuffles, then the block's waves in order
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wv][k] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < NC) {
        const int nw = (int)blockDim.x >> 6;
        double t = red[0][threadIdx.x];
        for (int w = 1; w < nw; ++w) t += red[w][threadIdx.x];
        // column-major: each sum's partials are contiguous for the final pass
        partials[(long long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// fixed-order sum of the partials, one block per column
__global__ void __launch_bounds__(256)
k_track_final(const double *__restrict__ partials, int npart, double *__restrict__ out) {
    __shared__ double red[256];
    const double *col = partials + (long long)blockIdx.x * npart;
    double v = 0.0;
    for (int b = threadIdx.x; b < npart; b += blockDim.x) v += col[b];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// the Q populations of one node in FP64 (f_in[x, y, :])
template <int SID>
__global__ void __launch_bounds__(64)
k_gather_node(const void *f, const StepParams<double> p, long long n, int Q,
              double *__restrict__ out) {
    const int i = threadIdx.x;
    if (i < Q) out[i] = load_d<SID>(f, p.plane, n, i, p);
}

// f[plane * i + n] = w_i: the rest state of weight-valued populations
__global__ void __launch_bounds__(256)
k_fill_weights(double *__restrict__ f, const StepParams<double> p, int Q) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.plane) return;
    for (int i = 0; i < Q; ++i) f[p.plane * i + n] = p.w[i];
}

template <class U>
__global__ void __launch_bounds__(256) k_copy_aux(const U *__restrict__ src, U *__restrict__ dst,
                                                  long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = i; k < n; k += stride) dst[k] = src[k];
}

// 16 bytes per lane, one element per lane for any copy below 2^36 bytes, with
// non-temporal stores: the fastest copy shape measured on MI355X (6.3 TB/s
// over 2 x 1.2 GB, against 6.1 for 8-byte lanes and 5.8 for an 8-byte
// grid-stride loop of 65,536 blocks)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(const u32x4 *__restrict__ src,
                                                u32x4 *__restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = i; k < n; k += stride) __builtin_nontemporal_store(src[k], &dst[k]);
}

// ---------------------------------------------------------------------------
template <int QID>
static int moments_q(int storage, const void *f, const StepParams<double> &p, double tau,
                     double *o0, double *o1, double *o2, double *o3, double *o4, double *o5,
                     hipStream_t s) {
    const int bx = p.nx >= 256 ? 256 : 64;
    dim3 grid((p.nx + bx - 1) / bx, p.ny, 1), block(bx, 1, 1);
    switch (storage) {
    case 0: k_moments<QID, 0><<<grid, block, 0, s>>>(f, p, tau, o0, o1, o2, o3, o4, o5); break;
    case 1: k_moments<QID, 1><<<grid, block, 0, s>>>(f, p, tau, o0, o1, o2, o3, o4, o5); break;
    case 2: k_moments<QID, 2><<<grid, block, 0, s>>>(f, p, tau, o0, o1, o2, o3, o4, o5); break;
    case 3: k_moments<QID, 3><<<grid, block, 0, s>>>(f, p, tau, o0, o1, o2, o3, o4, o5); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_moments(int qid, int storage, const void *f, const StepParams<double> &p, double tau,
                     double *rho, double *ux, double *uy, double *sxx, double *sxy, double *syy,
                     hipStream_t s) {
    switch (qid) {
    case 4: return moments_q<4>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 5: return moments_q<5>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 9: return moments_q<9>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 13: return moments_q<13>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 17: return moments_q<17>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 21: return moments_q<21>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    case 37: return moments_q<37>(storage, f, p, tau, rho, ux, uy, sxx, sxy, syy, s);
    default: return -1;
    }
}

template <int QID>
static int reduce_q(int storage, const void *f, const StepParams<double> &p, double *partials,
                    int npart, hipStream_t s) {
    dim3 grid(npart, 1, 1), block(256, 1, 1);
    switch (storage) {
    case 0: k_reduce_partial<QID, 0><<<grid, block, 0, s>>>(f, p, partials); break;
    case 1: k_reduce_partial<QID, 1><<<grid, block, 0, s>>>(f, p, partials); break;
    case 2: k_reduce_partial<QID, 2><<<grid, block, 0, s>>>(f, p, partials); break;
    case 3: k_reduce_partial<QID, 3><<<grid, block, 0, s>>>(f, p, partials); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_reduce(int qid, int storage, const void *f, const StepParams<double> &p,
                    double *partials, int npart, double *out5, hipStream_t s) {
    int rc;
    switch (qid) {
    case 4: rc = reduce_q<4>(storage, f, p, partials, npart, s); break;
    case 5: rc = reduce_q<5>(storage, f, p, partials, npart, s); break;
    case 9: rc = reduce_q<9>(storage, f, p, partials, npart, s); break;
    case 13: rc = reduce_q<13>(storage, f, p, partials, npart, s); break;
    case 17: rc = reduce_q<17>(storage, f, p, partials, npart, s); break;
    case 21: rc = reduce_q<21>(storage, f, p, partials, npart, s); break;
    case 37: rc = reduce_q<37>(storage, f, p, partials, npart, s); break;
    default: return -1;
    }
    if (rc) return rc;
    k_reduce_final<<<5, 256, 0, s>>>(partials, npart, out5);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <int QID, int... Is>
__device__ __forceinline__ void geq_rec_seq(const Mom4 &m, const StepParams<float> &p,
                                            float (&geq)[QT<QID>::Q],
                                            std::integer_sequence<int, Is...>) {
    ((geq[Is] = geq_rec<QID, Is>(m, p)), ...);
}

template <int QID>
__device__ __forceinline__ void geq_rec_all(const Mom4 &m, const StepParams<float> &p,
                                            float (&geq)[QT<QID>::Q]) {
    geq_rec_seq<QID>(m, p, geq, std::make_integer_sequence<int, QT<QID>::Q>{});
}

// Float64 (nx, ny, Q) <-> non-equilibrium storage (binary16 equilibrium record,
// feq in FP32 with the kernels' formula, 16-bit non-equilibrium parts)
template <int QID, class ST>
__global__ void __launch_bounds__(256)
k_neq_codec(double *__restrict__ hostlike, typename ST::store_t *dev, int to_device,
            const StepParams<float> pf) {
    constexpr int Q = QT<QID>::Q;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pf.plane) return;
    Rec16 *m = ST::recs(dev);
    auto *P = ST::pops(dev, pf.plane);
    if (to_device) {
        double f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = hostlike[pf.plane * i + n];
        // shifted populations g_i = f_i - w_i rho0 in FP64, moments from them
        double drho = 0.0, mx = 0.0, my = 0.0, g[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            g[i] = f[i] - (double)pf.w[i] * (double)pf.rho0;
            drho += g[i];
            mx += g[i] * QT<QID>::cx(i);
            my += g[i] * QT<QID>::cy(i);
        }
        const double rho = (double)pf.rho0 + drho;
        const Rec16 rec = ST::pack(make_rec<QID>((float)drho, (float)(mx / rho), (float)(my / rho), pf));
        m[n] = rec;
        float geq[Q];
        geq_rec_all<QID>(ST::unpack(rec), pf, geq);
#pragma unroll
        for (int i = 0; i < Q; ++i)
            P[pf.plane * i + n] = ST::from_f((float)(g[i] - (double)geq[i]), pf);
    } else {
        float geq[Q];
        geq_rec_all<QID>(ST::unpack(m[n]), pf, geq);
#pragma unroll
        for (int i = 0; i < Q; ++i)
            hostlike[pf.plane * i + n] = (double)pf.w[i] * (double)pf.rho0 + (double)geq[i] +
                                         (double)ST::to_f(P[pf.plane * i + n], pf);
    }
}

template <int QID>
static int neq_codec_q(int storage, double *tmp, void *dst, const StepParams<float> &pf,
                       int to_device, hipStream_t s) {
    const int blocks = (int)((pf.plane + 255) / 256);
    if (storage == 4)
        k_neq_codec<QID, StF16Neq><<<blocks, 256, 0, s>>>(tmp, static_cast<__half *>(dst),
                                                           to_device, pf);
    else
        k_neq_codec<QID, StBF16Neq><<<blocks, 256, 0, s>>>(
            tmp, static_cast<__hip_bfloat16 *>(dst), to_device, pf);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_upload(int qid, int storage, const double *src, void *dst,
                    const StepParams<double> &p, int to_device, hipStream_t s) {
    if (storage == 4 || storage == 5) {
        const StepParams<float> pf = to_float_params(p);
        double *tmp = const_cast<double *>(src);
        switch (qid) {
        case 4: return neq_codec_q<4>(storage, tmp, dst, pf, to_device, s);
        case 5: return neq_codec_q<5>(storage, tmp, dst, pf, to_device, s);
        case 9: return neq_codec_q<9>(storage, tmp, dst, pf, to_device, s);
        case 13: return neq_codec_q<13>(storage, tmp, dst, pf, to_device, s);
        case 17: return neq_codec_q<17>(storage, tmp, dst, pf, to_device, s);
        case 21: return neq_codec_q<21>(storage, tmp, dst, pf, to_device, s);
        case 37: return neq_codec_q<37>(storage, tmp, dst, pf, to_device, s);
        default: return -1;
        }
    }
    int Q = qid == 21 ? 25 : qid;
    const long long nodes = p.plane;
    const int blocks = (int)((nodes + 255) / 256);
    double *out = const_cast<double *>(src);
    switch (storage) {
    case 0: k_convert<0><<<blocks, 256, 0, s>>>(src, dst, Q, to_device, out, p); break;
    case 1: k_convert<1><<<blocks, 256, 0, s>>>(src, dst, Q, to_device, out, p); break;
    case 2: k_convert<2><<<blocks, 256, 0, s>>>(src, dst, Q, to_device, out, p); break;
    case 3: k_convert<3><<<blocks, 256, 0, s>>>(src, dst, Q, to_device, out, p); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_gather_node(int storage, const void *f, const StepParams<double> &p, long long n,
                         int Q, double *out, hipStream_t s) {
    if (Q < 1 || Q > 64 || n < 0 || n >= p.plane) return -1;
    switch (storage) {
    case 0: k_gather_node<0><<<1, 64, 0, s>>>(f, p, n, Q, out); break;
    case 1: k_gather_node<1><<<1, 64, 0, s>>>(f, p, n, Q, out); break;
    case 2: k_gather_node<2><<<1, 64, 0, s>>>(f, p, n, Q, out); break;
    case 3: k_gather_node<3><<<1, 64, 0, s>>>(f, p, n, Q, out); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_fill_weights(double *f, const StepParams<double> &p, int Q, hipStream_t s) {
    if (Q < 1 || Q > 37) return -1;
    const long long blocks = (p.plane + 255) / 256;
    k_fill_weights<<<(unsigned)blocks, 256, 0, s>>>(f, p, Q);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// device copy of nbytes (multiple of 8 for double storage, 4 float, 2 half)
int dispatch_copy(const void *src, void *dst, size_t nbytes, hipStream_t s) {
    const bool a16 = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
    if (nbytes % 16 == 0 && a16) {
        const long long n = (long long)(nbytes / 16);
        long long blocks = (n + 255) / 256;
        if (blocks > (1LL << 24)) blocks = 1LL << 24;
        if (blocks > 0)
            k_copy16<<<(unsigned)blocks, 256, 0, s>>>(static_cast<const u32x4 *>(src),
                                                      static_cast<u32x4 *>(dst), n);
    } else if (nbytes % 8 == 0) {
        const long long n = (long long)(nbytes / 8);
        long long blocks = (n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        k_copy_aux<double><<<(int)blocks, 256, 0, s>>>(static_cast<const double *>(src),
                                                        static_cast<double *>(dst), n);
    } else {
        const long long n = (long long)(nbytes / 2);
        long long blocks = (n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        k_copy_aux<unsigned short><<<(int)blocks, 256, 0, s>>>(
            static_cast<const unsigned short *>(src), static_cast<unsigned short *>(dst), n);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---------------------------------------------------------------------------
// Analytic initial conditions (initial_conditions/analytical_equilibrium.jl,
// analytical_offequilibrium.jl) per node, in the reference's FP64 order: the
// Hermite-sum equilibrium w_i (rho + sum_n a_eq[n] : H_n(c_i) / (n! cs^n))
// with the temperature terms of equilibrium_coefficient (hermite.jl:37-77,
// Val{4} with its index pairs as written there), plus the off-equilibrium
// part of the two off-equilibrium strategies.  cs^n comes from the host's
// libm pow (Julia's Float64 ^ Int).  One thread per node; the tensors are
// formed from their integer parts at run time (a one-off pass).
// ---------------------------------------------------------------------------
__device__ double herm_entry(int n, int cx, int cy, int k, double cs) {
    const int xi[2] = {cx, cy};
    const int a = k & 1, b = (k >> 1) & 1, g = (k >> 2) & 1, e = (k >> 3) & 1;
    auto dl = [](int i, int j) { return i == j ? 1 : 0; };
    if (n == 1) return (double)xi[k] * 1.0;
    if (n == 2) return (double)(xi[b] * xi[a]) - cs * (double)dl(a, b);
    if (n == 3) {
        const int A = xi[g] * xi[b] * xi[a];
        const int B = xi[a] * dl(b, g) + xi[b] * dl(a, g) + xi[g] * dl(a, b);
        return (double)A - cs * (double)B;
    }
    const int A = xi[e] * xi[g] * xi[b] * xi[a];
    const int B = xi[a] * xi[b] * dl(g, e) + xi[a] * xi[g] * dl(b, e) + xi[a] * xi[e] * dl(b, g) +
                  xi[b] * xi[g] * dl(a, e) + xi[b] * xi[e] * dl(a, g) + xi[g] * xi[e] * dl(a, b);
    const int C = dl(a, b) * dl(g, e) + dl(a, g) * dl(b, e) + dl(a, e) * dl(b, g);
    return ((double)A - cs * (double)B) + (cs * cs) * (double)C;
}

__device__ double eq_coefficient_T(int n, int k, double rho, const double (&u)[2], double T,
                                   double cs) {
    if (n == 1) return rho * u[k];
    const double ct = cs * (T - 1.0);
    const int a = k & 1, b = (k >> 1) & 1, c = (k >> 2) & 1, d = (k >> 3) & 1;
    auto dl = [](int i, int j) { return i == j ? 1.0 : 0.0; };
    double v;
    if (n == 2) {
        v = u[a] * u[b];
        if (a == b) v = v + ct;
    } else if (n == 3) {
        const double S = (u[a] * dl(b, c) + u[b] * dl(a, c)) + u[c] * dl(a, b);
        v = ((u[a] * u[b]) * u[c]) + ct * S;
    } else {
        const double S1 = (((((u[a] * u[b]) * dl(c, d) + (u[a] * u[c]) * dl(b, d)) +
                             (u[a] * u[d]) * dl(b, d)) +
                            (u[b] * u[c]) * dl(a, d)) +
                           (u[b] * u[d]) * dl(a, d)) +
                          (u[c] * u[d]) * dl(a, b);
        const int S0 = (int)(dl(a, b) * dl(c, d) + dl(a, c) * dl(b, d)) + (int)(dl(a, d) * dl(b, c));
        v = ((((u[a] * u[b]) * u[c]) * u[d]) + ct * S1) + ((cs * cs) * ((T - 1.0) * (T - 1.0))) * (double)S0;
    }
    return rho * v;
}

template <int QID>
__global__ void __launch_bounds__(256)
k_init_analytical(double *__restrict__ f, const StepParams<double> p, const InitArgs a) {
    constexpr int Q = QT<QID>::Q;
    constexpr int N = QT<QID>::ORDER / 2;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.plane) return;
    const double cs = p.cs_inv;
    const double rho = a.rho[n];
    const double u[2] = {a.ux[n], a.uy[n]};
    const double T = a.T ? a.T[n] : 1.0;
    double S[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.mode != 0)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c)
                S[c + 2 * b] = a.G[(c + 2 * b) * p.plane + n] + a.G[(b + 2 * c) * p.plane + n];
    for (int i = 0; i < Q; ++i) {
        const int cx = QT<QID>::cx(i), cy = QT<QID>::cy(i);
        double sum = 0.0;
        for (int m = 1; m <= N; ++m) {
            double d = 0.0;
            for (int k = 0; k < (1 << m); ++k) {
                const double t = eq_coefficient_T(m, k, rho, u, T, cs) * herm_entry(m, cx, cy, k, cs);
                d = (k == 0) ? t : d + t;
            }
            const double term = d / (a.fact[m] * a.csn[m]);
            sum = (m == 1) ? term : sum + term;
        }
        double fi = p.w[i] * (rho + sum);
        if (a.mode != 0) {
            double d2 = herm_entry(2, cx, cy, 0, cs) * S[0];
            for (int k = 1; k < 4; ++k) d2 = d2 + herm_entry(2, cx, cy, k, cs) * S[k];
            if (a.mode == 1)
                fi = fi - ((p.w[i] * (((p.s * (a.tau + 0.5)) * rho) * 1.0)) / 2.0) * d2;
            else
                fi = fi + ((((-a.factor) * p.w[i]) * 0.5) * ((a.tau + 0.5) * p.s)) * d2;
        }
        f[(long long)i * p.plane + n] = fi;
    }
}

int dispatch_init_analytical(int qid, double *f, const StepParams<double> &p, const InitArgs &a,
                             hipStream_t s) {
    const unsigned blocks = (unsigned)((p.plane + 255) / 256);
    switch (qid) {
        case 4: k_init_analytical<4><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 5: k_init_analytical<5><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 9: k_init_analytical<9><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 13: k_init_analytical<13><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 17: k_init_analytical<17><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 21: k_init_analytical<21><<<blocks, 256, 0, s>>>(f, p, a); break;
        case 37: k_init_analytical<37><<<blocks, 256, 0, s>>>(f, p, a); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// Clock probe: lane 0 of each block writes (XCC id, shader-clock counter,
// constant-rate counter).  Blocks are dealt round-robin over the XCDs, so
// 64 blocks sample every XCD; the counters of one XCD are compared between two
// probes.  s_getreg field: HW_REG_XCC_ID (20), bits [3:0].
// One row per block: where it ran, the shader-clock counter and the
// constant-rate counter.  The shader-clock counter is not synchronised
// between compute units, so the location key (XCC_ID bits 3:0, then HW_ID
// bits 15:8 = CU, SH and SE ids) lets two samples be paired per compute unit.
__global__ void __launch_bounds__(64) k_clock_probe(unsigned long long *out) {
    if (threadIdx.x != 0) return;
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;
    const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);
    const unsigned long long t = __builtin_amdgcn_s_memtime();
    const unsigned long long r = __builtin_amdgcn_s_memrealtime();
    out[3 * blockIdx.x] = (xcc << 8) | ((hw >> 8) & 0xffu);
    out[3 * blockIdx.x + 1] = t;
    out[3 * blockIdx.x + 2] = r;
}

int dispatch_clock_probe(unsigned long long *out, int blocks, hipStream_t s) {
    if (blocks < 1 || blocks > 1024) return -1;
    k_clock_probe<<<blocks, 64, 0, s>>>(out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <int QID>
static int track_q(int storage, const void *f, const StepParams<double> &p, const TrackArgs &a,
                   double *partials, int npart, hipStream_t s) {
    const dim3 grid(npart, 1, 1), block(256, 1, 1);
    const bool err = a.exp != nullptr;
#define LBM_TRACK_CASE(SID)                                                            \
    case SID:                                                                          \
        if (err) k_track<QID, SID, true><<<grid, block, 0, s>>>(f, p, a, partials);   \
        else k_track<QID, SID, false><<<grid, block, 0, s>>>(f, p, a, partials);      \
        break;
    switch (storage) {
        LBM_TRACK_CASE(0)
        LBM_TRACK_CASE(1)
        LBM_TRACK_CASE(2)
        LBM_TRACK_CASE(3)
    default: return -1;
    }
#undef LBM_TRACK_CASE
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int dispatch_track(int qid, int storage, const void *f, const StepParams<double> &p,
                   const TrackArgs &a, double *partials, int npart, double *out, hipStream_t s) {
    int rc;
    switch (qid) {
    case 4: rc = track_q<4>(storage, f, p, a, partials, npart, s); break;
    case 5: rc = track_q<5>(storage, f, p, a, partials, npart, s); break;
    case 9: rc = track_q<9>(storage, f, p, a, partials, npart, s); break;
    case 13: rc = track_q<13>(storage, f, p, a, partials, npart, s); break;
    case 17: rc = track_q<17>(storage, f, p, a, partials, npart, s); break;
    case 21: rc = track_q<21>(storage, f, p, a, partials, npart, s); break;
    case 37: rc = track_q<37>(storage, f, p, a, partials, npart, s); break;
    default: return -1;
    }
    if (rc) return rc;
    k_track_final<<<kTrackCols, 256, 0, s>>>(partials, npart, out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // namespace lbm
