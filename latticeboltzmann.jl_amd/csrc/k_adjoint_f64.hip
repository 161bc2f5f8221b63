// k_adjoint_f64.hip -- reverse-mode step kernels, FP64 populations, and the
// deterministic sum used for the loss and the relaxation-time gradients
// (-ffp-contract=off: the recomputed forward states feed the same arithmetic
// as the plain FP64 step).
#include "lbm_adjoint.h"

namespace lbm {
LBM_DEFINE_ADJOINT(f64, StF64)

// block b sums x[b*chunk .. (b+1)*chunk) in a fixed tree order
__global__ void __launch_bounds__(256)
k_sum_partial(const double *__restrict__ x, long long n, long long chunk, double *__restrict__ part) {
    __shared__ double sh[256];
    const long long b0 = (long long)blockIdx.x * chunk;
    long long b1 = b0 + chunk;
    if (b1 > n) b1 = n;
    double acc = 0.0;
    for (long long k = b0 + threadIdx.x; k < b1; k += blockDim.x) acc += x[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256)
k_sum_final(const double *__restrict__ part, int npart, double *__restrict__ out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < npart; k += blockDim.x) acc += part[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

int adjoint_sum(const double *x, long long n, double *partials, int npart, double *out,
                hipStream_t s) {
    if (n < 1 || npart < 1) return -1;
    const long long chunk = (n + npart - 1) / npart;
    const int nb = (int)((n + chunk - 1) / chunk);
    k_sum_partial<<<nb, 256, 0, s>>>(x, n, chunk, partials);
    k_sum_final<<<1, 256, 0, s>>>(partials, nb, out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}
}  // namespace lbm
