// lbm_internal.h -- types shared between the C-ABI host code and the kernel
// translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int kMaxBC = 16;

// One boundary condition with 0-based inclusive ranges.
struct BCDev {
    int kind;  // 0 bounce-back, 1 moving wall
    int dir;   // 0 N, 1 E, 2 S, 3 W
    int x0, x1, y0, y1;
};

// Per-launch parameters.  T is the arithmetic type (double for the plain
// path, float for the reduced-precision storages).
template <class T>
struct StepParams {
    int nx, ny;
    long long plane;  // nx * ny
    T w[37];
    // equilibrium constants (reference: maxwell_boltzmann_equilibrium.jl:43-66,
    // velocity_distribution_function/quadratures.jl:3-159)
    T s, neg_s, s2, three_s, six_s3, three_s2, s4;
    // SRT: omo = 1 - 1/tau, om = 1/tau; TRT: oms = 1/tau_s, oma = 1/tau_a
    T omo, om, oms, oma;
    T tau_force;  // tau (SRT), tau_a (TRT), taus[2] (MRT): velocity shift u + tau F
    int force_kind;  // 0 none, 1 const, 2 field
    T Fx, Fy;
    const double *Ffield;  // (nx, ny, 2) column-major
    int nbc;
    const BCDev *bcs;  // device array [nbc]
    const double *mw;  // device array [nbc * Q]: moving-wall term 2*a_1 per population
    // reduced-precision storage: f_i = w_i * (rho0 + scale_i * g_i)
    T rho0;
    T dscale, dscale_inv;
    // MRT (FP64 only)
    int mrt_n;               // div(order, 2)
    T mrt_omo[5], mrt_om[5]; // 1 - 1/taus[n], 1/taus[n] for n = 1..4 (index n)
    T mrt_sn[5];             // s^n
    T cs_inv;                // 1/s
    // iterative initialization (FP64 only): nonlinear equilibrium term (nx, ny, Q)
    const double *mei;
    // FP32 weights and s for the FP32-arithmetic storages (host side only;
    // f32_set = 0: round w and s instead)
    float w32[37];
    float s32;
    int f32_set;
};

// FP32 copy of the parameters for the FP32-arithmetic storages
inline StepParams<float> to_float_params(const StepParams<double> &d) {
    StepParams<float> f{};
    f.nx = d.nx;
    f.ny = d.ny;
    f.plane = d.plane;
    for (int i = 0; i < 37; ++i) f.w[i] = d.f32_set ? d.w32[i] : static_cast<float>(d.w[i]);
    f.s = d.f32_set ? d.s32 : static_cast<float>(d.s);
    f.neg_s = -f.s;
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
    f.cs_inv = static_cast<float>(d.cs_inv);
    f.mei = nullptr;
    return f;
}

// Batched lattices: one entry per lattice and one per tile of consecutive
// node slots (row-major within the lattice).  Rows are padded to `pitch`
// slots, a multiple of 128 bytes, so every row and plane starts on a cache
// line; the slots x = nx .. pitch - 1 of a row hold nothing.  Offsets are in
// elements of the pool.
struct BatchLattice {
    int nx, ny;
    long long plane;           // pitch * ny: the stride between populations
    long long off[2];          // ping-pong buffers
    double omo, om, oms, oma;  // SRT / TRT relaxation constants
    double inv_pitch;          // 1 / pitch (row of a slot index, corrected exactly)
    int coll;                  // 0 SRT, 1 TRT
    int pitch;                 // row stride in elements
};

struct BatchTile {
    int lat;
    int n0;
};

// FusedFirst: the first step of a run from the canonical state, collide +
// stream + BCs + collide in one pass (FP64, one-speed stencils; the host
// checks the tile shape)
enum class Op : int { Collide = 0, StreamBC = 1, Fused = 2, ApplyBC = 3, Copy = 4, Convert = 5,
                      FusedFirst = 6 };

struct Launch {
    Op op;
    int qid;
    int coll;     // 0 SRT, 1 TRT, 2 MRT, 3 iterative initialization (collide only)
    int storage;  // LBM_STORE_*
    const void *src;
    void *dst;
    int block;    // threads per block along x (0 = default)
    int block_y = 1;  // fused step: rows per block (block.x * block_y <= 256)
    int vec = 1;  // fused step: nodes per thread along x (1, 2, 4, 8)
    int nt = 0;   // fused step: non-temporal stores
};

// Reverse-mode kernels of the periodic SRT/TRT step (lbm_adjoint.h).
struct AdjLaunch {
    int what;            // 0 S0 = feq(rho0, u0), 1 seed lambda_T + loss, 2 reverse step, 3 dL/du0
    int qid, coll;       // coll 0 SRT, 1 TRT
    int direct;          // state stored directly (1) or pulled from a post-collision field (0)
    const void *state;   // populations read by the seed / reverse step
    void *pops;          // populations written by the initialisation
    const void *lam_in;  // lambda_{t+1} (reverse step) or lambda_0 (dL/du0)
    void *lam_out;       // lambda_t / lambda_T
    const double *u;     // u0 (init, dL/du0) or u_target (seed), (nx, ny, 2) column-major
    double *out;         // per-node loss (seed), dL/dtau accumulators (step), dL/du0
    double *aux;         // seed: u(S_T), (nx, ny, 2) column-major, or nullptr
};
int run_adjoint_f64(const AdjLaunch &L, const StepParams<double> &p, hipStream_t s);
int run_adjoint_f32(const AdjLaunch &L, const StepParams<float> &p, hipStream_t s);
// deterministic sum of n doubles (fixed order for fixed n and npart)
int adjoint_sum(const double *x, long long n, double *partials, int npart, double *out,
                hipStream_t s);

// Kernel dispatchers (lbm_dispatch.cpp): FP64 exact / reduced precision.
int dispatch_f64(const Launch &L, const StepParams<double> &p, hipStream_t s);
int dispatch_lp(const Launch &L, const StepParams<double> &p, hipStream_t s);

// per (storage, quadrature) launchers, one translation unit per group
#define LBM_DECLARE_RUN(TAG, T, QID) \
    int run_##TAG##_q##QID(const Launch &L, const StepParams<T> &p, hipStream_t s);
#define LBM_DECLARE_ALL_Q(TAG, T)                                                   \
    LBM_DECLARE_RUN(TAG, T, 4) LBM_DECLARE_RUN(TAG, T, 5) LBM_DECLARE_RUN(TAG, T, 9) \
    LBM_DECLARE_RUN(TAG, T, 13) LBM_DECLARE_RUN(TAG, T, 17) LBM_DECLARE_RUN(TAG, T, 21) \
    LBM_DECLARE_RUN(TAG, T, 37)
// batched periodic D2Q9 step: op 0 collide, 1 fused step, 2 stream
// (vec nodes per thread: 1, 2 or 4; tiles of vec * block nodes)
int run_batch_f64_q9(int op, const BatchLattice *lats, const BatchTile *tiles, int ntiles,
                     void *pool, int parity, int block, int grid, int vec,
                     const StepParams<double> &p, hipStream_t s);
int run_batch_f32_q9(int op, const BatchLattice *lats, const BatchTile *tiles, int ntiles,
                     void *pool, int parity, int block, int grid, int vec,
                     const StepParams<float> &p, hipStream_t s);

LBM_DECLARE_ALL_Q(f64, double)
LBM_DECLARE_ALL_Q(f32, float)
LBM_DECLARE_ALL_Q(f16dev, float)
LBM_DECLARE_ALL_Q(bf16dev, float)
LBM_DECLARE_ALL_Q(f16neq, float)
LBM_DECLARE_ALL_Q(bf16neq, float)
// moments and reductions (FP64 accumulation for every storage)
int dispatch_moments(int qid, int storage, const void *f, const StepParams<double> &p, double tau,
                     double *rho, double *ux, double *uy, double *sxx, double *sxy, double *syy,
                     hipStream_t s);
int dispatch_reduce(int qid, int storage, const void *f, const StepParams<double> &p,
                    double *partials, int npart, double *out5, hipStream_t s);
// on-device processing: TrackHydrodynamicErrors norms and stopping-criteria sums
constexpr int kTrackCols = 21;
struct TrackArgs {
    const double *exp;  // 8 expected planes (nx*ny each) or nullptr: velocity sums only
    double aff[16];     // expected_k = aff[2k] + aff[2k+1] * plane_k
    double tau;         // sigma relaxation time (s * nu)
    double u_max;       // dimensionless velocity u / u_max
    double sfac;        // dimensionless stress factor 1 / u_max^2
    double *u_old;      // previous velocity, 2 planes, updated in place; or nullptr
};
// columns of the sums (kTrackCols of them), in this order
enum TrackCol : int {
    kErrRho = 0, kErrU, kErrP, kErrSxx, kErrSxy, kErrSyy, kErrSyx,
    kTotU2, kTotP2, kTotSxx2, kTotSxy2, kTotSyy2, kTotSyx2,
    kMass, kMomentum, kEnergy, kSumUx, kSumUy, kDu2, kUold2, kNonFinite
};
int dispatch_track(int qid, int storage, const void *f, const StepParams<double> &p,
                   const TrackArgs &a, double *partials, int npart, double *out, hipStream_t s);
int dispatch_upload(int qid, int storage, const double *src, void *dst,
                    const StepParams<double> &p, int to_device, hipStream_t s);
int dispatch_copy(const void *src, void *dst, size_t nbytes, hipStream_t s);
// the Q populations of node n in FP64 (plain and DEV storages)
int dispatch_gather_node(int storage, const void *f, const StepParams<double> &p, long long n,
                         int Q, double *out, hipStream_t s);
// f[plane * i + n] = w_i for every node (plain FP64 storage)
int dispatch_fill_weights(double *f, const StepParams<double> &p, int Q, hipStream_t s);
// analytic initial conditions: device planes rho, ux, uy, T (null: 1) and
// G (4 planes, G[a + 2b]; null for mode 0); mode 0 equilibrium, 1
// off-equilibrium of the Taylor-Green vortex, 2 off-equilibrium (generic)
struct InitArgs {
    const double *rho, *ux, *uy, *T, *G;
    int mode;
    double tau, factor;
    double csn[5];   // cs^n, n = 1..4
    double fact[5];  // n!
};
// writes the FP64 populations (nx, ny, Q) of the lattice of p
int dispatch_init_analytical(int qid, double *f, const StepParams<double> &p, const InitArgs &a,
                             hipStream_t s);
// (XCC id, shader-clock counter, constant-rate counter) per block, 3 x blocks values
int dispatch_clock_probe(unsigned long long *out, int blocks, hipStream_t s);

}  // namespace lbm
