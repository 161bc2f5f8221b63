/*
 * lbm_hip.h -- C ABI of liblbm_hip.so, the MI355X (gfx950) collide-stream
 * step of LatticeBoltzmann.jl.
 *
 * Every entry point is plain C: opaque handles, pointers and sizes, an int
 * status (0 = ok, < 0 = error; lbm_last_error() gives the message).  A Julia
 * binding reaches them with ccall; Python reaches them with ctypes.  Reference
 * citations name the Julia function each entry point replaces (paths relative
 * to the LatticeBoltzmann.jl repository).
 *
 * Host arrays are the reference's layout: a Julia Array{Float64,3} of size
 * (nx, ny, Q), column-major, x fastest (src/initial_conditions.jl:13).  Device
 * populations keep that layout (one contiguous HBM plane per velocity), so an
 * FP64 upload or download is one memcpy.
 *
 * Ordering: all device work of one process is issued on one HIP stream owned
 * by the library; uploads/downloads synchronise that stream.  Handles are not
 * thread-safe; use one set of handles per host thread.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ------------------------------------------------------------ */
/* quadrature ids: src/quadratures.jl:30-38 (D2Q21 stores 25 populations) */
enum { LBM_D2Q4 = 4, LBM_D2Q5 = 5, LBM_D2Q9 = 9, LBM_D2Q13 = 13,
       LBM_D2Q17 = 17, LBM_D2Q21 = 21, LBM_D2Q37 = 37 };

/* population storage in HBM.  F64 is the plain (reference-exact) form; the
 * others are reduced-precision forms with FP32 collision arithmetic:
 *   F32       populations in binary32;
 *   F16DEV    deviations f_i - w_i rho0 in binary16;
 *   BF16DEV   the same in bfloat16;
 *   F16NEQ    per node an 8-byte record of four binary16 equilibrium
 *             coefficients (from rho - rho0 and u) and, per population, the
 *             non-equilibrium part f_i - feq_i(record) in binary16 (scaled by 2^16);
 *   BF16NEQ   the same with bfloat16 non-equilibrium parts (unscaled).
 * F32 / *DEV / *NEQ arithmetic runs on g_i = f_i - w_i rho0 (rho0 of the field).
 * The DEV storages round the whole deviation f_i - w_i rho0 to 16 bits: a flow
 * driven by a small body force or a moving wall stops accelerating once the
 * per-step change of a population falls below half a 16-bit ulp of its
 * deviation (measured: 9e-2 to 9.6e-1 relative velocity error on 8192 x 2048
 * Poiseuille / Couette channels after 20000 steps).  They are unsuitable for
 * forced or wall-driven flows; use F32 or a NEQ storage there. */
enum { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16DEV = 2, LBM_STORE_BF16DEV = 3,
       LBM_STORE_F16NEQ = 4, LBM_STORE_BF16NEQ = 5 };

/* collision models: src/collision_models/{srt,trt,mrt}.jl; LBM_MEI is the
 * collide-only IterativeInitializationCollisionModel
 * (collision_models/iterative_initialization.jl) */
enum { LBM_SRT = 0, LBM_TRT = 1, LBM_MRT = 2, LBM_MEI = 3 };

/* body force, evaluated on the host (the reference's per-node closure
 * srt.jl:52, trt.jl:77, mrt.jl:92 cannot cross the ABI) */
enum { LBM_FORCE_NONE = 0, LBM_FORCE_CONST = 1, LBM_FORCE_FIELD = 2 };

/* boundary conditions: src/boundary_conditions.jl:28-32, bounce_back.jl:2-6, moving_wall.jl:5-15 */
enum { LBM_BOUNCE_BACK = 0, LBM_MOVING_WALL = 1 };
enum { LBM_NORTH = 0, LBM_EAST = 1, LBM_SOUTH = 2, LBM_WEST = 3 };

typedef struct {
    int32_t kind;       /* LBM_BOUNCE_BACK or LBM_MOVING_WALL */
    int32_t direction;  /* LBM_NORTH .. LBM_WEST (moving wall: North only) */
    int32_t x0, x1;     /* bc.xs as a 1-based inclusive range */
    int32_t y0, y1;     /* bc.ys as a 1-based inclusive range */
    double ux, uy;      /* MovingWall u (moving_wall.jl:10-11) */
    double rho;         /* MovingWall rho */
} lbm_bc;

typedef struct lbm_field_s *lbm_field;   /* one population array f (nx, ny, Q) on the GPU */
typedef struct lbm_model_s *lbm_model;   /* quadrature + collision + force + BC list */

/* ---- library ---------------------------------------------------------- */
const char *lbm_last_error(void);
int lbm_version(void);
/* Number of visible GPUs; 0 without a GPU (does not create a context). */
int lbm_device_count(void);
/* Blocks until all work issued by the library has finished. */
int lbm_synchronize(void);

/* Measurement support; no counterpart in the reference.  Runs a probe of
 * LBM_CLOCK_PROBES blocks (enough to reach every compute unit) on the
 * library's stream, after the work already issued on it, and returns
 * out[3 * LBM_CLOCK_PROBES]: per probe block (location key, shader-clock
 * cycles, constant-rate ticks), the key being (XCC id << 8) | (SE, SH and CU
 * ids); *tick_mhz (may be NULL) is the constant counter's rate.  The
 * shader-clock counter runs per compute unit and its counters are not
 * aligned with each other, so two samples around a stretch of device work
 * are compared per location key: cycles over ticks is that compute unit's
 * average clock over the stretch.  Synchronous. */
#define LBM_CLOCK_PROBES 1024
int lbm_clock_sample(uint64_t *out, double *tick_mhz);

/* Quadrature constants as the library holds them (c as doubles, 2*Q).
 * Replaces field access q.abscissae / q.weights / q.speed_of_sound_squared /
 * order(q) / opposite(q, i) (src/quadratures.jl:11-19, quadratures/D2Q*.jl). */
int lbm_quadrature_info(int quadrature, int *Q, double *cx, double *cy, double *w,
                        double *s, int *order, int *opp /* 0-based */);

/* The FP32 weights (Q) and s = 1/c_s^2 that the FP32-arithmetic storages
 * compute with.  No counterpart in the reference (whose arithmetic is FP64):
 * on D2Q13/17/21/37 each weight group and s sit within a few ulps of their
 * FP32 rounding, chosen so that sum w = 1 and s sum w c_x^2 = 1 hold in FP32
 * to ~1e-9 (rounded one by one they hold to 1e-8 - 3e-8, which makes the
 * FP32 collision gain or lose momentum in proportion to u); D2Q4/5/9 use the
 * plain rounding. */
int lbm_quadrature_f32(int quadrature, float *w, float *s);

/* ---- population fields (device arrays) -------------------------------- */
/* Array{Float64}(undef, nx, ny, Q) on the device (initial_conditions.jl:13).
 * rho0 is the rest density about which the *DEV storages store deviations. */
int lbm_field_create(lbm_field *out, int quadrature, int nx, int ny, int storage, double rho0);
int lbm_field_destroy(lbm_field f);
int lbm_field_info(lbm_field f, int *quadrature, int *nx, int *ny, int *Q, int *storage);
/* copy a host (nx, ny, Q) column-major Float64 array in / out (converting for
 * reduced storages).  Synchronous. */
int lbm_field_upload(lbm_field f, const double *host);
int lbm_field_download(lbm_field f, double *host);
/* dst .= src (same shape and storage), a device copy kernel */
int lbm_field_copy(lbm_field dst, lbm_field src);
/* f[x, y, :] for 1-based (x, y): the Q populations of one node in FP64 */
int lbm_field_node(lbm_field f, int x, int y, double *out);

/* Analytic initial conditions, computed on the device: initialize(strategy,
 * q, problem) (initial_conditions.jl:7-22) for the analytic strategies, with
 * the problem's fields evaluated by the caller (the reference's per-node
 * lattice_density / lattice_velocity / lattice_temperature /
 * velocity_gradient calls, problems.jl:98-106) and passed as Float64 (nx, ny)
 * column-major planes:
 *   LBM_INIT_EQUILIBRIUM  AnalyticalEquilibrium
 *       (initial_conditions/analytical_equilibrium.jl:8-17): the Hermite-sum
 *       equilibrium of hermite_based_equilibrium! (hermite.jl:10-33) at
 *       temperature T;
 *   LBM_INIT_OFFEQ_TGV    AnalyticalEquilibriumAndOffEquilibrium for the
 *       Taylor-Green vortex (analytical_offequilibrium.jl:50-87): minus
 *       w_i (s (tau + 1/2) rho) / 2 * H_2(c_i) : (G + G');
 *   LBM_INIT_OFFEQ        the same for any other problem (:10-48): plus
 *       -factor w_i 0.5 ((tau + 0.5) s) H_2(c_i) : (G + G'), factor = the
 *       domain area.
 * T may be NULL (T = 1); grad = 4 planes of the 2 x 2 gradient as the
 * reference passes it (G[a + 2 b] column-major; the generic strategy's
 * u_max^2-scaled one); tau = s nu, the reference's tau before the 1/2.  The
 * result is stored in f's storage. */
enum { LBM_INIT_EQUILIBRIUM = 0, LBM_INIT_OFFEQ_TGV = 1, LBM_INIT_OFFEQ = 2 };
int lbm_initialize_analytical(lbm_field f, int strategy, const double *rho, const double *ux,
                              const double *uy, const double *T, const double *grad, double tau,
                              double factor);

/* ---- model (collision model, force, boundary conditions) --------------- */
int lbm_model_create(lbm_model *out, int quadrature);
int lbm_model_destroy(lbm_model m);
/* SRT(tau) -- srt.jl:1-5 */
int lbm_model_set_srt(lbm_model m, double tau);
/* TRT(tau_s, tau_a, force) field order -- trt.jl:1-6 (note the reference's
 * 2-argument TRT(a, b) means tau_a = a, tau_s = b; the binding must swap) */
int lbm_model_set_trt(lbm_model m, double tau_symmetric, double tau_asymmetric);
/* MRT(q, taus) -- mrt.jl:19-34; taus[n-1] relaxes Hermite order n */
int lbm_model_set_mrt(lbm_model m, const double *taus, int ntaus);
/* IterativeInitializationCollisionModel(q, tau, problem)
 * (iterative_initialization.jl:12-58): u = the lattice velocity (nx, ny, 2)
 * column-major, from which the nonlinear equilibrium term is formed; collide!
 * then relaxes towards w_i rho + that term.  FP64 storage; lbm_simulate runs
 * it as collide! + stream! + apply! per step. */
int lbm_model_set_mei(lbm_model m, double tau, const double *u, int nx, int ny);
/* LBM_FORCE_NONE; LBM_FORCE_CONST: values = {Fx, Fy}; LBM_FORCE_FIELD:
 * values = (nx, ny, 2) column-major lattice force, copied to the device. */
int lbm_model_set_force(lbm_model m, int kind, const double *values, int nx, int ny);
/* ordered list, applied in order (boundary_conditions.jl:6-16) */
int lbm_model_set_bcs(lbm_model m, const lbm_bc *bcs, int n);

/* Launch shape of the fused step kernel: threads per block along x (64, 128,
 * 256), lattice rows per block (1, 2, 4; block_x * rows <= 256), lattice
 * nodes per thread (1, 2, 4, 8; along x, except for the non-equilibrium
 * storages, where more than one are rows marched by one thread: 2..16) and
 * non-temporal stores (0/1);
 * 0 (or -1 for nontemporal) = automatic.  For MRT (FP64), nodes_per_thread =
 * 16 selects the matrix-core kernel (tiles of 16 nodes, FP64 MFMA over the
 * distinct Hermite components; periodic lattices with nx a multiple of 16,
 * other cases return an error; parity with the oracle to 1e-12 relative
 * instead of bit for bit).  It is opt-in: on MI355X it is slower than the
 * default vector kernel.  No counterpart in the reference (a CPU loop has no
 * launch shape). */
int lbm_model_set_launch(lbm_model m, int block_x, int rows_per_block, int nodes_per_thread,
                         int nontemporal);
/* Autotuner: times every launch shape of the fused step for this model on
 * scratch copies of f_stream (the caller's fields are not advanced), keeps
 * the fastest in the model and returns it.  No counterpart in the reference. */
int lbm_autotune(lbm_model m, lbm_field f_stream, int trial_steps, int *block_x,
                 int *rows_per_block, int *nodes_per_thread, int *nontemporal,
                 double *us_per_step);

/* ---- step operators ----------------------------------------------------- */
/* collide!(cm, q, f_in, f_out; time) -- collision_models.jl:19, srt.jl:18-62,
 * trt.jl:42-97, mrt.jl:49-118.  f_in and f_out must be distinct. */
int lbm_collide(lbm_model m, lbm_field f_in, lbm_field f_out);
/* stream!(q, f, f_new) -- stream.jl:18-30 (periodic pull, mod1 wrap) */
int lbm_stream(lbm_model m, lbm_field f, lbm_field f_new);
/* apply!(bcs, q, f_new, f_old; time) -- boundary_conditions.jl:6-20,
 * bounce_back.jl:8-72, moving_wall.jl:17-38.  Overwrites f_new in place. */
int lbm_apply_bcs(lbm_model m, lbm_field f_new, lbm_field f_old);
/* stream! then apply!, as one pass */
int lbm_stream_bc(lbm_model m, lbm_field f, lbm_field f_new);
/* simulate(model, t0:t0+nsteps-1) without processing -- lattice_boltzmann_model.jl:60-77:
 * nsteps x (collide! f_stream -> f_collision; stream! + apply!) with one fused
 * pull collide-stream kernel per step.  On return both buffers hold exactly
 * what the reference's model.f_stream / model.f_collision hold. */
int lbm_simulate(lbm_model m, lbm_field f_stream, lbm_field f_collision, int64_t nsteps);
/* As lbm_simulate, also returning the time of the fused kernels (HIP events on
 * the library stream): *ms_total over the fused launches of the run, *nlaunch
 * their number (nsteps - 1, or nsteps - 2 where the first step runs as one
 * collide + fused pass: D2Q4/5/9 on lattices of whole 64 x 8 tiles). */
int lbm_simulate_timed(lbm_model m, lbm_field f_stream, lbm_field f_collision, int64_t nsteps,
                       double *ms_total, int64_t *nlaunch);
/* Measurement support; no counterpart in the reference.  As lbm_simulate,
 * with a HIP event before every fused launch: launch_ms[k] (nsteps - 1
 * entries) is the time from the start of fused launch k to the start of the
 * next (the last: to the end of the last), *nlaunch the number of entries. */
int lbm_simulate_launch_times(lbm_model m, lbm_field f_stream, lbm_field f_collision,
                              int64_t nsteps, double *launch_ms, int64_t *nlaunch);

/* ---- batches of independent lattices ----------------------------------- */
/* Many independent periodic lattices (uneven sizes and relaxation times)
 * advanced together: every time step is one persistent kernel launch over a
 * tile table covering all lattices, and runs of fused steps are replayed
 * from a captured hipGraph.  D2Q9, storage LBM_STORE_F64 or LBM_STORE_F32.
 * No counterpart in the reference (it advances one lattice per process); each
 * lattice follows simulate(model, time) exactly as lbm_simulate does. */
typedef struct lbm_batch_s *lbm_batch;
/* tau_a == NULL: SRT(tau_s[k]); otherwise TRT(tau_s[k], tau_a[k]) */
int lbm_batch_create(lbm_batch *out, int quadrature, int storage, int n, const int *nx,
                     const int *ny, const double *tau_s, const double *tau_a, double rho0);
int lbm_batch_destroy(lbm_batch b);
int lbm_batch_info(lbm_batch b, int *n, int64_t *nodes, int64_t *bytes);
/* populations of lattice k, (nx_k, ny_k, Q) column-major Float64: upload sets
 * both f_stream and f_collision (f_collision = copy(f_stream)); role 0 =
 * f_stream, 1 = f_collision */
int lbm_batch_upload(lbm_batch b, int k, const double *f);
int lbm_batch_download(lbm_batch b, int k, int role, double *f);
/* threads per tile-block (64, 128, 256; 0 = 128) and grid size (0 = one block
 * per tile; smaller grids make the blocks stride over the tile table) */
int lbm_batch_set_launch(lbm_batch b, int block, int grid);
/* nodes per thread of the batch kernel (1, 2 or 4; default 1): a tile holds
 * block x nodes consecutive nodes of one lattice, a lane nodes of them that
 * lie block apart */
int lbm_batch_set_nodes_per_thread(lbm_batch b, int nodes);
/* the launch shape in use (after lbm_batch_autotune: the fastest one);
 * any output may be NULL */
int lbm_batch_get_launch(lbm_batch b, int *block, int *grid, int *nodes_per_thread);
/* times launch shapes on the batch (state saved and restored), keeps the fastest */
int lbm_batch_autotune(lbm_batch b, int trial_steps, int *block, int *grid, double *us_per_step);
/* nsteps of every lattice; graph_chunk > 0 replays runs of graph_chunk fused
 * steps from one captured hipGraph (0: direct launches).  *ms_fused (may be
 * NULL) receives the time of the fused steps (HIP events). */
int lbm_batch_simulate(lbm_batch b, int64_t nsteps, int graph_chunk, double *ms_fused);

/* ---- snapshots ----------------------------------------------------------- */
/* Storage for TakeSnapshots (take_snapshots.jl:3-29: snapshot = copy(f_in) at
 * the chosen steps), taken without stalling the step: each snapshot is
 * decoded / copied into one of device_slots staging buffers in HBM (which
 * bound the device memory held: device_slots x nx x ny x Q x 8 bytes) on the
 * library stream; a second HIP stream copies it into pinned host memory, where
 * every snapshot stays as an (nx, ny, Q) column-major Float64 array.  With a
 * directory (else NULL) each snapshot is also written there as
 * snapshot_<t, 10 digits>.npy (NumPy format 1.0: '<f8', Fortran order,
 * shape (nx, ny, Q)).  The library stream waits only to reuse a staging
 * buffer whose host copy is still running.  Plotting (take_snapshots.jl:
 * 31-202) is not done. */
typedef struct lbm_snapshots_s *lbm_snapshots;
int lbm_snapshots_create(lbm_snapshots *out, int quadrature, int nx, int ny, int device_slots,
                         const char *directory);
int lbm_snapshots_destroy(lbm_snapshots s);
/* enqueue copy(f) as the snapshot of step t; returns without waiting */
int lbm_snapshot_take(lbm_snapshots s, lbm_field f, int64_t t);
/* wait for every snapshot taken so far (and its file); *n (may be NULL) = count */
int lbm_snapshots_wait(lbm_snapshots s, int *n);
/* snapshot k, in the order taken: its step and (host may be NULL) its data */
int lbm_snapshot_get(lbm_snapshots s, int k, int64_t *t, double *host);

/* ---- reverse mode (gradients) ------------------------------------------ */
/* Differentiable run of a periodic, unforced SRT or TRT model:
 *   S_0 = feq(rho0, u0) (the equilibrium initial condition,
 *   initial_conditions.jl:7-22), T steps of collide! -> stream!
 *   (lattice_boltzmann_model.jl:60-70), loss L = 1/2 sum_x |u(S_T) - u_target|^2.
 * lbm_adjoint_gradient returns dL/du0 and dL/dtau by a hand-written reverse
 * pass (transposed stream = push, transposed collision Jacobian per node),
 * keeping every checkpoint_every-th post-collision field (0 = ceil(sqrt(T)))
 * and recomputing the states in between.  No counterpart in the reference
 * (it is differentiated only by running it); the model (collision and tau)
 * is read at each call and must outlive the handle.  Storage LBM_STORE_F64
 * or LBM_STORE_F32 (shifted populations, FP32 arithmetic).
 * u0, u_target, u_final, grad_u0: (nx, ny, 2) column-major Float64 host arrays. */
typedef struct lbm_adjoint_s *lbm_adjoint;
int lbm_adjoint_create(lbm_adjoint *out, lbm_model m, int storage, int nx, int ny, double rho0,
                       int64_t nsteps, int checkpoint_every);
int lbm_adjoint_destroy(lbm_adjoint a);
/* checkpoint interval, number of stored checkpoints, device bytes held */
int lbm_adjoint_info(lbm_adjoint a, int *checkpoint_every, int *n_checkpoints, int64_t *bytes);
/* forward only: loss and (u_final may be NULL) the velocity field u(S_T) */
int lbm_adjoint_loss(lbm_adjoint a, const double *u0, const double *u_target, double *loss,
                     double *u_final);
/* loss, dL/du0 (may be NULL) and dL/dtau: grad_tau[0] = dL/dtau (SRT) or
 * grad_tau[0..1] = dL/dtau_s, dL/dtau_a (TRT); *ms (may be NULL) = device time */
int lbm_adjoint_gradient(lbm_adjoint a, const double *u0, const double *u_target, double *loss,
                         double *grad_u0, double *grad_tau, double *ms);

/* ---- initialization --------------------------------------------------- */
/* initialize(IterativeInitializationMeiEtAl(tau, eps), q, problem)
 * (initial_conditions/mei_et_al.jl:13-41) with a model set by
 * lbm_model_set_mei and the problem's BC list: both fields start at f = w,
 * then up to max_steps steps of the model's collide! + stream! + apply!, each
 * followed by the DensityConvergence check of ProcessIterativeInitialization
 * (process_iterative_initialization.jl:13-22, density_convergence.jl:6-17:
 * stop when the density of node (nx, ny) changed by less than eps or more than
 * 100).  f_stream holds the result; *steps (may be NULL) the steps run. */
int lbm_mei_initialize(lbm_model m, lbm_field f_stream, lbm_field f_collision, double eps,
                       int64_t max_steps, int64_t *steps);

/* ---- moments ------------------------------------------------------------ */
/* density / velocity! / deviatoric_tensor per node (moments.jl:1-19,81-96),
 * as TrackHydrodynamicErrors uses them (track_hydrodynamic_errors.jl:139-177).
 * Outputs are (nx, ny) column-major Float64 host arrays; any may be NULL. */
int lbm_moments(lbm_field f, double tau, double *rho, double *ux, double *uy,
                double *sxx, double *sxy, double *syy);
/* FP64 global sums over the lattice (track_hydrodynamic_errors.jl:201-203):
 * out[0] = sum rho, out[1] = sum rho ux, out[2] = sum rho uy,
 * out[3] = sum rho |u|^2, out[4] = number of nodes with a non-finite population */
int lbm_reduce(lbm_field f, double *out5);
/* As lbm_reduce, allocating nothing (no hipMalloc/hipFree): no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid allocation).
 * Its partial sums live in a library-owned static device array of 2048 x 5 +
 * 8 doubles per GPU (80 KiB, held until the process ends), which covers any
 * lattice size (the grid is capped at 2048 blocks).  Storages F64, F32,
 * F16DEV, BF16DEV; the non-equilibrium storages return an error (they need a
 * decoded copy; use lbm_reduce).  Calls that overlap on one GPU share the
 * array and race.  Synchronous. */
int lbm_reduce_noalloc(lbm_field f, double *out5);

/* ---- on-device processing ----------------------------------------------- */
/* Analytic reference fields for the error norms of TrackHydrodynamicErrors:
 * eight Float64 (nx, ny) column-major planes, in this order:
 * rho, u_x, u_y, p, sigma_xx, sigma_xy, sigma_yx, sigma_yy
 * (the expected_* values of track_hydrodynamic_errors.jl:120-125).
 * They stay on the device; a call combines them with per-call affine factors
 * (lbm_track_params.affine), so fields that change in time only by a factor
 * (a decaying vortex) are uploaded once. */
typedef struct lbm_expected_s *lbm_expected;
int lbm_expected_create(lbm_expected *out, int nx, int ny);
int lbm_expected_destroy(lbm_expected e);
/* planes: 8 * nx * ny doubles, plane k at planes + k * nx * ny */
int lbm_expected_upload(lbm_expected e, const double *planes);

#define LBM_TRACK_NOUT 11
typedef struct {
    /* relaxation time of the deviatoric stress: s * nu, without the 1/2
     * (track_hydrodynamic_errors.jl:148-153, 176-177) */
    double tau;
    /* velocity scale: dimensionless u = u / u_max, sigma * (1 / u_max^2)
     * (problems.jl:127-134) */
    double u_max;
    /* cell area multiplying mass, momentum and energy (Δ_ of
     * track_hydrodynamic_errors.jl:85-94) */
    double delta;
    /* expected field k = affine[2k] + affine[2k+1] * plane k */
    double affine[16];
} lbm_track_params;

/* One row of next!(::TrackHydrodynamicErrors) (track_hydrodynamic_errors.jl:
 * 52-222) computed on the device: per node density, velocity!, the pressure
 * tr(P)/2 and deviatoric_tensor (moments.jl:81-96), scaled to dimensionless
 * units and compared with the expected fields, summed in FP64:
 * out[0] error_rho, out[1] error_u, out[2] error_p, out[3] error_sigma_xx,
 * out[4] error_sigma_xy, out[5] error_sigma_yy, out[6] error_sigma_yx,
 * out[7] mass, out[8] momentum, out[9] energy (the reference's NamedTuple
 * order, lines 205-221), out[10] number of nodes with a non-finite
 * population.  NaNs propagate into the norms as in the reference. */
int lbm_track_errors(lbm_field f, lbm_expected e, const lbm_track_params *prm, double *out);

/* Previous velocity field of VelocityConvergenceStoppingCriteria
 * (stopping_criteria.jl:57-69): two Float64 planes, zero at creation. */
typedef struct lbm_velocity_tracker_s *lbm_velocity_tracker;
int lbm_velocity_tracker_create(lbm_velocity_tracker *out, int nx, int ny);
int lbm_velocity_tracker_destroy(lbm_velocity_tracker t);

/* The sums behind the stopping criteria, in one pass over f:
 * out[0] = mean u_x, out[1] = mean u_y (MeanVelocityStoppingCriteria,
 * stopping_criteria.jl:17-39); with a tracker (else NULL):
 * out[2] = sum |u - u_old|^2, out[3] = sum |u_old|^2, after which u_old = u
 * (VelocityConvergenceStoppingCriteria, stopping_criteria.jl:71-99);
 * out[4] = number of nodes with a non-finite population. */
int lbm_velocity_stats(lbm_field f, lbm_velocity_tracker t, double *out5);

#ifdef __cplusplus
}
#endif
#endif
