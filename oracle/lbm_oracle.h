/*
 * lbm_oracle.h -- CPU restatement of LatticeBoltzmann.jl's collide-stream step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (liblbm_hip.so and the lbmhip package) never
 * links, loads or calls it.
 *
 * Layout: populations are a Julia Array{Float64,3} of size (nx, ny, Q),
 * column-major: f[x + nx*(y + ny*i)], x fastest, 0-based here where the
 * reference is 1-based (src/initial_conditions.jl:13).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define LBO_QMAX 37

typedef struct {
    int id;           /* 4, 5, 9, 13, 17, 21, 37 */
    int Q;            /* stored populations (D2Q21 stores 25) */
    int cx[LBO_QMAX];
    int cy[LBO_QMAX];
    double w[LBO_QMAX];
    double s;         /* q.speed_of_sound_squared (= 1/c_s^2, reference naming) */
    int order;        /* order(q) */
    int feq_order;    /* truncation order of the collide equilibrium */
    int opp[LBO_QMAX];
} lbo_quad;

enum { LBO_FORCE_NONE = 0, LBO_FORCE_CONST = 1, LBO_FORCE_FIELD = 2 };
enum { LBO_BOUNCE_BACK = 0, LBO_MOVING_WALL = 1 };
enum { LBO_NORTH = 0, LBO_EAST = 1, LBO_SOUTH = 2, LBO_WEST = 3 };

typedef struct {
    int kind;        /* LBO_BOUNCE_BACK / LBO_MOVING_WALL */
    int direction;   /* LBO_NORTH.. */
    int x0, x1;      /* 1-based inclusive range bc.xs */
    int y0, y1;      /* 1-based inclusive range bc.ys */
    double ux, uy;   /* moving-wall velocity */
    double rho;      /* moving-wall density */
} lbo_bc;

/* returns 0 on success, -1 on unknown quadrature */
int lbo_quadrature(int id, lbo_quad *out);

double lbo_density(const lbo_quad *q, const double *f);
void lbo_velocity(const lbo_quad *q, const double *f, double rho, double *u);
void lbo_equilibrium(const lbo_quad *q, double rho, const double *u, double *feq);
void lbo_hermite_equilibrium(const lbo_quad *q, double rho, const double *u, double *feq);
/* analytic initial conditions (mode 0: AnalyticalEquilibrium, 1: equilibrium +
 * off-equilibrium of the Taylor-Green vortex, 2: the same for other problems) */
void lbo_init_node(const lbo_quad *q, int mode, double rho, const double *u, double T,
                   const double *G, double tau, double factor, double *f);
void lbo_initialize(const lbo_quad *q, int nx, int ny, int mode, const double *rho,
                    const double *ux, const double *uy, const double *T, const double *G,
                    double tau, double factor, double *f);

/* force: kind NONE (ptr ignored), CONST (2 doubles), FIELD ((nx,ny,2) col-major) */
void lbo_collide_srt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau, int force_kind, const double *force);
void lbo_collide_trt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau_s, double tau_a, int force_kind, const double *force);
void lbo_collide_mrt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     const double *taus, int ntaus, int force_kind, const double *force);
void lbo_stream(const lbo_quad *q, int nx, int ny, const double *f, double *fnew);
/* returns -1 (and writes nothing) when a bounce-back range leaves the lattice */
int lbo_apply_bcs(const lbo_quad *q, int nx, int ny, const lbo_bc *bcs, int nbc,
                  double *fnew, const double *fold);

/* collision kind for lbo_simulate: 0 SRT (taus[0]), 1 TRT (taus[0]=tau_s, taus[1]=tau_a), 2 MRT */
int lbo_simulate(const lbo_quad *q, int nx, int ny, double *f_stream, double *f_collision,
                 int collision, const double *taus, int ntaus, int force_kind,
                 const double *force, const lbo_bc *bcs, int nbc, long nsteps);

/* moments: rho, ux, uy (nx*ny) and deviatoric tensor sxx, sxy, syy (nx*ny); any may be NULL */
void lbo_moments(const lbo_quad *q, int nx, int ny, const double *f, double tau,
                 double *rho, double *ux, double *uy, double *sxx, double *sxy, double *syy);

/* one row of TrackHydrodynamicErrors: out[11] = error_rho, error_u, error_p,
 * error_sigma_xx, _xy, _yy, _yx, mass, momentum, energy, non-finite nodes;
 * expected: 8 planes (rho, ux, uy, p, sxx, sxy, syx, syy), field k =
 * affine[2k] + affine[2k+1] * plane k */
void lbo_track_errors(const lbo_quad *q, int nx, int ny, const double *f,
                      const double *expected, const double *affine, double tau, double u_max,
                      double delta, double *out);
/* stopping-criteria sums: out[5] = mean ux, mean uy, sum |u-u_old|^2,
 * sum |u_old|^2, non-finite nodes; u_old (2 planes) may be NULL, else set to u */
void lbo_velocity_stats(const lbo_quad *q, int nx, int ny, const double *f, double *u_old,
                        double *out);

/* iterative initialization of Mei et al.: the nonlinear term of the lattice
 * velocity u (nx, ny, 2); its collision; the whole initialization (returns the
 * steps run, -1 on a bad BC range) */
void lbo_mei_nonlinear(const lbo_quad *q, int nx, int ny, const double *u, double *N);
void lbo_collide_mei(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau, const double *N);
long lbo_mei_initialize(const lbo_quad *q, int nx, int ny, const double *u, double tau,
                        double eps, long max_steps, const lbo_bc *bcs, int nbc,
                        double *f_stream, double *f_collision);

/* OpenMP thread count used by the node loops (1 = scalar) */
void lbo_set_threads(int n);
int lbo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
