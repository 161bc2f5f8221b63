/*
 * lbm_oracle.c -- FP64 CPU restatement of LatticeBoltzmann.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped,
 * never called by the product).  Built with -ffp-contract=off so that every
 * a*b+c is two IEEE roundings, as in the reference's Julia code (Julia does
 * not contract without @fastmath/muladd).  Each function follows the
 * evaluation order of the cited reference lines so that the HIP FP64 kernels,
 * which use the same order, can be compared bit for bit.
 *
 * Reference citations are relative to the LatticeBoltzmann.jl repository.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void lbo_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int lbo_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* Quadratures: src/quadratures.jl:11-19 (generic opposite) and              */
/* src/quadratures/D2Q{4,5,9,13,17,21,37}.jl                                 */
/* ------------------------------------------------------------------------ */

static void set_c(lbo_quad *q, const int *cx, const int *cy) {
    for (int i = 0; i < q->Q; ++i) {
        q->cx[i] = cx[i];
        q->cy[i] = cy[i];
    }
}

/* generic rule, quadratures.jl:11-19 (1-based: 1->1, even->i+1, odd->i-1) */
static void generic_opposite(lbo_quad *q) {
    for (int i = 0; i < q->Q; ++i) {
        int idx = i + 1;
        int o;
        if (idx == 1) o = 1;
        else if (idx % 2 == 0) o = idx + 1;
        else o = idx - 1;
        q->opp[i] = o - 1;
    }
}

int lbo_quadrature(int id, lbo_quad *q) {
    memset(q, 0, sizeof(*q));
    q->id = id;
    switch (id) {
    case 4: { /* D2Q4.jl:16-31 */
        static const int cx[] = {1, 0, -1, 0}, cy[] = {0, 1, 0, -1};
        q->Q = 4;
        set_c(q, cx, cy);
        for (int i = 0; i < 4; ++i) q->w[i] = 1.0 / 4.0;
        q->s = 2.0;
        q->order = 3;
        q->feq_order = 1;
        for (int i = 0; i < 4; ++i) q->opp[i] = (i + 1 <= 2) ? i + 2 : i - 2;
        return 0;
    }
    case 5: { /* D2Q5.jl:17-42 */
        static const int cx[] = {0, 1, 0, -1, 0}, cy[] = {0, 0, 1, 0, -1};
        q->Q = 5;
        set_c(q, cx, cy);
        q->w[0] = 4.0 / 6.0;
        for (int i = 1; i < 5; ++i) q->w[i] = 1.0 / 12.0;
        q->s = 6.0;
        q->order = 3;
        q->feq_order = 1;
        for (int i = 0; i < 5; ++i) {
            int idx = i + 1, o;
            if (idx == 1) o = 1;
            else if (idx <= 3) o = idx + 2;
            else o = idx - 2;
            q->opp[i] = o - 1;
        }
        return 0;
    }
    case 9: { /* D2Q9.jl:20-38 */
        static const int cx[] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
        static const int cy[] = {0, 1, 0, -1, -1, -1, 0, 1, 1};
        q->Q = 9;
        set_c(q, cx, cy);
        const double w[] = {4.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9,
                            1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9};
        for (int i = 0; i < 9; ++i) q->w[i] = w[i];
        q->s = 3.0;
        q->order = 5;
        q->feq_order = 2;
        for (int i = 0; i < 9; ++i) {
            int idx = i + 1, o;
            if (idx == 1) o = 1;
            else if (idx <= 5) o = idx + 4;
            else o = idx - 4;
            q->opp[i] = o - 1;
        }
        return 0;
    }
    case 13: { /* D2Q13.jl:10-29 */
        static const int cx[] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0};
        static const int cy[] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -2, 2};
        q->Q = 13;
        set_c(q, cx, cy);
        const double w0 = 3.0 / 8, w1 = 1.0 / 12, w2 = 1.0 / 16, w3 = 1.0 / 96;
        const double w[] = {w0, w1, w1, w1, w1, w2, w2, w2, w2, w3, w3, w3, w3};
        for (int i = 0; i < 13; ++i) q->w[i] = w[i];
        q->s = 2.0;
        q->order = 5;
        q->feq_order = 2;
        generic_opposite(q);
        return 0;
    }
    case 17: { /* D2Q17.jl:21-57 */
        static const int cx[] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 2, -2, 3, -3, 0, 0};
        static const int cy[] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 2, -2, -2, 2, 0, 0, 3, -3};
        q->Q = 17;
        set_c(q, cx, cy);
        const double sq = sqrt(193.0);
        const double w0 = (575.0 + 193.0 * sq) / 8100.0;
        const double w1 = (3355.0 - 91.0 * sq) / 18000.0;
        const double w2 = (655.0 + 17.0 * sq) / 27000.0;
        const double w3 = (685.0 - 49.0 * sq) / 54000.0;
        const double w4 = (1445.0 - 101.0 * sq) / 162000.0;
        const double w[] = {w0, w1, w1, w1, w1, w2, w2, w2, w2,
                            w3, w3, w3, w3, w4, w4, w4, w4};
        for (int i = 0; i < 17; ++i) q->w[i] = w[i];
        q->s = (125.0 + 5.0 * sqrt(193.0)) / 72.0;
        q->order = 7;
        q->feq_order = 3;
        generic_opposite(q);
        return 0;
    }
    case 21: { /* D2Q21.jl:15-71: 25 stored populations, the last 4 have w = 0 */
        static const int cx[] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0,
                                 2, -2, 2, -2, 3, -3, 0, 0, 3, -3, -3, 3};
        static const int cy[] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, 2, -2,
                                 2, -2, -2, 2, 0, 0, 3, -3, 3, -3, 3, -3};
        q->Q = 25;
        set_c(q, cx, cy);
        const double w0 = 91.0 / 324, w1 = 1.0 / 12, w2 = 2.0 / 27, w3 = 7.0 / 360,
                     w4 = 1.0 / 432, w5 = 1.0 / 1620, w6 = 0.0;
        const double w[] = {w0, w1, w1, w1, w1, w2, w2, w2, w2, w3, w3, w3, w3,
                            w4, w4, w4, w4, w5, w5, w5, w5, w6, w6, w6, w6};
        for (int i = 0; i < 25; ++i) q->w[i] = w[i];
        q->s = 3.0 / 2.0; /* D2Q21.jl:33 overrides :16 */
        q->order = 7;
        q->feq_order = 3;
        generic_opposite(q);
        return 0;
    }
    case 37: { /* D2Q37.jl:11-74 */
        static const int cx[] = {0, 1, -1, 0, 0, 1, -1, 1, -1, 2, -2, 0, 0, 2, -2, -2, 2, 1, -1,
                                 1, -1, 2, -2, 2, -2, 3, -3, 0, 0, 3, -3, 3, -3, 1, -1, -1, 1};
        static const int cy[] = {0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, 2, -2, 1, -1, 1, -1, 2, -2,
                                 -2, 2, 2, -2, -2, 2, 0, 0, 3, -3, 1, -1, -1, 1, 3, -3, 3, -3};
        q->Q = 37;
        set_c(q, cx, cy);
        const double g1 = 0.23315066913235250228650, g2 = 0.10730609154221900241246,
                     g3 = 0.05766785988879488203006, g4 = 0.01420821615845075026469,
                     g5 = 0.00535304900051377523273, g6 = 0.00101193759267357547541,
                     g7 = 0.00024530102775771734547, g8 = 0.00028341425299419821740;
        const double w[] = {g1, g2, g2, g2, g2, g3, g3, g3, g3, g4, g4, g4, g4,
                            g5, g5, g5, g5, g5, g5, g5, g5, g6, g6, g6, g6,
                            g7, g7, g7, g7, g8, g8, g8, g8, g8, g8, g8, g8};
        for (int i = 0; i < 37; ++i) q->w[i] = w[i];
        const double r = 1.19697977039307435897239;
        q->s = r * r; /* Julia literal_pow(^, r, Val(2)) = r*r */
        q->order = 9;
        q->feq_order = 4;
        generic_opposite(q);
        return 0;
    }
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* Moments: src/velocity_distribution_function/moments.jl:1-19               */
/* ------------------------------------------------------------------------ */

/* density = sum(f) (moments.jl:3): Base's left fold (f1+f2)+f3+... */
double lbo_density(const lbo_quad *q, const double *f) {
    double r = f[0] + f[1];
    for (int i = 2; i < q->Q; ++i) r = r + f[i];
    return r;
}

/* velocity! (moments.jl:5-19): u_d = 0; u_d += f_i*c_di; u_d /= rho */
void lbo_velocity(const lbo_quad *q, const double *f, double rho, double *u) {
    double u0 = 0.0, u1 = 0.0;
    for (int i = 0; i < q->Q; ++i) u0 = u0 + f[i] * (double)q->cx[i];
    u0 = u0 / rho;
    for (int i = 0; i < q->Q; ++i) u1 = u1 + f[i] * (double)q->cy[i];
    u1 = u1 / rho;
    u[0] = u0;
    u[1] = u1;
}

/* x^4 rounded once from the exact square: x*x = p + e exactly (fma), then
 * (p + e)^2 = p*p + 2 p e (+ e^2, below half an ulp). */
static double lbo_pow4(double x) {
    const double p = x * x;
    const double e = fma(x, x, -p);
    const double hi = p * p;
    const double lo = fma(p, p, -hi);
    return hi + (lo + (2.0 * p) * e);
}

/* ------------------------------------------------------------------------ */
/* Closed-form equilibrium used inside collide!                              */
/*  generic order 2: maxwell_boltzmann_equilibrium.jl:12-66                  */
/*  D2Q17/D2Q21 order 3: velocity_distribution_function/quadratures.jl:3-19,45-61 */
/*  D2Q37 order 4: velocity_distribution_function/quadratures.jl:63-124      */
/*  D2Q4/D2Q5 order 1: velocity_distribution_function/quadratures.jl:129-159 */
/* ------------------------------------------------------------------------ */
void lbo_equilibrium(const lbo_quad *q, double rho, const double *u, double *feq) {
    const double s = q->s;
    double usq = 0.0;
    usq = usq + u[0] * u[0];
    usq = usq + u[1] * u[1];
    for (int i = 0; i < q->Q; ++i) {
        const double ud = (double)q->cx[i] * u[0] + (double)q->cy[i] * u[1];
        const double rw = rho * q->w[i];
        const double a1 = s * ud;
        if (q->feq_order == 1) {
            feq[i] = rw * (1.0 + a1);
            continue;
        }
        const double a2 = ((s * s) * (ud * ud) + 0.0) + (-s) * usq;
        double acc = (1.0 + a1) + 0.5 * a2;
        if (q->feq_order >= 3) {
            const double a3 = (s * ud) * (((s * s) * (ud * ud) - (3.0 * s) * usq) + 0.0);
            acc = acc + (1.0 / 6.0) * a3;
        }
        if (q->feq_order >= 4) {
            /* cs^4 and u_dot_xi^4 are libm pow(x, 4.0) in Julia 1.3; both
             * sides of the parity check use lbo_pow4, which differs from a
             * libm pow by at most one ulp of x^4 */
            const double s2 = s * s, s3 = (s * s) * s, s4 = lbo_pow4(s);
            const double a4 = (s4 * lbo_pow4(ud) - ((6.0 * s3) * usq) * (ud * ud)) +
                              (3.0 * s2) * (usq * usq);
            acc = acc + (1.0 / 24.0) * a4;
        }
        feq[i] = rw * acc;
    }
}

/* ------------------------------------------------------------------------ */
/* Tensor Hermite polynomials scaled by 1/s: src/hermite_polynomials.jl:47-82 */  
/* A rank-n tensor over D=2 is flattened column-major (first index fastest). */
/* ------------------------------------------------------------------------ */
static int delta(int a, int b) { return a == b ? 1 : 0; }

static void hermite_tensor(int n, int cx, int cy, double cs, double *H) {
    const int xi[2] = {cx, cy};
    if (n == 1) {
        H[0] = (double)cx * 1.0;
        H[1] = (double)cy * 1.0;
        return;
    }
    if (n == 2) {
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
                H[a + 2 * b] = (double)(xi[b] * xi[a]) - cs * (double)delta(a, b);
        return;
    }
    if (n == 3) {
        for (int g = 0; g < 2; ++g)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    int A = xi[g] * xi[b] * xi[a];
                    int B = xi[a] * delta(b, g) + xi[b] * delta(a, g) + xi[g] * delta(a, b);
                    H[a + 2 * b + 4 * g] = (double)A - cs * (double)B;
                }
        return;
    }
    /* n == 4 */
    for (int e = 0; e < 2; ++e)
        for (int g = 0; g < 2; ++g)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    int A = xi[e] * xi[g] * xi[b] * xi[a];
                    int B = xi[a] * xi[b] * delta(g, e) + xi[a] * xi[g] * delta(b, e) +
                            xi[a] * xi[e] * delta(b, g) + xi[b] * xi[g] * delta(a, e) +
                            xi[b] * xi[e] * delta(a, g) + xi[g] * xi[e] * delta(a, b);
                    int C = delta(a, b) * delta(g, e) + delta(a, g) * delta(b, e) +
                            delta(a, e) * delta(b, g);
                    H[a + 2 * b + 4 * g + 8 * e] =
                        ((double)A - cs * (double)B) + (cs * cs) * (double)C;
                }
}

/* equilibrium_coefficient(Val{n}) at T = 1 (velocity_distribution_function/hermite.jl:37-77):
 * rho * (u_a u_b ...), the (T-1) terms vanish. */
static void eq_coefficient(int n, double rho, const double *u, double *a) {
    if (n == 1) {
        a[0] = rho * u[0];
        a[1] = rho * u[1];
        return;
    }
    int m = 1 << n;
    for (int k = 0; k < m; ++k) {
        double p = u[k & 1] * u[(k >> 1) & 1];
        if (n == 2) {
            a[k] = rho * (p + 0.0);
            continue;
        }
        p = p * u[(k >> 2) & 1];
        if (n == 4) p = p * u[(k >> 3) & 1];
        a[k] = rho * (p + 0.0);
    }
}

static int factorial_int(int n) {
    int r = 1;
    for (int k = 2; k <= n; ++k) r *= k;
    return r;
}

/* Generic Hermite-sum equilibrium hermite_based_equilibrium! (hermite.jl:10-33).
 * Used by the reference's initialization, and here as an independent
 * formulation that pins the closed-form lbo_equilibrium. */
void lbo_hermite_equilibrium(const lbo_quad *q, double rho, const double *u, double *feq) {
    const int N = q->order / 2;
    const double cs = 1.0 / q->s;
    double aeq[4][16];
    for (int n = 1; n <= N; ++n) eq_coefficient(n, rho, u, aeq[n - 1]);
    for (int i = 0; i < q->Q; ++i) {
        double sum = 0.0;
        for (int n = 1; n <= N; ++n) {
            double H[16];
            hermite_tensor(n, q->cx[i], q->cy[i], cs, H);
            double d = 0.0;
            for (int k = 0; k < (1 << n); ++k) d = d + aeq[n - 1][k] * H[k];
            double term = d / ((double)factorial_int(n) * pow(cs, (double)n));
            sum = (n == 1) ? term : sum + term;
        }
        feq[i] = q->w[i] * (rho + sum);
    }
}

/* equilibrium_coefficient(Val{n}, q, rho, u, T) with the temperature terms
 * (velocity_distribution_function/hermite.jl:37-77), entry k of the
 * column-major tensor (first index fastest).  Val{2}: u u' + cs (T - 1) I adds
 * to the diagonal only.  Val{4} keeps the reference's indices as written:
 * its third and fifth (T - 1) terms pair u_a u_d and u_b u_d with delta(b, d)
 * and delta(a, d). */
static void eq_coefficient_T(int n, double rho, const double *u, double T, double cs,
                             double *a) {
    const double ct = cs * (T - 1.0);
    if (n == 1) {
        a[0] = rho * u[0];
        a[1] = rho * u[1];
        return;
    }
    for (int k = 0; k < (1 << n); ++k) {
        const int ia = k & 1, ib = (k >> 1) & 1, ic = (k >> 2) & 1, id = (k >> 3) & 1;
        double v;
        if (n == 2) {
            v = u[ia] * u[ib];
            if (ia == ib) v = v + ct;
        } else if (n == 3) {
            const double S = (u[ia] * (double)delta(ib, ic) + u[ib] * (double)delta(ia, ic)) +
                             u[ic] * (double)delta(ia, ib);
            v = ((u[ia] * u[ib]) * u[ic]) + ct * S;
        } else {
            const double S1 =
                ((((u[ia] * u[ib]) * (double)delta(ic, id) + (u[ia] * u[ic]) * (double)delta(ib, id)) +
                  (u[ia] * u[id]) * (double)delta(ib, id)) +
                 (u[ib] * u[ic]) * (double)delta(ia, id)) +
                (u[ib] * u[id]) * (double)delta(ia, id);
            const double S1b = S1 + (u[ic] * u[id]) * (double)delta(ia, ib);
            const int S0 = (delta(ia, ib) * delta(ic, id) + delta(ia, ic) * delta(ib, id)) +
                           delta(ia, id) * delta(ib, ic);
            v = ((((u[ia] * u[ib]) * u[ic]) * u[id]) + ct * S1b) +
                ((cs * cs) * ((T - 1.0) * (T - 1.0))) * (double)S0;
        }
        a[k] = rho * v;
    }
}

/* initial_condition(strategy, q, problem, x, y) of one node, for the
 * analytic strategies:
 *   mode 0, AnalyticalEquilibrium (initial_conditions/analytical_equilibrium.jl:
 *     8-17): equilibrium(q, rho, u, T) = hermite_based_equilibrium!
 *     (hermite.jl:10-33, maxwell_boltzmann_equilibrium.jl:1-10), the Hermite
 *     sum w_i (rho + sum_n a_eq[n] : H_n(c_i) / (n! cs^n)) at temperature T;
 *   mode 1, AnalyticalEquilibriumAndOffEquilibrium for the Taylor-Green vortex
 *     (analytical_offequilibrium.jl:50-87): minus
 *     w_i (s tau_eff rho 1) / 2 * H_2(c_i) : (G + G'), tau_eff = tau + 1/2;
 *   mode 2, the same for any other problem (analytical_offequilibrium.jl:10-48):
 *     plus -factor w_i 0.5 ((tau + 0.5) s) H_2(c_i) : (G + G').
 * G is the 2 x 2 velocity gradient as the reference passes it (column-major,
 * G[a + 2 b]); tau = s nu, the reference's tau before the 1/2.  cs^n comes
 * from libm pow, as Julia's Float64 ^ Int; dot is a left fold in column-major
 * order. */
void lbo_init_node(const lbo_quad *q, int mode, double rho, const double *u, double T,
                   const double *G, double tau, double factor, double *f) {
    const int N = q->order / 2;
    const double cs = 1.0 / q->s;
    double aeq[4][16];
    for (int n = 1; n <= N; ++n) eq_coefficient_T(n, rho, u, T, cs, aeq[n - 1]);
    double S[4];
    if (mode != 0)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) S[a + 2 * b] = G[a + 2 * b] + G[b + 2 * a];
    for (int i = 0; i < q->Q; ++i) {
        double sum = 0.0;
        for (int n = 1; n <= N; ++n) {
            double H[16];
            hermite_tensor(n, q->cx[i], q->cy[i], cs, H);
            double d = 0.0;
            for (int k = 0; k < (1 << n); ++k) d = (k == 0) ? aeq[n - 1][k] * H[k] : d + aeq[n - 1][k] * H[k];
            const double term = d / ((double)factorial_int(n) * pow(cs, (double)n));
            sum = (n == 1) ? term : sum + term;
        }
        double fi = q->w[i] * (rho + sum);
        if (mode != 0) {
            double H2[4];
            hermite_tensor(2, q->cx[i], q->cy[i], cs, H2);
            double d2 = H2[0] * S[0];
            for (int k = 1; k < 4; ++k) d2 = d2 + H2[k] * S[k];
            if (mode == 1)
                fi = fi - ((q->w[i] * (((q->s * (tau + 0.5)) * rho) * 1.0)) / 2.0) * d2;
            else
                fi = fi + ((((-factor) * q->w[i]) * 0.5) * ((tau + 0.5) * q->s)) * d2;
        }
        f[i] = fi;
    }
}

/* initialize(strategy, q, problem) over a lattice (initial_conditions.jl:7-22):
 * planes rho, ux, uy, T (NULL: 1) and G (4 planes, NULL for mode 0), each
 * (nx, ny) column-major; f is (nx, ny, Q). */
void lbo_initialize(const lbo_quad *q, int nx, int ny, int mode, const double *rho,
                    const double *ux, const double *uy, const double *T, const double *G,
                    double tau, double factor, double *f) {
    const long plane = (long)nx * ny;
    double fn[37];
#pragma omp parallel for private(fn) schedule(static)
    for (long n = 0; n < plane; ++n) {
        const double u[2] = {ux[n], uy[n]};
        double g[4] = {0.0, 0.0, 0.0, 0.0};
        if (G)
            for (int k = 0; k < 4; ++k) g[k] = G[k * plane + n];
        lbo_init_node(q, mode, rho[n], u, T ? T[n] : 1.0, g, tau, factor, fn);
        for (int i = 0; i < q->Q; ++i) f[i * plane + n] = fn[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Collision operators                                                       */
/* ------------------------------------------------------------------------ */
#define IDX(x, y, i) ((size_t)(x) + (size_t)nx * ((size_t)(y) + (size_t)ny * (size_t)(i)))

static void node_force(int kind, const double *force, int nx, int ny, int x, int y, double *F) {
    if (kind == LBO_FORCE_CONST) {
        F[0] = force[0];
        F[1] = force[1];
    } else if (kind == LBO_FORCE_FIELD) {
        size_t n = (size_t)nx * (size_t)ny;
        F[0] = force[(size_t)x + (size_t)nx * (size_t)y];
        F[1] = force[n + (size_t)x + (size_t)nx * (size_t)y];
    } else {
        F[0] = F[1] = 0.0;
    }
}

/* collide!(::SRT) -- src/collision_models/srt.jl:18-62 */
void lbo_collide_srt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau, int force_kind, const double *force) {
    const int Q = q->Q;
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            double f[LBO_QMAX], feq[LBO_QMAX], u[2];
            for (int i = 0; i < Q; ++i) f[i] = fin[IDX(x, y, i)];
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            if (force_kind != LBO_FORCE_NONE) {
                double F[2];
                node_force(force_kind, force, nx, ny, x, y, F);
                double v[2] = {u[0] + tau * F[0], u[1] + tau * F[1]}; /* srt.jl:54 */
                lbo_equilibrium(q, rho, v, feq);
            } else {
                lbo_equilibrium(q, rho, u, feq);
            }
            for (int i = 0; i < Q; ++i)
                fout[IDX(x, y, i)] = (1.0 - 1.0 / tau) * f[i] + (1.0 / tau) * feq[i];
        }
}

/* collide!(::TRT) -- src/collision_models/trt.jl:42-97 */
void lbo_collide_trt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau_s, double tau_a, int force_kind, const double *force) {
    const int Q = q->Q;
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            double f[LBO_QMAX], feq[LBO_QMAX], u[2];
            for (int i = 0; i < Q; ++i) f[i] = fin[IDX(x, y, i)];
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            if (force_kind != LBO_FORCE_NONE) {
                double F[2];
                node_force(force_kind, force, nx, ny, x, y, F);
                double v[2] = {u[0] + tau_a * F[0], u[1] + tau_a * F[1]}; /* trt.jl:79 */
                lbo_equilibrium(q, rho, v, feq);
            } else {
                lbo_equilibrium(q, rho, u, feq);
            }
            for (int i = 0; i < Q; ++i) {
                const int j = q->opp[i];
                const double feq_s = 0.5 * (feq[i] + feq[j]);
                const double feq_a = 0.5 * (feq[i] - feq[j]);
                const double f_s = 0.5 * (f[i] + f[j]);
                const double f_a = 0.5 * (f[i] - f[j]);
                fout[IDX(x, y, i)] =
                    f[i] + ((-(1.0 / tau_s)) * (f_s - feq_s) - (1.0 / tau_a) * (f_a - feq_a));
            }
        }
}

/* collide_mrt! -- src/collision_models/mrt.jl:56-118 (regularized projection) */
void lbo_collide_mrt(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     const double *taus, int ntaus, int force_kind, const double *force) {
    const int Q = q->Q;
    const int N = q->order / 2; /* mrt.jl:68 */
    const double s = q->s, cs_inv = 1.0 / q->s;
    (void)ntaus;
    /* Hs[n][i]: hermite(Val{n}, c_i, q) -- mrt.jl:31 */
    static double Hs[5][LBO_QMAX][16];
    for (int n = 1; n <= N; ++n)
        for (int i = 0; i < Q; ++i) hermite_tensor(n, q->cx[i], q->cy[i], cs_inv, Hs[n][i]);
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            double f[LBO_QMAX], u[2];
            for (int i = 0; i < Q; ++i) f[i] = fin[IDX(x, y, i)];
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            if (force_kind != LBO_FORCE_NONE) {
                double F[2];
                node_force(force_kind, force, nx, ny, x, y, F);
                u[0] = u[0] + taus[1] * F[0]; /* mrt.jl:94 */
                u[1] = u[1] + taus[1] * F[1];
            }
            double acoll[5][16];
            for (int n = 2; n <= N; ++n) {
                const int m = 1 << n;
                double aeq[16], af[16];
                eq_coefficient(n, rho, u, aeq);               /* mrt.jl:99 */
                for (int k = 0; k < m; ++k) {                 /* mrt.jl:102 */
                    double a = f[0] * Hs[n][0][k] + f[1] * Hs[n][1][k];
                    for (int i = 2; i < Q; ++i) a = a + f[i] * Hs[n][i][k];
                    af[k] = a;
                }
                const double t = taus[n - 1];
                for (int k = 0; k < m; ++k)                   /* mrt.jl:104 */
                    acoll[n][k] = (1.0 - 1.0 / t) * af[k] + (1.0 / t) * aeq[k];
            }
            for (int i = 0; i < Q; ++i) {                     /* mrt.jl:106-115 */
                const double ru = u[0] * Hs[1][i][0] + u[1] * Hs[1][i][1];
                double hsum = 0.0;
                for (int n = 2; n <= N; ++n) {
                    double d = 0.0;
                    for (int k = 0; k < (1 << n); ++k) d = d + acoll[n][k] * Hs[n][i][k];
                    double e = (pow(s, (double)n) * d) / (double)factorial_int(n);
                    hsum = (n == 2) ? e : hsum + e;
                }
                fout[IDX(x, y, i)] = q->w[i] * ((rho + (s * rho) * ru) + hsum);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* stream! -- src/stream.jl:19-30,69-74 (pull with mod1 wrap)               */
/* ------------------------------------------------------------------------ */
static int mod0(int v, int n) {
    int r = v % n;
    return r < 0 ? r + n : r;
}

void lbo_stream(const lbo_quad *q, int nx, int ny, const double *f, double *fnew) {
    for (int i = 0; i < q->Q; ++i) {
        const int cx = q->cx[i], cy = q->cy[i];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int y = 0; y < ny; ++y) {
            const int sy = mod0(y - cy, ny);
            for (int x = 0; x < nx; ++x) fnew[IDX(x, y, i)] = f[IDX(mod0(x - cx, nx), sy, i)];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* apply!(bcs, ...) -- src/boundary_conditions.jl:6-20; BounceBack           */
/* boundary_conditions/bounce_back.jl:8-72; MovingWall{North}                */
/* boundary_conditions/moving_wall.jl:17-38                                  */
/* ------------------------------------------------------------------------ */
int lbo_apply_bcs(const lbo_quad *q, int nx, int ny, const lbo_bc *bcs, int nbc,
                  double *fnew, const double *fold) {
    /* the reference indexes f_new[x_idx, y_idx, :] without @inbounds: a range
     * outside the lattice is a BoundsError, reported here as -1 */
    for (int b = 0; b < nbc; ++b) {
        const lbo_bc *bc = &bcs[b];
        if (bc->kind != LBO_BOUNCE_BACK) continue;
        if ((bc->x1 >= bc->x0 && (bc->x0 < 1 || bc->x1 > nx)) ||
            (bc->y1 >= bc->y0 && (bc->y0 < 1 || bc->y1 > ny)))
            return -1;
    }
    for (int b = 0; b < nbc; ++b) {
        const lbo_bc *bc = &bcs[b];
        if (bc->kind == LBO_BOUNCE_BACK) {
            for (int i = 0; i < q->Q; ++i) {
                const int j = q->opp[i];
                if (bc->direction == LBO_NORTH || bc->direction == LBO_SOUTH) {
                    for (int yi = bc->y0; yi <= bc->y1; ++yi) {
                        int hit = bc->direction == LBO_NORTH ? (yi + q->cy[j] > ny)
                                                             : (yi + q->cy[j] < 1);
                        if (!hit) continue;
                        for (int xi = bc->x0; xi <= bc->x1; ++xi)
                            fnew[IDX(xi - 1, yi - 1, i)] = fold[IDX(xi - 1, yi - 1, j)];
                    }
                } else {
                    for (int xi = bc->x0; xi <= bc->x1; ++xi) {
                        int hit = bc->direction == LBO_EAST ? (xi + q->cx[j] > nx)
                                                            : (xi + q->cx[j] < 1);
                        if (!hit) continue;
                        for (int yi = bc->y0; yi <= bc->y1; ++yi)
                            fnew[IDX(xi - 1, yi - 1, i)] = fold[IDX(xi - 1, yi - 1, j)];
                    }
                }
            }
        } else { /* moving wall, North only (the reference defines no other) */
            const double s = q->s;
            const double a1eq[2] = {bc->rho * bc->ux, bc->rho * bc->uy};
            for (int i = 0; i < q->Q; ++i) {
                const double H1[2] = {(double)q->cx[i] * 1.0, (double)q->cy[i] * 1.0};
                const double dot = a1eq[0] * H1[0] + a1eq[1] * H1[1];
                const double a1 = (q->w[i] * s) * dot;
                const int j = q->opp[i];
                for (int yi = 1; yi <= ny; ++yi) {
                    if (yi + q->cy[j] <= ny) continue;
                    for (int xi = 1; xi <= nx; ++xi)
                        fnew[IDX(xi - 1, yi - 1, i)] = fold[IDX(xi - 1, yi - 1, j)] + 2.0 * a1;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* simulate(model, time) step loop -- src/lattice_boltzmann_model.jl:60-77   */
/* collide f_stream -> f_collision; stream f_collision -> f_stream; BCs      */
/* ------------------------------------------------------------------------ */
int lbo_simulate(const lbo_quad *q, int nx, int ny, double *f_stream, double *f_collision,
                 int collision, const double *taus, int ntaus, int force_kind,
                 const double *force, const lbo_bc *bcs, int nbc, long nsteps) {
    for (long t = 0; t < nsteps; ++t) {
        if (collision == 0)
            lbo_collide_srt(q, nx, ny, f_stream, f_collision, taus[0], force_kind, force);
        else if (collision == 1)
            lbo_collide_trt(q, nx, ny, f_stream, f_collision, taus[0], taus[1], force_kind,
                            force);
        else
            lbo_collide_mrt(q, nx, ny, f_stream, f_collision, taus, ntaus, force_kind, force);
        lbo_stream(q, nx, ny, f_collision, f_stream);
        if (lbo_apply_bcs(q, nx, ny, bcs, nbc, f_stream, f_collision) != 0) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Moment fields: density/velocity! (moments.jl:1-19) and deviatoric_tensor  */
/* (moments.jl:81-96) as used   by TrackHydrodynamicErrors                   */
/* (processing_methods/track_hydrodynamic_errors.jl:139-177)                 */
/* ------------------------------------------------------------------------ */
void lbo_moments(const lbo_quad *q, int nx, int ny, const double *fin, double tau,
                 double *rho_o, double *ux_o, double *uy_o, double *sxx_o, double *sxy_o,
                 double *syy_o) {
    const int Q = q->Q;
    const double cs = 1.0 / q->s;
    double H2[LBO_QMAX][4];
    for (int i = 0; i < Q; ++i) hermite_tensor(2, q->cx[i], q->cy[i], cs, H2[i]);
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            double f[LBO_QMAX], u[2];
            for (int i = 0; i < Q; ++i) f[i] = fin[IDX(x, y, i)];
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            const size_t n = (size_t)x + (size_t)nx * (size_t)y;
            if (rho_o) rho_o[n] = rho;
            if (ux_o) ux_o[n] = u[0];
            if (uy_o) uy_o[n] = u[1];
            if (!(sxx_o || sxy_o || syy_o)) continue;
            double abar[4];
            for (int k = 0; k < 4; ++k) {
                double a = f[0] * H2[0][k] + f[1] * H2[1][k];
                for (int i = 2; i < Q; ++i) a = a + f[i] * H2[i][k];
                abar[k] = a;
            }
            double aeq[4];
            eq_coefficient(2, rho, u, aeq);
            const double den = 1.0 + 1.0 / (2.0 * tau);
            double sg[4];
            for (int k = 0; k < 4; ++k) sg[k] = (abar[k] - aeq[k]) / den;
            const double tr2 = (sg[0] + sg[3]) / 2.0;
            if (sxx_o) sxx_o[n] = sg[0] - tr2;
            if (sxy_o) sxy_o[n] = sg[2];
            if (syy_o) syy_o[n] = sg[3] - tr2;
        }
}

/* ------------------------------------------------------------------------ */
/* Processing: one row of next!(::TrackHydrodynamicErrors)                  */
/* (processing_methods/track_hydrodynamic_errors.jl:52-222) and the sums of  */
/* the stopping criteria (processing_methods/stopping_criteria/              */
/* stopping_criteria.jl:17-115).  The sums run sequentially in the           */
/* reference's loop order `for x_idx in 1:nx, y_idx in 1:ny` (x outer).     */
/* expected: 8 planes rho, u_x, u_y, p, sigma_xx, sigma_xy, sigma_yx,        */
/* sigma_yy; expected field k = affine[2k] + affine[2k+1] * plane k.         */
/* ------------------------------------------------------------------------ */
void lbo_track_errors(const lbo_quad *q, int nx, int ny, const double *fin,
                      const double *expected, const double *affine, double tau, double u_max,
                      double delta, double *out) {
    const int Q = q->Q;
    const double cs = 1.0 / q->s;
    const size_t plane = (size_t)nx * (size_t)ny;
    double H2[LBO_QMAX][4];
    for (int i = 0; i < Q; ++i) hermite_tensor(2, q->cx[i], q->cy[i], cs, H2[i]);
    double err_rho = 0.0, err_u = 0.0, err_p = 0.0;
    double err_sxx = 0.0, err_sxy = 0.0, err_syy = 0.0, err_syx = 0.0;
    double tot_u2 = 0.0, tot_sxx = 0.0, tot_syy = 0.0, tot_sxy = 0.0, tot_syx = 0.0, tot_p2 = 0.0;
    double mass = 0.0, momentum = 0.0, energy = 0.0, nonfinite = 0.0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            const size_t n = (size_t)x + (size_t)nx * (size_t)y;
            double e[8];
            for (int k = 0; k < 8; ++k) e[k] = affine[2 * k] + affine[2 * k + 1] * expected[plane * k + n];
            /* lines 127-132 */
            tot_u2 += e[1] * e[1] + e[2] * e[2];
            tot_sxx += e[4] * e[4];
            tot_syy += e[7] * e[7];
            tot_sxy += e[5] * e[5];
            tot_syx += e[6] * e[6];
            tot_p2 += e[3] * e[3];
            double f[LBO_QMAX], u[2];
            int finite = 1;
            for (int i = 0; i < Q; ++i) {
                f[i] = fin[IDX(x, y, i)];
                finite = finite && isfinite(f[i]);
            }
            if (!finite) nonfinite += 1.0;
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            /* a_bar_2, a_eq_2 (lines 149-150) */
            double abar[4], aeq[4];
            for (int k = 0; k < 4; ++k) {
                double a = f[0] * H2[0][k] + f[1] * H2[1][k];
                for (int i = 2; i < Q; ++i) a = a + f[i] * H2[i][k];
                abar[k] = a;
            }
            eq_coefficient(2, rho, u, aeq);
            /* a_2 (line 153), P = a_2 - rho (u u' - I) (line 173), p = tr(P)/D (line 181) */
            const double c = 1.0 / (2.0 * tau);
            const double den = 1.0 + 1.0 / (2.0 * tau);
            double P[4];
            for (int k = 0; k < 4; ++k) {
                const int a = k & 1, b = k >> 1;
                const double a2 = (abar[k] + c * aeq[k]) / den;
                P[k] = a2 - rho * (a == b ? u[a] * u[b] - 1.0 : u[a] * u[b]);
            }
            const double p = ((0.0 + P[0]) + P[3]) / 2.0;
            /* sigma_lb = deviatoric_tensor(q, tau, f, rho, u) (line 177, moments.jl:81-96) */
            double sg[4];
            for (int k = 0; k < 4; ++k) sg[k] = (abar[k] - aeq[k]) / den;
            const double tr2 = ((0.0 + sg[0]) + sg[3]) / 2.0;
            sg[0] = sg[0] - tr2;
            sg[3] = sg[3] - tr2;
            /* dimensionless_* (problems.jl:126-134), lines 184-187 */
            const double d1 = u[0] / u_max, d2 = u[1] / u_max;
            const double sfac = 1.0 / (u_max * u_max);
            for (int k = 0; k < 4; ++k) sg[k] = sg[k] * sfac;
            /* lines 192-202, cell weight Δ = 1 */
            err_p += (p - e[3]) * (p - e[3]);
            err_rho += (rho - e[0]) * (rho - e[0]);
            err_u += (d1 - e[1]) * (d1 - e[1]) + (d2 - e[2]) * (d2 - e[2]);
            err_sxx += (e[4] - sg[0]) * (e[4] - sg[0]);
            err_sxy += (e[5] - sg[2]) * (e[5] - sg[2]);
            err_syx += (e[6] - sg[1]) * (e[6] - sg[1]);
            err_syy += (e[7] - sg[3]) * (e[7] - sg[3]);
            mass += rho;
            momentum += rho * (d1 + d2);
            energy += rho * (d1 * d1 + d2 * d2);
        }
    /* lines 205-221 */
    out[0] = sqrt(err_rho);
    out[1] = sqrt(err_u / tot_u2);
    out[2] = sqrt(err_p / tot_p2);
    out[3] = sqrt(err_sxx / tot_sxx);
    out[4] = sqrt(err_sxy / tot_sxy);
    out[5] = sqrt(err_syy / tot_syy);
    out[6] = sqrt(err_syx / tot_syx);
    out[7] = delta * mass;
    out[8] = delta * momentum;
    out[9] = delta * energy;
    out[10] = nonfinite;
}

/* MeanVelocityStoppingCriteria (stopping_criteria.jl:17-39): out[0..1] mean
 * u_x, u_y; VelocityConvergenceStoppingCriteria (stopping_criteria.jl:71-99)
 * with u_old (2 planes, may be NULL; updated to u): out[2] sum |u - u_old|^2,
 * out[3] sum |u_old|^2; out[4] nodes with a non-finite population. */
void lbo_velocity_stats(const lbo_quad *q, int nx, int ny, const double *fin, double *u_old,
                        double *out) {
    const int Q = q->Q;
    const size_t plane = (size_t)nx * (size_t)ny;
    double sx = 0.0, sy = 0.0, du2 = 0.0, old2 = 0.0, nonfinite = 0.0;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            const size_t n = (size_t)x + (size_t)nx * (size_t)y;
            double f[LBO_QMAX], u[2];
            int finite = 1;
            for (int i = 0; i < Q; ++i) {
                f[i] = fin[IDX(x, y, i)];
                finite = finite && isfinite(f[i]);
            }
            if (!finite) nonfinite += 1.0;
            const double rho = lbo_density(q, f);
            lbo_velocity(q, f, rho, u);
            sx += u[0];
            sy += u[1];
            if (u_old) {
                const double o1 = u_old[n], o2 = u_old[plane + n];
                du2 += (u[0] - o1) * (u[0] - o1) + (u[1] - o2) * (u[1] - o2);
                old2 += o1 * o1 + o2 * o2;
                u_old[n] = u[0];
                u_old[plane + n] = u[1];
            }
        }
    out[0] = sx / (double)plane;
    out[1] = sy / (double)plane;
    out[2] = du2;
    out[3] = old2;
    out[4] = nonfinite;
}

/* ------------------------------------------------------------------------ */
/* Iterative initialization of Mei et al.                                   */
/* (initial_conditions/mei_et_al.jl:13-41,                                  */
/* collision_models/iterative_initialization.jl:12-58,                      */
/* processing_methods/process_iterative_initialization.jl:1-22)             */
/* ------------------------------------------------------------------------ */
/* nonlinear_term[x, y, i] = w_i rho_0 (a_H_1 + a_H_2 / 2) of the lattice      */
/* velocity u (nx, ny, 2) (iterative_initialization.jl:12-40)                */
void lbo_mei_nonlinear(const lbo_quad *q, int nx, int ny, const double *u, double *N) {
    const int Q = q->Q;
    const size_t plane = (size_t)nx * (size_t)ny;
    const double cs = q->s;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            const size_t n = (size_t)x + (size_t)nx * (size_t)y;
            const double u0 = u[n], u1 = u[plane + n];
            const double usq = u0 * u0 + u1 * u1;
            for (int i = 0; i < Q; ++i) {
                const double ud = (0.0 + (double)q->cx[i] * u0) + (double)q->cy[i] * u1;
                const double a1 = cs * ud;
                const double a2 = (cs * cs) * (ud * ud) - cs * usq;
                N[IDX(x, y, i)] = (q->w[i] * 1.0) * (a1 + a2 / 2.0);
            }
        }
}

/* collide!(::IterativeInitializationCollisionModel) (iterative_initialization.jl:41-58):
 * f_out = (1 - 1/tau) f_in + (1/tau) (w_i rho + N_i) */
void lbo_collide_mei(const lbo_quad *q, int nx, int ny, const double *fin, double *fout,
                     double tau, const double *N) {
    const int Q = q->Q;
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            double f[LBO_QMAX];
            for (int i = 0; i < Q; ++i) f[i] = fin[IDX(x, y, i)];
            const double rho = lbo_density(q, f);
            for (int i = 0; i < Q; ++i) {
                const double feq = q->w[i] * rho + N[IDX(x, y, i)];
                fout[IDX(x, y, i)] = (1.0 - 1.0 / tau) * f[i] + (1.0 / tau) * feq;
            }
        }
}

/* initialize(::IterativeInitializationMeiEtAl, q, problem) (mei_et_al.jl:13-41):
 * f = w everywhere, then simulate(model, 1:max_steps) with the collision above,
 * stream! and the BC list; after each step next!(ProcessIterativeInitialization)
 * checks DensityConvergence (density_convergence.jl:6-17), whose loop visits
 * only node (nx, ny): stop when |rho - rho_old| there is < eps or > 100.
 * Returns the number of steps run; f_stream holds the result. */
long lbo_mei_initialize(const lbo_quad *q, int nx, int ny, const double *u, double tau,
                        double eps, long max_steps, const lbo_bc *bcs, int nbc,
                        double *f_stream, double *f_collision) {
    const int Q = q->Q;
    const size_t plane = (size_t)nx * (size_t)ny;
    double *N = (double *)malloc(plane * (size_t)Q * sizeof(double));
    if (!N) return -1;
    lbo_mei_nonlinear(q, nx, ny, u, N);
    for (int i = 0; i < Q; ++i)
        for (size_t n = 0; n < plane; ++n) {
            f_stream[plane * (size_t)i + n] = q->w[i];
            f_collision[plane * (size_t)i + n] = q->w[i];
        }
    double rho = 0.0, rho_old = 0.0;
    const size_t corner = (size_t)(nx - 1) + (size_t)nx * (size_t)(ny - 1);
    long t = 1;
    for (; t <= max_steps; ++t) {
        lbo_collide_mei(q, nx, ny, f_stream, f_collision, tau, N);
        lbo_stream(q, nx, ny, f_collision, f_stream);
        if (lbo_apply_bcs(q, nx, ny, bcs, nbc, f_stream, f_collision) != 0) {
            free(N);
            return -1;
        }
        double f[LBO_QMAX];
        for (int i = 0; i < Q; ++i) f[i] = f_stream[plane * (size_t)i + corner];
        rho_old = rho;
        rho = lbo_density(q, f);
        const double d = fabs(rho - rho_old);
        if (d < eps || d > 100.0) break;
    }
    free(N);
    return t > max_steps ? max_steps : t;
}
