// lbm_capi.cpp -- C ABI of liblbm_hip.so (declared in include/lbm_hip.h).
//
// Host-side logic: handles, validation, parameter packing and the step
// schedule of lbm_simulate.  All device work goes to one HIP stream owned by
// the library.  Compiled with -ffp-contract=off so that the constants it
// precomputes (1/tau, s*s, moving-wall terms, ...) round exactly as the
// reference's per-node expressions do.
#include "lbm_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "lbm_internal.h"
#include "lbm_quad.h"

namespace {

thread_local std::string g_err;
hipStream_t g_stream = nullptr;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(-3, "%s failed: %s", #expr, hipGetErrorString(e_));               \
    } while (0)

int ensure_stream() {
    if (g_stream) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(-3, "no HIP device visible (liblbm_hip needs an MI355X / gfx950 GPU)");
    HIP_TRY(hipStreamCreate(&g_stream));
    return 0;
}

// x^4 rounded once from an exact two-product (the same sequence as the kernels)
double pow4(double x) {
    double p = x * x;
    double e = std::fma(x, x, -p);
    double hi = p * p;
    double lo = std::fma(p, p, -hi);
    return hi + (lo + (2.0 * p) * e);
}

struct HostQuad {
    int id = 0, Q = 0, order = 0, maxc = 0;
    int cx[37], cy[37], opp[37];
    double w[37];
    double s = 0.0;
    // the FP32 storages' weights and s (see tune_f32); f32_set = false: rounded
    float w32[37] = {};
    float s32 = 0.0f;
    bool f32_set = false;
};

template <int QID>
void fill_tables(HostQuad &q) {
    using QTq = lbm::QT<QID>;
    q.Q = QTq::Q;
    q.order = QTq::ORDER;
    q.maxc = QTq::MAXC;
    for (int i = 0; i < QTq::Q; ++i) {
        q.cx[i] = QTq::cx(i);
        q.cy[i] = QTq::cy(i);
        q.opp[i] = QTq::opp(i);
    }
}

// Weights and speed_of_sound_squared with the reference's expressions
// (src/quadratures/D2Q{4,5,9,13,17,21,37}.jl).
static bool host_quad_exact(int qid, HostQuad &q) {
    q = HostQuad{};
    q.id = qid;
    switch (qid) {
    case 4:
        fill_tables<4>(q);
        for (int i = 0; i < 4; ++i) q.w[i] = 1.0 / 4.0;
        q.s = 2.0;
        return true;
    case 5:
        fill_tables<5>(q);
        q.w[0] = 4.0 / 6.0;
        for (int i = 1; i < 5; ++i) q.w[i] = 1.0 / 12.0;
        q.s = 6.0;
        return true;
    case 9: {
        fill_tables<9>(q);
        const double w[9] = {4.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9,
                             1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9};
        for (int i = 0; i < 9; ++i) q.w[i] = w[i];
        q.s = 3.0;
        return true;
    }
    case 13: {
        fill_tables<13>(q);
        const double a = 3.0 / 8, b = 1.0 / 12, c = 1.0 / 16, d = 1.0 / 96;
        const double w[13] = {a, b, b, b, b, c, c, c, c, d, d, d, d};
        for (int i = 0; i < 13; ++i) q.w[i] = w[i];
        q.s = 2.0;
        return true;
    }
    case 17: {
        fill_tables<17>(q);
        const double sq = std::sqrt(193.0);
        const double g[5] = {(575.0 + 193.0 * sq) / 8100.0, (3355.0 - 91.0 * sq) / 18000.0,
                             (655.0 + 17.0 * sq) / 27000.0, (685.0 - 49.0 * sq) / 54000.0,
                             (1445.0 - 101.0 * sq) / 162000.0};
        q.w[0] = g[0];
        for (int i = 1; i < 17; ++i) q.w[i] = g[1 + (i - 1) / 4];
        q.s = (125.0 + 5.0 * std::sqrt(193.0)) / 72.0;
        return true;
    }
    case 21: {
        fill_tables<21>(q);
        const double g[7] = {91.0 / 324, 1.0 / 12, 2.0 / 27, 7.0 / 360, 1.0 / 432, 1.0 / 1620, 0.0};
        q.w[0] = g[0];
        for (int i = 1; i < 25; ++i) q.w[i] = g[1 + (i - 1) / 4];
        q.s = 3.0 / 2.0;
        return true;
    }
    case 37: {
        fill_tables<37>(q);
        const double g[8] = {0.23315066913235250228650, 0.10730609154221900241246,
                             0.05766785988879488203006, 0.01420821615845075026469,
                             0.00535304900051377523273, 0.00101193759267357547541,
                             0.00024530102775771734547, 0.00028341425299419821740};
        const int grp[37] = {0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4,
                             4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7, 7, 7};
        for (int i = 0; i < 37; ++i) q.w[i] = g[grp[i]];
        const double r = 1.19697977039307435897239;
        q.s = r * r;
        return true;
    }
    default:
        return false;
    }
}

// FP32 weights and s for the FP32-arithmetic storages.  Rounded one by one,
// the discrete identities that make the equilibrium conserve mass and
// momentum, sum_i w_i = 1 and s sum_i w_i c_ix^2 = 1, hold in FP32 only to
// 1e-8 - 3e-8 on the sets with irrational weights (D2Q17: -1.3e-8 and
// -2.8e-8).  The collision turns a momentum defect e into a push of
// omega e rho u per step, which a uniformly accelerated flow integrates to a
// relative velocity error of about omega e t / 2 (-3.5e-4 after 20,000 steps
// at tau = 0.8 on D2Q17; -5.3e-4 measured).  Each group of equal weights and s
// move here by at most K ulps from their rounding, to the combination that
// minimises max(|sum w - 1|, |s sum w c_x^2 - 1|) (sums of floats times
// integers, exact in double); it is kept when it at least halves the defect
// of the plain rounding (D2Q13/17/21/37; D2Q4/5/9 keep the rounded values,
// s = 3 stays exact on D2Q9).
static void tune_f32(HostQuad &q) {
    double gw[37];
    int gn[37], gm[37], grp[37], ng = 0;
    for (int i = 0; i < q.Q; ++i) {
        grp[i] = -1;
        if (q.w[i] == 0.0) continue;
        int g = 0;
        while (g < ng && gw[g] != q.w[i]) ++g;
        if (g == ng) {
            gw[ng] = q.w[i];
            gn[ng] = gm[ng] = 0;
            ++ng;
        }
        grp[i] = g;
        gn[g] += 1;
        gm[g] += q.cx[i] * q.cx[i];
    }
    const int K = ng <= 6 ? 3 : 2, span = 2 * K + 1, KS = 2;
    // candidates: the rounding stepped -K .. K ulps (index K = the rounding)
    float cand[37][7], scand[2 * KS + 1];
    for (int g = 0; g < ng; ++g) {
        const float r = static_cast<float>(gw[g]);
        cand[g][K] = r;
        for (int k = 1; k <= K; ++k) {
            cand[g][K + k] = std::nextafter(cand[g][K + k - 1], 2.0f);
            cand[g][K - k] = std::nextafter(cand[g][K - k + 1], 0.0f);
        }
    }
    const float sr = static_cast<float>(q.s);
    scand[KS] = sr;
    for (int k = 1; k <= KS; ++k) {
        scand[KS + k] = std::nextafter(scand[KS + k - 1], 1e30f);
        scand[KS - k] = std::nextafter(scand[KS - k + 1], 0.0f);
    }
    auto defect = [&](const int *k, int ks, int &moves) {
        double m = -1.0, e = 0.0;
        moves = ks > KS ? ks - KS : KS - ks;
        for (int g = 0; g < ng; ++g) {
            m += gn[g] * (double)cand[g][k[g]];
            e += gm[g] * (double)cand[g][k[g]];
            moves += k[g] > K ? k[g] - K : K - k[g];
        }
        e = (double)scand[ks] * e - 1.0;
        return std::fmax(std::fabs(m), std::fabs(e));
    };
    int k[37], best_k[37], moves = 0, best_moves = 0, best_ks = KS;
    for (int g = 0; g < ng; ++g) k[g] = best_k[g] = K;
    const double rounded = defect(k, KS, moves);
    double best = rounded;
    for (int ks = 0; ks <= 2 * KS; ++ks) {
        for (int g = 0; g < ng; ++g) k[g] = 0;
        while (true) {
            const double d = defect(k, ks, moves);
            if (d < best || (d == best && moves < best_moves)) {
                best = d;
                best_moves = moves;
                best_ks = ks;
                for (int g = 0; g < ng; ++g) best_k[g] = k[g];
            }
            int g = 0;
            while (g < ng && ++k[g] == span) k[g++] = 0;
            if (g == ng) break;
        }
    }
    const bool use = best <= 0.5 * rounded;
    for (int i = 0; i < 37; ++i)
        q.w32[i] = i < q.Q ? (use && grp[i] >= 0 ? cand[grp[i]][best_k[grp[i]]]
                                                 : static_cast<float>(q.w[i]))
                           : 0.0f;
    q.s32 = use ? scand[best_ks] : sr;
    q.f32_set = true;
}

// the constants of every quadrature, built once per process on first use
// (the FP32 search takes about 40 ms in all)
bool host_quad(int qid, HostQuad &q) {
    struct Table {
        HostQuad q[38];
        bool ok[38] = {};
        Table() {
            for (int id : {4, 5, 9, 13, 17, 21, 37}) {
                ok[id] = host_quad_exact(id, q[id]);
                if (ok[id]) tune_f32(q[id]);
            }
        }
    };
    static const Table table;
    if (qid < 0 || qid > 37 || !table.ok[qid]) {
        q = HostQuad{};
        return false;
    }
    q = table.q[qid];
    return true;
}

// the quadrature constants of a launch (FP64 values; the FP32 storages take
// w32 / s32)
static void quad_consts(lbm::StepParams<double> &p, const HostQuad &q) {
    for (int i = 0; i < 37; ++i) p.w[i] = i < q.Q ? q.w[i] : 0.0;
    const double s = q.s;
    p.s = s;
    p.neg_s = -s;
    p.s2 = s * s;
    p.three_s = 3.0 * s;
    p.six_s3 = 6.0 * ((s * s) * s);
    p.three_s2 = 3.0 * (s * s);
    p.s4 = pow4(s);
    p.cs_inv = 1.0 / s;
    for (int i = 0; i < 37; ++i) p.w32[i] = q.w32[i];
    p.s32 = q.s32;
    p.f32_set = q.f32_set ? 1 : 0;
}

int storage_bytes(int storage) {
    switch (storage) {
    case LBM_STORE_F64: return 8;
    case LBM_STORE_F32: return 4;
    case LBM_STORE_F16DEV:
    case LBM_STORE_BF16DEV:
    case LBM_STORE_F16NEQ:
    case LBM_STORE_BF16NEQ: return 2;
    default: return 0;
    }
}

}  // namespace

struct lbm_field_s {
    int qid = 0, nx = 0, ny = 0, Q = 0, storage = 0;
    double rho0 = 1.0, dscale = 1.0;
    void *d = nullptr;
    size_t bytes = 0;
};

struct lbm_model_s {
    int qid = 0;
    HostQuad q;
    int coll = LBM_SRT;
    double tau = 1.0, tau_s = 1.0, tau_a = 1.0;
    std::vector<double> taus;
    int force_kind = LBM_FORCE_NONE;
    double F[2] = {0.0, 0.0};
    double *dF = nullptr;
    int fnx = 0, fny = 0;
    std::vector<lbm_bc> bcs;
    lbm::BCDev *dB = nullptr;
    double *dMw = nullptr;
    double *dMei = nullptr;  // iterative initialization: nonlinear term (nx, ny, Q)
    int mei_nx = 0, mei_ny = 0;
    double *dPart = nullptr;  // reduction partials
    size_t npart = 0;
    double *dOut = nullptr;
    // launch shape of the fused step (0 = automatic)
    int launch_block = 0, launch_rows = 0, launch_vec = 0, launch_nt = -1;
};

namespace {

int check_field(lbm_field f, const char *what) {
    if (!f || !f->d) return fail(-1, "%s: null field", what);
    return 0;
}

int check_pair(lbm_model m, lbm_field a, lbm_field b, const char *what, bool uses_force,
               bool uses_bcs) {
    if (!m) return fail(-1, "%s: null model", what);
    if (int rc = check_field(a, what)) return rc;
    if (int rc = check_field(b, what)) return rc;
    if (a == b || a->d == b->d)
        return fail(-1, "%s: input and output populations must be distinct arrays", what);
    if (a->qid != m->qid || b->qid != m->qid)
        return fail(-1, "%s: field quadrature D2Q%d/D2Q%d does not match the model's D2Q%d", what,
                    a->qid, b->qid, m->qid);
    if (a->nx != b->nx || a->ny != b->ny || a->storage != b->storage)
        return fail(-1, "%s: fields differ in shape or storage", what);
    if (uses_force && m->force_kind == LBM_FORCE_FIELD && (m->fnx != a->nx || m->fny != a->ny))
        return fail(-1, "%s: force field is %dx%d but the lattice is %dx%d", what, m->fnx,
                    m->fny, a->nx, a->ny);
    for (const lbm_bc &b : m->bcs) {
        if (!uses_bcs) break;
        if (b.kind == LBM_MOVING_WALL) continue;  // acts on the whole lattice
        const bool xe = b.x1 < b.x0, ye = b.y1 < b.y0;
        if ((!xe && (b.x0 < 1 || b.x1 > a->nx)) || (!ye && (b.y0 < 1 || b.y1 > a->ny)))
            return fail(-1, "%s: boundary range x %d:%d, y %d:%d outside the %dx%d lattice", what,
                        b.x0, b.x1, b.y0, b.y1, a->nx, a->ny);
    }
    if (uses_force && m->coll == LBM_MRT && a->storage != LBM_STORE_F64)
        return fail(-2, "%s: MRT is implemented for FP64 storage only", what);
    if (uses_force && m->coll == LBM_MEI) {
        if (a->storage != LBM_STORE_F64)
            return fail(-2, "%s: the iterative initialization runs on FP64 storage only", what);
        if (m->mei_nx != a->nx || m->mei_ny != a->ny)
            return fail(-1, "%s: initialization velocity is %dx%d but the lattice is %dx%d", what,
                        m->mei_nx, m->mei_ny, a->nx, a->ny);
    }
    return 0;
}

lbm::StepParams<double> make_params(const lbm_model_s *m, const lbm_field_s *f, bool with_bcs) {
    lbm::StepParams<double> p{};
    const HostQuad &q = m->q;
    p.nx = f->nx;
    p.ny = f->ny;
    p.plane = (long long)f->nx * (long long)f->ny;
    quad_consts(p, q);
    if (m->coll == LBM_SRT) {
        p.omo = 1.0 - 1.0 / m->tau;
        p.om = 1.0 / m->tau;
        p.tau_force = m->tau;
    } else if (m->coll == LBM_TRT) {
        p.oms = 1.0 / m->tau_s;
        p.oma = 1.0 / m->tau_a;
        p.tau_force = m->tau_a;
    } else if (m->coll == LBM_MEI) {
        p.omo = 1.0 - 1.0 / m->tau;
        p.om = 1.0 / m->tau;
        p.mei = m->dMei;
    } else {
        p.mrt_n = q.order / 2;
        for (int n = 1; n <= 4; ++n) {
            const double t = (size_t)(n - 1) < m->taus.size() ? m->taus[n - 1] : 1.0;
            p.mrt_omo[n] = 1.0 - 1.0 / t;
            p.mrt_om[n] = 1.0 / t;
            p.mrt_sn[n] = std::pow(q.s, (double)n);
        }
        p.tau_force = m->taus.size() > 1 ? m->taus[1] : 1.0;
    }
    p.force_kind = m->force_kind;
    p.Fx = m->F[0];
    p.Fy = m->F[1];
    p.Ffield = m->dF;
    p.nbc = with_bcs ? (int)m->bcs.size() : 0;
    p.bcs = m->dB;
    p.mw = m->dMw;
    p.rho0 = f->rho0;
    p.dscale = f->dscale;
    p.dscale_inv = 1.0 / f->dscale;
    return p;
}

// non-equilibrium storages: more than one node per thread counts lattice rows
// (k_fused_neq_rows), any nx
bool rows_mode(int storage) {
    return storage == LBM_STORE_F16NEQ || storage == LBM_STORE_BF16NEQ;
}

// largest nodes-per-thread the kernels instantiate (mirrors max_vec in lbm_kernels.h)
int max_vec(int Q, int storage) {
    if (rows_mode(storage)) return 16;
    const int eb = storage_bytes(storage);
    const int cap = eb >= 8 ? 2 : (eb == 4 ? 4 : 8);
    return Q <= 9 ? cap : (Q <= 17 ? (cap < 4 ? cap : 4) : 2);
}

// automatic launch shape of the fused step.  Measured on MI355X (D2Q9 TRT
// 4096^2, D2Q17 / D2Q37 8192 x 2048): non-temporal stores are fastest for
// every storage; one node per thread for 8- and 4-byte populations and for
// 16-bit deviations on the wide stencils, four for 16-bit deviations on
// D2Q9 (one dwordx2 per plane per lane); for the non-equilibrium storages
// the row kernels (records in an LDS window) with the rows per thread below.
// lbm_autotune refines this per lattice.
void auto_launch(const lbm_model_s *m, const lbm_field_s *f, int &block, int &rows, int &vec,
                 int &nt) {
    // 8-byte populations: 128 x 2 blocks (2-3% faster than 256 x 1 for D2Q9 and
    // D2Q17 FP64 on MI355X); everything else 256 x 1
    const bool f64 = storage_bytes(f->storage) == 8 && m->coll != LBM_MRT;
    // 16-bit deviations on the 21- and 37-velocity stencils: 64-thread blocks
    // (D2Q37 FP16-DEV 8192 x 2048: 508 us against 534 us with 256)
    const bool wide16 = storage_bytes(f->storage) == 2 && f->Q > 17 && !rows_mode(f->storage);
    // MRT, non-temporal stores: 128 x 2 on D2Q4/5/9/13 (D2Q9 4096^2: 423 us),
    // 256 x 1 on D2Q17 (8192 x 2048: 816 us), 64 x 1 on D2Q21/37, whose
    // kernel holds 229 VGPRs (D2Q37 8192 x 2048: 2.19 ms against 2.42 with 256)
    const bool mrt = m->coll == LBM_MRT;
    const int mrt_block = f->Q > 17 ? 64 : (f->Q > 13 ? 256 : 128);
    block = m->launch_block > 0 ? m->launch_block
                                : (mrt ? mrt_block : (f64 ? 128 : (wide16 ? 64 : 256)));
    // (lattices below 256^2 are latency-bound: one row per block there, FP64
    // D2Q9 SRT 64^2 3.8 us per step against 4.2 us with two)
    const bool small = (long long)f->nx * f->ny < 65536;
    rows = m->launch_rows > 0 ? m->launch_rows
                              : (!small && (f64 || (mrt && mrt_block == 128)) ? 2 : 1);
    if (block * rows > 256) rows = 256 / block;
    vec = m->launch_vec;
    nt = m->launch_nt;
    // non-equilibrium storages, rows per thread (MI355X, FP16-NEQ):
    // * D2Q4/5/9 on even widths run two columns per lane: 2 rows (D2Q9 4096^2,
    //   256 threads: 166-169 us against 181 with 4 and 189 with 8);
    // * D2Q4/5/9 on odd widths (one column per lane): 4 rows (4095^2: 224 us
    //   against 227 with 2 and 245 with 8);
    // * 8192 x 2048, 256 threads (D2Q17/21 after the third-order record
    //   factoring):
    //   - D2Q13: 8 rows (253 us against 259 with 16 and 263 with 4);
    //   - D2Q17: 16 rows (366 us against 372 with 8 and 407 with 4; 128
    //     threads the same, 64 threads 527);
    //   - D2Q21: 8 rows (543 us against 544 with 16, 559 with 2, 575 with 4);
    //   - D2Q37, coefficient window: 8 rows (747 us against 788 with 4, 875
    //     with 2, 737 with 16).
    if (vec <= 0 && rows_mode(f->storage) && m->coll != LBM_MRT) {
        if (f->Q == 17) vec = 16;
        else if (f->Q > 9) vec = 8;
        else vec = f->nx % 2 == 0 && f->nx > 4 ? 2 : 4;
    }
    if (vec <= 0) {
        vec = 1;
        const bool neq = f->storage == LBM_STORE_F16NEQ || f->storage == LBM_STORE_BF16NEQ;
        // several nodes per lane only for the 9-velocity stencils: on wider
        // ones one node per lane with per-population stores is faster (D2Q17
        // FP16-DEV 8192 x 2048: 216 us against 238 us with 4 nodes; D2Q37: 508
        // against 636 us with 2)
        if (m->coll != LBM_MRT && storage_bytes(f->storage) == 2 && !neq && f->Q <= 9) {
            const int want = max_vec(f->Q, f->storage) < 4 ? max_vec(f->Q, f->storage) : 4;
            for (int v = want; v > 1; v >>= 1)
                if (f->nx % v == 0 && f->nx / v >= 64) {
                    vec = v;
                    break;
                }
        }
    }
    if (nt < 0) nt = 1;
}

int run(lbm::Op op, lbm_model m, lbm_field src, lbm_field dst, bool with_bcs) {
    lbm::Launch L{};
    L.op = op;
    L.qid = m->qid;
    L.coll = m->coll;
    L.storage = src->storage;
    L.src = src->d;
    L.dst = dst->d;
    if (op == lbm::Op::Fused || op == lbm::Op::FusedFirst)
        auto_launch(m, src, L.block, L.block_y, L.vec, L.nt);
    const lbm::StepParams<double> p = make_params(m, src, with_bcs);
    const int rc = src->storage == LBM_STORE_F64 ? lbm::dispatch_f64(L, p, g_stream)
                                                 : lbm::dispatch_lp(L, p, g_stream);
    if (rc) return fail(-4, "kernel launch failed (code %d): %s", rc,
                        hipGetErrorString(hipGetLastError()));
    return 0;
}

void free_dev(void *&p) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
}

}  // namespace

// Scratch of lbm_reduce_noalloc: a file-scope device array (each GPU holds
// its own copy for the life of the process), sized for lbm_reduce's largest
// grid -- kReduceRows partial rows of 5 sums, then the 5 results (and 3
// spare).  Host code must not pass the array's name to a kernel (in host code
// it is a stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device with that device current.  Calls that
// overlap on one GPU share it and race.
namespace {
constexpr int kReduceRows = 2048;
constexpr size_t kScratchDoubles = (size_t)kReduceRows * 5 + 8;
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) double scratch[kScratchDoubles];

static int scratch_ptr(double **out) {
    if (int rc = ensure_stream()) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(-3, "device index %d out of range", dev);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(-3, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return 0;
}

static int lbm_field_copy_raw(void *dst, const void *src, size_t bytes) {
    if (int rc = lbm::dispatch_copy(src, dst, bytes, g_stream)) return fail(-4, "copy kernel failed (%d)", rc);
    return 0;
}

extern "C" {

const char *lbm_last_error(void) { return g_err.c_str(); }

int lbm_version(void) { return 10000; }

int lbm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lbm_synchronize(void) {
    if (!g_stream) return 0;
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int lbm_clock_sample(uint64_t *out, double *tick_mhz) {
    if (!out) return fail(-1, "lbm_clock_sample: null output");
    if (int rc = ensure_stream()) return rc;
    // one probe buffer per device, allocated on first use with that device
    // current and kept until the process ends
    static unsigned long long *bufs[64] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(-3, "device index %d out of range", dev);
    constexpr size_t bytes = 3 * (size_t)LBM_CLOCK_PROBES * sizeof(unsigned long long);
    if (!bufs[dev]) HIP_TRY(hipMalloc(&bufs[dev], bytes));
    unsigned long long *d = bufs[dev];
    if (lbm::dispatch_clock_probe(d, LBM_CLOCK_PROBES, g_stream))
        return fail(-4, "clock probe launch failed");
    HIP_TRY(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    if (tick_mhz) {
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        *tick_mhz = khz / 1000.0;
    }
    return 0;
}

int lbm_quadrature_info(int quadrature, int *Q, double *cx, double *cy, double *w, double *s,
                        int *order, int *opp) {
    HostQuad q;
    if (!host_quad(quadrature, q)) return fail(-1, "unknown quadrature D2Q%d", quadrature);
    if (Q) *Q = q.Q;
    for (int i = 0; i < q.Q; ++i) {
        if (cx) cx[i] = (double)q.cx[i];
        if (cy) cy[i] = (double)q.cy[i];
        if (w) w[i] = q.w[i];
        if (opp) opp[i] = q.opp[i];
    }
    if (s) *s = q.s;
    if (order) *order = q.order;
    return 0;
}

int lbm_quadrature_f32(int quadrature, float *w, float *s) {
    HostQuad q;
    if (!host_quad(quadrature, q)) return fail(-1, "unknown quadrature D2Q%d", quadrature);
    for (int i = 0; i < q.Q; ++i)
        if (w) w[i] = q.w32[i];
    if (s) *s = q.s32;
    return 0;
}

// ---- fields -----------------------------------------------------------------
int lbm_field_create(lbm_field *out, int quadrature, int nx, int ny, int storage, double rho0) {
    if (!out) return fail(-1, "lbm_field_create: null output handle");
    *out = nullptr;
    HostQuad q;
    if (!host_quad(quadrature, q)) return fail(-1, "unknown quadrature D2Q%d", quadrature);
    if (nx < 1 || ny < 1) return fail(-1, "lattice size %dx%d must be positive", nx, ny);
    if (ny > 65535) return fail(-1, "lattice height %d exceeds 65535 (one grid row per y)", ny);
    const int eb = storage_bytes(storage);
    if (!eb) return fail(-1, "unknown storage %d", storage);
    if ((long long)nx * (long long)ny > (1LL << 31) - 1)
        return fail(-1, "lattice of %lld nodes exceeds 2^31-1", (long long)nx * ny);
    if (int rc = ensure_stream()) return rc;
    lbm_field f = new lbm_field_s;
    f->qid = quadrature;
    f->nx = nx;
    f->ny = ny;
    f->Q = q.Q;
    f->storage = storage;
    f->rho0 = rho0;
    // binary16 storages scale the stored deviation by a power of two (exact)
    // so that small values stay out of the subnormal range: |f - w rho0| up to
    // 255 (x 2^8), |f - feq| up to 1 (x 2^16); bfloat16 has FP32's range.
    f->dscale = storage == LBM_STORE_F16DEV ? 256.0 : (storage == LBM_STORE_F16NEQ ? 65536.0 : 1.0);
    f->bytes = (size_t)nx * (size_t)ny * (size_t)q.Q * (size_t)eb;
    if (storage == LBM_STORE_F16NEQ || storage == LBM_STORE_BF16NEQ)
        f->bytes += (size_t)nx * (size_t)ny * 8;  // 8-byte equilibrium records
    hipError_t e = hipMalloc(&f->d, f->bytes);
    if (e != hipSuccess) {
        delete f;
        return fail(-3, "hipMalloc(%zu) failed: %s", (size_t)nx * ny * q.Q * eb,
                    hipGetErrorString(e));
    }
    e = hipMemsetAsync(f->d, 0, f->bytes, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    if (e != hipSuccess) {
        (void)hipFree(f->d);
        delete f;
        return fail(-3, "clearing a new field failed: %s", hipGetErrorString(e));
    }
    *out = f;
    return 0;
}

int lbm_field_destroy(lbm_field f) {
    if (!f) return 0;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    free_dev(f->d);
    delete f;
    return 0;
}

int lbm_field_info(lbm_field f, int *quadrature, int *nx, int *ny, int *Q, int *storage) {
    if (!f) return fail(-1, "lbm_field_info: null field");
    if (quadrature) *quadrature = f->qid;
    if (nx) *nx = f->nx;
    if (ny) *ny = f->ny;
    if (Q) *Q = f->Q;
    if (storage) *storage = f->storage;
    return 0;
}

// parameters that only depend on the field (quadrature constants, storage)
static lbm::StepParams<double> field_params(const lbm_field_s *f) {
    lbm::StepParams<double> p{};
    HostQuad q;
    host_quad(f->qid, q);
    p.nx = f->nx;
    p.ny = f->ny;
    p.plane = (long long)f->nx * f->ny;
    quad_consts(p, q);
    p.rho0 = f->rho0;
    p.dscale = f->dscale;
    p.dscale_inv = 1.0 / f->dscale;
    return p;
}

static bool is_neq(const lbm_field_s *f) {
    return f->storage == LBM_STORE_F16NEQ || f->storage == LBM_STORE_BF16NEQ;
}

static int transfer(lbm_field f, double *host, bool up) {
    if (int rc = check_field(f, up ? "lbm_field_upload" : "lbm_field_download")) return rc;
    if (!host) return fail(-1, "null host array");
    const size_t nbytes64 = (size_t)f->nx * f->ny * f->Q * sizeof(double);
    if (f->storage == LBM_STORE_F64) {
        HIP_TRY(hipMemcpyAsync(up ? f->d : (void *)host, up ? (const void *)host : f->d, nbytes64,
                               up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, g_stream));
        HIP_TRY(hipStreamSynchronize(g_stream));
        return 0;
    }
    double *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, nbytes64));
    const lbm::StepParams<double> p = field_params(f);
    hipError_t e = hipSuccess;
    if (up) e = hipMemcpyAsync(tmp, host, nbytes64, hipMemcpyHostToDevice, g_stream);
    int rc = 0;
    if (e == hipSuccess) rc = lbm::dispatch_upload(f->qid, f->storage, tmp, f->d, p, up ? 1 : 0,
                                                   g_stream);
    if (e == hipSuccess && !rc && !up)
        e = hipMemcpyAsync(host, tmp, nbytes64, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(-3, "transfer failed: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "conversion kernel failed (%d)", rc);
    return 0;
}

int lbm_field_upload(lbm_field f, const double *host) {
    return transfer(f, const_cast<double *>(host), true);
}

int lbm_field_download(lbm_field f, double *host) { return transfer(f, host, false); }

int lbm_initialize_analytical(lbm_field f, int strategy, const double *rho, const double *ux,
                              const double *uy, const double *T, const double *grad, double tau,
                              double factor) {
    if (int rc = check_field(f, "lbm_initialize_analytical")) return rc;
    if (strategy < LBM_INIT_EQUILIBRIUM || strategy > LBM_INIT_OFFEQ)
        return fail(-1, "lbm_initialize_analytical: unknown strategy %d", strategy);
    if (!rho || !ux || !uy) return fail(-1, "lbm_initialize_analytical: null density or velocity");
    if (strategy != LBM_INIT_EQUILIBRIUM && !grad)
        return fail(-1, "lbm_initialize_analytical: the off-equilibrium start needs the gradient");
    const size_t plane = (size_t)f->nx * f->ny;
    const int nin = 3 + (T ? 1 : 0) + (strategy != LBM_INIT_EQUILIBRIUM ? 4 : 0);
    double *din = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&din, nin * plane * sizeof(double)));
    hipError_t e = hipSuccess;
    int rc = 0;
    lbm::InitArgs a{};
    size_t off = 0;
    auto put = [&](const double *h, int nplanes) -> const double * {
        double *d = din + off * plane;
        if (e == hipSuccess)
            e = hipMemcpyAsync(d, h, nplanes * plane * sizeof(double), hipMemcpyHostToDevice,
                               g_stream);
        off += (size_t)nplanes;
        return d;
    };
    a.rho = put(rho, 1);
    a.ux = put(ux, 1);
    a.uy = put(uy, 1);
    a.T = T ? put(T, 1) : nullptr;
    a.G = strategy != LBM_INIT_EQUILIBRIUM ? put(grad, 4) : nullptr;
    a.mode = strategy;
    a.tau = tau;
    a.factor = factor;
    const lbm::StepParams<double> p = field_params(f);
    // cs^n as Julia's Float64 ^ Int (libm pow), n! as Float64
    for (int n = 1; n <= 4; ++n) {
        a.csn[n] = std::pow(p.cs_inv, (double)n);
        double fa = 1.0;
        for (int k = 2; k <= n; ++k) fa *= k;
        a.fact[n] = fa;
    }
    if (e == hipSuccess && f->storage != LBM_STORE_F64)
        e = hipMalloc(&dout, (size_t)f->Q * plane * sizeof(double));
    if (e == hipSuccess) {
        double *target = f->storage == LBM_STORE_F64 ? static_cast<double *>(f->d) : dout;
        rc = lbm::dispatch_init_analytical(f->qid, target, p, a, g_stream);
        if (!rc && dout) rc = lbm::dispatch_upload(f->qid, f->storage, dout, f->d, p, 1, g_stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(-3, "lbm_initialize_analytical: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "lbm_initialize_analytical: kernel failed (%d)", rc);
    return 0;
}

int lbm_field_copy(lbm_field dst, lbm_field src) {
    if (int rc = check_field(dst, "lbm_field_copy")) return rc;
    if (int rc = check_field(src, "lbm_field_copy")) return rc;
    if (dst->bytes != src->bytes || dst->storage != src->storage || dst->qid != src->qid)
        return fail(-1, "lbm_field_copy: fields differ in shape or storage");
    if (dst == src) return 0;
    if (int rc = lbm::dispatch_copy(src->d, dst->d, src->bytes, g_stream))
        return fail(-4, "copy kernel failed (%d)", rc);
    return 0;
}

// ---- models -----------------------------------------------------------------
int lbm_model_create(lbm_model *out, int quadrature) {
    if (!out) return fail(-1, "lbm_model_create: null output handle");
    *out = nullptr;
    HostQuad q;
    if (!host_quad(quadrature, q)) return fail(-1, "unknown quadrature D2Q%d", quadrature);
    if (int rc = ensure_stream()) return rc;
    lbm_model m = new lbm_model_s;
    m->qid = quadrature;
    m->q = q;
    hipError_t e = hipMalloc(&m->dOut, 8 * sizeof(double));
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    if (e != hipSuccess) {
        free_dev(reinterpret_cast<void *&>(m->dOut));
        delete m;
        return fail(-3, "model allocation failed: %s", hipGetErrorString(e));
    }
    *out = m;
    return 0;
}

int lbm_model_destroy(lbm_model m) {
    if (!m) return 0;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    free_dev(reinterpret_cast<void *&>(m->dF));
    free_dev(reinterpret_cast<void *&>(m->dB));
    free_dev(reinterpret_cast<void *&>(m->dMw));
    free_dev(reinterpret_cast<void *&>(m->dMei));
    free_dev(reinterpret_cast<void *&>(m->dPart));
    free_dev(reinterpret_cast<void *&>(m->dOut));
    delete m;
    return 0;
}

int lbm_model_set_srt(lbm_model m, double tau) {
    if (!m) return fail(-1, "null model");
    if (!(tau > 0.0) || !std::isfinite(tau)) return fail(-1, "SRT tau must be positive");
    m->coll = LBM_SRT;
    m->tau = tau;
    return 0;
}

int lbm_model_set_trt(lbm_model m, double tau_symmetric, double tau_asymmetric) {
    if (!m) return fail(-1, "null model");
    if (!(tau_symmetric > 0.0) || !(tau_asymmetric > 0.0))
        return fail(-1, "TRT relaxation times must be positive");
    m->coll = LBM_TRT;
    m->tau_s = tau_symmetric;
    m->tau_a = tau_asymmetric;
    return 0;
}

int lbm_model_set_mrt(lbm_model m, const double *taus, int ntaus) {
    if (!m) return fail(-1, "null model");
    const int N = m->q.order / 2;
    const int need = N > 2 ? N : 2;
    if (!taus || ntaus < need)
        return fail(-1, "MRT on D2Q%d needs at least %d relaxation times", m->qid, need);
    for (int k = 0; k < ntaus; ++k)
        if (!(taus[k] > 0.0)) return fail(-1, "MRT relaxation times must be positive");
    m->coll = LBM_MRT;
    m->taus.assign(taus, taus + ntaus);
    return 0;
}

int lbm_model_set_mei(lbm_model m, double tau, const double *u, int nx, int ny) {
    if (!m) return fail(-1, "null model");
    if (!(tau > 0.0) || !std::isfinite(tau)) return fail(-1, "initialization tau must be positive");
    if (!u) return fail(-1, "initialization velocity is null");
    if (nx < 1 || ny < 1) return fail(-1, "initialization velocity size must be positive");
    if ((long long)nx * (long long)ny > (1LL << 31) - 1)
        return fail(-1, "lattice of %lld nodes exceeds 2^31-1", (long long)nx * ny);
    if (int rc = ensure_stream()) return rc;
    // nonlinear_term (iterative_initialization.jl:12-40) in the reference's order
    const HostQuad &q = m->q;
    const size_t plane = (size_t)nx * (size_t)ny;
    std::vector<double> N(plane * (size_t)q.Q);
    const double cs = q.s;
    for (size_t n = 0; n < plane; ++n) {
        const double u0 = u[n], u1 = u[plane + n];
        const double usq = u0 * u0 + u1 * u1;
        for (int i = 0; i < q.Q; ++i) {
            const double ud = (0.0 + (double)q.cx[i] * u0) + (double)q.cy[i] * u1;
            const double a1 = cs * ud;
            const double a2 = (cs * cs) * (ud * ud) - cs * usq;
            N[plane * (size_t)i + n] = (q.w[i] * 1.0) * (a1 + a2 / 2.0);
        }
    }
    const size_t bytes = N.size() * sizeof(double);
    if (m->mei_nx != nx || m->mei_ny != ny || !m->dMei) {
        free_dev(reinterpret_cast<void *&>(m->dMei));
        m->mei_nx = m->mei_ny = 0;
        HIP_TRY(hipMalloc(&m->dMei, bytes));
    }
    HIP_TRY(hipMemcpyAsync(m->dMei, N.data(), bytes, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    m->mei_nx = nx;
    m->mei_ny = ny;
    m->coll = LBM_MEI;
    m->tau = tau;
    return 0;
}

int lbm_model_set_force(lbm_model m, int kind, const double *values, int nx, int ny) {
    if (!m) return fail(-1, "null model");
    if (kind == LBM_FORCE_NONE) {
        m->force_kind = kind;
        return 0;
    }
    if (!values) return fail(-1, "force values are null");
    if (kind == LBM_FORCE_CONST) {
        m->force_kind = kind;
        m->F[0] = values[0];
        m->F[1] = values[1];
        return 0;
    }
    if (kind != LBM_FORCE_FIELD) return fail(-1, "unknown force kind %d", kind);
    if (nx < 1 || ny < 1) return fail(-1, "force field size must be positive");
    const size_t bytes = (size_t)nx * ny * 2 * sizeof(double);
    if (m->fnx != nx || m->fny != ny || !m->dF) {
        free_dev(reinterpret_cast<void *&>(m->dF));
        HIP_TRY(hipMalloc(&m->dF, bytes));
    }
    HIP_TRY(hipMemcpyAsync(m->dF, values, bytes, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    m->fnx = nx;
    m->fny = ny;
    m->force_kind = kind;
    return 0;
}

int lbm_model_set_bcs(lbm_model m, const lbm_bc *bcs, int n) {
    if (!m) return fail(-1, "null model");
    if (n < 0 || n > lbm::kMaxBC) return fail(-1, "at most %d boundary conditions", lbm::kMaxBC);
    if (n > 0 && !bcs) return fail(-1, "null boundary-condition list");
    std::vector<lbm::BCDev> dev(n > 0 ? n : 1);
    std::vector<double> mw((size_t)(n > 0 ? n : 1) * m->q.Q, 0.0);
    for (int b = 0; b < n; ++b) {
        const lbm_bc &c = bcs[b];
        if (c.kind != LBM_BOUNCE_BACK && c.kind != LBM_MOVING_WALL)
            return fail(-1, "unknown boundary kind %d", c.kind);
        if (c.direction < LBM_NORTH || c.direction > LBM_WEST)
            return fail(-1, "unknown boundary direction %d", c.direction);
        if (c.kind == LBM_MOVING_WALL && c.direction != LBM_NORTH)
            return fail(-2, "MovingWall is defined for North only (moving_wall.jl:17)");
        dev[b] = lbm::BCDev{c.kind, c.direction, c.x0 - 1, c.x1 - 1, c.y0 - 1, c.y1 - 1};
        if (c.kind == LBM_MOVING_WALL) {
            // a_1 = w_i * s * dot(rho_w u_w, H1(c_i)); f += 2 a_1 (moving_wall.jl:20-24, 33)
            const double a1eq[2] = {c.rho * c.ux, c.rho * c.uy};
            for (int i = 0; i < m->q.Q; ++i) {
                const double H1[2] = {(double)m->q.cx[i] * 1.0, (double)m->q.cy[i] * 1.0};
                const double dot = a1eq[0] * H1[0] + a1eq[1] * H1[1];
                const double a1 = (m->q.w[i] * m->q.s) * dot;
                mw[(size_t)b * m->q.Q + i] = 2.0 * a1;
            }
        }
    }
    free_dev(reinterpret_cast<void *&>(m->dB));
    free_dev(reinterpret_cast<void *&>(m->dMw));
    HIP_TRY(hipMalloc(&m->dB, dev.size() * sizeof(lbm::BCDev)));
    HIP_TRY(hipMalloc(&m->dMw, mw.size() * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(m->dB, dev.data(), dev.size() * sizeof(lbm::BCDev),
                           hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipMemcpyAsync(m->dMw, mw.data(), mw.size() * sizeof(double), hipMemcpyHostToDevice,
                           g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    m->bcs.assign(bcs, bcs + n);
    return 0;
}

// ---- operators --------------------------------------------------------------
int lbm_collide(lbm_model m, lbm_field f_in, lbm_field f_out) {
    if (int rc = check_pair(m, f_in, f_out, "lbm_collide", true, false)) return rc;
    return run(lbm::Op::Collide, m, f_in, f_out, false);
}

int lbm_stream(lbm_model m, lbm_field f, lbm_field f_new) {
    if (int rc = check_pair(m, f, f_new, "lbm_stream", false, false)) return rc;
    return run(lbm::Op::StreamBC, m, f, f_new, false);
}

int lbm_apply_bcs(lbm_model m, lbm_field f_new, lbm_field f_old) {
    if (int rc = check_pair(m, f_old, f_new, "lbm_apply_bcs", false, true)) return rc;
    if (m->bcs.empty()) return 0;
    return run(lbm::Op::ApplyBC, m, f_old, f_new, true);
}

int lbm_stream_bc(lbm_model m, lbm_field f, lbm_field f_new) {
    if (int rc = check_pair(m, f, f_new, "lbm_stream_bc", false, true)) return rc;
    return run(lbm::Op::StreamBC, m, f, f_new, true);
}

static int simulate_impl(lbm_model m, lbm_field S, lbm_field C, int64_t nsteps, double *ms,
                         int64_t *nlaunch, double *per_launch = nullptr) {
    if (int rc = check_pair(m, S, C, "lbm_simulate", true, true)) return rc;
    if (nsteps < 0) return fail(-1, "negative step count");
    if (ms) *ms = 0.0;
    if (nlaunch) *nlaunch = 0;
    if (nsteps == 0) return 0;
    if (m->coll == LBM_MEI) {
        // collide-only model: collide! then stream! + apply! each step, in the
        // reference's roles (no fused kernel)
        for (int64_t k = 0; k < nsteps; ++k) {
            if (int rc = run(lbm::Op::Collide, m, S, C, false)) return rc;
            if (int rc = run(lbm::Op::StreamBC, m, C, S, true)) return rc;
        }
        if (nlaunch) *nlaunch = 2 * nsteps;
        return 0;
    }
    // step 1 starts from the canonical state S.  One-speed stencils on
    // lattices of whole 64 x 8 tiles (every storage) take
    // C_1 = collide(BC(stream(collide(S_0)))) in one pass (k_fused_first*:
    // one pass over HBM instead of collide! plus the first fused step);
    // otherwise C_0 = collide(S_0) first.  Runs that
    // time every launch keep the two-pass start, so that their launches are
    // all of one kernel.
    const bool first_fused =
        nsteps >= 2 && !per_launch && (m->qid == 4 || m->qid == 5 || m->qid == 9) &&
        S->nx % 64 == 0 && S->ny % 8 == 0 &&
        (m->coll <= LBM_TRT || (S->storage == LBM_STORE_F64 && m->coll == LBM_MRT));
    if (first_fused) {
        if (int rc = run(lbm::Op::FusedFirst, m, S, C, true)) return rc;
    } else {
        if (int rc = run(lbm::Op::Collide, m, S, C, false)) return rc;
    }
    const int64_t k_first = first_fused ? 2 : 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, g_stream));
    }
    // per-launch events (measurement runs only): one before every fused launch
    std::vector<hipEvent_t> ev;
    if (per_launch) {
        ev.assign((size_t)(nsteps > 1 ? nsteps : 1), nullptr);
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    }
    // the remaining steps: C_k = collide(BC(stream(C_{k-1}))) in one pull pass each
    lbm_field cur = C, other = S;
    for (int64_t k = k_first; k < nsteps; ++k) {
        if (per_launch) HIP_TRY(hipEventRecord(ev[(size_t)k - 1], g_stream));
        if (int rc = run(lbm::Op::Fused, m, cur, other, true)) return rc;
        std::swap(cur, other);
    }
    if (per_launch) {
        HIP_TRY(hipEventRecord(ev.back(), g_stream));
        HIP_TRY(hipEventSynchronize(ev.back()));
        for (int64_t k = 1; k + 1 < nsteps; ++k) {
            float t = 0.0f;
            HIP_TRY(hipEventElapsedTime(&t, ev[(size_t)k - 1], ev[(size_t)k]));
            per_launch[k - 1] = (double)t;
        }
        if (nsteps > 1) {
            float t = 0.0f;
            HIP_TRY(hipEventElapsedTime(&t, ev[(size_t)nsteps - 2], ev.back()));
            per_launch[nsteps - 2] = (double)t;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (nlaunch) *nlaunch = nsteps - 1;
    }
    if (ms) {
        HIP_TRY(hipEventRecord(e1, g_stream));
    }
    // S_n = BC(stream(C_{n-1}), C_{n-1})
    if (int rc = run(lbm::Op::StreamBC, m, cur, other, true)) return rc;
    if (cur != C) {
        // the buffers ended in swapped roles: exchange the device arrays so the
        // handles keep the reference's roles (f_stream holds S_n, f_collision C_{n-1})
        std::swap(S->d, C->d);
    }
    if (ms) {
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t;
        if (nlaunch) *nlaunch = nsteps - k_first;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

int lbm_model_set_launch(lbm_model m, int block_x, int rows_per_block, int nodes_per_thread,
                         int nontemporal) {
    if (!m) return fail(-1, "null model");
    if (block_x != 0 && block_x != 64 && block_x != 128 && block_x != 256)
        return fail(-1, "block_x must be 0 (automatic), 64, 128 or 256");
    if (rows_per_block != 0 && rows_per_block != 1 && rows_per_block != 2 && rows_per_block != 4)
        return fail(-1, "rows_per_block must be 0 (automatic), 1, 2 or 4");
    if (block_x * (rows_per_block > 0 ? rows_per_block : 1) > 256)
        return fail(-1, "block_x * rows_per_block must not exceed 256");
    if (nodes_per_thread != 0 && nodes_per_thread != 1 && nodes_per_thread != 2 &&
        nodes_per_thread != 4 && nodes_per_thread != 8 && nodes_per_thread != 16)
        return fail(-1, "nodes_per_thread must be 0 (automatic), 1, 2, 4, 8 or 16");
    m->launch_block = block_x;
    m->launch_rows = rows_per_block;
    m->launch_vec = nodes_per_thread;
    m->launch_nt = nontemporal < 0 ? -1 : (nontemporal ? 1 : 0);
    return 0;
}

int lbm_autotune(lbm_model m, lbm_field f_stream, int trial_steps, int *block_x,
                 int *rows_per_block, int *nodes_per_thread, int *nontemporal,
                 double *us_per_step) {
    if (!m) return fail(-1, "null model");
    if (int rc = check_field(f_stream, "lbm_autotune")) return rc;
    if (f_stream->qid != m->qid) return fail(-1, "lbm_autotune: quadrature mismatch");
    if (m->coll == LBM_MEI)
        return fail(-2, "lbm_autotune: the iterative initialization model has no fused step");
    if (trial_steps < 1) trial_steps = 1;
    // scratch copies: tuning must not advance the caller's state
    lbm_field A = nullptr, B = nullptr;
    int rc = lbm_field_create(&A, f_stream->qid, f_stream->nx, f_stream->ny, f_stream->storage,
                              f_stream->rho0);
    if (!rc) rc = lbm_field_create(&B, f_stream->qid, f_stream->nx, f_stream->ny,
                                   f_stream->storage, f_stream->rho0);
    // the model/field checks of lbm_simulate (force field shape, BC ranges,
    // MRT storage) before any kernel reads the model's buffers
    if (!rc) rc = check_pair(m, A, B, "lbm_autotune", true, true);
    if (!rc) rc = lbm_field_copy(A, f_stream);
    const int save_b = m->launch_block, save_r = m->launch_rows, save_v = m->launch_vec,
              save_n = m->launch_nt;
    double best = 1e300;
    int bb = 0, br = 1, bv = 1, bn = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(-3, "hipEventCreate failed");
    const int mv = (m->coll == LBM_MRT) ? 1 : max_vec(f_stream->Q, f_stream->storage);
    // (threads along x, rows) per block
    const int shapes[6][2] = {{64, 1}, {128, 1}, {256, 1}, {64, 4}, {128, 2}, {64, 2}};
    const bool rows = rows_mode(f_stream->storage);
    for (int vec = 1; !rc && vec <= mv; vec <<= 1) {
        if (!rows && f_stream->nx % vec) continue;
        for (int nt = 0; !rc && nt <= (m->coll == LBM_MRT ? 0 : 1); ++nt) {
            for (int bi = 0; !rc && bi < 6; ++bi) {
                if (m->coll == LBM_MRT && shapes[bi][1] != 1) continue;
                if (rows && vec > 1 && shapes[bi][1] != 1) continue;  // one row per block
                m->launch_block = shapes[bi][0];
                m->launch_rows = shapes[bi][1];
                m->launch_vec = vec;
                m->launch_nt = nt;
                rc = run(lbm::Op::Fused, m, A, B, true);  // warm-up
                if (rc) break;
                (void)hipEventRecord(e0, g_stream);
                lbm_field cur = B, other = A;
                for (int k = 0; k < trial_steps && !rc; ++k) {
                    rc = run(lbm::Op::Fused, m, cur, other, true);
                    std::swap(cur, other);
                }
                (void)hipEventRecord(e1, g_stream);
                if (hipEventSynchronize(e1) != hipSuccess) rc = fail(-3, "autotune sync failed");
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double us = 1e3 * (double)ms / trial_steps;
                if (!rc && us < best) {
                    best = us;
                    bb = shapes[bi][0];
                    br = shapes[bi][1];
                    bv = vec;
                    bn = nt;
                }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    lbm_field_destroy(A);
    lbm_field_destroy(B);
    if (rc) {
        m->launch_block = save_b;
        m->launch_rows = save_r;
        m->launch_vec = save_v;
        m->launch_nt = save_n;
        return rc;
    }
    m->launch_block = bb;
    m->launch_rows = br;
    m->launch_vec = bv;
    m->launch_nt = bn;
    if (block_x) *block_x = bb;
    if (rows_per_block) *rows_per_block = br;
    if (nodes_per_thread) *nodes_per_thread = bv;
    if (nontemporal) *nontemporal = bn;
    if (us_per_step) *us_per_step = best;
    return 0;
}

int lbm_simulate_launch_times(lbm_model m, lbm_field f_stream, lbm_field f_collision,
                              int64_t nsteps, double *launch_ms, int64_t *nlaunch) {
    if (!launch_ms) return fail(-1, "lbm_simulate_launch_times: null output");
    return simulate_impl(m, f_stream, f_collision, nsteps, nullptr, nlaunch, launch_ms);
}

int lbm_simulate(lbm_model m, lbm_field f_stream, lbm_field f_collision, int64_t nsteps) {
    return simulate_impl(m, f_stream, f_collision, nsteps, nullptr, nullptr);
}

int lbm_simulate_timed(lbm_model m, lbm_field f_stream, lbm_field f_collision, int64_t nsteps,
                       double *ms_total, int64_t *nlaunch) {
    double ms = 0.0;
    int64_t nl = 0;
    int rc = simulate_impl(m, f_stream, f_collision, nsteps, &ms, &nl);
    if (ms_total) *ms_total = ms;
    if (nlaunch) *nlaunch = nl;
    return rc;
}

// ---- moments ----------------------------------------------------------------
// NEQ storages: decode into a temporary FP64 field (caller frees *tmp)
static int as_f64(lbm_field f, const void **data, int *storage, double **tmp) {
    *tmp = nullptr;
    *data = f->d;
    *storage = f->storage;
    if (!is_neq(f)) return 0;
    const size_t nb = (size_t)f->nx * f->ny * f->Q * sizeof(double);
    HIP_TRY(hipMalloc(tmp, nb));
    const lbm::StepParams<double> p = field_params(f);
    if (int rc = lbm::dispatch_upload(f->qid, f->storage, *tmp, f->d, p, 0, g_stream)) {
        (void)hipFree(*tmp);
        *tmp = nullptr;
        return fail(-4, "decode kernel failed (%d)", rc);
    }
    *data = *tmp;
    *storage = LBM_STORE_F64;
    return 0;
}

int lbm_moments(lbm_field f, double tau, double *rho, double *ux, double *uy, double *sxx,
                double *sxy, double *syy) {
    if (int rc = check_field(f, "lbm_moments")) return rc;
    const lbm::StepParams<double> p = field_params(f);
    const void *fdata;
    int fstorage;
    double *ftmp;
    if (int rc = as_f64(f, &fdata, &fstorage, &ftmp)) return rc;
    double *host[6] = {rho, ux, uy, sxx, sxy, syy};
    double *dev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const size_t nb = (size_t)p.plane * sizeof(double);
    hipError_t e = hipSuccess;
    for (int k = 0; k < 6 && e == hipSuccess; ++k)
        if (host[k]) e = hipMalloc(&dev[k], nb);
    int rc = 0;
    if (e == hipSuccess)
        rc = lbm::dispatch_moments(f->qid, fstorage, fdata, p, tau, dev[0], dev[1], dev[2],
                                   dev[3], dev[4], dev[5], g_stream);
    for (int k = 0; k < 6 && e == hipSuccess && !rc; ++k)
        if (host[k]) e = hipMemcpyAsync(host[k], dev[k], nb, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    for (int k = 0; k < 6; ++k) free_dev(reinterpret_cast<void *&>(dev[k]));
    free_dev(reinterpret_cast<void *&>(ftmp));
    if (e != hipSuccess) return fail(-3, "lbm_moments: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "moments kernel failed (%d)", rc);
    return 0;
}

int lbm_reduce(lbm_field f, double *out5) {
    if (int rc = check_field(f, "lbm_reduce")) return rc;
    if (!out5) return fail(-1, "lbm_reduce: null output");
    const lbm::StepParams<double> p = field_params(f);
    const void *fdata;
    int fstorage;
    double *ftmp;
    if (int rc = as_f64(f, &fdata, &fstorage, &ftmp)) return rc;
    // a fixed grid for a given lattice size: the summation order is reproducible
    const long long blocks = (p.plane + 255) / 256;
    const int npart = (int)(blocks < 2048 ? blocks : 2048);
    double *part = nullptr, *out = nullptr;
    HIP_TRY(hipMalloc(&part, (size_t)npart * 5 * sizeof(double)));
    hipError_t e = hipMalloc(&out, 8 * sizeof(double));
    int rc = 0;
    if (e == hipSuccess)
        rc = lbm::dispatch_reduce(f->qid, fstorage, fdata, p, part, npart, out, g_stream);
    if (e == hipSuccess && !rc)
        e = hipMemcpyAsync(out5, out, 5 * sizeof(double), hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(part);
    if (out) (void)hipFree(out);
    free_dev(reinterpret_cast<void *&>(ftmp));
    if (e != hipSuccess) return fail(-3, "lbm_reduce: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "reduction kernel failed (%d)", rc);
    return 0;
}

int lbm_reduce_noalloc(lbm_field f, double *out5) {
    if (int rc = check_field(f, "lbm_reduce_noalloc")) return rc;
    if (!out5) return fail(-1, "lbm_reduce_noalloc: null output");
    if (is_neq(f))
        return fail(-2, "lbm_reduce_noalloc: non-equilibrium storages are decoded into a copy of "
                        "the field first; use lbm_reduce");
    const lbm::StepParams<double> p = field_params(f);
    // lbm_reduce's grid: at most kReduceRows blocks, whatever the lattice size
    const long long blocks = (p.plane + 255) / 256;
    const int npart = (int)(blocks < kReduceRows ? blocks : kReduceRows);
    if (npart < 1 || (size_t)npart * 5 + 8 > kScratchDoubles)
        return fail(-1, "lbm_reduce_noalloc: %d partial rows exceed the %d-row scratch", npart,
                    kReduceRows);
    double *sc = nullptr;
    if (int rc = scratch_ptr(&sc)) return rc;
    // every partial row and output entry is written before it is read: the
    // scratch needs no clearing between calls
    double *part = sc, *out = sc + (size_t)kReduceRows * 5;
    if (int rc = lbm::dispatch_reduce(f->qid, f->storage, f->d, p, part, npart, out, g_stream))
        return fail(-4, "reduction kernel failed (%d)", rc);
    HIP_TRY(hipMemcpyAsync(out5, out, 5 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int lbm_field_node(lbm_field f, int x, int y, double *out) {
    if (int rc = check_field(f, "lbm_field_node")) return rc;
    if (!out) return fail(-1, "lbm_field_node: null output");
    if (x < 1 || x > f->nx || y < 1 || y > f->ny)
        return fail(-1, "lbm_field_node: node (%d, %d) outside the %dx%d lattice", x, y, f->nx,
                    f->ny);
    const lbm::StepParams<double> p = field_params(f);
    const void *fdata;
    int fstorage;
    double *ftmp;
    if (int rc = as_f64(f, &fdata, &fstorage, &ftmp)) return rc;
    double *dout = nullptr;
    hipError_t e = hipMalloc(&dout, 64 * sizeof(double));
    int rc = 0;
    const long long n = (long long)(x - 1) + (long long)f->nx * (y - 1);
    if (e == hipSuccess) rc = lbm::dispatch_gather_node(fstorage, fdata, p, n, f->Q, dout, g_stream);
    if (e == hipSuccess && !rc)
        e = hipMemcpyAsync(out, dout, (size_t)f->Q * sizeof(double), hipMemcpyDeviceToHost,
                           g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    free_dev(reinterpret_cast<void *&>(dout));
    free_dev(reinterpret_cast<void *&>(ftmp));
    if (e != hipSuccess) return fail(-3, "lbm_field_node: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "lbm_field_node: gather kernel failed (%d)", rc);
    return 0;
}

int lbm_mei_initialize(lbm_model m, lbm_field f_stream, lbm_field f_collision, double eps,
                       int64_t max_steps, int64_t *steps) {
    if (int rc = check_pair(m, f_stream, f_collision, "lbm_mei_initialize", true, true)) return rc;
    if (m->coll != LBM_MEI)
        return fail(-1, "lbm_mei_initialize: the model has no initialization velocity "
                        "(lbm_model_set_mei)");
    if (max_steps < 1) return fail(-1, "lbm_mei_initialize: max_steps must be positive");
    const lbm::StepParams<double> p = field_params(f_stream);
    // f = w everywhere (mei_et_al.jl:21-25), in both roles
    int rc = lbm::dispatch_fill_weights(static_cast<double *>(f_stream->d), p, f_stream->Q, g_stream);
    if (!rc) rc = lbm::dispatch_fill_weights(static_cast<double *>(f_collision->d), p,
                                             f_stream->Q, g_stream);
    if (rc) return fail(-4, "lbm_mei_initialize: fill kernel failed (%d)", rc);
    double *dnode = nullptr;
    HIP_TRY(hipMalloc(&dnode, 64 * sizeof(double)));
    const long long corner = (long long)(f_stream->nx - 1) + (long long)f_stream->nx * (f_stream->ny - 1);
    double rho = 0.0, rho_old = 0.0, node[64];
    int64_t t = 1;
    hipError_t e = hipSuccess;
    for (; t <= max_steps; ++t) {
        if ((rc = run(lbm::Op::Collide, m, f_stream, f_collision, false))) break;
        if ((rc = run(lbm::Op::StreamBC, m, f_collision, f_stream, true))) break;
        // next!(ProcessIterativeInitialization): DensityConvergence at node (nx, ny)
        if ((rc = lbm::dispatch_gather_node(LBM_STORE_F64, f_stream->d, p, corner, f_stream->Q,
                                            dnode, g_stream))) {
            rc = fail(-4, "lbm_mei_initialize: gather kernel failed (%d)", rc);
            break;
        }
        e = hipMemcpyAsync(node, dnode, (size_t)f_stream->Q * sizeof(double),
                           hipMemcpyDeviceToHost, g_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
        if (e != hipSuccess) break;
        double r = node[0] + node[1];
        for (int i = 2; i < f_stream->Q; ++i) r = r + node[i];
        rho_old = rho;
        rho = r;
        const double d = std::fabs(rho - rho_old);
        if (d < eps || d > 100.0) break;
    }
    free_dev(reinterpret_cast<void *&>(dnode));
    if (e != hipSuccess) return fail(-3, "lbm_mei_initialize: %s", hipGetErrorString(e));
    if (rc) return rc;
    if (steps) *steps = t > max_steps ? max_steps : t;
    return 0;
}

// ---- on-device processing -----------------------------------------------------
struct lbm_expected_s {
    int nx = 0, ny = 0;
    double *d = nullptr;  // 8 planes
};

struct lbm_velocity_tracker_s {
    int nx = 0, ny = 0;
    double *d = nullptr;  // u_x, u_y planes
};

static int planes_create(int nx, int ny, int nplanes, double **d, const char *what) {
    if (nx < 1 || ny < 1) return fail(-1, "%s: lattice size %dx%d must be positive", what, nx, ny);
    if ((long long)nx * (long long)ny > (1LL << 31) - 1)
        return fail(-1, "%s: lattice of %lld nodes exceeds 2^31-1", what, (long long)nx * ny);
    if (int rc = ensure_stream()) return rc;
    const size_t nb = (size_t)nx * (size_t)ny * (size_t)nplanes * sizeof(double);
    HIP_TRY(hipMalloc(d, nb));
    hipError_t e = hipMemsetAsync(*d, 0, nb, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    if (e != hipSuccess) {
        (void)hipFree(*d);
        *d = nullptr;
        return fail(-3, "%s: clearing device planes failed: %s", what, hipGetErrorString(e));
    }
    return 0;
}

// one pass of k_track over f; sums[kTrackCols] on the host
static int track_sums(lbm_field f, const lbm::TrackArgs &a, double *sums, const char *what) {
    const lbm::StepParams<double> p = field_params(f);
    const void *fdata;
    int fstorage;
    double *ftmp;
    if (int rc = as_f64(f, &fdata, &fstorage, &ftmp)) return rc;
    // fixed grid for a given lattice size: the summation order is reproducible
    const long long blocks = (p.plane + 255) / 256;
    const int npart = (int)(blocks < 2048 ? blocks : 2048);
    double *part = nullptr, *out = nullptr;
    hipError_t e = hipMalloc(&part, (size_t)npart * lbm::kTrackCols * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&out, lbm::kTrackCols * sizeof(double));
    int rc = 0;
    if (e == hipSuccess)
        rc = lbm::dispatch_track(f->qid, fstorage, fdata, p, a, part, npart, out, g_stream);
    if (e == hipSuccess && !rc)
        e = hipMemcpyAsync(sums, out, lbm::kTrackCols * sizeof(double), hipMemcpyDeviceToHost,
                           g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    free_dev(reinterpret_cast<void *&>(part));
    free_dev(reinterpret_cast<void *&>(out));
    free_dev(reinterpret_cast<void *&>(ftmp));
    if (e != hipSuccess) return fail(-3, "%s: %s", what, hipGetErrorString(e));
    if (rc) return fail(-4, "%s: processing kernel failed (%d)", what, rc);
    return 0;
}

int lbm_expected_create(lbm_expected *out, int nx, int ny) {
    if (!out) return fail(-1, "lbm_expected_create: null output handle");
    *out = nullptr;
    double *d = nullptr;
    if (int rc = planes_create(nx, ny, 8, &d, "lbm_expected_create")) return rc;
    lbm_expected e = new lbm_expected_s;
    e->nx = nx;
    e->ny = ny;
    e->d = d;
    *out = e;
    return 0;
}

int lbm_expected_destroy(lbm_expected e) {
    if (!e) return 0;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    free_dev(reinterpret_cast<void *&>(e->d));
    delete e;
    return 0;
}

int lbm_expected_upload(lbm_expected e, const double *planes) {
    if (!e || !e->d) return fail(-1, "lbm_expected_upload: null handle");
    if (!planes) return fail(-1, "lbm_expected_upload: null host planes");
    const size_t nb = (size_t)e->nx * (size_t)e->ny * 8 * sizeof(double);
    HIP_TRY(hipMemcpyAsync(e->d, planes, nb, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int lbm_track_errors(lbm_field f, lbm_expected e, const lbm_track_params *prm, double *out) {
    if (int rc = check_field(f, "lbm_track_errors")) return rc;
    if (!e || !e->d) return fail(-1, "lbm_track_errors: null expected fields");
    if (!prm || !out) return fail(-1, "lbm_track_errors: null parameters or output");
    if (e->nx != f->nx || e->ny != f->ny)
        return fail(-1, "lbm_track_errors: expected fields are %dx%d, the lattice %dx%d", e->nx,
                    e->ny, f->nx, f->ny);
    lbm::TrackArgs a{};
    a.exp = e->d;
    for (int k = 0; k < 16; ++k) a.aff[k] = prm->affine[k];
    a.tau = prm->tau;
    a.u_max = prm->u_max;
    a.sfac = 1.0 / (prm->u_max * prm->u_max);
    a.u_old = nullptr;
    double s[lbm::kTrackCols];
    if (int rc = track_sums(f, a, s, "lbm_track_errors")) return rc;
    // the norms of the pushed row (track_hydrodynamic_errors.jl:205-221)
    out[0] = std::sqrt(s[lbm::kErrRho]);
    out[1] = std::sqrt(s[lbm::kErrU] / s[lbm::kTotU2]);
    out[2] = std::sqrt(s[lbm::kErrP] / s[lbm::kTotP2]);
    out[3] = std::sqrt(s[lbm::kErrSxx] / s[lbm::kTotSxx2]);
    out[4] = std::sqrt(s[lbm::kErrSxy] / s[lbm::kTotSxy2]);
    out[5] = std::sqrt(s[lbm::kErrSyy] / s[lbm::kTotSyy2]);
    out[6] = std::sqrt(s[lbm::kErrSyx] / s[lbm::kTotSyx2]);
    out[7] = prm->delta * s[lbm::kMass];
    out[8] = prm->delta * s[lbm::kMomentum];
    out[9] = prm->delta * s[lbm::kEnergy];
    out[10] = s[lbm::kNonFinite];
    return 0;
}

int lbm_velocity_tracker_create(lbm_velocity_tracker *out, int nx, int ny) {
    if (!out) return fail(-1, "lbm_velocity_tracker_create: null output handle");
    *out = nullptr;
    double *d = nullptr;
    if (int rc = planes_create(nx, ny, 2, &d, "lbm_velocity_tracker_create")) return rc;
    lbm_velocity_tracker t = new lbm_velocity_tracker_s;
    t->nx = nx;
    t->ny = ny;
    t->d = d;
    *out = t;
    return 0;
}

int lbm_velocity_tracker_destroy(lbm_velocity_tracker t) {
    if (!t) return 0;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    free_dev(reinterpret_cast<void *&>(t->d));
    delete t;
    return 0;
}

int lbm_velocity_stats(lbm_field f, lbm_velocity_tracker t, double *out5) {
    if (int rc = check_field(f, "lbm_velocity_stats")) return rc;
    if (!out5) return fail(-1, "lbm_velocity_stats: null output");
    if (t && (t->nx != f->nx || t->ny != f->ny))
        return fail(-1, "lbm_velocity_stats: tracker is %dx%d, the lattice %dx%d", t->nx, t->ny,
                    f->nx, f->ny);
    lbm::TrackArgs a{};
    a.exp = nullptr;
    a.u_old = t ? t->d : nullptr;
    double s[lbm::kTrackCols];
    if (int rc = track_sums(f, a, s, "lbm_velocity_stats")) return rc;
    // u_mean /= divide_by (stopping_criteria.jl:39), divide_by = nx * ny
    const double nodes = (double)f->nx * (double)f->ny;
    out5[0] = s[lbm::kSumUx] / nodes;
    out5[1] = s[lbm::kSumUy] / nodes;
    out5[2] = t ? s[lbm::kDu2] : 0.0;
    out5[3] = t ? s[lbm::kUold2] : 0.0;
    out5[4] = s[lbm::kNonFinite];
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches of independent lattices
// ---------------------------------------------------------------------------
struct lbm_batch_s {
    int qid = 9, storage = LBM_STORE_F32, n = 0;
    double rho0 = 1.0;
    HostQuad q;
    std::vector<lbm::BatchLattice> lats;
    std::vector<int> nx, ny;
    lbm::BatchLattice *dLats = nullptr;
    lbm::BatchTile *dTiles = nullptr;
    int ntiles = 0, tile_block = 0;
    void *pool = nullptr;
    size_t pool_bytes = 0;
    int64_t nodes = 0;
    int parity_stream = 0;  // buffer holding f_stream
    int block = 128, grid = 0, vec = 1;  // threads per block, grid, nodes per thread
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int g_chunk = 0, g_parity = -1, g_block = 0, g_grid = 0, g_vec = 0;
};

namespace {

int batch_elem(const lbm_batch_s *b) { return storage_bytes(b->storage); }

void batch_drop_graph(lbm_batch b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_chunk = 0;
    b->g_parity = -1;
}

// the tile table for tiles of `block` nodes.  A captured graph holds the old
// table's address in its kernel nodes, so it is dropped with the table.
int batch_tiles(lbm_batch b, int block) {
    if (b->tile_block == block && b->dTiles) return 0;
    batch_drop_graph(b);
    std::vector<lbm::BatchTile> tiles;
    for (int k = 0; k < b->n; ++k)
        for (long long n0 = 0; n0 < b->lats[k].plane; n0 += block)
            tiles.push_back(lbm::BatchTile{k, (int)n0});
    free_dev(reinterpret_cast<void *&>(b->dTiles));
    HIP_TRY(hipMalloc(&b->dTiles, tiles.size() * sizeof(lbm::BatchTile)));
    HIP_TRY(hipMemcpyAsync(b->dTiles, tiles.data(), tiles.size() * sizeof(lbm::BatchTile),
                           hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    b->ntiles = (int)tiles.size();
    b->tile_block = block;
    return 0;
}

// default grid: one block per tile; lbm_batch_autotune also times persistent
// grids of 1/4 to 4 resident waves of blocks, which stride over the table.
int batch_grid(const lbm_batch_s *b, int block) {
    (void)block;
    if (b->grid > 0) return b->grid;
    return b->ntiles;
}

lbm::StepParams<double> batch_params(const lbm_batch_s *b) {
    lbm::StepParams<double> p{};
    const HostQuad &q = b->q;
    quad_consts(p, q);
    p.rho0 = b->rho0;
    p.dscale = p.dscale_inv = 1.0;
    return p;
}

int batch_launch(lbm_batch b, int op, int parity) {
    if (int rc = batch_tiles(b, b->block * b->vec)) return rc;
    const int grid = batch_grid(b, b->block);
    const lbm::StepParams<double> p = batch_params(b);
    int rc;
    if (b->storage == LBM_STORE_F64)
        rc = lbm::run_batch_f64_q9(op, b->dLats, b->dTiles, b->ntiles, b->pool, parity, b->block,
                                   grid, b->vec, p, g_stream);
    else
        rc = lbm::run_batch_f32_q9(op, b->dLats, b->dTiles, b->ntiles, b->pool, parity, b->block,
                                   grid, b->vec, lbm::to_float_params(p), g_stream);
    if (rc) return fail(-4, "batch kernel launch failed (%d)", rc);
    return 0;
}

// capture `chunk` fused steps starting at `parity` (chunk even: parity returns)
int batch_graph(lbm_batch b, int chunk, int parity) {
    const int grid = batch_grid(b, b->block);
    if (b->gexec && b->g_chunk == chunk && b->g_parity == parity && b->g_block == b->block &&
        b->g_grid == grid && b->g_vec == b->vec)
        return 0;
    batch_drop_graph(b);
    if (int rc = batch_tiles(b, b->block * b->vec)) return rc;
    HIP_TRY(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    int rc = 0, cur = parity;
    for (int k = 0; k < chunk && !rc; ++k) {
        rc = batch_launch(b, 1, cur);
        cur ^= 1;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(g_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(-3, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(-3, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->graph = g;
    b->g_chunk = chunk;
    b->g_parity = parity;
    b->g_block = b->block;
    b->g_grid = grid;
    b->g_vec = b->vec;
    return 0;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_batch *out, int quadrature, int storage, int n, const int *nx,
                     const int *ny, const double *tau_s, const double *tau_a, double rho0) {
    if (!out) return fail(-1, "lbm_batch_create: null output handle");
    *out = nullptr;
    if (quadrature != LBM_D2Q9) return fail(-2, "batches are implemented for D2Q9");
    if (storage != LBM_STORE_F64 && storage != LBM_STORE_F32)
        return fail(-2, "batches use LBM_STORE_F64 or LBM_STORE_F32");
    if (n < 1 || !nx || !ny || !tau_s) return fail(-1, "lbm_batch_create: empty batch");
    if (int rc = ensure_stream()) return rc;
    lbm_batch b = new lbm_batch_s;
    b->qid = quadrature;
    b->storage = storage;
    b->n = n;
    b->rho0 = rho0;
    host_quad(quadrature, b->q);
    const int eb = storage_bytes(storage);
    long long off = 0;
    for (int k = 0; k < n; ++k) {
        // node byte offsets within a lattice plane are 32-bit in the kernel
        if (nx[k] < 1 || ny[k] < 1 || nx[k] > (1 << 27) ||
            (long long)(nx[k] + 31) * ny[k] > (1LL << 28)) {
            delete b;
            return fail(-1, "lattice %d: size %dx%d not supported", k, nx[k], ny[k]);
        }
        const double ts = tau_s[k], ta = tau_a ? tau_a[k] : tau_s[k];
        if (!(ts > 0.0) || !(ta > 0.0)) {
            delete b;
            return fail(-1, "lattice %d: relaxation times must be positive", k);
        }
        lbm::BatchLattice L{};
        L.nx = nx[k];
        L.ny = ny[k];
        const int line = 128 / eb;  // elements per 128-byte line
        L.pitch = (nx[k] + line - 1) / line * line;
        L.plane = (long long)L.pitch * ny[k];
        L.off[0] = off;
        off += L.plane * b->q.Q;
        L.off[1] = off;
        off += L.plane * b->q.Q;
        L.omo = 1.0 - 1.0 / ts;
        L.om = 1.0 / ts;
        L.oms = 1.0 / ts;
        L.oma = 1.0 / ta;
        L.inv_pitch = 1.0 / (double)L.pitch;
        L.coll = tau_a ? 1 : 0;
        b->lats.push_back(L);
        b->nx.push_back(nx[k]);
        b->ny.push_back(ny[k]);
        b->nodes += (long long)nx[k] * ny[k];
    }
    b->pool_bytes = (size_t)off * eb;
    hipError_t e = hipMalloc(&b->pool, b->pool_bytes);
    if (e == hipSuccess) e = hipMalloc(&b->dLats, b->lats.size() * sizeof(lbm::BatchLattice));
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->dLats, b->lats.data(), b->lats.size() * sizeof(lbm::BatchLattice),
                           hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->pool, 0, b->pool_bytes, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    if (e != hipSuccess) {
        free_dev(b->pool);
        free_dev(reinterpret_cast<void *&>(b->dLats));
        delete b;
        return fail(-3, "batch allocation failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

int lbm_batch_destroy(lbm_batch b) {
    if (!b) return 0;
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    batch_drop_graph(b);
    free_dev(b->pool);
    free_dev(reinterpret_cast<void *&>(b->dLats));
    free_dev(reinterpret_cast<void *&>(b->dTiles));
    delete b;
    return 0;
}

int lbm_batch_info(lbm_batch b, int *n, int64_t *nodes, int64_t *bytes) {
    if (!b) return fail(-1, "null batch");
    if (n) *n = b->n;
    if (nodes) *nodes = b->nodes;
    if (bytes) *bytes = (int64_t)b->pool_bytes;
    return 0;
}

static int batch_xfer(lbm_batch b, int k, int buf, double *host, bool up) {
    if (!b) return fail(-1, "null batch");
    if (k < 0 || k >= b->n) return fail(-1, "lattice index %d outside 0..%d", k, b->n - 1);
    if (!host) return fail(-1, "null host array");
    const lbm::BatchLattice &L = b->lats[k];
    // the host array is dense (nx, ny, Q); the pool's rows are `pitch` long:
    // ny * Q rows of nx values either way
    const size_t nb64 = (size_t)L.plane * b->q.Q * sizeof(double);
    const size_t hrow = (size_t)L.nx * sizeof(double), drow = (size_t)L.pitch * sizeof(double);
    const size_t rows = (size_t)L.ny * b->q.Q;
    const int eb = batch_elem(b);
    char *dst = static_cast<char *>(b->pool) + (size_t)L.off[buf] * eb;
    if (b->storage == LBM_STORE_F64) {
        if (up)
            HIP_TRY(hipMemcpy2DAsync(dst, drow, host, hrow, hrow, rows, hipMemcpyHostToDevice,
                                     g_stream));
        else
            HIP_TRY(hipMemcpy2DAsync(host, hrow, dst, drow, hrow, rows, hipMemcpyDeviceToHost,
                                     g_stream));
        HIP_TRY(hipStreamSynchronize(g_stream));
        return 0;
    }
    // FP32: through an FP64 buffer of the pool's shape (its padding slots are
    // converted along and never read back)
    double *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, nb64));
    lbm::StepParams<double> p = batch_params(b);
    p.nx = L.pitch;
    p.ny = L.ny;
    p.plane = L.plane;
    hipError_t e = hipSuccess;
    if (up) e = hipMemcpy2DAsync(tmp, drow, host, hrow, hrow, rows, hipMemcpyHostToDevice, g_stream);
    int rc = 0;
    if (e == hipSuccess) rc = lbm::dispatch_upload(b->qid, b->storage, tmp, dst, p, up ? 1 : 0, g_stream);
    if (e == hipSuccess && !rc && !up)
        e = hipMemcpy2DAsync(host, hrow, tmp, drow, hrow, rows, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(-3, "batch transfer failed: %s", hipGetErrorString(e));
    if (rc) return fail(-4, "conversion kernel failed (%d)", rc);
    return 0;
}

int lbm_batch_upload(lbm_batch b, int k, const double *f) {
    if (int rc = batch_xfer(b, k, 0, const_cast<double *>(f), true)) return rc;
    if (int rc = batch_xfer(b, k, 1, const_cast<double *>(f), true)) return rc;
    return 0;
}

int lbm_batch_download(lbm_batch b, int k, int role, double *f) {
    if (!b) return fail(-1, "null batch");
    if (role != 0 && role != 1) return fail(-1, "role must be 0 (f_stream) or 1 (f_collision)");
    const int buf = role == 0 ? b->parity_stream : (b->parity_stream ^ 1);
    return batch_xfer(b, k, buf, f, false);
}

int lbm_batch_set_launch(lbm_batch b, int block, int grid) {
    if (!b) return fail(-1, "null batch");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return fail(-1, "block must be 0 (automatic), 64, 128 or 256");
    if (grid < 0) return fail(-1, "grid must be >= 0");
    b->block = block ? block : 128;
    b->grid = grid;
    return 0;
}

int lbm_batch_set_nodes_per_thread(lbm_batch b, int nodes) {
    if (!b) return fail(-1, "null batch");
    if (nodes != 1 && nodes != 2 && nodes != 4)
        return fail(-1, "nodes per thread must be 1, 2 or 4");
    b->vec = nodes;
    return 0;
}

int lbm_batch_get_launch(lbm_batch b, int *block, int *grid, int *nodes_per_thread) {
    if (!b) return fail(-1, "null batch");
    if (block) *block = b->block;
    if (grid) *grid = b->grid;
    if (nodes_per_thread) *nodes_per_thread = b->vec;
    return 0;
}

int lbm_batch_simulate(lbm_batch b, int64_t nsteps, int graph_chunk, double *ms_fused) {
    if (!b) return fail(-1, "null batch");
    if (nsteps < 0 || graph_chunk < 0) return fail(-1, "negative step count");
    if (ms_fused) *ms_fused = 0.0;
    if (nsteps == 0) return 0;
    const int chunk = graph_chunk & ~1;  // even: the parity returns after a replay
    if (int rc = batch_launch(b, 0, b->parity_stream)) return rc;  // C_0 = collide(S_0)
    int cur = b->parity_stream ^ 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms_fused) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, g_stream));
    }
    int64_t left = nsteps - 1;
    if (chunk > 0 && left >= chunk) {
        if (int rc = batch_graph(b, chunk, cur)) return rc;
        while (left >= chunk) {
            HIP_TRY(hipGraphLaunch(b->gexec, g_stream));
            left -= chunk;
        }
    }
    for (; left > 0; --left) {
        if (int rc = batch_launch(b, 1, cur)) return rc;
        cur ^= 1;
    }
    if (ms_fused) HIP_TRY(hipEventRecord(e1, g_stream));
    if (int rc = batch_launch(b, 2, cur)) return rc;  // S_n = stream(C_{n-1})
    b->parity_stream = cur ^ 1;
    if (ms_fused) {
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        *ms_fused = (double)t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int lbm_batch_autotune(lbm_batch b, int trial_steps, int *block, int *grid, double *us_per_step) {
    if (!b) return fail(-1, "null batch");
    if (trial_steps < 1) trial_steps = 1;
    // tuning advances the pool: save it and restore afterwards
    void *save = nullptr;
    HIP_TRY(hipMalloc(&save, b->pool_bytes));
    int rc = lbm_field_copy_raw(save, b->pool, b->pool_bytes);
    const int saved_parity = b->parity_stream, sb = b->block, sg = b->grid, sv = b->vec;
    double best = 1e300;
    int bb = 256, bg = 0, bv = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks[3] = {64, 128, 256};
    const int vecs[3] = {1, 2, 4};
    for (int vi = 0; !rc && vi < 3; ++vi)
        for (int bi = 0; !rc && bi < 3; ++bi) {
            const int per_cu = 2048 / blocks[bi];
            if (int r2 = batch_tiles(b, blocks[bi] * vecs[vi])) {
                rc = r2;
                break;
            }
            // persistent grids of 1/4 .. 4 resident waves of blocks, and one block per tile
            const int grids[5] = {cus * per_cu / 4, cus * per_cu / 2, cus * per_cu,
                                  cus * per_cu * 4, b->ntiles};
            for (int gi = 0; !rc && gi < 5; ++gi) {
                b->block = blocks[bi];
                b->grid = grids[gi];
                b->vec = vecs[vi];
                double ms = 0.0;
                rc = lbm_batch_simulate(b, trial_steps + 1, 0, &ms);
                const double us = 1e3 * ms / trial_steps;
                if (!rc && us < best) {
                    best = us;
                    bb = blocks[bi];
                    bg = grids[gi];
                    bv = vecs[vi];
                }
            }
        }
    if (!rc) rc = lbm_field_copy_raw(b->pool, save, b->pool_bytes);
    (void)hipStreamSynchronize(g_stream);
    (void)hipFree(save);
    b->parity_stream = saved_parity;
    batch_drop_graph(b);
    if (rc) {
        b->block = sb;
        b->grid = sg;
        b->vec = sv;
        return rc;
    }
    b->block = bb;
    b->grid = bg;
    b->vec = bv;
    if (block) *block = bb;
    if (grid) *grid = bg;
    if (us_per_step) *us_per_step = best;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// reverse mode: gradients of L = 1/2 sum |u(S_T) - u_target|^2 with respect to
// u0 (S_0 = feq(rho0, u0)) and the relaxation time(s).
//
// Checkpoints.  The forward pass keeps, every K steps, the post-collision
// field C_{aK-1} from which the pull kernel re-forms S_{aK} (segment a starts
// there; segment 0 starts from S_0).  The reverse pass walks the segments
// backwards: it recomputes C_{t0} .. C_{t1-2} of the segment into K - 1
// buffers, then runs the reverse steps t1-1 .. t0, each reading S_t by a pull
// from C_{t-1}.  Storage: ceil(T/K) + K + 2 population fields; work: two
// forward passes and one reverse pass.  K = ceil(sqrt(T)) by default.
// ---------------------------------------------------------------------------
struct lbm_adjoint_s {
    lbm_model m = nullptr;
    lbm_field_s geo;  // geometry and storage of the populations (no data)
    int64_t nsteps = 0;
    int K = 1, nseg = 0;
    size_t state_bytes = 0, lam_bytes = 0, total_bytes = 0;
    void *S0 = nullptr;
    void *pp[2] = {nullptr, nullptr};
    std::vector<void *> ck;   // ck[a] = C_{aK-1}, a = 1 .. nseg-1
    std::vector<void *> seg;  // recomputed C_{t0} .. C_{t1-2} of one segment
    void *lam[2] = {nullptr, nullptr};
    double *du0 = nullptr, *dut = nullptr, *dgrad = nullptr, *dlossn = nullptr, *dgtau = nullptr,
           *duT = nullptr, *dpart = nullptr, *dout = nullptr;
};

namespace {

constexpr int kAdjParts = 1024;

void adjoint_free(lbm_adjoint a) {
    free_dev(a->S0);
    free_dev(a->pp[0]);
    free_dev(a->pp[1]);
    for (void *&p : a->ck) free_dev(p);
    for (void *&p : a->seg) free_dev(p);
    free_dev(a->lam[0]);
    free_dev(a->lam[1]);
    for (double **p : {&a->du0, &a->dut, &a->dgrad, &a->dlossn, &a->dgtau, &a->duT, &a->dpart,
                       &a->dout}) {
        void *v = *p;
        free_dev(v);
        *p = nullptr;
    }
}

int adjoint_check_model(lbm_model m, int qid, const char *what) {
    if (!m) return fail(-1, "%s: null model", what);
    if (m->qid != qid)
        return fail(-1, "%s: the model is D2Q%d but the adjoint was made for D2Q%d", what, m->qid, qid);
    if (m->coll != LBM_SRT && m->coll != LBM_TRT)
        return fail(-2, "%s: the adjoint supports SRT and TRT collisions", what);
    if (m->force_kind != LBM_FORCE_NONE)
        return fail(-2, "%s: the adjoint supports unforced flows only", what);
    if (!m->bcs.empty())
        return fail(-2, "%s: the adjoint supports periodic lattices only (no boundary conditions)",
                    what);
    return 0;
}

int adj_launch(lbm_adjoint a, const lbm::StepParams<double> &p, lbm::AdjLaunch L) {
    L.qid = a->geo.qid;
    L.coll = a->m->coll == LBM_SRT ? 0 : 1;
    const int rc = a->geo.storage == LBM_STORE_F64
                       ? lbm::run_adjoint_f64(L, p, g_stream)
                       : lbm::run_adjoint_f32(L, lbm::to_float_params(p), g_stream);
    if (rc) return fail(-4, "adjoint kernel launch failed (code %d): %s", rc,
                        hipGetErrorString(hipGetLastError()));
    return 0;
}

// forward step kernels on raw buffers of the adjoint's geometry
int adj_forward_op(lbm_adjoint a, lbm::Op op, const void *src, void *dst) {
    lbm::Launch L{};
    L.op = op;
    L.qid = a->geo.qid;
    L.coll = a->m->coll;
    L.storage = a->geo.storage;
    L.src = src;
    L.dst = dst;
    if (op == lbm::Op::Fused) auto_launch(a->m, &a->geo, L.block, L.block_y, L.vec, L.nt);
    const lbm::StepParams<double> p = make_params(a->m, &a->geo, false);
    const int rc = a->geo.storage == LBM_STORE_F64 ? lbm::dispatch_f64(L, p, g_stream)
                                                   : lbm::dispatch_lp(L, p, g_stream);
    if (rc) return fail(-4, "kernel launch failed (code %d): %s", rc,
                        hipGetErrorString(hipGetLastError()));
    return 0;
}

// post-collision buffer of forward step t: checkpoints at t = aK - 1
void *adj_cbuf(lbm_adjoint a, int64_t t) {
    if ((t + 1) % a->K == 0 && t + 1 < a->nsteps) return a->ck[(size_t)((t + 1) / a->K)];
    return a->pp[t & 1];
}

// S_0 from u0, the forward pass with checkpoints, lambda_T and the loss
int adjoint_forward(lbm_adjoint a, const double *u0, const double *ut, bool want_uT) {
    const size_t vb = (size_t)a->geo.nx * a->geo.ny * 2 * sizeof(double);
    HIP_TRY(hipMemcpyAsync(a->du0, u0, vb, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipMemcpyAsync(a->dut, ut, vb, hipMemcpyHostToDevice, g_stream));
    const lbm::StepParams<double> p = make_params(a->m, &a->geo, false);
    lbm::AdjLaunch L{};
    L.what = 0;
    L.pops = a->S0;
    L.u = a->du0;
    if (int rc = adj_launch(a, p, L)) return rc;
    const int64_t T = a->nsteps;
    if (T > 0) {
        if (int rc = adj_forward_op(a, lbm::Op::Collide, a->S0, adj_cbuf(a, 0))) return rc;
        for (int64_t t = 1; t < T; ++t)
            if (int rc = adj_forward_op(a, lbm::Op::Fused, adj_cbuf(a, t - 1), adj_cbuf(a, t)))
                return rc;
    }
    L = lbm::AdjLaunch{};
    L.what = 1;
    L.direct = T == 0 ? 1 : 0;
    L.state = T == 0 ? a->S0 : adj_cbuf(a, T - 1);
    L.u = a->dut;
    L.lam_out = a->lam[0];
    L.out = a->dlossn;
    L.aux = want_uT ? a->duT : nullptr;
    if (int rc = adj_launch(a, p, L)) return rc;
    if (lbm::adjoint_sum(a->dlossn, (long long)a->geo.nx * a->geo.ny, a->dpart, kAdjParts,
                         a->dout, g_stream))
        return fail(-4, "reduction kernel failed");
    return 0;
}

}  // namespace

extern "C" {

int lbm_adjoint_create(lbm_adjoint *out, lbm_model m, int storage, int nx, int ny, double rho0,
                       int64_t nsteps, int checkpoint_every) {
    if (!out) return fail(-1, "lbm_adjoint_create: null output handle");
    *out = nullptr;
    if (!m) return fail(-1, "lbm_adjoint_create: null model");
    if (int rc = adjoint_check_model(m, m->qid, "lbm_adjoint_create")) return rc;
    if (storage != LBM_STORE_F64 && storage != LBM_STORE_F32)
        return fail(-2, "lbm_adjoint_create: storage must be LBM_STORE_F64 or LBM_STORE_F32");
    if (nx < 1 || ny < 1) return fail(-1, "lattice size %dx%d must be positive", nx, ny);
    if (ny > 65535) return fail(-1, "lattice height %d exceeds 65535", ny);
    if ((long long)nx * (long long)ny > (1LL << 31) - 1)
        return fail(-1, "lattice of %lld nodes exceeds 2^31-1", (long long)nx * ny);
    if (nsteps < 0) return fail(-1, "negative step count");
    if (checkpoint_every < 0) return fail(-1, "checkpoint interval must be >= 0");
    if (!(rho0 > 0.0)) return fail(-1, "rho0 must be positive");
    if (int rc = ensure_stream()) return rc;
    lbm_adjoint a = new lbm_adjoint_s;
    a->m = m;
    a->geo.qid = m->qid;
    a->geo.nx = nx;
    a->geo.ny = ny;
    a->geo.Q = m->q.Q;
    a->geo.storage = storage;
    a->geo.rho0 = rho0;
    a->geo.dscale = 1.0;
    a->nsteps = nsteps;
    int K = checkpoint_every;
    if (K == 0) {
        K = 1;
        while ((int64_t)K * K < nsteps) ++K;
    }
    a->K = K;
    a->nseg = nsteps > 0 ? (int)((nsteps + K - 1) / K) : 0;
    const size_t plane = (size_t)nx * (size_t)ny;
    const size_t eb = storage == LBM_STORE_F64 ? 8 : 4;
    a->state_bytes = plane * (size_t)m->q.Q * eb;
    a->lam_bytes = a->state_bytes;
    const int64_t seglen = nsteps < K ? nsteps : K;
    a->ck.assign(a->nseg > 0 ? (size_t)a->nseg : 1, nullptr);
    a->seg.assign(seglen > 1 ? (size_t)(seglen - 1) : 0, nullptr);
    auto alloc = [&](void **p, size_t bytes) -> int {
        const hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(-3, "hipMalloc(%zu) for the adjoint failed: %s", bytes, hipGetErrorString(e));
        }
        a->total_bytes += bytes;
        return 0;
    };
    int rc = alloc(&a->S0, a->state_bytes);
    for (int k = 0; !rc && k < 2; ++k) rc = alloc(&a->pp[k], a->state_bytes);
    for (size_t k = 1; !rc && k < a->ck.size(); ++k) rc = alloc(&a->ck[k], a->state_bytes);
    for (size_t k = 0; !rc && k < a->seg.size(); ++k) rc = alloc(&a->seg[k], a->state_bytes);
    for (int k = 0; !rc && k < 2; ++k) rc = alloc(&a->lam[k], a->lam_bytes);
    void *v = nullptr;
    if (!rc && !(rc = alloc(&v, plane * 2 * sizeof(double)))) a->du0 = (double *)v;
    if (!rc && !(rc = alloc(&v, plane * 2 * sizeof(double)))) a->dut = (double *)v;
    if (!rc && !(rc = alloc(&v, plane * 2 * sizeof(double)))) a->dgrad = (double *)v;
    if (!rc && !(rc = alloc(&v, plane * 2 * sizeof(double)))) a->dgtau = (double *)v;
    if (!rc && !(rc = alloc(&v, plane * 2 * sizeof(double)))) a->duT = (double *)v;
    if (!rc && !(rc = alloc(&v, plane * sizeof(double)))) a->dlossn = (double *)v;
    if (!rc && !(rc = alloc(&v, kAdjParts * sizeof(double)))) a->dpart = (double *)v;
    if (!rc && !(rc = alloc(&v, 4 * sizeof(double)))) a->dout = (double *)v;
    if (rc) {
        adjoint_free(a);
        delete a;
        return rc;
    }
    *out = a;
    return 0;
}

int lbm_adjoint_destroy(lbm_adjoint a) {
    if (!a) return 0;
    (void)hipStreamSynchronize(g_stream);
    adjoint_free(a);
    delete a;
    return 0;
}

int lbm_adjoint_info(lbm_adjoint a, int *checkpoint_every, int *n_checkpoints, int64_t *bytes) {
    if (!a) return fail(-1, "null adjoint");
    if (checkpoint_every) *checkpoint_every = a->K;
    if (n_checkpoints) *n_checkpoints = a->nseg > 0 ? a->nseg - 1 : 0;
    if (bytes) *bytes = (int64_t)a->total_bytes;
    return 0;
}

int lbm_adjoint_loss(lbm_adjoint a, const double *u0, const double *u_target, double *loss,
                     double *u_final) {
    if (!a) return fail(-1, "null adjoint");
    if (!u0 || !u_target || !loss) return fail(-1, "lbm_adjoint_loss: null argument");
    if (int rc = adjoint_check_model(a->m, a->geo.qid, "lbm_adjoint_loss")) return rc;
    if (int rc = adjoint_forward(a, u0, u_target, u_final != nullptr)) return rc;
    HIP_TRY(hipMemcpyAsync(loss, a->dout, sizeof(double), hipMemcpyDeviceToHost, g_stream));
    if (u_final)
        HIP_TRY(hipMemcpyAsync(u_final, a->duT, (size_t)a->geo.nx * a->geo.ny * 2 * sizeof(double),
                               hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return 0;
}

int lbm_adjoint_gradient(lbm_adjoint a, const double *u0, const double *u_target, double *loss,
                         double *grad_u0, double *grad_tau, double *ms) {
    if (!a) return fail(-1, "null adjoint");
    if (!u0 || !u_target) return fail(-1, "lbm_adjoint_gradient: null argument");
    if (int rc = adjoint_check_model(a->m, a->geo.qid, "lbm_adjoint_gradient")) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, g_stream));
    }
    if (int rc = adjoint_forward(a, u0, u_target, false)) return rc;
    const lbm::StepParams<double> p = make_params(a->m, &a->geo, false);
    const long long plane = (long long)a->geo.nx * a->geo.ny;
    HIP_TRY(hipMemsetAsync(a->dgtau, 0, (size_t)plane * 2 * sizeof(double), g_stream));
    int cur = 0;  // lam[cur] holds lambda_{t+1}
    for (int s = a->nseg - 1; s >= 0; --s) {
        const int64_t t0 = (int64_t)s * a->K;
        const int64_t t1 = t0 + a->K < a->nsteps ? t0 + a->K : a->nsteps;
        // recompute C_{t0} .. C_{t1-2}
        for (int64_t j = 0; t0 + j <= t1 - 2; ++j) {
            int rc;
            if (j > 0) rc = adj_forward_op(a, lbm::Op::Fused, a->seg[j - 1], a->seg[j]);
            else if (s == 0) rc = adj_forward_op(a, lbm::Op::Collide, a->S0, a->seg[0]);
            else rc = adj_forward_op(a, lbm::Op::Fused, a->ck[s], a->seg[0]);
            if (rc) return rc;
        }
        for (int64_t t = t1 - 1; t >= t0; --t) {
            lbm::AdjLaunch L{};
            L.what = 2;
            if (t == t0) {
                L.direct = s == 0 ? 1 : 0;
                L.state = s == 0 ? a->S0 : a->ck[s];
            } else {
                L.direct = 0;
                L.state = a->seg[t - 1 - t0];
            }
            L.lam_in = a->lam[cur];
            L.lam_out = a->lam[cur ^ 1];
            L.out = a->dgtau;
            if (int rc = adj_launch(a, p, L)) return rc;
            cur ^= 1;
        }
    }
    lbm::AdjLaunch L{};
    L.what = 3;
    L.u = a->du0;
    L.lam_in = a->lam[cur];
    L.out = a->dgrad;
    if (int rc = adj_launch(a, p, L)) return rc;
    const int ntau = a->m->coll == LBM_SRT ? 1 : 2;
    for (int k = 0; k < ntau; ++k)
        if (lbm::adjoint_sum(a->dgtau + k * plane, plane, a->dpart, kAdjParts, a->dout + 1 + k,
                             g_stream))
            return fail(-4, "reduction kernel failed");
    if (ms) HIP_TRY(hipEventRecord(e1, g_stream));
    double h[3] = {0.0, 0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(h, a->dout, 3 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    if (grad_u0)
        HIP_TRY(hipMemcpyAsync(grad_u0, a->dgrad, (size_t)plane * 2 * sizeof(double),
                               hipMemcpyDeviceToHost, g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    if (loss) *loss = h[0];
    if (grad_tau)
        for (int k = 0; k < ntau; ++k) grad_tau[k] = h[1 + k];
    if (ms) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// snapshots: TakeSnapshots (take_snapshots.jl:3-29) without stalling the step
// ---------------------------------------------------------------------------
// A snapshot is copy(f_in) in the reference's (nx, ny, Q) Float64 layout.
// lbm_snapshot_take decodes / copies the field into one of a few staging
// buffers in HBM on the library stream (a pass over the field), then a second
// stream copies that buffer into pinned host memory and, when a directory was
// given, a host callback on that stream writes the .npy file.  The library
// stream runs on meanwhile; it waits only to reuse a staging buffer whose
// copy to the host has not finished.  The staging buffers bound the HBM held
// (slots x nx x ny x Q x 8 bytes); the host keeps every snapshot, as the
// reference's vector of arrays does.
struct SnapshotHost {
    int64_t t = 0;
    double *host = nullptr;  // pinned
    std::string path;        // empty: not written
    size_t bytes = 0;
    int nx = 0, ny = 0, Q = 0;
    int write_rc = 0;
};

struct lbm_snapshots_s {
    int qid = 0, nx = 0, ny = 0, Q = 0, slots = 0;
    size_t bytes = 0;
    std::string dir;
    hipStream_t copy = nullptr;
    std::vector<double *> staging;
    std::vector<hipEvent_t> staged, freed;
    std::vector<bool> used;
    std::vector<SnapshotHost *> shots;
};

namespace {

// .npy version 1.0: little-endian Float64, Fortran order, shape (nx, ny, Q)
int write_npy(const SnapshotHost &s) {
    FILE *fp = std::fopen(s.path.c_str(), "wb");
    if (!fp) return -1;
    char dict[160];
    std::snprintf(dict, sizeof dict,
                  "{'descr': '<f8', 'fortran_order': True, 'shape': (%d, %d, %d), }", s.nx, s.ny,
                  s.Q);
    std::string header(dict);
    // magic (6) + version (2) + header length (2) + header, padded to 64 bytes
    size_t total = 10 + header.size() + 1;
    header.append((64 - total % 64) % 64, ' ');
    header.push_back('\n');
    const unsigned short hl = (unsigned short)header.size();
    const unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0,
                                   (unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
    bool ok = std::fwrite(pre, 1, 10, fp) == 10 &&
              std::fwrite(header.data(), 1, header.size(), fp) == header.size() &&
              std::fwrite(s.host, 1, s.bytes, fp) == s.bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : -1;
}

void write_npy_cb(void *arg) {
    SnapshotHost *s = static_cast<SnapshotHost *>(arg);
    s->write_rc = write_npy(*s);
}

}  // namespace

extern "C" {

int lbm_snapshots_create(lbm_snapshots *out, int quadrature, int nx, int ny, int device_slots,
                         const char *directory) {
    if (!out) return fail(-1, "lbm_snapshots_create: null output handle");
    *out = nullptr;
    HostQuad q;
    if (!host_quad(quadrature, q)) return fail(-1, "unknown quadrature D2Q%d", quadrature);
    if (nx < 1 || ny < 1) return fail(-1, "lattice size must be positive");
    if (device_slots < 1 || device_slots > 64) return fail(-1, "device_slots must be 1..64");
    if (int rc = ensure_stream()) return rc;
    auto *s = new lbm_snapshots_s;
    s->qid = quadrature;
    s->nx = nx;
    s->ny = ny;
    s->Q = q.Q;
    s->slots = device_slots;
    s->bytes = (size_t)nx * ny * q.Q * sizeof(double);
    if (directory) s->dir = directory;
    hipError_t e = hipStreamCreate(&s->copy);
    for (int k = 0; k < device_slots && e == hipSuccess; ++k) {
        double *d = nullptr;
        e = hipMalloc(&d, s->bytes);
        if (e == hipSuccess) s->staging.push_back(d);
        hipEvent_t a = nullptr, b = nullptr;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
        if (e == hipSuccess) s->staged.push_back(a);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
        if (e == hipSuccess) s->freed.push_back(b);
        s->used.push_back(false);
    }
    if (e != hipSuccess) {
        lbm_snapshots_destroy(s);
        return fail(-3, "lbm_snapshots_create: %s", hipGetErrorString(e));
    }
    *out = s;
    return 0;
}

int lbm_snapshots_destroy(lbm_snapshots s) {
    if (!s) return 0;
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    if (g_stream) (void)hipStreamSynchronize(g_stream);
    for (double *d : s->staging) (void)hipFree(d);
    for (hipEvent_t e : s->staged) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->freed) (void)hipEventDestroy(e);
    for (SnapshotHost *h : s->shots) {
        if (h->host) (void)hipHostFree(h->host);
        delete h;
    }
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return 0;
}

int lbm_snapshot_take(lbm_snapshots s, lbm_field f, int64_t t) {
    if (!s) return fail(-1, "null snapshots handle");
    if (int rc = check_field(f, "lbm_snapshot_take")) return rc;
    if (f->qid != s->qid || f->nx != s->nx || f->ny != s->ny)
        return fail(-1, "lbm_snapshot_take: field is D2Q%d %dx%d, the snapshots D2Q%d %dx%d",
                    f->qid, f->nx, f->ny, s->qid, s->nx, s->ny);
    const int j = (int)(s->shots.size() % (size_t)s->slots);
    auto *h = new SnapshotHost;
    h->t = t;
    h->bytes = s->bytes;
    h->nx = s->nx;
    h->ny = s->ny;
    h->Q = s->Q;
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&h->host), s->bytes);
    if (e != hipSuccess) {
        delete h;
        return fail(-3, "lbm_snapshot_take: pinned host buffer: %s", hipGetErrorString(e));
    }
    if (!s->dir.empty()) {
        char name[64];
        std::snprintf(name, sizeof name, "/snapshot_%010lld.npy", (long long)t);
        h->path = s->dir + name;
    }
    s->shots.push_back(h);
    // the staging buffer is reused only after its previous copy to the host
    if (s->used[j]) HIP_TRY(hipStreamWaitEvent(g_stream, s->freed[j], 0));
    s->used[j] = true;
    if (f->storage == LBM_STORE_F64) {
        if (int rc = lbm::dispatch_copy(f->d, s->staging[j], s->bytes, g_stream))
            return fail(-4, "snapshot copy kernel failed (%d)", rc);
    } else {
        const lbm::StepParams<double> p = field_params(f);
        if (int rc = lbm::dispatch_upload(f->qid, f->storage, s->staging[j], f->d, p, 0, g_stream))
            return fail(-4, "snapshot decode kernel failed (%d)", rc);
    }
    HIP_TRY(hipEventRecord(s->staged[j], g_stream));
    HIP_TRY(hipStreamWaitEvent(s->copy, s->staged[j], 0));
    HIP_TRY(hipMemcpyAsync(h->host, s->staging[j], s->bytes, hipMemcpyDeviceToHost, s->copy));
    if (!h->path.empty()) HIP_TRY(hipLaunchHostFunc(s->copy, write_npy_cb, h));
    HIP_TRY(hipEventRecord(s->freed[j], s->copy));
    return 0;
}

int lbm_snapshots_wait(lbm_snapshots s, int *n) {
    if (!s) return fail(-1, "null snapshots handle");
    HIP_TRY(hipStreamSynchronize(s->copy));
    for (const SnapshotHost *h : s->shots)
        if (h->write_rc) return fail(-3, "writing %s failed", h->path.c_str());
    if (n) *n = (int)s->shots.size();
    return 0;
}

int lbm_snapshot_get(lbm_snapshots s, int k, int64_t *t, double *host) {
    if (!s) return fail(-1, "null snapshots handle");
    if (k < 0 || (size_t)k >= s->shots.size())
        return fail(-1, "snapshot %d of %d", k, (int)s->shots.size());
    HIP_TRY(hipStreamSynchronize(s->copy));
    const SnapshotHost *h = s->shots[(size_t)k];
    if (t) *t = h->t;
    if (host) std::memcpy(host, h->host, h->bytes);
    return 0;
}

}  // extern "C"
