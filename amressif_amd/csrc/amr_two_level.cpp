// Two-level composite time step (C-ABI amressif_amr_*): a coarse level that is
// a full single-level context plus one refined patch (ratio 2) over the coarse
// rows [jc0, n0), advanced with one dt (no subcycling):
//   1. coarse predictor: the single-level BDF2 + RK4 dual-time loop
//      (main.cpp:879-1040) over the whole coarse domain;
//   2. fine predictor: the same loop on the patch, south ghosts and interface
//      faces from the coarse predictor (kernels_amr.hip);
//   3. coarse faces and cells under the patch <- averages of the fine ones;
//   4. composite projection: FGMRES on the composite Poisson operator,
//      preconditioned by a coarse V-cycle on the restricted residual plus
//      red-black Gauss-Seidel on the patch; update of both levels and p;
//   5. average down again.
// In the mixed mode the coarse level runs the mixed single-level momentum
// form and the preconditioner's V-cycle runs in FP32; the fine predictor, the
// FGMRES vectors, the composite residual and every reduction are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "internal.hpp"


using amr::Ctx;

struct amressif_amr {
    amressif_ctx *coarse = nullptr;
    amr::AmrGeom a{};
    hipStream_t stream = nullptr;
    double dt = 0.0;
    int mixed = 0, m = 30, last_P = 4;
    int host_gmres = 0;  // AMRESSIF_AMR_HOST_GMRES=1: the Arnoldi process on the host (for A/B runs)
    // fine level (faces pitch P1, cells pitch P1, ghosted cells pitch Pg1)
    double *U1[2]{}, *Up1[2]{}, *C1[2]{}, *F1[2]{}, *F2[2]{}, *p1 = nullptr, *XU = nullptr, *XV = nullptr;
    double *DXU = nullptr, *DXV = nullptr;  // cells of the residual (the mixed step's refinement form)
    amr::Ctl *ctlF = nullptr, *ctlF_host = nullptr;
    double *part = nullptr, *red = nullptr, *red_host = nullptr, *coef = nullptr;
    // composite vectors [coarse segment | fine segment]
    size_t off1 = 0, vs = 0;
    double *V = nullptr, *Z = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *vcur = nullptr, *zcur = nullptr;
    amr::GmState *gst = nullptr;  // the composite FGMRES's Arnoldi state (device)
    int last_gm = 8;              // Arnoldi steps of the previous solve (first batch size)
    double *dstat = nullptr;      // per step: composite ||D(u*)||inf, ||D(u^{n+1})||inf (device)
    int dstat_cap = 0;
    std::vector<void *> allocs;
    std::string err;
};

namespace {

int afail(amressif_amr *A, int code, const std::string &msg)
{
    if (A) A->err = msg;
    return code;
}

#define ACHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return afail(A, AMRESSIF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int round_up(long long x, int m) { return (int)((x + m - 1) / m * m); }

template <class T>
hipError_t aalloc(amressif_amr *A, T **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 256);
    if (e != hipSuccess) return e;
    A->allocs.push_back(q);
    e = hipMemsetAsync(q, 0, count * sizeof(T) + 256, A->stream);
    *p = (T *)q;
    return e;
}

// reduce ncol columns of per-block partials to the host
int reduce(amressif_amr *A, int ncol, int is_max, double *out)
{
    ACHK(amr::launch_reduce_cols(A->stream, A->part, amr::amr_flat_blocks(A->a), ncol, is_max, A->red));
    ACHK(hipMemcpyAsync(A->red_host, A->red, sizeof(double) * ncol, hipMemcpyDeviceToHost, A->stream));
    ACHK(hipStreamSynchronize(A->stream));
    for (int t = 0; t < ncol; ++t) out[t] = A->red_host[t];
    return 0;
}

// ------------------------------------------------ fine momentum predictor
// The mixed step's form on the patch (as the coarse level's mixed iteration,
// in FP64): the residual of the implicit BDF2 system and the correction by the
// two-term Neumann series of its inverse Jacobian, u += r/c + R'(u) r / c^2
// (c = 1.5/dt), to the same stopping test.  Iteration k reads u^k from U1 (k
// even) or F1 and writes the other; F2 holds r.  Launches past convergence
// copy u^k, so after an odd count the converged state is in F1 and the two
// pointers are exchanged.
int fine_predictor_refine(amressif_amr *A, int *iters, int *flags)
{
    const amr::AmrGeom &a = A->a;
    Ctx &cc = A->coarse->c;
    hipStream_t s = A->stream;
    const double cr = 1.0 / a.c15, cj = 1.0 / (a.c15 * a.c15);
    double *bu[2] = {A->U1[0], A->F1[0]}, *bv[2] = {A->U1[1], A->F1[1]};
    int launched = 0, target = std::max(1, A->last_P), nb = 0;
    const int cap = cc.in.IterNum + 1;
    for (;;) {
        while (launched < target && launched < cap) {
            const int q = launched & 1;
            ACHK(amr::amr_fine_cells(a, s, bu[q], bv[q], cc.U[0], cc.U[1], A->XU, A->XV));
            ACHK(amr::amr_fine_resid(a, s, bu[q], bv[q], A->C1[0], A->C1[1], A->p1, A->XU, A->XV, A->F2[0], A->F2[1],
                                     A->ctlF));
            ACHK(amr::amr_fine_dcells(a, s, A->F2[0], A->F2[1], A->DXU, A->DXV, A->ctlF));
            ACHK(amr::amr_fine_tangent(a, s, bu[q], bv[q], A->XU, A->XV, A->F2[0], A->F2[1], A->DXU, A->DXV,
                                       bu[q ^ 1], bv[q ^ 1], cr, cj, A->part, A->ctlF, &nb));
            ACHK(amr::amr_fine_fin(a, s, A->part, nb, A->ctlF));
            ++launched;
        }
        ACHK(hipMemcpyAsync(A->ctlF_host, A->ctlF, sizeof(amr::Ctl), hipMemcpyDeviceToHost, s));
        ACHK(hipStreamSynchronize(s));
        if (A->ctlF_host->mom_done || launched >= cap) break;
        target = launched + 2;
    }
    if (launched & 1) {
        std::swap(A->U1[0], A->F1[0]);
        std::swap(A->U1[1], A->F1[1]);
    }
    A->last_P = std::max(1, A->ctlF_host->mom_count);
    *iters = A->ctlF_host->mom_count;
    *flags = A->ctlF_host->flags;
    return 0;
}

int fine_predictor(amressif_amr *A, int *iters, int *flags)
{
    const amr::AmrGeom &a = A->a;
    Ctx &cc = A->coarse->c;
    hipStream_t s = A->stream;
    ACHK(amr::amr_fine_ctl_reset(s, A->ctlF, cc.in.momentum_tolerance, cc.in.IterNum));
    for (int d = 0; d < 2; ++d) ACHK(amr::amr_fine_bdf2(a, s, A->U1[d], A->Up1[d], A->C1[d], d, A->dt));
    ACHK(amr::amr_fine_interface(a, s, A->U1[1], cc.U[1]));  // fixed interface faces from the coarse u*
    if (A->mixed) return fine_predictor_refine(A, iters, flags);
    const double *in_u[4] = {A->U1[0], A->F1[0], A->F2[0], A->F1[0]};
    const double *in_v[4] = {A->U1[1], A->F1[1], A->F2[1], A->F1[1]};
    double *out_u[4] = {A->F1[0], A->F2[0], A->F1[0], A->U1[0]};
    double *out_v[4] = {A->F1[1], A->F2[1], A->F1[1], A->U1[1]};
    int launched = 0, target = std::max(1, A->last_P), nb = 0;
    const int cap = cc.in.IterNum + 1;
    for (;;) {
        while (launched < target && launched < cap) {
            for (int st = 0; st < 4; ++st) {
                ACHK(amr::amr_fine_cells(a, s, in_u[st], in_v[st], cc.U[0], cc.U[1], A->XU, A->XV));
                ACHK(amr::amr_fine_stage(a, s, in_u[st], in_v[st], A->U1[0], A->U1[1], A->C1[0], A->C1[1], A->p1,
                                         A->XU, A->XV, out_u[st], out_v[st], cc.rk[st], st == 3 ? 1 : 0, A->part,
                                         A->ctlF, &nb));
            }
            ACHK(amr::amr_fine_fin(a, s, A->part, nb, A->ctlF));
            ++launched;
        }
        ACHK(hipMemcpyAsync(A->ctlF_host, A->ctlF, sizeof(amr::Ctl), hipMemcpyDeviceToHost, s));
        ACHK(hipStreamSynchronize(s));
        if (A->ctlF_host->mom_done || launched >= cap) break;
        target = launched + 2;
    }
    A->last_P = std::max(1, A->ctlF_host->mom_count);
    *iters = A->ctlF_host->mom_count;
    *flags = A->ctlF_host->flags;
    return 0;
}

// -------------------------------------------------- composite projection
// z = M v: coarse V-cycle on the restricted v, prolonged, then red-black
// Gauss-Seidel on the patch
int precondition(amressif_amr *A, double *v, double *z)
{
    const amr::AmrGeom &a = A->a;
    Ctx &cc = A->coarse->c;
    hipStream_t s = A->stream;
    amr::Level &L0 = cc.lev[0];
    void *rc = A->mixed ? (void *)L0.r32 : (void *)L0.r;
    const void *e = A->mixed ? (const void *)L0.e32b : (const void *)L0.e2;
    ACHK(amr::amr_restrict(a, s, v, v + A->off1, rc, A->mixed));
    ACHK(amr::launch_mg_start(cc));
    ACHK(amr::launch_vcycle(cc, A->mixed));
    ACHK(amr::amr_prolong(a, s, e, z, z + A->off1, A->mixed));
    for (int sweep = 0; sweep < 2; ++sweep)
        for (int color = 0; color < 2; ++color) ACHK(amr::amr_fine_gs(a, s, z, z + A->off1, v + A->off1, color));
    return 0;
}

// FGMRES(m) on L_c x = b (b centred) with the Arnoldi process on the device
// (kernels_amr.hip): batches of Arnoldi steps, each followed by a cycle edge,
// and one host wait per batch.  The coarse context's control block is the
// solve's: its V-cycle kernels honour the same mg_done / gm_stop flags.
// Returns the steps taken and ||r||inf / ||b||inf.
int composite_solve(amressif_amr *A, int *iters, double *res_rel)
{
    const amr::AmrGeom &a = A->a;
    Ctx &cc = A->coarse->c;
    hipStream_t s = A->stream;
    const amr::CgmBufs g{cc.ctl, A->gst, A->V, A->Z, A->x, A->b, A->vcur, A->zcur, A->part, A->vs, A->off1, A->m};
    amr::Level &L0 = cc.lev[0];
    void *rc = A->mixed ? (void *)L0.r32 : (void *)L0.r;
    const void *e = A->mixed ? (const void *)L0.e32b : (const void *)L0.e2;
    ACHK(amr::launch_fill(cc, A->x, (int64_t)A->vs, 0.0));
    ACHK(amr::amr_cgm_begin(s, cc.ctl, A->gst));
    ACHK(amr::amr_cgm_edge(a, s, g));  // r = b, ||b||inf, the first cycle's start
    int target = std::max(1, A->last_gm);
    for (;;) {
        for (int k = 0; k < target; ++k) {
            // z = M V_j: coarse V-cycle on the restricted V_j, prolonged, two
            // red-black Gauss-Seidel sweeps on the patch
            ACHK(amr::amr_restrict(a, s, A->vcur, A->vcur + A->off1, rc, A->mixed, cc.ctl));
            ACHK(amr::launch_vcycle(cc, A->mixed));
            ACHK(amr::amr_prolong(a, s, e, A->zcur, A->zcur + A->off1, A->mixed, cc.ctl));
            for (int sweep = 0; sweep < 2; ++sweep)
                for (int color = 0; color < 2; ++color)
                    ACHK(amr::amr_fine_gs(a, s, A->zcur, A->zcur + A->off1, A->vcur + A->off1, color, cc.ctl));
            ACHK(amr::amr_cgm_apply(a, s, g));
            ACHK(amr::amr_cgm_orthonormalise(a, s, g));
        }
        ACHK(amr::amr_cgm_edge(a, s, g));
        if (int rc2 = amr_poll(A->coarse)) return afail(A, rc2, amressif_last_error(A->coarse));
        if (cc.ctl_host->mg_done) break;
        target = 2;
    }
    A->last_gm = std::max(1, cc.ctl_host->mg_iters);
    if (cc.ctl_host->flags & AMRESSIF_FLAG_NONFINITE)
        return afail(A, AMRESSIF_ERR_NONFINITE, "composite Poisson solve diverged");
    *iters = cc.ctl_host->mg_iters;
    *res_rel = cc.ctl_host->b_norm > 0.0 ? cc.ctl_host->res_norm / cc.ctl_host->b_norm : 0.0;
    return 0;
}

// The same solve with the Arnoldi process on the host (one wait per
// dot-product pass); kept for the A/B comparison (AMRESSIF_AMR_HOST_GMRES=1).
int composite_solve_host(amressif_amr *A, int *iters, double *res_rel)
{
    const amr::AmrGeom &a = A->a;
    Ctx &cc = A->coarse->c;
    hipStream_t s = A->stream;
    const size_t o1 = A->off1, vs = A->vs;
    const int m = A->m;
    ACHK(hipMemsetAsync(A->x, 0, vs * sizeof(double), s));
    double nr[2];
    ACHK(amr::amr_apply(a, s, A->x, A->x + o1, A->b, A->b + o1, A->r, A->r + o1, A->part));
    if (int rc = reduce(A, 2, 1, nr)) return rc;  // column 0 max, column 1 sum
    const double bnorm = nr[0];
    const double thr = std::max(cc.in.poisson_tol_abs, cc.in.poisson_tol_rel * bnorm);
    std::vector<double> H((m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), h(m + 1), h2(m + 1);
    int total = 0;
    const int cap = std::max(1, cc.in.mg_max_iter);
    double rinf = nr[0], r2 = std::sqrt(nr[1]);
    while (rinf > thr && total < cap) {
        const double beta = r2;
        const double target = thr * beta / rinf;
        // V_0 = r / beta
        ACHK(hipMemcpyAsync(A->V, A->r, vs * sizeof(double), hipMemcpyDeviceToDevice, s));
        ACHK(amr::amr_combine(a, s, A->V, 1.0 / beta, A->V, vs, 0, A->coef, o1));
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < cap; ++j) {
            double *vj = A->V + (size_t)j * vs, *zj = A->Z + (size_t)j * vs, *w = A->V + (size_t)(j + 1) * vs;
            if (int rc = precondition(A, vj, zj)) return rc;
            ACHK(amr::amr_apply(a, s, zj, zj + o1, nullptr, nullptr, w, w + o1, A->part));
            // classical Gram-Schmidt, twice
            for (int pass = 0; pass < 2; ++pass) {
                ACHK(amr::amr_dots(a, s, w, A->V, vs, j + 1, o1, A->part));
                if (int rc = reduce(A, j + 1, 0, h2.data())) return rc;
                for (int t = 0; t <= j; ++t) {
                    H[t * m + j] = pass ? H[t * m + j] + h2[t] : h2[t];
                    h[t] = -h2[t];
                }
                ACHK(hipMemcpyAsync(A->coef, h.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
                ACHK(amr::amr_combine(a, s, w, 1.0, A->V, vs, j + 1, A->coef, o1));
            }
            double hn2;
            ACHK(amr::amr_dots(a, s, w, w, 0, 1, o1, A->part));
            if (int rc = reduce(A, 1, 0, &hn2)) return rc;
            const double hn = std::sqrt(hn2);
            if (hn > 0.0) ACHK(amr::amr_combine(a, s, w, 1.0 / hn, A->V, vs, 0, A->coef, o1));
            for (int t = 0; t < j; ++t) {
                const double p = H[t * m + j], q = H[(t + 1) * m + j];
                H[t * m + j] = cs[t] * p + sn[t] * q;
                H[(t + 1) * m + j] = -sn[t] * p + cs[t] * q;
            }
            const double p = H[j * m + j], rr = std::sqrt(p * p + hn * hn);
            cs[j] = rr > 0.0 ? p / rr : 1.0;
            sn[j] = rr > 0.0 ? hn / rr : 0.0;
            H[j * m + j] = rr;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++total;
            if (std::fabs(g[j + 1]) <= target || hn == 0.0) { ++j; break; }
        }
        for (int t = j - 1; t >= 0; --t) {
            double sum = g[t];
            for (int q = t + 1; q < j; ++q) sum -= H[t * m + q] * y[q];
            y[t] = sum / H[t * m + t];
        }
        ACHK(hipMemcpyAsync(A->coef, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
        ACHK(amr::amr_combine(a, s, A->x, 1.0, A->Z, vs, j, A->coef, o1));
        ACHK(amr::amr_apply(a, s, A->x, A->x + o1, A->b, A->b + o1, A->r, A->r + o1, A->part));
        if (int rc = reduce(A, 2, 1, nr)) return rc;
        rinf = nr[0];
        r2 = std::sqrt(nr[1]);
        if (!(rinf == rinf)) return afail(A, AMRESSIF_ERR_NONFINITE, "composite Poisson solve diverged");
    }
    *iters = total;
    *res_rel = bnorm > 0.0 ? rinf / bnorm : 0.0;
    return 0;
}

int center(amressif_amr *A, double *x)
{
    ACHK(amr::amr_wsum(A->a, A->stream, x, x + A->off1, A->part));
    ACHK(amr::launch_reduce_cols(A->stream, A->part, amr::amr_flat_blocks(A->a), 1, 0, A->red));
    ACHK(amr::amr_shift(A->a, A->stream, x, x + A->off1, A->red));  // total area is 1: the sum is the mean
    return 0;
}

int check_view(amressif_amr *A, const amr_array4 *v, int lo0, int lo1, int hi0, int hi1, int ncomp, const char *what)
{
    if (!v || !v->ptr) return afail(A, AMRESSIF_ERR_ARG, std::string(what) + ": null view");
    if (v->dtype != 0) return afail(A, AMRESSIF_ERR_ARG, std::string(what) + ": double views only");
    if (v->lo[0] > lo0 || v->lo[1] > lo1 || v->hi[0] < hi0 || v->hi[1] < hi1 || v->ncomp < ncomp)
        return afail(A, AMRESSIF_ERR_ARG, std::string(what) + ": view does not cover the patch");
    return 0;
}

}  // namespace

extern "C" {

int amressif_amr_create(const amressif_inputs *in, int patch_jlo, void *hip_stream, amressif_amr **out)
{
    if (!in || !out) return AMRESSIF_ERR_ARG;
    auto *A = new amressif_amr();
    *out = A;
    A->stream = (hipStream_t)hip_stream;
    const int n0 = in->n_cell;
    if (patch_jlo < 2 || patch_jlo >= n0 - 1) return afail(A, AMRESSIF_ERR_ARG, "patch_jlo must be in [2, n_cell - 1)");
    if (in->PSEUDO_TIMESTEPPING != 1)
        return afail(A, AMRESSIF_ERR_UNSUPPORTED, "the two-level step uses the dual-time predictor (PSEUDO_TIMESTEPPING = 1)");
    if (amr_gen_wanted(in)) return afail(A, AMRESSIF_ERR_UNSUPPORTED, "the two-level step is 2D with wall boundaries");
    // one dt for both levels, from the fine spacing (main.cpp:284-294 on the finest level)
    amressif_inputs ci = *in;
    const double h0 = (in->hi_phy_dim[0] - in->lo_phy_dim[0]) / n0, h1 = 0.5 * h0;
    if (std::fabs((in->hi_phy_dim[1] - in->lo_phy_dim[1]) / n0 - h0) > 1e-14 * h0)
        return afail(A, AMRESSIF_ERR_UNSUPPORTED, "the two-level step needs square cells");
    A->dt = in->fixed_dt != -1.0 ? in->fixed_dt : in->cfl / (2.0 * (2.0 / (h1 * h1)));
    ci.fixed_dt = A->dt;
    if (int rc = amressif_create(&ci, hip_stream, &A->coarse)) return afail(A, rc, amressif_last_error(A->coarse));
    Ctx &cc = A->coarse->c;
    A->mixed = in->precision == AMRESSIF_MIXED;
    if (const char *ev = std::getenv("AMRESSIF_AMR_HOST_GMRES")) A->host_gmres = std::atoi(ev) ? 1 : 0;
    amr::AmrGeom &a = A->a;
    a.n0 = n0;
    a.P0 = cc.g.P;
    a.jc0 = patch_jlo;
    a.nx = 2 * n0;
    a.ny = 2 * (n0 - patch_jlo);
    a.P1 = round_up(a.nx + 1, 64);
    a.Pg1 = round_up(a.nx + 4, 64);
    a.h0 = h0;
    a.h1 = h1;
    a.iren = 1.0 / in->ren;
    a.c15 = 1.5 / A->dt;
    a.bc_w = cc.g.bc_lo0;
    a.bc_e = cc.g.bc_hi0;
    a.bc_n = cc.g.bc_hi1;
    a.inflow0 = cc.g.inflow0;
    a.inflow1 = cc.g.inflow1;
    const size_t fsz = (size_t)a.P1 * (a.ny + 1), gsz = (size_t)a.Pg1 * (a.ny + 4);
#define AALLOC(p, cnt) ACHK(aalloc(A, p, cnt))
    for (int d = 0; d < 2; ++d) {
        AALLOC(&A->U1[d], fsz); AALLOC(&A->Up1[d], fsz); AALLOC(&A->C1[d], fsz);
        AALLOC(&A->F1[d], fsz); AALLOC(&A->F2[d], fsz);
    }
    AALLOC(&A->p1, fsz);
    AALLOC(&A->XU, gsz);
    AALLOC(&A->XV, gsz);
    if (A->mixed) {
        AALLOC(&A->DXU, gsz);
        AALLOC(&A->DXV, gsz);
    }
    AALLOC(&A->ctlF, 1);
    ACHK(hipHostMalloc((void **)&A->ctlF_host, sizeof(amr::Ctl), hipHostMallocDefault));
    const int nblk = amr::amr_flat_blocks(a);
    A->m = std::max(4, std::min(in->gmres_restart > 0 ? in->gmres_restart : 30, 60));
    AALLOC(&A->part, (size_t)nblk * (A->m + 2));
    AALLOC(&A->red, (size_t)A->m + 2);
    AALLOC(&A->coef, (size_t)A->m + 2);
    ACHK(hipHostMalloc((void **)&A->red_host, sizeof(double) * (A->m + 2), hipHostMallocDefault));
    A->off1 = round_up((long long)a.P0 * n0, 64);
    A->vs = round_up((long long)A->off1 + (long long)a.P1 * (a.ny + 1), 64);
    AALLOC(&A->V, A->vs * (size_t)(A->m + 1));
    AALLOC(&A->Z, A->vs * (size_t)A->m);
    AALLOC(&A->x, A->vs);
    AALLOC(&A->b, A->vs);
    AALLOC(&A->r, A->vs);
    AALLOC(&A->vcur, A->vs);
    AALLOC(&A->zcur, A->vs);
    AALLOC(&A->gst, 1);
#undef AALLOC
    ACHK(hipStreamSynchronize(A->stream));
    return AMRESSIF_OK;
}

int amressif_amr_destroy(amressif_amr *A)
{
    if (!A) return AMRESSIF_OK;
    if (A->stream) hipStreamSynchronize(A->stream);
    for (void *p : A->allocs) hipFree(p);
    if (A->ctlF_host) hipHostFree(A->ctlF_host);
    if (A->red_host) hipHostFree(A->red_host);
    if (A->coarse) amressif_destroy(A->coarse);
    delete A;
    return AMRESSIF_OK;
}

const char *amressif_amr_last_error(const amressif_amr *A) { return A ? A->err.c_str() : "null context"; }

int amressif_amr_dt(const amressif_amr *A, double *dt)
{
    if (!A || !dt) return AMRESSIF_ERR_ARG;
    *dt = A->dt;
    return AMRESSIF_OK;
}

// state in: coarse faces and p (userCtx comp 0 view), fine faces and p; the
// previous-step faces are set equal (BDF2 history 0 at the first step)
int amressif_amr_load_state(amressif_amr *A, const amr_array4 coarse_vel[2], const amr_array4 *coarse_ctx,
                            const amr_array4 fine_vel[2], const amr_array4 *fine_p)
{
    if (!A || !coarse_vel || !fine_vel) return afail(A, AMRESSIF_ERR_ARG, "null argument");
    Ctx &cc = A->coarse->c;
    const amr::AmrGeom &a = A->a;
    if (int rc = amressif_load_state(A->coarse, coarse_vel, coarse_vel, coarse_ctx))
        return afail(A, rc, amressif_last_error(A->coarse));
    (void)cc;
    if (int rc = check_view(A, &fine_vel[0], 0, 0, a.nx, a.ny - 1, 1, "fine u")) return rc;
    if (int rc = check_view(A, &fine_vel[1], 0, 0, a.nx - 1, a.ny, 1, "fine v")) return rc;
    for (int d = 0; d < 2; ++d) {
        ACHK(amr::launch_pack_rect(A->stream, fine_vel[d], A->U1[d], a.P1, d == 0 ? a.nx + 1 : a.nx,
                                   d == 0 ? a.ny : a.ny + 1));
        ACHK(hipMemcpyAsync(A->Up1[d], A->U1[d], sizeof(double) * (size_t)a.P1 * (a.ny + 1),
                            hipMemcpyDeviceToDevice, A->stream));
    }
    if (fine_p) {
        if (int rc = check_view(A, fine_p, 0, 0, a.nx - 1, a.ny - 1, 1, "fine p")) return rc;
        ACHK(amr::launch_pack_rect(A->stream, *fine_p, A->p1, a.P1, a.nx, a.ny));
    } else {
        ACHK(hipMemsetAsync(A->p1, 0, sizeof(double) * (size_t)a.P1 * (a.ny + 1), A->stream));
    }
    ACHK(amr::amr_average_down(a, A->stream, A->U1[0], A->U1[1], A->p1, cc.U[0], cc.U[1], cc.p));
    ACHK(amr::amr_average_down(a, A->stream, A->Up1[0], A->Up1[1], A->p1, cc.Uprev[0], cc.Uprev[1], cc.p));
    ACHK(hipStreamSynchronize(A->stream));
    return AMRESSIF_OK;
}

int amressif_amr_store_state(amressif_amr *A, amr_array4 coarse_vel[2], amr_array4 *coarse_ctx,
                             amr_array4 fine_vel[2], amr_array4 *fine_p)
{
    if (!A) return AMRESSIF_ERR_ARG;
    const amr::AmrGeom &a = A->a;
    if (coarse_vel || coarse_ctx) {
        if (int rc = amressif_store_state(A->coarse, coarse_vel, nullptr, coarse_ctx, nullptr))
            return afail(A, rc, amressif_last_error(A->coarse));
    }
    if (fine_vel) {
        if (int rc = check_view(A, &fine_vel[0], 0, 0, a.nx, a.ny - 1, 1, "fine u")) return rc;
        if (int rc = check_view(A, &fine_vel[1], 0, 0, a.nx - 1, a.ny, 1, "fine v")) return rc;
        for (int d = 0; d < 2; ++d)
            ACHK(amr::launch_unpack_rect(A->stream, A->U1[d], fine_vel[d], a.P1, d == 0 ? a.nx + 1 : a.nx,
                                         d == 0 ? a.ny : a.ny + 1));
    }
    if (fine_p) {
        if (int rc = check_view(A, fine_p, 0, 0, a.nx - 1, a.ny - 1, 1, "fine p")) return rc;
        ACHK(amr::launch_unpack_rect(A->stream, A->p1, *fine_p, a.P1, a.nx, a.ny));
    }
    ACHK(hipStreamSynchronize(A->stream));
    return AMRESSIF_OK;
}

int amressif_amr_advance(amressif_amr *A, int nsteps, amressif_amr_stats *stats)
{
    if (!A || nsteps < 0) return afail(A, AMRESSIF_ERR_ARG, "bad argument");
    Ctx &cc = A->coarse->c;
    const amr::AmrGeom &a = A->a;
    hipStream_t s = A->stream;
    if (int rc = amr_set_ctl(A->coarse)) return afail(A, rc, amressif_last_error(A->coarse));
    if (A->dstat_cap < nsteps) {
        ACHK(aalloc(A, &A->dstat, (size_t)2 * nsteps));
        A->dstat_cap = nsteps;
    }
    for (int step = 0; step < nsteps; ++step) {
        amressif_amr_stats st{};
        // 1. coarse predictor over the whole coarse domain
        if (int rc = amr_momentum_predictor(A->coarse)) return afail(A, rc, amressif_last_error(A->coarse));
        st.coarse_iters = cc.ctl_host->mom_count;
        st.flags = cc.ctl_host->flags;
        // 2. fine predictor with C/F data from the coarse u*
        int fi = 0, ff = 0;
        if (int rc = fine_predictor(A, &fi, &ff)) return rc;
        st.fine_iters = fi;
        st.flags |= ff;
        // 3. coarse faces under the patch from the fine predictor
        ACHK(amr::amr_average_down(a, s, A->U1[0], A->U1[1], A->p1, cc.U[0], cc.U[1], cc.p));
        // 4. composite projection (the divergence norms stay on the device
        // until the call's end)
        const int nblk = amr::amr_flat_blocks(a);
        ACHK(amr::amr_rhs(a, s, A->dt, cc.U[0], cc.U[1], A->U1[0], A->U1[1], A->b, A->b + A->off1, A->part));
        ACHK(amr::launch_reduce_cols(s, A->part, nblk, 1, 1, A->dstat + 2 * step));
        if (int rc = center(A, A->b)) return rc;
        int it = 0;
        double res = 0.0;
        if (int rc = A->host_gmres ? composite_solve_host(A, &it, &res) : composite_solve(A, &it, &res)) return rc;
        st.poisson_iters = it;
        st.poisson_res = res;
        if (int rc = center(A, A->x)) return rc;
        ACHK(amr::amr_project(a, s, A->dt, A->x, A->x + A->off1, cc.U[0], cc.U[1], A->U1[0], A->U1[1], cc.p, A->p1));
        // 5. average down and the composite divergence of the result
        ACHK(amr::amr_average_down(a, s, A->U1[0], A->U1[1], A->p1, cc.U[0], cc.U[1], cc.p));
        ACHK(amr::amr_rhs(a, s, A->dt, cc.U[0], cc.U[1], A->U1[0], A->U1[1], A->r, A->r + A->off1, A->part));
        ACHK(amr::launch_reduce_cols(s, A->part, nblk, 1, 1, A->dstat + 2 * step + 1));
        if (stats) stats[step] = st;
        if (st.flags & AMRESSIF_FLAG_NONFINITE)
            return afail(A, AMRESSIF_ERR_NONFINITE, "non-finite pseudo-iteration norm at step " + std::to_string(step));
    }
    if (stats && nsteps > 0) {
        std::vector<double> d((size_t)2 * nsteps);
        ACHK(hipMemcpyAsync(d.data(), A->dstat, sizeof(double) * d.size(), hipMemcpyDeviceToHost, s));
        ACHK(hipStreamSynchronize(s));
        for (int step = 0; step < nsteps; ++step) {
            stats[step].div_star = d[2 * step];
            stats[step].div_new = d[2 * step + 1];
        }
    }
    ACHK(hipStreamSynchronize(s));
    return AMRESSIF_OK;
}

}  // extern "C"
