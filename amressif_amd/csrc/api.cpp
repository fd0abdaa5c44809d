// C-ABI entry points (include/amressif.h) and the host-side step loop.
//
// The step loop mirrors the reference's time loop (Source/src1_rever/main.cpp:
// 875-1202) with the per-iteration host syncs replaced by device-side loop
// control: pseudo-iterations and V-cycles are enqueued in batches of guarded
// kernels (each returns at once when the control block says the loop is
// done), and the host polls the control block once per batch.  The batch size
// is the previous step's count, so a step usually costs one poll per loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "gen.hpp"
#include "internal.hpp"


// u* of consecutive recorded steps, 2 face arrays (u then v) per step
struct amressif_tape {
    amressif_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;  // the context's stream (the tape may outlive the context)
    int cap = 0, count = 0;
    size_t fsz = 0;  // doubles per face array
    double *store = nullptr;
    double *slot(int s) const { return store + (size_t)s * 2 * fsz; }
};

using amr::Ctx;

namespace amr {
template <class T>
hipError_t dev_alloc_public(Ctx &c, T **ptr, size_t count)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
    if (e != hipSuccess) return e;
    c.allocs.push_back(p);
    c.bytes += (int64_t)(count * sizeof(T) + 256);
    *ptr = (T *)p;
    return hipSuccess;
}

}  // namespace amr

namespace {

int fail(amressif_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->c.err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(ctx, AMRESSIF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// main.cpp:810-835 (BCRec mapping) and the cont2cart branches
// (utilities.cpp:87-269): returns 0, AMRESSIF_ERR_BC where the reference
// aborts, AMRESSIF_ERR_UNSUPPORTED for paths not built here.
int check_bc(const amressif_inputs &in, std::string &msg)
{
    for (int d = 0; d < 2; ++d) {
        const int lo = in.phy_bc_lo[d], hi = in.phy_bc_hi[d];
        if (!(lo == 111 || std::abs(lo) == 135 || std::abs(lo) == 165 || std::abs(lo) == 195)) {
            msg = "Invalid bc_lo";
            return AMRESSIF_ERR_BC;
        }
        // the reference's hi-side test reads phy_bc_lo for the 135 case (main.cpp:827)
        if (!(hi == 111 || std::abs(lo) == 135 || std::abs(hi) == 165 || std::abs(hi) == 195)) {
            msg = "Invalid bc_hi";
            return AMRESSIF_ERR_BC;
        }
        if (lo == 111 || hi == 111) {
            msg = "periodic boundaries (111) are not supported by this library";
            return AMRESSIF_ERR_UNSUPPORTED;
        }
    }
    const int W = in.phy_bc_lo[0], E = in.phy_bc_hi[0], S = in.phy_bc_lo[1], N = in.phy_bc_hi[1];
    if (W == 195 || W == -195) { msg = "west 195: Not yet implemented"; return AMRESSIF_ERR_BC; }
    if (E == -165 || E == 195 || E == -195) { msg = "east code: Not yet implemented"; return AMRESSIF_ERR_BC; }
    if (S == -165 || S == 195 || S == -195) { msg = "south code: Not yet implemented"; return AMRESSIF_ERR_BC; }
    if (N == -165 || N == 195 || N == -195) { msg = "north code: Not yet implemented"; return AMRESSIF_ERR_BC; }
    if (!(E == 135 || E == 165)) {
        msg = "east code leaves the reference's ghost cells unfilled (utilities.cpp:137-184)";
        return AMRESSIF_ERR_UNSUPPORTED;
    }
    return 0;
}

// dense pseudo-inverse of the Neumann 5-point Laplacian on the coarsest
// level: pinv(L) = J - (J - L)^{-1}, J = 11^T/m, via Cholesky of J - L.
bool neumann_pinv(int nx, int ny, double ax, double ay, std::vector<double> &out)
{
    const int m = nx * ny;
    std::vector<double> A((size_t)m * m, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int r = i + nx * j;
            double d = 0.0;
            if (i > 0) { A[(size_t)r * m + r - 1] -= ax; d += ax; }
            if (i < nx - 1) { A[(size_t)r * m + r + 1] -= ax; d += ax; }
            if (j > 0) { A[(size_t)r * m + r - nx] -= ay; d += ay; }
            if (j < ny - 1) { A[(size_t)r * m + r + nx] -= ay; d += ay; }
            A[(size_t)r * m + r] += d;
        }
    for (auto &a : A) a += 1.0 / m;
    // in-place Cholesky A = R^T R (upper R)
    for (int k = 0; k < m; ++k) {
        double d = A[(size_t)k * m + k];
        for (int s = 0; s < k; ++s) d -= A[(size_t)s * m + k] * A[(size_t)s * m + k];
        if (!(d > 0)) return false;
        d = std::sqrt(d);
        A[(size_t)k * m + k] = d;
        for (int j = k + 1; j < m; ++j) {
            double v = A[(size_t)k * m + j];
            for (int s = 0; s < k; ++s) v -= A[(size_t)s * m + k] * A[(size_t)s * m + j];
            A[(size_t)k * m + j] = v / d;
        }
    }
    out.assign((size_t)m * m, 0.0);
    std::vector<double> col(m);
    for (int c = 0; c < m; ++c) {
        std::fill(col.begin(), col.end(), 0.0);
        col[c] = 1.0;
        for (int i = 0; i < m; ++i) {
            double v = col[i];
            for (int s = 0; s < i; ++s) v -= A[(size_t)s * m + i] * col[s];
            col[i] = v / A[(size_t)i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double v = col[i];
            for (int s = i + 1; s < m; ++s) v -= A[(size_t)i * m + s] * col[s];
            col[i] = v / A[(size_t)i * m + i];
        }
        for (int i = 0; i < m; ++i) out[(size_t)i * m + c] = 1.0 / m - col[i];
    }
    return true;
}

// ---- Array4 validation: the view must cover the box the entry point uses
int check_view(amressif_ctx *ctx, const amr_array4 *a, int lo0, int lo1, int hi0, int hi1, int ncomp,
               const char *what)
{
    if (!a || !a->ptr) return fail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": null view");
    if (a->dtype != 0) return fail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": double views only");
    if (a->lo[0] > lo0 || a->lo[1] > lo1 || a->hi[0] < hi0 || a->hi[1] < hi1)
        return fail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": view does not cover the box");
    if (a->hi[2] < a->lo[2] || a->ncomp < ncomp)
        return fail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": bad component/z extent");
    return 0;
}

int check_face(amressif_ctx *ctx, const amr_array4 *a, int dir, const char *what)
{
    const int n = ctx->c.g.n;
    return dir == 0 ? check_view(ctx, a, 0, 0, n, n - 1, 1, what) : check_view(ctx, a, 0, 0, n - 1, n, 1, what);
}

int check_cell(amressif_ctx *ctx, const amr_array4 *a, int ncomp, const char *what)
{
    const int n = ctx->c.g.n;
    return check_view(ctx, a, 0, 0, n - 1, n - 1, ncomp, what);
}

int poll(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    HIPCHK(hipMemcpyAsync(c.ctl_host, c.ctl, sizeof(amr::Ctl), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(amr::kprof_collect(c));
    return 0;
}

int ensure_module_scratch(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    if (c.Mc[0]) return 0;
    const size_t fsz = (size_t)(c.g.n + 1) * c.g.P;
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 2; ++d) HIPCHK(amr::dev_alloc_public(c, &c.M[k][d], fsz));
    for (int d = 0; d < 2; ++d) HIPCHK(amr::dev_alloc_public(c, &c.Mc[d], fsz));
    HIPCHK(amr::dev_alloc_public(c, &c.dscalar, 8));
    return 0;
}

}  // namespace

int amr_module_scratch(amressif_ctx *ctx) { return ensure_module_scratch(ctx); }

namespace {

// ---- FGMRES(m) with a V-cycle preconditioner on c.b (already centred):
// cycles of guarded Arnoldi steps, polled in batches; the true residual is
// tested at every cycle start
int gmres_loop(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    const int mixed = c.in.precision == AMRESSIF_MIXED;
    HIPCHK(amr::launch_mg_start(c));
    HIPCHK(amr::launch_gm_reset(c));
    HIPCHK(amr::launch_fill(c, c.phiv[0], (int64_t)(c.g.n + 1) * c.g.P, 0.0));
    double *x = c.phiv[0];
    const int cap = std::max(1, c.in.mg_max_iter);
    for (int cycle = 0; cycle <= cap; ++cycle) {
        HIPCHK(amr::launch_gm_cycle_start(c, x, c.b, mixed));
        int j = 0, target = std::max(1, std::min(c.gm_m, c.last_V));
        for (;;) {
            while (j < target && j < c.gm_m) {
                HIPCHK(amr::launch_gm_iteration(c, mixed));
                ++j;
            }
            if (int rc = poll(ctx)) return rc;
            if (c.ctl_host->mg_done || c.ctl_host->gm_stop || j >= c.gm_m) break;
            target = j + 2;
        }
        if (c.ctl_host->mg_done) break;
        HIPCHK(amr::launch_gm_cycle_end(c, x));
    }
    c.last_V = std::max(1, std::min(c.gm_m, c.ctl_host->mg_iters));
    c.phi = x;
    return 0;
}

// ---- the Poisson solve on c.phi / c.b, shared by the step and the module
// entry point
int mg_loop(amressif_ctx *ctx, bool warm, bool center, const double *us, const double *vs);
// warm: start from the linear extrapolation 2 phi^{n-1} - phi^{n-2} of the
// previous steps' solutions (the previous solution alone right after a state
// load); the stopping test stays relative to ||b||inf.  cold: from zero, as
// poisson.cpp:63.  center: b still holds its mean (ctl->b_mean), removed here.
int poisson_loop(amressif_ctx *ctx, bool warm = false, bool center = false)
{
    if (ctx->c.in.poisson_solver == 1) {
        if (center) HIPCHK(amr::launch_center_b(ctx->c, ctx->c.b));
        return gmres_loop(ctx);
    }
    return mg_loop(ctx, warm, center, nullptr, nullptr);
}

// the step's projection solve: right-hand side (1.5/dt) D(u*) of the
// predictor faces us, vs (fn_rhs.cpp:71-125), centred, then the solve; the
// multigrid iteration forms the right-hand side inside its first pass
int poisson_loop_step(amressif_ctx *ctx, bool warm, const double *us, const double *vs)
{
    Ctx &c = ctx->c;
    if (c.in.poisson_solver == 1) {
        int nb = 0;
        HIPCHK(amr::launch_poisson_rhs(c, us, vs, c.b, &nb));
        HIPCHK(amr::launch_fin_prhs(c, nb));
        return poisson_loop(ctx, warm, true);
    }
    return mg_loop(ctx, warm, false, us, vs);
}

// warm-start buffers of one solve (see mg_prepare)
struct WarmStart {
    int mode = 0;  // k_poisson_start: 0 from zero, 1 phi^{n-1}, 2 extrapolation
    const double *phi1 = nullptr, *phi2 = nullptr;
};

WarmStart mg_prepare(amressif_ctx *ctx, bool warm, int *rc)
{
    Ctx &c = ctx->c;
    WarmStart w;
    *rc = 0;
    if (!warm) return w;
    if (!c.phih) {
        const hipError_t e = amr::dev_alloc_public(c, &c.phih, (size_t)(c.g.n + 1) * c.g.P);
        if (e != hipSuccess) {
            *rc = fail(ctx, AMRESSIF_ERR_HIP, hipGetErrorString(e));
            return w;
        }
    }
    // three buffers, rotated instead of copied: X = phi^{n-1} (the last
    // solution, one of the solve's pair), Y = phi^{n-2}, F = the other of
    // the pair.  The guess goes into F and the solve ping-pongs F <-> Y;
    // X becomes phi^{n-2} of the next step.
    double *X = c.phi, *Y = c.phih, *F = c.phi == c.phiv[0] ? c.phiv[1] : c.phiv[0];
    w.mode = c.phih_valid ? 2 : 1;
    w.phi1 = X;
    w.phi2 = Y;
    c.phiv[0] = F;
    c.phiv[1] = Y;
    c.phih = X;
    c.phih_valid = 1;
    return w;
}

// The multigrid iteration from the prepared start.  require_mom: enqueued
// before the host saw the pseudo-iteration loop end; *skipped reports that it
// had not (every kernel of the phase then did nothing) and the caller runs it
// again once the loop has ended.
// Work enqueued behind each batch of V-cycles before the host waits (the
// step's projection, gated on the device by the solve's end), and the undo of
// its host-side effects when the wait shows the solve has not ended.
struct SpecTail {
    std::function<int()> enqueue;
    std::function<void()> undo;
};

int mg_run(amressif_ctx *ctx, const WarmStart &w, bool center, const double *us, const double *vs,
           bool require_mom, bool *skipped, const SpecTail *tail = nullptr, bool *tail_done = nullptr)
{
    if (tail_done) *tail_done = false;
    Ctx &c = ctx->c;
    const int mixed = c.in.precision == AMRESSIF_MIXED;
    int nb = 0;
    if (!(require_mom && c.mp_spec)) HIPCHK(amr::launch_mg_start(c, require_mom ? 1 : 0, us, vs));
    {
        // per cell: u*, v* in (or b), the warm-start iterates in, b (when formed or
        // centred), phi0 and the residual (V-cycle precision) out
        const double t = mixed ? 4.0 : 8.0, warm = w.mode == 2 ? 16.0 : w.mode == 1 ? 8.0 : 0.0;
        const double per = (us ? 16.0 : 8.0) + warm + (us || center ? 8.0 : 0.0) + 8.0 + t;
        HIPCHK(amr::kprof_begin(c, amr::KP_PSTART, amr::KL_PHASE, 0, per * (double)c.g.n * c.g.n));
    }
    HIPCHK(amr::launch_poisson_start(c, w.mode, center ? 1 : 0, w.phi1, w.phi2, c.phiv[0], mixed, us, vs));
    HIPCHK(amr::kprof_end(c));
    int launched = 0, target = std::max(1, c.last_V);
    if (c.mg_batch > 0) target = std::min(target, c.mg_batch);
    const int cap = std::max(1, c.in.mg_max_iter);
    for (;;) {
        while (launched < target && launched < cap) {
            const int par = launched & 1;  // executed cycles always follow this order
            if (c.kp) c.kp->mg_idx = launched;
            HIPCHK(amr::launch_vcycle_step(c, mixed, par, c.phiv[par], c.phiv[par ^ 1], &nb));
            HIPCHK(amr::launch_fin_resid(c, nb));
            ++launched;
        }
        if (tail) {
            // the control block is copied right behind the batch; the host waits
            // for that copy only, so it decides while the device runs the tail
            HIPCHK(hipMemcpyAsync(c.ctl_host, c.ctl, sizeof(amr::Ctl), hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipEventRecord(c.poll_ev, c.stream));
            if (int rc = tail->enqueue()) return rc;
            HIPCHK(hipEventSynchronize(c.poll_ev));
            HIPCHK(amr::kprof_collect(c));
        } else if (int rc = poll(ctx)) {
            return rc;
        }
        if (require_mom && c.ctl_host->mg_skip) {
            *skipped = true;
            if (tail) tail->undo();
            amr::kprof_close(c, amr::KL_MG, 0);  // every kernel of the phase exited at once
            amr::kprof_close(c, amr::KL_PHASE, 0);
            return 0;
        }
        if (c.ctl_host->mg_done) {
            if (tail_done) *tail_done = tail != nullptr;
            break;
        }
        if (tail) tail->undo();
        if (launched >= cap) break;
        target = launched + 2;
    }
    amr::kprof_close(c, amr::KL_MG, c.ctl_host->mg_iters);
    amr::kprof_close(c, amr::KL_PHASE, 1);
    c.last_V = std::max(1, c.ctl_host->mg_iters);
    c.phi = c.phiv[c.ctl_host->mg_iters & 1];
    return 0;
}

int mg_loop(amressif_ctx *ctx, bool warm, bool center, const double *us, const double *vs)
{
    int rc = 0;
    const WarmStart w = mg_prepare(ctx, warm, &rc);
    if (rc) return rc;
    return mg_run(ctx, w, center, us, vs, false, nullptr);
}

// One pseudo-iteration.  FP64: the RK4 dual-time sweep, four guarded stage
// launches; the stencil input ping-pongs U -> S1 -> S2 -> S1 and the last stage
// updates U (the pseudo base u^k) in place.  Mixed: the FP64 residual r of U,
// then the FP32 tangent-linear correction of U in place: by default the
// two-term Neumann series of the inverse Jacobian (iterative refinement of the
// implicit BDF2 system), or the RK4 recursion to first order in the Jacobian.
int advance_impl(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats, amressif_tape *tape);
int launch_pseudo_iteration(amressif_ctx *ctx, int it, int *nb);
}  // namespace

// BDF2 history and the RK4 dual-time pseudo-iterations of one step
// (main.cpp:879-884, 898-1040): c.U holds the converged predictor u* after it,
// the control block copy c.ctl_host its count and flags
// reset_done: the previous step's k_fin_project has already cleared the
// pseudo-iteration control
int amr_momentum_begin(amressif_ctx *ctx, bool bdf_ready, int *launched, bool reset_done)
{
    Ctx &c = ctx->c;
    if (!reset_done) HIPCHK(amr::launch_ctl_reset_step(c));
    c.mp_count = 0;
    if (bdf_ready) {
        // the previous projection formed C and copied u^n to S2 (with the
        // mixed iteration in one pass it copied nothing: U and Uprev already
        // hold u^n and u^{n-1})
        if (!c.mp_on) {
            std::swap(c.Uprev[0], c.S2[0]);
            std::swap(c.Uprev[1], c.S2[1]);
        }
    } else {
        HIPCHK(amr::launch_bdf2(c));
        c.c_gradp = 0;
    }
    // p^n is the same in every stage of the step: its gradient term sits in C
    // (put there by the previous projection, or here) and the stage kernels
    // read no p
    if (!c.c_gradp) {
        HIPCHK(amr::launch_fold_gradp(c));
        c.c_gradp = 1;
    }
    int nb = 0;
    if (c.in.PSEUDO_TIMESTEPPING == 0) {
        // main.cpp:900-926, momentum.cpp:96-191: one explicit update from u^n,
        // u* = u^n + (dt/1.5)(R(u^n) - G p + (u^n - u^{n-1})/(2 dt)); in the form of
        // a stage with rk = dt/1.5 whose stencil input and base are both u^n
        amr::StageLaunch a{};
        a.us = a.ub = c.U[0];
        a.vs = a.vb = c.U[1];
        a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
        a.uo = c.S1[0]; a.vo = c.S1[1];
        a.rk = c.g.dt / 1.5;
        a.mode = 1;
        a.foldp = 1;
        a.expl = 1;
        HIPCHK(amr::launch_stage(c, a, &nb));
        HIPCHK(amr::launch_fin_explicit(c, nb));
        std::swap(c.U[0], c.S1[0]);
        std::swap(c.U[1], c.S1[1]);
        *launched = 1;
        return 0;
    }
    const int target = std::min(std::max(1, c.last_P), c.in.IterNum + 1);
    for (*launched = 0; *launched < target; ++*launched) {
        if (int rc = launch_pseudo_iteration(ctx, *launched, &nb)) return rc;
        // with the one-pass mixed iteration, the batch's last finalizer also
        // starts the Poisson phase enqueued behind it (mg_run then skips k_mg_start)
        HIPCHK(amr::launch_fin_pseudo(c, nb, c.mp_spec && *launched == target - 1 ? 1 : 0));
    }
    return 0;
}

int amr_momentum_finish(amressif_ctx *ctx, int launched)
{
    Ctx &c = ctx->c;
    const int cap = c.in.IterNum + 1;
    int target = launched, nb = 0;
    for (;;) {
        while (launched < target && launched < cap) {
            if (int rc = launch_pseudo_iteration(ctx, launched, &nb)) return rc;
            HIPCHK(amr::launch_fin_pseudo(c, nb));
            ++launched;
        }
        if (int rc = poll(ctx)) return rc;
        if (c.ctl_host->mom_done || launched >= cap) break;
        target = launched + 2;
    }
    c.last_P = std::max(1, c.ctl_host->mom_count);
    amr::kprof_close(c, amr::KL_MOM, c.ctl_host->mom_count);
    return 0;
}

int amr_momentum_predictor(amressif_ctx *ctx, bool bdf_ready, bool reset_done)
{
    int launched = 0;
    if (int rc = amr_momentum_begin(ctx, bdf_ready, &launched, reset_done)) return rc;
    return amr_momentum_finish(ctx, launched);
}

int amr_poll(amressif_ctx *ctx) { return poll(ctx); }
int amr_set_ctl(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                               c.in.mg_max_iter));
    return 0;
}

namespace {

// the output of the step's k-th k_mixed_pair launch: S1 for even k, S2 for odd
double *const *mp_out(const Ctx &c, int k) { return k % 2 == 0 ? c.S1 : c.S2; }

int launch_pseudo_iteration(amressif_ctx *ctx, int it, int *nb)
{
    Ctx &c = ctx->c;
    if (c.in.precision == AMRESSIF_MIXED) {
        amr::MixedStageLaunch a{};
        a.u64 = c.U[0]; a.v64 = c.U[1]; a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
        a.foldp = 1;  // C holds -G p (amr_momentum_begin)
        a.uk = c.uk32[0]; a.vk = c.uk32[1]; a.ru = c.r32[0]; a.rv = c.r32[1];
        const double c15 = 1.5 / c.g.dt;
        if (c.in.mixed_momentum == 1) {
            // Newton correction with J^{-1} ~ -(1/c15)(I + R'/c15): u += r/c15 + R'(r)/c15^2
            a.cr = 1.0 / c15;
            a.cj = 1.0 / (c15 * c15);
        } else {
            // RK4 dual-time recursion to first order in J = R' - c15
            a.cj = c.rk[2] * c.rk[3];
            a.cr = c.rk[3] - a.cj * c15;
        }
        const double cells = (double)c.g.n * c.g.n;
        if (c.mp_on) {
            // one pass: u^k from U (the first launch, which leaves u^n in U) or
            // the previous launch's output, u^{k+1} into S1 or S2
            amr::MixedPairLaunch m{};
            double *const *in = it == 0 ? c.U : mp_out(c, it - 1);
            double *const *out = mp_out(c, it);
            m.u64 = in[0]; m.v64 = in[1]; m.cu = c.C[0]; m.cv = c.C[1];
            m.uo = out[0]; m.vo = out[1];
            m.cr = a.cr; m.cj = a.cj;
            // 48 B per cell: u^k, v^k, c_u, c_v in; u^{k+1}, v^{k+1} out
            HIPCHK(amr::kprof_begin(c, amr::KP_MPAIR, amr::KL_MOM, it, 48.0 * cells));
            HIPCHK(amr::launch_mixed_pair(c, m, nb));
            HIPCHK(amr::kprof_end(c));
            c.mp_count = it + 1;
            return 0;
        }
        a.stage = 0;
        // algorithmic bytes per cell: residual 48 (u^k, v^k, c_u, c_v in; FP32 base and
        // residual out), tangent 48 (FP32 base and residual in; u^k, v^k in and out)
        HIPCHK(amr::kprof_begin(c, amr::KP_RESID, amr::KL_MOM, it, 48.0 * cells));
        HIPCHK(amr::launch_stage_mixed(c, a, nb));
        HIPCHK(amr::kprof_end(c));
        a.stage = 1;
        HIPCHK(amr::kprof_begin(c, amr::KP_TANGENT, amr::KL_MOM, it, 48.0 * cells));
        HIPCHK(amr::launch_stage_mixed(c, a, nb));
        HIPCHK(amr::kprof_end(c));
        return 0;
    }
    amr::StageLaunch a{};
    a.ub = c.U[0]; a.vb = c.U[1]; a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
    a.foldp = 1;  // C holds -G p (amr_momentum_begin)
    if (c.stage_pair && amr::stage_pair_fits(c)) {
        // stages 1-2 from u^k into S2, stages 3-4 from S2 into S1, then S1 becomes
        // u^k (a launch skipped after convergence copies u^k to S1)
        const double cells = (double)c.g.n * c.g.n;
        a.us = c.U[0]; a.vs = c.U[1]; a.uo = c.S2[0]; a.vo = c.S2[1];
        a.rk = c.rk[0]; a.rk2 = c.rk[1]; a.pair = 1;
        // 64 B per cell: u*, v*, u^k, v^k, c_u, c_v in; u*, v* out
        HIPCHK(amr::kprof_begin(c, amr::KP_PAIR, amr::KL_MOM, it, 64.0 * cells));
        HIPCHK(amr::launch_stage_pair(c, a, nb));
        HIPCHK(amr::kprof_end(c));
        a.us = c.S2[0]; a.vs = c.S2[1]; a.uo = c.S1[0]; a.vo = c.S1[1];
        a.rk = c.rk[2]; a.rk2 = c.rk[3]; a.pair = 2;
        HIPCHK(amr::kprof_begin(c, amr::KP_PAIR_LAST, amr::KL_MOM, it, 64.0 * cells));
        HIPCHK(amr::launch_stage_pair(c, a, nb));
        HIPCHK(amr::kprof_end(c));
        std::swap(c.U[0], c.S1[0]);
        std::swap(c.U[1], c.S1[1]);
        return 0;
    }
    const double *in_u[4] = {c.U[0], c.S1[0], c.S2[0], c.S1[0]};
    const double *in_v[4] = {c.U[1], c.S1[1], c.S2[1], c.S1[1]};
    double *out_u[4] = {c.S1[0], c.S2[0], c.S1[0], c.U[0]};
    double *out_v[4] = {c.S1[1], c.S2[1], c.S1[1], c.U[1]};
    for (int sub = 0; sub < 4; ++sub) {
        a.us = in_u[sub]; a.vs = in_v[sub]; a.uo = out_u[sub]; a.vo = out_v[sub];
        a.rk = c.rk[sub];
        a.mode = sub == 3 ? 1 : 0;
        // 64 B per cell: u*, v*, u^k, v^k, c_u, c_v in; u*, v* out
        HIPCHK(amr::kprof_begin(c, sub == 3 ? amr::KP_STAGE_LAST : amr::KP_STAGE, amr::KL_MOM, it,
                                64.0 * (double)c.g.n * c.g.n));
        HIPCHK(amr::launch_stage(c, a, nb));
        HIPCHK(amr::kprof_end(c));
    }
    return 0;
}

}  // namespace

namespace amr {

hipError_t kprof_begin(const Ctx &c, int id, int loop, int idx, double bytes)
{
    KProf *k = c.kp;
    if (!k || !k->on) return hipSuccess;
    hipEvent_t e[2];
    for (int q = 0; q < 2; ++q) {
        if (k->pool.empty()) {
            // timing-only events: no system-scope fence (no cache writeback
            // between the bracketed kernels); they are read after a stream sync
            hipError_t err = hipEventCreateWithFlags(&e[q], hipEventDisableSystemFence);
            if (err != hipSuccess) return err;
        } else {
            e[q] = k->pool.back();
            k->pool.pop_back();
        }
    }
    k->recs.push_back({e[0], e[1], id, loop, idx, loop == KL_ONCE ? 1 : 0, bytes});
    return hipEventRecord(e[0], c.stream);
}

hipError_t kprof_end(const Ctx &c)
{
    KProf *k = c.kp;
    if (!k || !k->on) return hipSuccess;
    return hipEventRecord(k->recs.back().e1, c.stream);
}

void kprof_close(const Ctx &c, int loop, int count)
{
    KProf *k = c.kp;
    if (!k) return;
    for (KProf::Rec &r : k->recs)
        if (r.state == 0 && r.loop == loop) r.state = r.idx < count ? 1 : 2;
}

// records whose loop has closed, once their end event has completed
hipError_t kprof_collect(const Ctx &c)
{
    KProf *k = c.kp;
    if (!k || k->recs.empty()) return hipSuccess;
    std::vector<KProf::Rec> keep;
    for (const KProf::Rec &r : k->recs) {
        if (r.state == 0 || hipEventQuery(r.e1) != hipSuccess) {
            (void)hipGetLastError();  // hipErrorNotReady is not an error here
            keep.push_back(r);
            continue;
        }
        if (r.state == 1) {
            float ms = 0.f;
            hipError_t err = hipEventElapsedTime(&ms, r.e0, r.e1);
            if (err != hipSuccess) return err;
            k->sum_ms[r.id] += ms;
            k->sum_bytes[r.id] += r.bytes;
            k->cnt[r.id] += 1;
        }
        k->pool.push_back(r.e0);
        k->pool.push_back(r.e1);
    }
    k->recs.swap(keep);
    return hipSuccess;
}

}  // namespace amr

extern "C" {

void amressif_inputs_default(amressif_inputs *in)
{
    std::memset(in, 0, sizeof(*in));
    in->n_cell = -1;
    in->max_grid_size[0] = in->max_grid_size[1] = in->max_grid_size[2] = -1;
    in->lo_phy_dim[0] = in->lo_phy_dim[1] = in->lo_phy_dim[2] = 0.0;
    in->hi_phy_dim[0] = in->hi_phy_dim[1] = in->hi_phy_dim[2] = 1.0;
    in->IterNum = -1;
    in->nsteps = 1;
    in->cfl = 0.9;
    in->fixed_dt = -1.0;
    in->ren = -1.0;
    in->vis = -1.0;
    in->target_resolution = -1;
    in->momentum_tolerance = 1.e-10;
    in->PSEUDO_TIMESTEPPING = 1;
    in->plot_int = -1;
    in->chk_int = -1;
    in->txt_int = 0;
    in->chk_out = 0;
    in->poisson_tol_rel = 1.0e-10;
    in->poisson_tol_abs = 0.0;
    in->mg_max_iter = 2000;
    in->mg_bottom_max = 4;
    in->precision = AMRESSIF_FP64;
    in->poisson_solver = 0;
    in->gmres_restart = 30;
    in->mixed_momentum = 1;
    in->poisson_warm_start = -1;
    in->spacedim = 2;
}

double amressif_dt(const amressif_inputs *in)
{
    double coeff = 0.0;
    for (int d = 0; d < (in->spacedim == 3 ? 3 : 2); ++d) {
        const double h = (in->hi_phy_dim[d] - in->lo_phy_dim[d]) / in->n_cell;
        coeff += 1. / (h * h);
    }
    double dt = in->cfl / (2.0 * coeff);
    if (in->fixed_dt != -1.0) dt = in->fixed_dt;
    return dt;
}

int amressif_parse_inputs(const char *path, int argc, const char *const *argv, amressif_inputs *out, char *err,
                          int errlen)
{
    auto seterr = [&](const std::string &m) {
        if (err && errlen > 0) {
            std::snprintf(err, (size_t)errlen, "%s", m.c_str());
        }
        return AMRESSIF_ERR_PARSE;
    };
    amressif_inputs_default(out);
    std::map<std::string, std::vector<std::string>> kv;
    auto add_line = [&](std::string line) -> bool {
        const auto h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        const auto eq = line.find('=');
        if (eq == std::string::npos) {
            for (char ch : line)
                if (!std::isspace((unsigned char)ch)) return false;
            return true;
        }
        std::string key = line.substr(0, eq);
        key.erase(0, key.find_first_not_of(" \t\r"));
        key.erase(key.find_last_not_of(" \t\r") + 1);
        std::istringstream vs(line.substr(eq + 1));
        std::vector<std::string> vals;
        std::string tok;
        while (vs >> tok) vals.push_back(tok);
        if (key.empty()) return false;
        kv[key] = vals;
        return true;
    };
    if (path && path[0]) {
        std::ifstream f(path);
        if (!f) return seterr(std::string("cannot open inputs file ") + path);
        std::string line;
        int ln = 0;
        while (std::getline(f, line)) {
            ++ln;
            if (!add_line(line)) return seterr("malformed line " + std::to_string(ln) + ": " + line);
        }
    }
    for (int a = 0; a < argc; ++a)
        if (!add_line(argv[a])) return seterr(std::string("malformed override: ") + argv[a]);

    auto num = [&](const std::string &key, const std::string &s, double &v) -> bool {
        char *end = nullptr;
        v = std::strtod(s.c_str(), &end);
        return end && *end == '\0' && !s.empty();
    };
    std::string bad;
    auto get_int = [&](const char *key, int *dst, int count, bool required) -> bool {
        auto it = kv.find(key);
        if (it == kv.end()) {
            if (required) bad = std::string("missing required key ") + key;
            return !required;
        }
        if ((int)it->second.size() < count) { bad = std::string("too few values for ") + key; return false; }
        for (int k = 0; k < count; ++k) {
            double v;
            if (!num(key, it->second[k], v) || v != std::floor(v)) { bad = std::string("bad integer for ") + key; return false; }
            dst[k] = (int)v;
        }
        return true;
    };
    auto get_real = [&](const char *key, double *dst, int count, bool required) -> bool {
        auto it = kv.find(key);
        if (it == kv.end()) {
            if (required) bad = std::string("missing required key ") + key;
            return !required;
        }
        if ((int)it->second.size() < count) { bad = std::string("too few values for ") + key; return false; }
        for (int k = 0; k < count; ++k)
            if (!num(key, it->second[k], dst[k])) { bad = std::string("bad number for ") + key; return false; }
        return true;
    };
    // the array keys take AMREX_SPACEDIM values (extra values ignored)
    if (!get_int("amressif.spacedim", &out->spacedim, 1, false)) return seterr(bad);
    if (out->spacedim != 2 && out->spacedim != 3) return seterr("amressif.spacedim must be 2 or 3");
    const int nd = out->spacedim;
    bool ok = get_int("n_cell", &out->n_cell, 1, true) && get_int("max_grid_size_x", &out->max_grid_size[0], 1, true) &&
              get_int("max_grid_size_y", &out->max_grid_size[1], 1, true) &&
              get_int("max_grid_size_z", &out->max_grid_size[2], 1, true) &&
              get_real("lo_phy_dim", out->lo_phy_dim, nd, false) && get_real("hi_phy_dim", out->hi_phy_dim, nd, false) &&
              get_int("IterNum", &out->IterNum, 1, true) && get_int("nsteps", &out->nsteps, 1, false) &&
              get_real("cfl", &out->cfl, 1, false) && get_real("fixed_dt", &out->fixed_dt, 1, false) &&
              get_real("ren", &out->ren, 1, true) && get_real("vis", &out->vis, 1, true) &&
              get_int("phy_bc_lo", out->phy_bc_lo, nd, false) && get_int("phy_bc_hi", out->phy_bc_hi, nd, false) &&
              get_real("inflow_waveform", out->inflow_waveform, nd, false) &&
              get_int("target_resolution", &out->target_resolution, 1, false) &&
              get_real("momentum_tolerance", &out->momentum_tolerance, 1, false) &&
              get_int("PSEUDO_TIMESTEPPING", &out->PSEUDO_TIMESTEPPING, 1, false) &&
              get_int("plot_int", &out->plot_int, 1, false) && get_int("chk_int", &out->chk_int, 1, false) &&
              get_int("txt_int", &out->txt_int, 1, false) && get_int("chk_out", &out->chk_out, 1, false) &&
              get_real("amressif.poisson_tol_rel", &out->poisson_tol_rel, 1, false) &&
              get_real("amressif.poisson_tol_abs", &out->poisson_tol_abs, 1, false) &&
              get_int("amressif.mg_max_iter", &out->mg_max_iter, 1, false) &&
              get_int("amressif.mg_bottom_max", &out->mg_bottom_max, 1, false) &&
              get_int("amressif.precision", &out->precision, 1, false) &&
              get_int("amressif.poisson_solver", &out->poisson_solver, 1, false) &&
              get_int("amressif.gmres_restart", &out->gmres_restart, 1, false) &&
              get_int("amressif.mixed_momentum", &out->mixed_momentum, 1, false) &&
              get_int("amressif.poisson_warm_start", &out->poisson_warm_start, 1, false);
    if (!ok) return seterr(bad);
    return AMRESSIF_OK;
}

int amressif_create(const amressif_inputs *in, void *hip_stream, amressif_ctx **out)
{
    if (!in || !out) return AMRESSIF_ERR_ARG;
    *out = nullptr;
    auto *ctx = new amressif_ctx();
    Ctx &c = ctx->c;
    c.in = *in;
    c.stream = (hipStream_t)hip_stream;
    // tuning overrides of the stage kernels' decomposition (launch shape only)
    if (const char *e = std::getenv("AMRESSIF_STAGE_ROWS")) c.stage_rows = std::max(4, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_TANGENT_ROWS"))
        c.tangent_rows = std::max(4, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_STAGE_ORDER")) c.stage_order = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_STAGE_ALT")) c.stage_alt = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_STAGE_FAST")) c.stage_fast = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_STAGE_PAIR")) c.stage_pair = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_PAIR_ROWS")) c.pair_rows = std::max(4, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_MIXED_PAIR")) c.mixed_pair = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MPAIR_ROWS")) c.mpair_rows = std::max(4, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_EDGE_FIRST")) c.stage_edge_first = std::atoi(e) ? 1 : 0;
    // and of the multigrid smoothers (kernel choice and rows per wave: 4, 8, 16 or 32)
    if (const char *e = std::getenv("AMRESSIF_MG_SMOOTHER")) c.mg_smoother = std::max(0, std::min(2, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_TAIL_DCT")) c.tail_dct = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MG_LATE")) c.mg_late = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MG_POST3")) c.mg_post3 = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MG_PRE3")) c.mg_pre3 = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MG_UP64")) c.mg_up64_single = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_MG_BATCH")) c.mg_batch = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("AMRESSIF_DCT_MFMA64")) c.dct_mfma64 = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_START_ROWS")) c.start_rows = std::max(0, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_PROJECT_ROWS"))
        c.project_rows = std::max(0, std::min(256, std::atoi(e)));
    if (const char *e = std::getenv("AMRESSIF_TAIL_EXT")) c.tail_ext = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_TAIL_WAVES")) c.tail_waves = std::atoi(e) == 4 ? 4 : 16;
    auto rows_ok = [](int r) { return r == 4 || r == 8 || r == 16 || r == 32; };
    for (int t = 0; t < 2; ++t)
        if (const char *e = std::getenv(t ? "AMRESSIF_MG_ROWS32" : "AMRESSIF_MG_ROWS64")) {
            int v[3];
            if (std::sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3 && rows_ok(v[0]) && rows_ok(v[1]) && rows_ok(v[2]))
                for (int m = 0; m < 3; ++m) c.mg_rows[t][m] = v[m];
        }
    if (const char *e = std::getenv("AMRESSIF_MG_ROWS_COARSE"))
        if (rows_ok(std::atoi(e))) c.mg_rows_coarse = std::atoi(e);
    auto bail = [&](int code, const std::string &m) {
        // keep the context so the caller can read the message; it owns nothing yet
        ctx->c.err = m;
        *out = ctx;
        return code;
    };
    const int n = in->n_cell;
    if (n < 4 || n > 16384) return bail(AMRESSIF_ERR_ARG, "n_cell must be in [4, 16384]");
    if (in->ren <= 0) return bail(AMRESSIF_ERR_ARG, "ren must be > 0");
    if (in->IterNum < 0) return bail(AMRESSIF_ERR_ARG, "IterNum must be >= 0");
    if (in->PSEUDO_TIMESTEPPING != 0 && in->PSEUDO_TIMESTEPPING != 1)
        return bail(AMRESSIF_ERR_ARG, "PSEUDO_TIMESTEPPING must be 0 (explicit) or 1 (dual-time RK4)");
    if (in->precision != AMRESSIF_FP64 && in->precision != AMRESSIF_MIXED)
        return bail(AMRESSIF_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    if (in->mg_bottom_max < 2 || in->mg_bottom_max > 32) return bail(AMRESSIF_ERR_ARG, "mg_bottom_max must be in [2, 32]");
    if (in->poisson_warm_start < -1 || in->poisson_warm_start > 1)
        return bail(AMRESSIF_ERR_ARG, "poisson_warm_start must be -1 (auto), 0 or 1");
    if (in->mixed_momentum != 0 && in->mixed_momentum != 1)
        return bail(AMRESSIF_ERR_ARG, "mixed_momentum must be 0 (dual-time form) or 1 (iterative refinement)");
    if (in->poisson_solver != 0 && in->poisson_solver != 1)
        return bail(AMRESSIF_ERR_ARG, "poisson_solver must be 0 (multigrid) or 1 (FGMRES)");
    if (in->poisson_solver == 1 && (in->gmres_restart < 1 || in->gmres_restart > amr::kGmMax))
        return bail(AMRESSIF_ERR_ARG, "gmres_restart must be in [1, 64]");
    if (in->spacedim != 2 && in->spacedim != 3) return bail(AMRESSIF_ERR_ARG, "spacedim must be 2 or 3");
    if (amr_gen_wanted(in)) {
        // 3D, or a periodic direction: the general path owns the state
        const int rc = amr_gen_create(ctx);
        *out = ctx;
        return rc;
    }
    {
        std::string msg;
        if (int rc = check_bc(*in, msg)) return bail(rc, msg);
    }
    amr::Geometry &g = c.g;
    g.n = n;
    g.P = round_up(n + 1, 64);
    g.dx = (in->hi_phy_dim[0] - in->lo_phy_dim[0]) / n;
    g.dy = (in->hi_phy_dim[1] - in->lo_phy_dim[1]) / n;
    if (!(g.dx > 0) || !(g.dy > 0)) return bail(AMRESSIF_ERR_ARG, "hi_phy_dim must exceed lo_phy_dim");
    g.dt = amressif_dt(in);
    g.bc_lo0 = in->phy_bc_lo[0]; g.bc_hi0 = in->phy_bc_hi[0];
    g.bc_lo1 = in->phy_bc_lo[1]; g.bc_hi1 = in->phy_bc_hi[1];
    g.inflow0 = in->inflow_waveform[0]; g.inflow1 = in->inflow_waveform[1];
    g.ren = in->ren;
    const double d_tau = 0.4 * g.dt;  // main.cpp:861-869
    c.rk[0] = d_tau * 0.25;
    c.rk[1] = d_tau * (1.0 / 3.0);
    c.rk[2] = d_tau * 0.5;
    c.rk[3] = d_tau * 1.0;

    const size_t fsz = (size_t)(n + 1) * g.P;
    auto A = [&](auto **p, size_t cnt) { return amr::dev_alloc_public(c, p, cnt); };
#define ALLOC(p, cnt)                                                               \
    do {                                                                            \
        hipError_t e_ = A(p, cnt);                                                  \
        if (e_ != hipSuccess) return bail(AMRESSIF_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)
    for (int d = 0; d < 2; ++d) {
        ALLOC(&c.U[d], fsz); ALLOC(&c.Uprev[d], fsz); ALLOC(&c.C[d], fsz);
        ALLOC(&c.S1[d], fsz); ALLOC(&c.S2[d], fsz);
    }
    ALLOC(&c.p, fsz); ALLOC(&c.p_alt, fsz); ALLOC(&c.phiv[0], fsz); ALLOC(&c.phiv[1], fsz); ALLOC(&c.b, fsz);
    c.phi = c.phiv[0];
    if (in->precision == AMRESSIF_MIXED)
        for (int d = 0; d < 2; ++d) {
            ALLOC(&c.uk32[d], fsz); ALLOC(&c.r32[d], fsz);
        }
    // multigrid hierarchy (level 0 shares the fine pitch)
    {
        int nx = n, ny = n;
        double hx = g.dx, hy = g.dy;
        int l = 0;
        for (;;) {
            amr::Level &L = c.lev[l];
            L.nx = nx; L.ny = ny;
            L.pitch = l == 0 ? g.P : round_up(nx, nx >= 64 ? 64 : 16);
            L.ax = 1.0 / (hx * hx); L.ay = 1.0 / (hy * hy);
            const size_t lsz = (size_t)L.pitch * (ny + 1);
            if (in->precision == AMRESSIF_MIXED) {
                ALLOC(&L.e32, lsz); ALLOC(&L.e32b, lsz); ALLOC(&L.r32, lsz);
                if (l == 0) { ALLOC(&L.e, lsz); ALLOC(&L.r, lsz); ALLOC(&L.r32b, lsz); }  // e, r: diagnostics only
            } else {
                ALLOC(&L.e, lsz); ALLOC(&L.e2, lsz); ALLOC(&L.r, lsz);
                if (l == 0) ALLOC(&L.rb, lsz);
            }
            ++l;
            if (nx > in->mg_bottom_max && ny > in->mg_bottom_max && nx % 2 == 0 && ny % 2 == 0 && l < amr::kMaxLevels) {
                nx /= 2; ny /= 2; hx *= 2; hy *= 2;
            } else break;
        }
        c.nlev = l;
        const amr::Level &B = c.lev[l - 1];
        c.mb = B.nx * B.ny;
        if (c.mb > 1024) {
            // an odd coarsest level too large for the dense pseudo-inverse: its
            // exact DCT solve as four dense products (kernels_mg.hip)
            if (B.nx > 2048 || B.ny > 2048)
                return bail(AMRESSIF_ERR_UNSUPPORTED, "coarsest multigrid level wider than 2048 cells (n_cell = m 2^k "
                                                      "with an odd m > 2048)");
            const int bx = B.nx, by = B.ny;
            const double pi = 3.14159265358979323846;
            std::vector<double> tab((size_t)bx * bx + (size_t)by * by + (size_t)by * bx);
            auto fill = [&](double *Cm, int m) {
                for (int i = 0; i < m; ++i)
                    for (int k = 0; k < m; ++k)
                        Cm[(size_t)i * m + k] = std::sqrt((k == 0 ? 1.0 : 2.0) / m) * std::cos(pi * k * (2 * i + 1) / (2.0 * m));
            };
            fill(tab.data(), bx);
            fill(tab.data() + (size_t)bx * bx, by);
            double *Dinv = tab.data() + (size_t)bx * bx + (size_t)by * by;
            for (int mm = 0; mm < by; ++mm)
                for (int k = 0; k < bx; ++k) {
                    const double lam = -B.ax * (2.0 - 2.0 * std::cos(pi * k / bx)) - B.ay * (2.0 - 2.0 * std::cos(pi * mm / by));
                    Dinv[(size_t)mm * bx + k] = (k == 0 && mm == 0) ? 0.0 : 1.0 / lam;
                }
            std::vector<float> tab32(tab.begin(), tab.end());
            ALLOC(&c.big_tab, tab.size()); ALLOC(&c.big_tab32, tab.size());
            ALLOC(&c.big_work, (size_t)2 * bx * by); ALLOC(&c.big_work32, (size_t)2 * bx * by);
            hipError_t e1 = hipMemcpy(c.big_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice);
            hipError_t e2 = hipMemcpy(c.big_tab32, tab32.data(), tab32.size() * 4, hipMemcpyHostToDevice);
            if (e1 != hipSuccess || e2 != hipSuccess) return bail(AMRESSIF_ERR_HIP, "bottom DCT table upload failed");
            c.tail_level = l - 1;
            c.tail_ext = 0;
        } else {
        std::vector<double> pinv;
        if (!neumann_pinv(B.nx, B.ny, B.ax, B.ay, pinv)) return bail(AMRESSIF_ERR_ARG, "bottom factorisation failed");
        std::vector<float> pinv32(pinv.begin(), pinv.end());
        ALLOC(&c.pinv, pinv.size()); ALLOC(&c.pinv32, pinv.size());
        hipError_t e1 = hipMemcpy(c.pinv, pinv.data(), pinv.size() * 8, hipMemcpyHostToDevice);
        hipError_t e2 = hipMemcpy(c.pinv32, pinv32.data(), pinv32.size() * 4, hipMemcpyHostToDevice);
        if (e1 != hipSuccess || e2 != hipSuccess) return bail(AMRESSIF_ERR_HIP, "pinv upload failed");
        if (amr::configure_tail(c) != hipSuccess) return bail(AMRESSIF_ERR_UNSUPPORTED, "multigrid tail does not fit LDS");
        const amr::Level &T = c.lev[c.tail_level];
        if (c.tail_dct && T.nx <= 64 && T.ny <= 64) {
            // orthonormal DCT-II tables of the top tail level and the inverse
            // eigenvalues of its Neumann operator (zero for the constant mode)
            const int K = 64;
            std::vector<double> tab(3 * K * K, 0.0);
            const double pi = 3.14159265358979323846;
            auto fill = [&](double *C, int m) {
                for (int i = 0; i < m; ++i)
                    for (int k = 0; k < m; ++k)
                        C[i * K + k] = std::sqrt((k == 0 ? 1.0 : 2.0) / m) * std::cos(pi * k * (2 * i + 1) / (2.0 * m));
            };
            fill(tab.data(), T.nx);
            fill(tab.data() + K * K, T.ny);
            for (int mm = 0; mm < T.ny; ++mm)
                for (int k = 0; k < T.nx; ++k) {
                    const double lam = -T.ax * (2.0 - 2.0 * std::cos(pi * k / T.nx)) -
                                       T.ay * (2.0 - 2.0 * std::cos(pi * mm / T.ny));
                    tab[2 * K * K + mm * K + k] = (k == 0 && mm == 0) ? 0.0 : 1.0 / lam;
                }
            std::vector<float> tab32(tab.begin(), tab.end());
            ALLOC(&c.dct_tab, tab.size()); ALLOC(&c.dct_tab32, tab.size());
            hipError_t e3 = hipMemcpy(c.dct_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice);
            hipError_t e4 = hipMemcpy(c.dct_tab32, tab32.data(), tab32.size() * 4, hipMemcpyHostToDevice);
            if (e3 != hipSuccess || e4 != hipSuccess) return bail(AMRESSIF_ERR_HIP, "DCT table upload failed");
            const amr::Level *L1 = c.tail_level >= 2 ? &c.lev[c.tail_level - 1] : nullptr;
            c.tail_ext = c.tail_ext && in->precision == AMRESSIF_MIXED && L1 && L1->nx == 2 * T.nx &&
                         L1->ny == 2 * T.ny && L1->nx <= 128 && (L1->ny == 32 || L1->ny == 64 || L1->ny == 128);
        } else {
            c.tail_ext = 0;
        }
        }
    }
    if (in->poisson_solver == 1) {  // FGMRES: m+1 Arnoldi vectors, m preconditioned vectors
        c.gm_m = in->gmres_restart;
        c.gm_vs = round_up((int)fsz, 64);
        ALLOC(&c.gmV, c.gm_vs * (size_t)(c.gm_m + 1));
        ALLOC(&c.gmZ, c.gm_vs * (size_t)c.gm_m);
        ALLOC(&c.gm_part, amr::gm_partials_doubles(c, c.gm_m));
        ALLOC(&c.gm, 1);
    }
    ALLOC(&c.partials, (size_t)amr::kMaxPartials * 4);
    ALLOC(&c.ctl, 1);
    {
        hipError_t e = hipHostMalloc((void **)&c.ctl_host, sizeof(amr::Ctl), hipHostMallocDefault);
        if (e != hipSuccess) return bail(AMRESSIF_ERR_HIP, hipGetErrorString(e));
        e = hipEventCreateWithFlags(&c.poll_ev, hipEventDisableTiming);
        if (e != hipSuccess) return bail(AMRESSIF_ERR_HIP, hipGetErrorString(e));
        std::memset(c.ctl_host, 0, sizeof(amr::Ctl));
        e = hipMemcpy(c.ctl, c.ctl_host, sizeof(amr::Ctl), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(AMRESSIF_ERR_HIP, hipGetErrorString(e));
    }
    // zero the state (a cavity at rest) so advance works before load_state
    {
        double *zero[] = {c.U[0], c.U[1], c.Uprev[0], c.Uprev[1], c.p, c.p_alt, c.phiv[0], c.phiv[1]};
        for (double *z : zero) {
            const hipError_t e = hipMemset(z, 0, fsz * 8);
            if (e != hipSuccess) return bail(AMRESSIF_ERR_HIP, hipGetErrorString(e));
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) return bail(AMRESSIF_ERR_HIP, "device sync failed");
#undef ALLOC
    *out = ctx;
    return AMRESSIF_OK;
}

int amressif_destroy(amressif_ctx *ctx)
{
    if (!ctx) return AMRESSIF_OK;
    Ctx &c = ctx->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (void *p : c.allocs) (void)hipFree(p);
    if (c.ctl_host) (void)hipHostFree(c.ctl_host);
    if (c.poll_ev) (void)hipEventDestroy(c.poll_ev);
    if (c.log) (void)hipFree(c.log);
    if (c.kp) {
        for (const amr::KProf::Rec &r : c.kp->recs) { (void)hipEventDestroy(r.e0); (void)hipEventDestroy(r.e1); }
        for (hipEvent_t e : c.kp->pool) (void)hipEventDestroy(e);
        delete c.kp;
    }
    delete c.adj;
    delete c.gen;
    delete ctx;
    return AMRESSIF_OK;
}

const char *amressif_last_error(const amressif_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int64_t amressif_workspace_bytes(const amressif_ctx *ctx) { return ctx ? ctx->c.bytes : 0; }

int amressif_set_lid_profile(amressif_ctx *ctx, const double *dev_lid)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_set_lid_profile: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    if (dev_lid && ctx->c.g.bc_hi1 != 165)
        return fail(ctx, AMRESSIF_ERR_ARG, "a lid profile needs the north inlet code 165");
    ctx->c.lid = dev_lid;
    return AMRESSIF_OK;
}

int amressif_load_state(amressif_ctx *ctx, const amr_array4 velCont[2], const amr_array4 velContPrev[2],
                        const amr_array4 *userCtx)
{
    if (ctx && ctx->c.gen) return amr_gen_load_state(ctx, velCont, velContPrev, userCtx);
    if (!ctx || !velCont || !velContPrev) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
        if (int rc = check_face(ctx, &velContPrev[d], d, "velContPrev")) return rc;
        HIPCHK(amr::launch_pack_face(c, velCont[d], c.U[d], d));
        HIPCHK(amr::launch_pack_face(c, velContPrev[d], c.Uprev[d], d));
    }
    if (userCtx) {
        if (int rc = check_cell(ctx, userCtx, 2, "userCtx")) return rc;
        HIPCHK(amr::launch_pack_cell(c, *userCtx, 0, c.p));
        HIPCHK(amr::launch_pack_cell(c, *userCtx, 1, c.phi));
    } else {
        HIPCHK(amr::launch_fill(c, c.p, (int64_t)(c.g.n + 1) * c.g.P, 0.0));
        HIPCHK(amr::launch_fill(c, c.phi, (int64_t)(c.g.n + 1) * c.g.P, 0.0));
    }
    HIPCHK(amr::launch_mg_start(c));  // phi_mean = 0 for the loaded phi
    c.phih_valid = 0;                  // no solution history for the warm start
    c.pred_valid = 0;
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_store_state(const amressif_ctx *cctx, amr_array4 velCont[2], amr_array4 velContPrev[2],
                         amr_array4 *userCtx, amr_array4 *velCart)
{
    if (cctx && cctx->c.gen)
        return amr_gen_store_state(const_cast<amressif_ctx *>(cctx), velCont, velContPrev, userCtx, velCart);
    amressif_ctx *ctx = const_cast<amressif_ctx *>(cctx);
    if (!ctx) return AMRESSIF_ERR_ARG;
    Ctx &c = ctx->c;
    for (int d = 0; d < 2; ++d) {
        if (velCont) {
            if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
            HIPCHK(amr::launch_unpack_face(c, c.U[d], velCont[d], d));
        }
        if (velContPrev) {
            if (int rc = check_face(ctx, &velContPrev[d], d, "velContPrev")) return rc;
            HIPCHK(amr::launch_unpack_face(c, c.Uprev[d], velContPrev[d], d));
        }
    }
    if (userCtx) {
        if (int rc = check_cell(ctx, userCtx, 2, "userCtx")) return rc;
        HIPCHK(amr::launch_unpack_cell_shift(c, c.p, *userCtx, 0, 1, nullptr));
        HIPCHK(amr::launch_unpack_cell_shift(c, c.phi, *userCtx, 1, 1, &c.ctl->phi_mean));
    }
    if (velCart) {
        if (int rc = check_cell(ctx, velCart, 2, "velCart")) return rc;
        HIPCHK(amr::launch_cont2cart_out(c, c.U[0], c.U[1], *velCart));
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

}  // extern "C"

namespace {

// Valid cell box of one fab of a multi-box MultiFab: kind 0 / 1 = x / y faces
// (nodal in that direction), 2 = cells.  The valid boxes must tile the domain.
struct CellBox {
    int lo0, lo1, hi0, hi1;
};

int check_mf(amressif_ctx *ctx, const amr_multifab *mf, int kind, int ncomp, int max_grow, const char *what,
             std::vector<CellBox> &boxes)
{
    const int n = ctx->c.g.n;
    const std::string w(what);
    if (!mf || mf->nbox < 1 || !mf->fab) return fail(ctx, AMRESSIF_ERR_ARG, w + ": empty MultiFab");
    if (mf->ngrow < 0 || mf->ngrow > max_grow)
        return fail(ctx, AMRESSIF_ERR_ARG, w + ": ngrow must be in [0, " + std::to_string(max_grow) + "]");
    std::vector<unsigned char> cover((size_t)n * n, 0);
    boxes.clear();
    for (int b = 0; b < mf->nbox; ++b) {
        const amr_array4 &a = mf->fab[b];
        if (!a.ptr) return fail(ctx, AMRESSIF_ERR_ARG, w + ": null fab");
        if (a.dtype != 0) return fail(ctx, AMRESSIF_ERR_ARG, w + ": double views only");
        if (a.hi[2] < a.lo[2] || a.ncomp < ncomp) return fail(ctx, AMRESSIF_ERR_ARG, w + ": bad component/z extent");
        CellBox cb{a.lo[0] + mf->ngrow, a.lo[1] + mf->ngrow, a.hi[0] - mf->ngrow, a.hi[1] - mf->ngrow};
        if (kind == 0) cb.hi0 -= 1;  // the high nodal face column
        if (kind == 1) cb.hi1 -= 1;
        if (cb.lo0 < 0 || cb.lo1 < 0 || cb.hi0 >= n || cb.hi1 >= n || cb.hi0 < cb.lo0 || cb.hi1 < cb.lo1)
            return fail(ctx, AMRESSIF_ERR_ARG, w + ": a valid box lies outside the domain");
        for (int j = cb.lo1; j <= cb.hi1; ++j)
            for (int i = cb.lo0; i <= cb.hi0; ++i)
                if (cover[(size_t)i + (size_t)n * j]++)
                    return fail(ctx, AMRESSIF_ERR_ARG, w + ": valid boxes overlap");
        boxes.push_back(cb);
    }
    for (unsigned char x : cover)
        if (!x) return fail(ctx, AMRESSIF_ERR_ARG, w + ": valid boxes do not cover the domain");
    return 0;
}

}  // namespace

extern "C" {

int amressif_load_state_mf(amressif_ctx *ctx, const amr_multifab velCont[2], const amr_multifab velContPrev[2],
                           const amr_multifab *userCtx)
{
    if (ctx && ctx->c.gen) return amr_gen_load_state_mf(ctx, velCont, velContPrev, userCtx);
    if (!ctx || !velCont || !velContPrev) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    // validate everything before the first copy: a failed load leaves the state as it was
    std::vector<CellBox> bx[5];
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_mf(ctx, &velCont[d], d, 1, 0, "velCont", bx[2 * d])) return rc;
        if (int rc = check_mf(ctx, &velContPrev[d], d, 1, 0, "velContPrev", bx[2 * d + 1])) return rc;
    }
    if (userCtx)
        if (int rc = check_mf(ctx, userCtx, 2, 2, 1 << 20, "userCtx", bx[4])) return rc;
    for (int d = 0; d < 2; ++d)
        for (int k = 0; k < 2; ++k) {
            const amr_multifab &mf = k == 0 ? velCont[d] : velContPrev[d];
            double *dst = k == 0 ? c.U[d] : c.Uprev[d];
            for (int b = 0; b < mf.nbox; ++b) {
                const CellBox &q = bx[2 * d + k][b];
                HIPCHK(amr::launch_pack_window(c, mf.fab[b], 0, dst, q.lo0, q.lo1, q.hi0 + (d == 0), q.hi1 + (d == 1)));
            }
        }
    if (userCtx) {
        for (int b = 0; b < userCtx->nbox; ++b) {
            const CellBox &q = bx[4][b];
            HIPCHK(amr::launch_pack_window(c, userCtx->fab[b], 0, c.p, q.lo0, q.lo1, q.hi0, q.hi1));
            HIPCHK(amr::launch_pack_window(c, userCtx->fab[b], 1, c.phi, q.lo0, q.lo1, q.hi0, q.hi1));
        }
    } else {
        HIPCHK(amr::launch_fill(c, c.p, (int64_t)(c.g.n + 1) * c.g.P, 0.0));
        HIPCHK(amr::launch_fill(c, c.phi, (int64_t)(c.g.n + 1) * c.g.P, 0.0));
    }
    HIPCHK(amr::launch_mg_start(c));  // phi_mean = 0 for the loaded phi
    c.phih_valid = 0;                  // no solution history for the warm start
    c.pred_valid = 0;
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_store_state_mf(const amressif_ctx *cctx, amr_multifab velCont[2], amr_multifab velContPrev[2],
                            amr_multifab *userCtx, amr_multifab *velCart)
{
    if (cctx && cctx->c.gen)
        return amr_gen_store_state_mf(const_cast<amressif_ctx *>(cctx), velCont, velContPrev, userCtx, velCart);
    amressif_ctx *ctx = const_cast<amressif_ctx *>(cctx);
    if (!ctx) return AMRESSIF_ERR_ARG;
    Ctx &c = ctx->c;
    std::vector<CellBox> bx[4], cb;
    for (int d = 0; d < 2; ++d) {
        if (velCont)
            if (int rc = check_mf(ctx, &velCont[d], d, 1, 0, "velCont", bx[2 * d])) return rc;
        if (velContPrev)
            if (int rc = check_mf(ctx, &velContPrev[d], d, 1, 0, "velContPrev", bx[2 * d + 1])) return rc;
    }
    if (userCtx)
        if (int rc = check_mf(ctx, userCtx, 2, 2, 1, "userCtx", cb)) return rc;
    if (velCart)
        if (int rc = check_mf(ctx, velCart, 2, 2, 2, "velCart", cb)) return rc;
    for (int d = 0; d < 2; ++d)
        for (int k = 0; k < 2; ++k) {
            amr_multifab *mf = k == 0 ? (velCont ? &velCont[d] : nullptr) : (velContPrev ? &velContPrev[d] : nullptr);
            if (!mf) continue;
            const double *src = k == 0 ? c.U[d] : c.Uprev[d];
            for (int b = 0; b < mf->nbox; ++b) {
                const CellBox &q = bx[2 * d + k][b];
                HIPCHK(amr::launch_unpack_window(c, src, mf->fab[b], 0, q.lo0, q.lo1, q.hi0 + (d == 0),
                                                 q.hi1 + (d == 1)));
            }
        }
    // cell views are written whole, ghosts by global index: the neighbouring
    // box's values inside the domain, the wall rules outside it
    if (userCtx)
        for (int b = 0; b < userCtx->nbox; ++b) {
            HIPCHK(amr::launch_unpack_cell_shift(c, c.p, userCtx->fab[b], 0, 1, nullptr));
            HIPCHK(amr::launch_unpack_cell_shift(c, c.phi, userCtx->fab[b], 1, 1, &c.ctl->phi_mean));
        }
    if (velCart)
        for (int b = 0; b < velCart->nbox; ++b) HIPCHK(amr::launch_cont2cart_out(c, c.U[0], c.U[1], velCart->fab[b]));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

}  // extern "C"

int amr_fail(amressif_ctx *ctx, int code, const std::string &msg) { return fail(ctx, code, msg); }

int amr_state_to_host(amressif_ctx *ctx, double *u, double *v, double *up, double *vp, double *p, double *phi)
{
    Ctx &c = ctx->c;
    const size_t n = (size_t)c.g.n, P = (size_t)c.g.P;
    struct Xfer {
        double *h;
        const double *d;
        size_t w, rows;
    } xs[] = {{u, c.U[0], n + 1, n}, {v, c.U[1], n, n + 1}, {up, c.Uprev[0], n + 1, n},
              {vp, c.Uprev[1], n, n + 1}, {p, c.p, n, n}, {phi, c.phi, n, n}};
    for (const Xfer &x : xs)
        if (x.h)
            HIPCHK(hipMemcpy2DAsync(x.h, x.w * 8, x.d, P * 8, x.w * 8, x.rows, hipMemcpyDeviceToHost, c.stream));
    if (phi) HIPCHK(hipMemcpyAsync(c.ctl_host, c.ctl, sizeof(amr::Ctl), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (phi) {
        const double m = c.ctl_host->phi_mean;
        for (size_t k = 0; k < n * n; ++k) phi[k] -= m;
    }
    return AMRESSIF_OK;
}

int amr_state_from_host(amressif_ctx *ctx, const double *u, const double *v, const double *up, const double *vp,
                        const double *p, const double *phi)
{
    Ctx &c = ctx->c;
    const size_t n = (size_t)c.g.n, P = (size_t)c.g.P;
    struct Xfer {
        const double *h;
        double *d;
        size_t w, rows;
    } xs[] = {{u, c.U[0], n + 1, n}, {v, c.U[1], n, n + 1}, {up, c.Uprev[0], n + 1, n},
              {vp, c.Uprev[1], n, n + 1}, {p, c.p, n, n}, {phi, c.phi, n, n}};
    for (const Xfer &x : xs) {
        if (x.h)
            HIPCHK(hipMemcpy2DAsync(x.d, P * 8, x.h, x.w * 8, x.w * 8, x.rows, hipMemcpyHostToDevice, c.stream));
        else
            HIPCHK(amr::launch_fill(c, x.d, (int64_t)(n + 1) * (int64_t)P, 0.0));
    }
    HIPCHK(amr::launch_mg_start(c));  // phi_mean = 0 for the loaded phi
    c.phih_valid = 0;                  // no solution history for the warm start
    c.pred_valid = 0;
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

extern "C" {

int amressif_advance(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats)
{
    if (ctx && ctx->c.gen) return amr_gen_advance(ctx, nsteps, stats);
    return advance_impl(ctx, nsteps, stats, nullptr);
}

int amressif_step_fwd(amressif_ctx *ctx, int nsteps, amressif_tape *tape, amressif_step_stats *stats)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_step_fwd: not available for 3D or periodic inputs");
    if (!ctx || !tape || tape->ctx != ctx || nsteps < 0) return fail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    if (tape->count + nsteps > tape->cap)
        return fail(ctx, AMRESSIF_ERR_ARG, "tape holds " + std::to_string(tape->cap) + " steps");
    return advance_impl(ctx, nsteps, stats, tape);
}

}  // extern "C"

namespace {
int advance_impl(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats, amressif_tape *tape)
{
    if (!ctx || nsteps < 0) return fail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    if (nsteps == 0) return AMRESSIF_OK;
    if (c.log_cap < nsteps) {
        if (c.log) HIPCHK(hipFree(c.log));
        c.log = nullptr;
        HIPCHK(hipMalloc((void **)&c.log, sizeof(amressif_step_stats) * (size_t)nsteps));
        c.log_cap = nsteps;
    }
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                               c.in.mg_max_iter));
    int status = AMRESSIF_OK;
    const bool mixed = c.in.precision == AMRESSIF_MIXED;
    const bool warm = c.in.poisson_warm_start == 1 || (c.in.poisson_warm_start == -1 && mixed);
    bool bdf_ready = false;  // C and S2 hold the next step's BDF2 constant and u^n
    // Without a tape and with the multigrid solver, the Poisson phase is
    // enqueued right behind the first batch of pseudo-iterations, with no
    // host wait between them; its kernels do nothing if that batch did not
    // end the loop, and the phase then runs again after the loop has ended.
    const bool speculate = !tape && c.in.poisson_solver != 1;
    // the mixed iteration in one pass (k_mixed_pair) chains its output through
    // S1 and S2 and leaves U = u^n untouched, so the projection needs no copy of
    // u^{n+1}: it writes u^{n+1} to the free one of S1, S2 and the buffers rotate
    struct MpScope {
        Ctx &c;
        ~MpScope() { c.mp_on = 0; c.mp_spec = 0; }
    } mp_scope{c};
    c.mp_on = speculate && mixed && c.in.mixed_momentum == 1 && c.in.PSEUDO_TIMESTEPPING == 1 && c.mixed_pair &&
              amr::mixed_pair_fits(c);
    // where this step's predictor u* is
    auto ustar = [&]() -> double *const * { return c.mp_on ? mp_out(c, c.mp_count - 1) : c.U; };
    for (int s = 0; s < nsteps; ++s) {
        int nb = 0;
        bool projected = false;  // the projection ran as the Poisson phase's gated tail
        if (speculate) {
            int launched = 0, rc = 0;
            c.mp_spec = c.mp_on;
            if ((rc = amr_momentum_begin(ctx, bdf_ready, &launched, s > 0))) return rc;
            bdf_ready = false;
            const WarmStart w = mg_prepare(ctx, warm, &rc);
            if (rc) return rc;
            // the projection is enqueued behind every batch of V-cycles, before the
            // host waits: it runs on the device only if the solve has ended, and
            // the host undoes its buffer swaps if it has not
            const bool fold_next = s + 1 < nsteps;
            const int saved_gradp = c.c_gradp;
            double *saved[4][2] = {};  // U, Uprev, S1, S2 before a rotation (undo)
            SpecTail tail;
            tail.enqueue = [&]() -> int {
                // per cell: u*, v*, phi, p in, u, v, p out (56 B); with the next step's
                // BDF2 constant also u^n in, c out and (without the one-pass mixed
                // iteration) the copy of u^{n+1} out: 88 / 104 B
                const double fb = fold_next ? (c.mp_on ? 88.0 : 104.0) : 56.0;
                HIPCHK(amr::kprof_begin(c, amr::KP_PROJECT, amr::KL_PROJ, 0, fb * (double)c.g.n * c.g.n));
                if (c.mp_on) {
                    double *X[2] = {ustar()[0], ustar()[1]};
                    double *const *Yp = X[0] == c.S1[0] ? c.S2 : c.S1;
                    double *Y[2] = {Yp[0], Yp[1]};
                    HIPCHK(amr::launch_project_ex(c, X[0], X[1], Y[0], Y[1], c.p, c.phiv[0], 1, &nb, nullptr, nullptr,
                                                  fold_next ? 1 : 0, c.p_alt, c.phiv[1], 1, c.U, 1));
                    HIPCHK(amr::kprof_end(c));
                    HIPCHK(amr::launch_fin_project(c, nb, s, fold_next ? 1 : 0, 1));
                    for (int d = 0; d < 2; ++d) {
                        saved[0][d] = c.U[d]; saved[1][d] = c.Uprev[d]; saved[2][d] = c.S1[d]; saved[3][d] = c.S2[d];
                        // U <- u^{n+1}, Uprev <- u^n, S1 <- u*, S2 <- the free array
                        c.Uprev[d] = saved[0][d];
                        c.S2[d] = saved[1][d];
                        c.U[d] = Y[d];
                        c.S1[d] = X[d];
                    }
                } else {
                    HIPCHK(amr::launch_project_ex(c, c.U[0], c.U[1], c.S1[0], c.S1[1], c.p, c.phiv[0], 1, &nb, nullptr,
                                                  nullptr, fold_next ? 1 : 0, c.p_alt, c.phiv[1], 1));
                    HIPCHK(amr::kprof_end(c));
                    HIPCHK(amr::launch_fin_project(c, nb, s, fold_next ? 1 : 0, 1));
                    std::swap(c.U[0], c.S1[0]);
                    std::swap(c.U[1], c.S1[1]);
                }
                std::swap(c.p, c.p_alt);
                c.c_gradp = fold_next && c.project_rows > 0;
                return 0;
            };
            tail.undo = [&]() {
                if (c.mp_on) {
                    for (int d = 0; d < 2; ++d) {
                        c.U[d] = saved[0][d]; c.Uprev[d] = saved[1][d]; c.S1[d] = saved[2][d]; c.S2[d] = saved[3][d];
                    }
                } else {
                    std::swap(c.U[0], c.S1[0]);
                    std::swap(c.U[1], c.S1[1]);
                }
                std::swap(c.p, c.p_alt);
                c.c_gradp = saved_gradp;
                amr::kprof_close(c, amr::KL_PROJ, 0);
            };
            bool skipped = false;
            rc = mg_run(ctx, w, false, ustar()[0], ustar()[1], true, &skipped, &tail, &projected);
            c.mp_spec = 0;
            if (rc) return rc;
            if (skipped) {
                if ((rc = amr_momentum_finish(ctx, launched))) return rc;
                if ((rc = mg_run(ctx, w, false, ustar()[0], ustar()[1], false, nullptr, &tail, &projected)))
                    return rc;
            } else {
                c.last_P = std::max(1, c.ctl_host->mom_count);
                amr::kprof_close(c, amr::KL_MOM, c.ctl_host->mom_count);
            }
            if (projected) {
                amr::kprof_close(c, amr::KL_PROJ, 1);
                bdf_ready = fold_next;
            }
        } else {
            if (int rc = amr_momentum_predictor(ctx, bdf_ready, s > 0)) return rc;
            bdf_ready = false;
        }
        const int flags = c.ctl_host->flags;  // (copied before the projection clears them)
        if (tape) {  // record the converged predictor u* of this step
            const size_t bytes = tape->fsz * sizeof(double);
            HIPCHK(hipMemcpyAsync(tape->slot(tape->count), c.U[0], bytes, hipMemcpyDeviceToDevice, c.stream));
            HIPCHK(hipMemcpyAsync(tape->slot(tape->count) + tape->fsz, c.U[1], bytes, hipMemcpyDeviceToDevice,
                                  c.stream));
            tape->count += 1;
        }
        // ---- projection: Poisson RHS, MLMG-equivalent solve, update
        if (!speculate)
            if (int rc = poisson_loop_step(ctx, warm, c.U[0], c.U[1])) return rc;
        // between two steps of this call the projection also forms the next
        // step's BDF2 constant (main.cpp:879-884); the last step leaves
        // C and velContPrev as the reference does
        if (!projected) {
            const bool fold = s + 1 < nsteps && !(flags & AMRESSIF_FLAG_NONFINITE);
            // per cell: u*, v*, phi, p in, u, v, p out (56 B); with the next step's
            // BDF2 constant also u^n in, c and the copy of u^{n+1} out (104 B)
            // p^{n+1} goes to the other pressure array (p^n stays readable for the
            // neighbour terms), and the next constant takes -G p^{n+1}
            HIPCHK(amr::kprof_begin(c, amr::KP_PROJECT, amr::KL_ONCE, 0,
                                    (fold ? (c.mp_on ? 88.0 : 104.0) : 56.0) * (double)c.g.n * c.g.n));
            if (c.mp_on) {  // u* in S1 or S2, u^{n+1} into the other, then the buffers rotate
                double *X[2] = {ustar()[0], ustar()[1]};
                double *const *Yp = X[0] == c.S1[0] ? c.S2 : c.S1;
                double *Y[2] = {Yp[0], Yp[1]};
                HIPCHK(amr::launch_project_ex(c, X[0], X[1], Y[0], Y[1], c.p, c.phi, 1, &nb, nullptr, nullptr,
                                              fold ? 1 : 0, c.p_alt, nullptr, 0, c.U, 1));
                for (int d = 0; d < 2; ++d) {
                    double *un = c.U[d], *um = c.Uprev[d];
                    c.Uprev[d] = un;
                    c.S2[d] = um;
                    c.U[d] = Y[d];
                    c.S1[d] = X[d];
                }
            } else {
                HIPCHK(amr::launch_project_ex(c, c.U[0], c.U[1], c.S1[0], c.S1[1], c.p, c.phi, 1, &nb, nullptr,
                                              nullptr, fold ? 1 : 0, c.p_alt));
                std::swap(c.U[0], c.S1[0]);
                std::swap(c.U[1], c.S1[1]);
            }
            HIPCHK(amr::kprof_end(c));
            std::swap(c.p, c.p_alt);
            c.c_gradp = fold && c.project_rows > 0;
            // (between two steps of the call it also clears the next step's control)
            HIPCHK(amr::launch_fin_project(c, nb, s, s + 1 < nsteps ? 1 : 0));
            bdf_ready = fold;
        }
        if (flags & AMRESSIF_FLAG_NONFINITE) {
            status = fail(ctx, AMRESSIF_ERR_NONFINITE, "non-finite pseudo-iteration norm at step " + std::to_string(s));
            nsteps = s + 1;
            break;
        }
        c.pred_valid = 1;  // S1 now holds this step's predictor u* (the projection's input)
    }
    if (stats) HIPCHK(hipMemcpyAsync(stats, c.log, sizeof(amressif_step_stats) * (size_t)nsteps,
                                     hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(amr::kprof_collect(c));
    return status;
}
}  // namespace

extern "C" {

// ------------------------------------------------------ module entry points
int amressif_cont2cart(amressif_ctx *ctx, const amr_array4 velCont[2], amr_array4 *velCart)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_cont2cart: not available for 3D or periodic inputs");
    if (!ctx || !velCont) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
        HIPCHK(amr::launch_pack_face(c, velCont[d], c.M[0][d], d));
    }
    if (int rc = check_cell(ctx, velCart, 2, "velCart")) return rc;
    HIPCHK(amr::launch_cont2cart_out(c, c.M[0][0], c.M[0][1], *velCart));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_momentum_rhs(amressif_ctx *ctx, const amr_array4 velStar[2], const amr_array4 *userCtx,
                          amr_array4 rhs[2])
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_momentum_rhs: not available for 3D or periodic inputs");
    if (!ctx || !velStar || !userCtx || !rhs) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velStar[d], d, "velStar")) return rc;
        if (int rc = check_face(ctx, &rhs[d], d, "rhs")) return rc;
        HIPCHK(amr::launch_pack_face(c, velStar[d], c.M[0][d], d));
    }
    if (int rc = check_cell(ctx, userCtx, 1, "userCtx")) return rc;
    HIPCHK(amr::launch_pack_cell(c, *userCtx, 0, c.Mc[0]));
    amr::StageLaunch a{};
    a.us = c.M[0][0]; a.vs = c.M[0][1]; a.p = c.Mc[0];
    a.uo = c.M[1][0]; a.vo = c.M[1][1];
    a.mode = 2;
    int nb = 0;
    HIPCHK(amr::launch_stage(c, a, &nb));
    for (int d = 0; d < 2; ++d) HIPCHK(amr::launch_unpack_face(c, c.M[1][d], rhs[d], d));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_rk_stage(amressif_ctx *ctx, int sub, amr_array4 velStar[2], const amr_array4 velCont[2],
                      const amr_array4 velContPrev[2], const amr_array4 velContPrevPrev[2], const amr_array4 *userCtx)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_rk_stage: not available for 3D or periodic inputs");
    if (!ctx || !velStar || !velCont || !velContPrev || !velContPrevPrev || !userCtx)
        return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    if (sub < 0 || sub > 3) return fail(ctx, AMRESSIF_ERR_ARG, "sub must be 0..3");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velStar[d], d, "velStar")) return rc;
        if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
        if (int rc = check_face(ctx, &velContPrev[d], d, "velContPrev")) return rc;
        if (int rc = check_face(ctx, &velContPrevPrev[d], d, "velContPrevPrev")) return rc;
        HIPCHK(amr::launch_pack_face(c, velStar[d], c.M[0][d], d));
        HIPCHK(amr::launch_pack_face(c, velCont[d], c.M[1][d], d));
        HIPCHK(amr::launch_pack_face(c, velContPrev[d], c.M[2][d], d));
        HIPCHK(amr::launch_pack_face(c, velContPrevPrev[d], c.M[3][d], d));
        HIPCHK(amr::launch_bdf2_const(c, c.M[2][d], c.M[3][d], c.M[2][d], d));  // in place: c from (u^n, u^{n-1})
    }
    if (int rc = check_cell(ctx, userCtx, 1, "userCtx")) return rc;
    HIPCHK(amr::launch_pack_cell(c, *userCtx, 0, c.Mc[0]));
    amr::StageLaunch a{};
    a.us = c.M[0][0]; a.vs = c.M[0][1];
    a.ub = c.M[1][0]; a.vb = c.M[1][1];
    a.cu = c.M[2][0]; a.cv = c.M[2][1];
    a.p = c.Mc[0];
    a.uo = c.M[3][0]; a.vo = c.M[3][1];
    a.rk = c.rk[sub];
    a.mode = 0;
    HIPCHK(amr::launch_ctl_reset_step(c));  // clears mom_done so the guarded kernel runs
    int nb = 0;
    HIPCHK(amr::launch_stage(c, a, &nb));
    for (int d = 0; d < 2; ++d) HIPCHK(amr::launch_unpack_face(c, c.M[3][d], velStar[d], d));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_fluxes(amressif_ctx *ctx, const amr_array4 velStar[2], amr_array4 *fluxConvect, amr_array4 *fluxViscous,
                    amr_array4 fluxHalfN1[2], amr_array4 fluxHalfN2[2])
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_fluxes: not available for 3D or periodic inputs");
    if (!ctx || !velStar) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    if ((fluxHalfN1 == nullptr) != (fluxHalfN2 == nullptr))
        return fail(ctx, AMRESSIF_ERR_ARG, "fluxHalfN1 and fluxHalfN2 go together");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velStar[d], d, "velStar")) return rc;
        if (fluxHalfN1) {
            if (int rc = check_face(ctx, &fluxHalfN1[d], d, "fluxHalfN1")) return rc;
            if (int rc = check_face(ctx, &fluxHalfN2[d], d, "fluxHalfN2")) return rc;
        }
        HIPCHK(amr::launch_pack_face(c, velStar[d], c.M[0][d], d));
    }
    if (fluxConvect)
        if (int rc = check_cell(ctx, fluxConvect, 2, "fluxConvect")) return rc;
    if (fluxViscous)
        if (int rc = check_cell(ctx, fluxViscous, 2, "fluxViscous")) return rc;
    // M[1]: fluxHalfN1 (x-, y-faces), M[2]: fluxHalfN2, Mc: fluxConvect, M[3]: fluxViscous
    const bool half = fluxHalfN1 != nullptr;
    HIPCHK(amr::launch_flux_fields(c, c.M[0][0], c.M[0][1], half ? c.M[1][0] : nullptr, half ? c.M[2][0] : nullptr,
                                   half ? c.M[1][1] : nullptr, half ? c.M[2][1] : nullptr,
                                   fluxConvect ? c.Mc[0] : nullptr, fluxConvect ? c.Mc[1] : nullptr,
                                   fluxViscous ? c.M[3][0] : nullptr, fluxViscous ? c.M[3][1] : nullptr));
    if (half)
        for (int d = 0; d < 2; ++d) {
            HIPCHK(amr::launch_unpack_face(c, c.M[1][d], fluxHalfN1[d], d));
            HIPCHK(amr::launch_unpack_face(c, c.M[2][d], fluxHalfN2[d], d));
        }
    for (int comp = 0; comp < 2; ++comp) {
        if (fluxConvect) HIPCHK(amr::launch_unpack_cell(c, c.Mc[comp], *fluxConvect, comp, 0));
        if (fluxViscous) HIPCHK(amr::launch_unpack_cell(c, c.M[3][comp], *fluxViscous, comp, 0));
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_pseudo_norm(amressif_ctx *ctx, const amr_array4 velCont[2], const amr_array4 velStar[2], double *norm)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_pseudo_norm: not available for 3D or periodic inputs");
    if (!ctx || !velCont || !velStar || !norm) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
        if (int rc = check_face(ctx, &velStar[d], d, "velStar")) return rc;
        HIPCHK(amr::launch_pack_face(c, velCont[d], c.M[0][d], d));
        HIPCHK(amr::launch_pack_face(c, velStar[d], c.M[1][d], d));
    }
    HIPCHK(amr::launch_norm(c, c.M[0][0], c.M[0][1], c.M[1][0], c.M[1][1], c.dscalar));
    HIPCHK(hipMemcpyAsync(norm, c.dscalar, 8, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_poisson_rhs(amressif_ctx *ctx, const amr_array4 vel[2], amr_array4 *rhs, double *div_linf,
                         double *rhs_sum)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_poisson_rhs: not available for 3D or periodic inputs");
    if (!ctx || !vel || !rhs) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &vel[d], d, "vel")) return rc;
        HIPCHK(amr::launch_pack_face(c, vel[d], c.M[0][d], d));
    }
    if (int rc = check_cell(ctx, rhs, 1, "rhs")) return rc;
    int nb = 0;
    HIPCHK(amr::launch_poisson_rhs(c, c.M[0][0], c.M[0][1], c.Mc[0], &nb));
    HIPCHK(amr::launch_fin_prhs(c, nb));
    HIPCHK(amr::launch_unpack_cell(c, c.Mc[0], *rhs, 0, 0));
    if (int rc = poll(ctx)) return rc;
    if (div_linf) *div_linf = c.ctl_host->div_star;
    if (rhs_sum) *rhs_sum = c.ctl_host->rhs_sum;
    return AMRESSIF_OK;
}

int amressif_poisson_solve(amressif_ctx *ctx, amr_array4 *sol, const amr_array4 *rhs, double tol_rel, double tol_abs,
                           int *iters, double *res_rel)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_poisson_solve: not available for 3D or periodic inputs");
    if (!ctx || !sol || !rhs) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = check_cell(ctx, rhs, 1, "rhs")) return rc;
    if (int rc = check_cell(ctx, sol, 1, "sol")) return rc;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, tol_rel, tol_abs, c.in.IterNum, c.in.mg_max_iter));
    HIPCHK(amr::launch_pack_cell(c, *rhs, 0, c.b));
    HIPCHK(amr::launch_mean_b(c, c.b));
    HIPCHK(amr::launch_center_b(c, c.b));
    if (int rc = poisson_loop(ctx)) return rc;
    HIPCHK(amr::launch_unpack_cell_shift(c, c.phi, *sol, 0, 0, &c.ctl->phi_mean));
    if (int rc = poll(ctx)) return rc;
    if (iters) *iters = c.ctl_host->mg_iters;
    if (res_rel) *res_rel = c.ctl_host->b_norm > 0 ? c.ctl_host->res_norm / c.ctl_host->b_norm : c.ctl_host->res_norm;
    // restore the step tolerances
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                               c.in.mg_max_iter));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_project(amressif_ctx *ctx, const amr_array4 velStar[2], amr_array4 *userCtx, amr_array4 velCont[2],
                     double *div_linf)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_project: not available for 3D or periodic inputs");
    if (!ctx || !velStar || !userCtx || !velCont) return fail(ctx, AMRESSIF_ERR_ARG, "null argument");
    Ctx &c = ctx->c;
    if (int rc = ensure_module_scratch(ctx)) return rc;
    for (int d = 0; d < 2; ++d) {
        if (int rc = check_face(ctx, &velStar[d], d, "velStar")) return rc;
        if (int rc = check_face(ctx, &velCont[d], d, "velCont")) return rc;
        HIPCHK(amr::launch_pack_face(c, velStar[d], c.M[0][d], d));
    }
    if (int rc = check_cell(ctx, userCtx, 2, "userCtx")) return rc;
    HIPCHK(amr::launch_pack_cell(c, *userCtx, 0, c.Mc[0]));
    HIPCHK(amr::launch_pack_cell(c, *userCtx, 1, c.Mc[1]));
    int nb = 0;
    HIPCHK(amr::launch_project_ex(c, c.M[0][0], c.M[0][1], c.M[1][0], c.M[1][1], c.Mc[0], c.Mc[1], 0, &nb));
    HIPCHK(amr::launch_fin_project(c, nb, -1));
    for (int d = 0; d < 2; ++d) HIPCHK(amr::launch_unpack_face(c, c.M[1][d], velCont[d], d));
    HIPCHK(amr::launch_unpack_cell(c, c.Mc[0], *userCtx, 0, 1));
    if (int rc = poll(ctx)) return rc;
    if (div_linf) *div_linf = c.ctl_host->div_new;
    return AMRESSIF_OK;
}

int amressif_kernel_timing(amressif_ctx *ctx, int enable)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_kernel_timing: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.kp) c.kp = new amr::KProf();
    if (enable) {
        HIPCHK(hipStreamSynchronize(c.stream));
        HIPCHK(amr::kprof_collect(c));
        for (int k = 0; k < amr::KP_COUNT; ++k) { c.kp->sum_ms[k] = c.kp->sum_bytes[k] = 0.0; c.kp->cnt[k] = 0; }
    }
    c.kp->on = enable != 0;
    return AMRESSIF_OK;
}

int amressif_kernel_time(const amressif_ctx *ctx, int which, double *avg_ms, int64_t *launches,
                         double *bytes_per_launch)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_kernel_time: not available for 3D or periodic inputs");
    if (!ctx || !avg_ms || !launches || !bytes_per_launch) return AMRESSIF_ERR_ARG;
    if (which < 0 || which >= amr::KP_COUNT) return AMRESSIF_ERR_ARG;
    const amr::KProf *k = ctx->c.kp;
    *launches = k ? k->cnt[which] : 0;
    *avg_ms = *launches ? k->sum_ms[which] / (double)*launches : 0.0;
    *bytes_per_launch = *launches ? k->sum_bytes[which] / (double)*launches : 0.0;
    return AMRESSIF_OK;
}

int amressif_time_kernel(amressif_ctx *ctx, int which, int reps, int64_t copy_bytes, double *avg_ms,
                         double *bytes_per_launch)
{
    if (ctx && ctx->c.gen) {
        if (which < 7 || which > 9 || reps < 1 || !avg_ms)
            return fail(ctx, AMRESSIF_ERR_UNSUPPORTED, "time_kernel: 3D or periodic contexts time kernels 7..9");
        return amr_gen_time_kernel(ctx, which - 7, reps, avg_ms, bytes_per_launch);
    }
    if (!ctx || !avg_ms || reps < 1) return fail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    Ctx &c = ctx->c;
    c.pred_valid = 0;  // the timed kernels overwrite the step buffers
    const double cells = (double)c.g.n * (double)c.g.n;
    void *src = nullptr, *dst = nullptr;
    if (which == 2) {
        if (copy_bytes < 16) return fail(ctx, AMRESSIF_ERR_ARG, "copy_bytes must be >= 16");
        HIPCHK(hipMalloc(&src, (size_t)copy_bytes));
        HIPCHK(hipMalloc(&dst, (size_t)copy_bytes));
        HIPCHK(hipMemsetAsync(src, 0, (size_t)copy_bytes, c.stream));
    } else if (which < 0 || (which > 6 && which != 10)) {
        return fail(ctx, AMRESSIF_ERR_ARG, "which must be 0..6 or 10");
    }
    if (which == 6 && !amr::stage_pair_fits(c)) return fail(ctx, AMRESSIF_ERR_ARG, "no two-stage kernel at this size");
    if ((which == 3 || which == 4 || which == 10) && c.in.precision != AMRESSIF_MIXED)
        return fail(ctx, AMRESSIF_ERR_ARG, "which = 3, 4, 10 need a mixed-precision context");
    if (which == 10 && !amr::mixed_pair_fits(c)) return fail(ctx, AMRESSIF_ERR_ARG, "no one-pass mixed kernel at this size");
    auto launch = [&]() -> hipError_t {
        if (which == 10) {  // the mixed iteration in one pass, u^k from U into S2
            amr::MixedPairLaunch m{};
            m.u64 = c.U[0]; m.v64 = c.U[1]; m.cu = c.C[0]; m.cv = c.C[1];
            m.uo = c.S2[0]; m.vo = c.S2[1];
            m.cr = 1.0e-12; m.cj = 1.0e-24;
            m.unguarded = 1;
            int nb = 0;
            return amr::launch_mixed_pair(c, m, &nb);
        }
        if (which == 3 || which == 4) {  // the two kernels of a mixed pseudo-iteration
            amr::MixedStageLaunch a{};
            a.stage = which - 3;
            a.u64 = which == 3 ? c.U[0] : c.S2[0];  // the residual only reads it
            a.v64 = which == 3 ? c.U[1] : c.S2[1];
            a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
            a.foldp = 1;  // the step's form (C holds -G p)
            a.uk = c.uk32[0]; a.vk = c.uk32[1]; a.ru = c.r32[0]; a.rv = c.r32[1];
            a.cj = 0.0;  // keeps the timed in-place update of S2 a no-op in value
            a.cr = 0.0;
            a.unguarded = 1;
            int nb = 0;
            return amr::launch_stage_mixed(c, a, &nb);
        }
        if (which == 0) {
            amr::StageLaunch a{};
            // distinct stencil input and base arrays, as inside a step (S1 holds the step's last stage)
            a.us = c.S1[0]; a.vs = c.S1[1]; a.ub = c.U[0]; a.vb = c.U[1];
            a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
            a.uo = c.S2[0]; a.vo = c.S2[1];
            a.rk = c.rk[0];
            a.mode = 0;
            a.foldp = 1;  // the step's form (C holds -G p)
            a.unguarded = 1;
            int nb = 0;
            return amr::launch_stage(c, a, &nb);
        }
        if (which == 6) {  // stages 1-2 in one pass, as inside a step
            amr::StageLaunch a{};
            a.us = c.S1[0]; a.vs = c.S1[1]; a.ub = c.U[0]; a.vb = c.U[1];
            a.cu = c.C[0]; a.cv = c.C[1]; a.p = c.p;
            a.uo = c.S2[0]; a.vo = c.S2[1];
            a.rk = c.rk[0]; a.rk2 = c.rk[1]; a.pair = 1;
            a.foldp = 1;
            a.unguarded = 1;
            int nb = 0;
            return amr::launch_stage_pair(c, a, &nb);
        }
        if (which == 1 || which == 5) return amr::launch_smooth_fine(c, which, c.S2[0]);
        return amr::launch_copy_bytes(c, src, dst, copy_bytes);
    };
    HIPCHK(launch());  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r) HIPCHK(launch());
    HIPCHK(hipEventRecord(e1, c.stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    *avg_ms = (double)ms / reps;
    if (bytes_per_launch) {
        // algorithmic bytes per cell (t = 4 or 8: the V-cycle precision):
        // 0 FP64 stage: reads u*, v*, u^k, v^k, c_u, c_v (with -G p), writes u*, v* (8 x 8 B);
        // 1 finest down-leg smoother: reads r, writes e (2 t), the coarse r (t / 4);
        // 2 copy: read + write of copy_bytes;
        // 3 mixed residual: reads u^k, v^k, c_u, c_v (with -G p; 4 x 8 B), writes the FP32 base
        //   and residual (4 x 4 B);
        // 4 tangent-linear update: reads the FP32 base and residual (4 x 4 B), reads and
        //   writes u^k, v^k (4 x 8 B);
        // 5 finest up-leg smoother + FP64 outer step: reads e, r (2 t), the coarse e (t / 4),
        //   phi, b (2 x 8 B), writes phi (8 B) and the next r (t);
        // 6 two FP64 stages in one pass: the stage's 64 B (the first stage stays in registers);
        // 10 mixed iteration in one pass: reads u^k, v^k, c_u, c_v, writes u^{k+1}, v^{k+1} (6 x 8 B)
        const double t = c.in.precision == AMRESSIF_MIXED ? 4.0 : 8.0;
        const double per_cell[7] = {64.0, 2.25 * t, 0.0, 48.0, 48.0, 3.25 * t + 24.0, 64.0};
        *bytes_per_launch = which == 2 ? 2.0 * (double)copy_bytes : which == 10 ? 48.0 * cells : per_cell[which] * cells;
    }
    return AMRESSIF_OK;
}

}  // extern "C"

// ----------------------------------------------------- discrete adjoint
namespace {

int ensure_adjoint(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    if (c.adj) return 0;
    const int n = c.g.n;
    if (n < 4) return fail(ctx, AMRESSIF_ERR_UNSUPPORTED, "the adjoint needs n_cell >= 4");
    c.Pg = round_up(n + 4, 64);
    const size_t gsz = (size_t)c.Pg * (n + 4), fsz = (size_t)(n + 1) * c.g.P;
    auto *w = new amr::AdjWork();
    c.adj = w;
    std::vector<std::pair<double **, size_t>> bufs = {
        {&w->XU, gsz}, {&w->XV, gsz}, {&w->DU, gsz}, {&w->DV, gsz}, {&w->F[0], fsz}, {&w->F[1], fsz},
        {&w->F[2], fsz}, {&w->F[3], fsz}, {&w->Q[0], fsz}, {&w->Q[1], fsz}, {&w->XbU, fsz}, {&w->XbV, fsz},
        {&w->lidg, (size_t)2 * n}, {&w->L[0][0], fsz}, {&w->L[0][1], fsz}, {&w->L[1][0], fsz}, {&w->L[1][1], fsz},
        {&c.Du[0], fsz}, {&c.Du[1], fsz}, {&c.Dup[0], fsz}, {&c.Dup[1], fsz}, {&c.Dp, fsz}, {&c.Dlid, (size_t)n},
        {&w->phi_lin[0], fsz}, {&w->phi_lin[1], fsz}, {&w->phi_mean_save, (size_t)1}};
    for (auto &b : bufs) {
        HIPCHK(amr::dev_alloc_public(c, b.first, b.second));
        HIPCHK(amr::launch_fill(c, *b.first, (int64_t)b.second, 0.0));
    }
    return 0;
}

// The tangent and adjoint projections run the step's Poisson solve on their
// own pair of iterates: the forward state's phi (userCtx comp 1), its mean and
// the warm start's history stay as the forward step left them.
struct LinPhi {
    double *phiv[2] = {}, *phi = nullptr;
    int last_V = 0;
};
int lin_phi_enter(amressif_ctx *ctx, LinPhi &keep)
{
    Ctx &c = ctx->c;
    keep.phiv[0] = c.phiv[0];
    keep.phiv[1] = c.phiv[1];
    keep.phi = c.phi;
    keep.last_V = c.last_V;
    HIPCHK(hipMemcpyAsync(c.adj->phi_mean_save, &c.ctl->phi_mean, sizeof(double), hipMemcpyDeviceToDevice,
                          c.stream));
    c.phiv[0] = c.adj->phi_lin[0];
    c.phiv[1] = c.adj->phi_lin[1];
    c.phi = c.phiv[0];
    return 0;
}
int lin_phi_leave(amressif_ctx *ctx, const LinPhi &keep)
{
    Ctx &c = ctx->c;
    c.phiv[0] = keep.phiv[0];
    c.phiv[1] = keep.phiv[1];
    c.phi = keep.phi;
    c.last_V = keep.last_V;
    HIPCHK(hipMemcpyAsync(&c.ctl->phi_mean, c.adj->phi_mean_save, sizeof(double), hipMemcpyDeviceToDevice,
                          c.stream));
    return 0;
}

// Richardson solve of the tangent (adjoint = false) or adjoint system; the
// initial guess is in L[0], the solution is returned in *xu, *xv.
int lin_solve(amressif_ctx *ctx, bool adjoint, const double *dlid, double **xu, double **xv, amressif_lin_stats *st)
{
    Ctx &c = ctx->c;
    amr::AdjWork &w = *c.adj;
    const int cap = std::max(1, c.adj_cap);
    HIPCHK(amr::launch_lin_start(c, c.adj_tol, cap));
    int launched = 0, target = std::max(1, c.last_L);
    for (;;) {
        while (launched < target && launched < cap) {
            const int par = launched & 1;
            if (adjoint)
                HIPCHK(amr::launch_adj_sweep(c, w, w.L[par][0], w.L[par][1], w.L[par ^ 1][0], w.L[par ^ 1][1]));
            else
                HIPCHK(amr::launch_tl_sweep(c, w, w.L[par][0], w.L[par][1], w.L[par ^ 1][0], w.L[par ^ 1][1], dlid));
            ++launched;
        }
        if (int rc = poll(ctx)) return rc;
        if (c.ctl_host->lin_done || launched >= cap) break;
        target = launched + 2;
    }
    const int cnt = c.ctl_host->lin_count;
    c.last_L = std::max(1, cnt);
    *xu = w.L[cnt & 1][0];
    *xv = w.L[cnt & 1][1];
    if (st) {
        st->lin_iters = cnt;
        st->lin_rel = c.ctl_host->lin_rel;
    }
    return 0;
}

bool lid_carries_g(const Ctx &c) { return c.g.bc_hi1 != 135 && c.g.bc_hi1 != -135; }

int check_tape(amressif_ctx *ctx, const amressif_tape *tape)
{
    if (!tape || tape->ctx != ctx) return fail(ctx, AMRESSIF_ERR_ARG, "tape belongs to another context");
    if (!ctx->c.adj) return fail(ctx, AMRESSIF_ERR_ARG, "no dual state: create a tape first");
    return 0;
}

}  // namespace

extern "C" {

int amressif_tape_create(amressif_ctx *ctx, int max_steps, amressif_tape **out)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_tape_create: not available for 3D or periodic inputs");
    if (ctx && ctx->c.in.PSEUDO_TIMESTEPPING == 0)
        return fail(ctx, AMRESSIF_ERR_UNSUPPORTED, "the adjoint is that of the dual-time step (PSEUDO_TIMESTEPPING = 1)");
    if (!ctx || !out || max_steps < 1) return fail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    *out = nullptr;
    if (int rc = ensure_adjoint(ctx)) return rc;
    Ctx &c = ctx->c;
    auto *t = new amressif_tape();
    t->ctx = ctx;
    t->stream = c.stream;
    t->cap = max_steps;
    t->fsz = (size_t)(c.g.n + 1) * c.g.P;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)max_steps * 2 * t->fsz * sizeof(double));
    if (e != hipSuccess) {
        delete t;
        return fail(ctx, AMRESSIF_ERR_HIP, std::string("tape allocation: ") + hipGetErrorString(e));
    }
    t->store = (double *)p;
    *out = t;
    return AMRESSIF_OK;
}

int amressif_tape_destroy(amressif_tape *tape)
{
    if (!tape) return AMRESSIF_OK;
    (void)hipStreamSynchronize(tape->stream);
    if (tape->store) (void)hipFree(tape->store);
    delete tape;
    return AMRESSIF_OK;
}

int amressif_tape_reset(amressif_tape *tape)
{
    if (!tape) return AMRESSIF_ERR_ARG;
    tape->count = 0;
    return AMRESSIF_OK;
}

int amressif_tape_steps(const amressif_tape *tape) { return tape ? tape->count : 0; }

int amressif_set_adjoint_solver(amressif_ctx *ctx, double tol, int max_sweeps)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_set_adjoint_solver: not available for 3D or periodic inputs");
    if (!ctx || !(tol >= 0.0) || max_sweeps < 1) return fail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    ctx->c.adj_tol = tol;
    ctx->c.adj_cap = max_sweeps;
    return AMRESSIF_OK;
}

int amressif_load_dual(amressif_ctx *ctx, const amr_array4 du[2], const amr_array4 du_prev[2], const amr_array4 *dp,
                       const double *dlid)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_load_dual: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    if (int rc = ensure_adjoint(ctx)) return rc;
    Ctx &c = ctx->c;
    const int64_t fsz = (int64_t)(c.g.n + 1) * c.g.P;
    for (int d = 0; d < 2; ++d) {
        if (du) {
            if (int rc = check_face(ctx, &du[d], d, "du")) return rc;
            HIPCHK(amr::launch_pack_face(c, du[d], c.Du[d], d));
        } else {
            HIPCHK(amr::launch_fill(c, c.Du[d], fsz, 0.0));
        }
        if (du_prev) {
            if (int rc = check_face(ctx, &du_prev[d], d, "du_prev")) return rc;
            HIPCHK(amr::launch_pack_face(c, du_prev[d], c.Dup[d], d));
        } else {
            HIPCHK(amr::launch_fill(c, c.Dup[d], fsz, 0.0));
        }
    }
    if (dp) {
        if (int rc = check_cell(ctx, dp, 1, "dp")) return rc;
        HIPCHK(amr::launch_pack_cell(c, *dp, 0, c.Dp));
    } else {
        HIPCHK(amr::launch_fill(c, c.Dp, fsz, 0.0));
    }
    if (dlid)
        HIPCHK(hipMemcpyAsync(c.Dlid, dlid, sizeof(double) * c.g.n, hipMemcpyDeviceToDevice, c.stream));
    else
        HIPCHK(amr::launch_fill(c, c.Dlid, c.g.n, 0.0));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_store_dual(amressif_ctx *ctx, amr_array4 du[2], amr_array4 du_prev[2], amr_array4 *dp, double *dlid)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_store_dual: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.adj) return fail(ctx, AMRESSIF_ERR_ARG, "no dual state");
    for (int d = 0; d < 2; ++d) {
        if (du) {
            if (int rc = check_face(ctx, &du[d], d, "du")) return rc;
            HIPCHK(amr::launch_unpack_face(c, c.Du[d], du[d], d));
        }
        if (du_prev) {
            if (int rc = check_face(ctx, &du_prev[d], d, "du_prev")) return rc;
            HIPCHK(amr::launch_unpack_face(c, c.Dup[d], du_prev[d], d));
        }
    }
    if (dp) {
        if (int rc = check_cell(ctx, dp, 1, "dp")) return rc;
        HIPCHK(amr::launch_unpack_cell_shift(c, c.Dp, *dp, 0, 1, nullptr));
    }
    if (dlid) HIPCHK(hipMemcpyAsync(dlid, c.Dlid, sizeof(double) * c.g.n, hipMemcpyDeviceToDevice, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_step_tlm(amressif_ctx *ctx, const amressif_tape *tape, int first, int nsteps, amressif_lin_stats *stats)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_step_tlm: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    if (int rc = check_tape(ctx, tape)) return rc;
    if (first < 0 || nsteps < 0 || first + nsteps > tape->count)
        return fail(ctx, AMRESSIF_ERR_ARG, "steps outside the tape");
    Ctx &c = ctx->c;
    amr::AdjWork &w = *c.adj;
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                               c.in.mg_max_iter));
    HIPCHK(amr::launch_ctl_reset_step(c));
    LinPhi keep;
    if (int rc = lin_phi_enter(ctx, keep)) return rc;
    for (int s = 0; s < nsteps; ++s) {
        amressif_lin_stats st{};
        const double *ustar = tape->slot(first + s);
        w.base_u = ustar;
        w.base_v = ustar + tape->fsz;
        HIPCHK(amr::launch_cells(c, w.base_u, w.base_v, w.XU, w.XV, 0, nullptr, 0));
        // du* from (A_u - 1.5/dt) du* = G dp - dc - A_g dg
        HIPCHK(amr::launch_tl_rhs(c, w, c.Du[0], c.Du[1], c.Dup[0], c.Dup[1], c.Dp, w.L[0][0], w.L[0][1]));
        double *xu = nullptr, *xv = nullptr;
        if (int rc = lin_solve(ctx, false, lid_carries_g(c) ? c.Dlid : nullptr, &xu, &xv, &st)) return rc;
        // projection of du*: the step's own Poisson RHS, solve and update
        int nb = 0;
        HIPCHK(amr::launch_poisson_rhs(c, xu, xv, c.b, &nb));
        HIPCHK(amr::launch_fin_prhs(c, nb));
        HIPCHK(amr::launch_center_b(c, c.b));
        if (int rc = poisson_loop(ctx)) return rc;
        std::swap(c.Du[0], c.Dup[0]);  // du_prev <- du; du is rewritten below
        std::swap(c.Du[1], c.Dup[1]);
        HIPCHK(amr::launch_project_ex(c, xu, xv, c.Du[0], c.Du[1], c.Dp, c.phi, 1, &nb));
        st.vcycles = c.ctl_host->mg_iters;
        st.flags = c.ctl_host->flags;
        if (stats) stats[s] = st;
    }
    if (int rc = lin_phi_leave(ctx, keep)) return rc;
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amressif_step_bwd(amressif_ctx *ctx, amressif_tape *tape, int nsteps, amressif_lin_stats *stats)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_step_bwd: not available for 3D or periodic inputs");
    if (!ctx) return AMRESSIF_ERR_ARG;
    if (int rc = check_tape(ctx, tape)) return rc;
    if (nsteps < 0 || nsteps > tape->count) return fail(ctx, AMRESSIF_ERR_ARG, "more steps than the tape holds");
    Ctx &c = ctx->c;
    amr::AdjWork &w = *c.adj;
    HIPCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                               c.in.mg_max_iter));
    HIPCHK(amr::launch_ctl_reset_step(c));
    LinPhi keep;
    if (int rc = lin_phi_enter(ctx, keep)) return rc;
    for (int s = 0; s < nsteps; ++s) {
        amressif_lin_stats st{};
        const double *ustar = tape->slot(tape->count - 1);
        w.base_u = ustar;
        w.base_v = ustar + tape->fsz;
        HIPCHK(amr::launch_cells(c, w.base_u, w.base_v, w.XU, w.XV, 0, nullptr, 0));
        // projection: phi_bar = C p_bar + (dt/1.5) D u_bar, b_bar = L^+ phi_bar (same solve)
        HIPCHK(amr::launch_adj_phibar(c, c.Du[0], c.Du[1], c.Dp, c.b));
        HIPCHK(amr::launch_mean_b(c, c.b));
        HIPCHK(amr::launch_center_b(c, c.b));
        if (int rc = poisson_loop(ctx)) return rc;
        st.vcycles = c.ctl_host->mg_iters;
        // momentum: (A_u - 1.5/dt)^T s = u_bar - (1.5/dt) G b_bar
        HIPCHK(amr::launch_adj_rhs(c, w, c.Du[0], c.Du[1], c.phi, w.L[0][0], w.L[0][1]));
        double *su = nullptr, *sv = nullptr;
        if (int rc = lin_solve(ctx, true, nullptr, &su, &sv, &st)) return rc;
        if (lid_carries_g(c)) HIPCHK(amr::launch_adj_lid(c, w, su, sv, c.Dlid));
        HIPCHK(amr::launch_adj_finish(c, su, sv, c.Du[0], c.Du[1], c.Dup[0], c.Dup[1], c.Dp));
        st.flags = c.ctl_host->flags;
        tape->count -= 1;
        if (stats) stats[s] = st;
    }
    if (int rc = lin_phi_leave(ctx, keep)) return rc;
    HIPCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

}  // extern "C"
