// Host side of the general path (gen.hpp): context set-up, state load/store
// and the time step on n^dim cells (dim 2 or 3) with periodic directions.
// The C-ABI entry points of api.cpp hand a context to these functions when
// its inputs are 3D or name a periodic direction.
//
// The step follows main.cpp:875-1202 as the single-level path does: the BDF2
// history with -G p folded into the constant, the RK4 dual-time loop (or the
// explicit update of PSEUDO_TIMESTEPPING = 0) under device-side control with
// one host poll per batch of pseudo-iterations, the Poisson RHS, the V-cycle
// iteration to ||r||inf <= tol_rel ||b||inf (poisson.cpp:41-186; one poll per
// batch of cycles), and the projection with p += phi - mean(phi).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <cstdlib>

#include "gen.hpp"

namespace {

using amr::Ctx;
namespace G = amr::gen;

int gfail(amressif_ctx *ctx, int code, const std::string &msg) { return amr_fail(ctx, code, msg); }

#define GCHK(expr)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return gfail(ctx, AMRESSIF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t alloc(Ctx &c, T **ptr, size_t count)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
    if (e != hipSuccess) return e;
    c.allocs.push_back(p);
    c.bytes += (int64_t)(count * sizeof(T) + 256);
    *ptr = (T *)p;
    return hipSuccess;
}

// cont2cart's wall rule of one side (utilities.cpp:87-346): the kind of
// gen.hpp, or -1 where the reference aborts, -3 where it leaves the ghosts
// unfilled
int side_kind(int d, int hi, int code)
{
    if (code == 135) return 0;
    if (code == 195 || code == -195) return -1;
    if (d == 0 && !hi) {
        if (code == -135) return 1;
        if (code == 165 || code == -165) return 2;
        return -3;
    }
    if (d == 0 && hi) {
        if (code == -165) return -1;
        if (code == 1) return 1;
        if (code == 165) return 2;
        return -3;
    }
    if (code == -165) return -1;
    if (code == -135) return 1;
    if (code == 165) return (d == 1 && hi) ? 3 : 2;
    return -3;
}

// pinv(L) = J - (J - L)^{-1}, J = 11^T/m, by Cholesky of J - L (SPD: L's
// null space is the constants with Neumann and periodic sides alike); stored
// transposed (column j of row i at j m + i) for the device's column reads
bool bottom_pinv(const G::Geom &g, const G::Level &L, std::vector<double> &out)
{
    const int n = L.n, dim = g.dim;
    const int m = (int)L.ncell;
    std::vector<double> A((size_t)m * m, 0.0);
    auto at = [&](int i, int j) -> double & { return A[(size_t)i * m + j]; };
    for (int q = 0; q < m; ++q) {
        const int x[3] = {q % n, (q / n) % n, dim == 3 ? q / (n * n) : 0};
        int st = 1;
        for (int d = 0; d < dim; ++d) {
            for (int s = -1; s <= 1; s += 2) {
                int y = x[d] + s;
                if (y < 0 || y >= n) {
                    if (!g.per[d]) continue;  // Neumann: the mirror ghost is the cell itself
                    y = (y + n) % n;
                }
                const int nb = q + (y - x[d]) * st;
                at(q, nb) += L.a[d];
                at(q, q) -= L.a[d];
            }
            st *= n;
        }
    }
    // J - L
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) at(i, j) = 1.0 / m - at(i, j);
    for (int j = 0; j < m; ++j) {
        double s = at(j, j);
        for (int k = 0; k < j; ++k) s -= at(j, k) * at(j, k);
        if (!(s > 0)) return false;
        const double l = std::sqrt(s);
        at(j, j) = l;
        for (int i = j + 1; i < m; ++i) {
            double t = at(i, j);
            for (int k = 0; k < j; ++k) t -= at(i, k) * at(j, k);
            at(i, j) = t / l;
        }
    }
    out.assign((size_t)m * m, 0.0);
    std::vector<double> col(m);
    for (int j = 0; j < m; ++j) {
        std::fill(col.begin(), col.end(), 0.0);
        col[j] = 1.0;
        for (int i = 0; i < m; ++i) {
            double t = col[i];
            for (int k = 0; k < i; ++k) t -= at(i, k) * col[k];
            col[i] = t / at(i, i);
        }
        for (int i = m - 1; i >= 0; --i) {
            double t = col[i];
            for (int k = i + 1; k < m; ++k) t -= at(k, i) * col[k];
            col[i] = t / at(i, i);
        }
        // pinv(i, j) for every i, stored at j m + i
        for (int i = 0; i < m; ++i) out[(size_t)j * m + i] = 1.0 / m - col[i];
    }
    return true;
}

int poll(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    GCHK(hipMemcpyAsync(c.ctl_host, c.ctl, sizeof(amr::Ctl), hipMemcpyDeviceToHost, c.stream));
    GCHK(hipStreamSynchronize(c.stream));
    return 0;
}

int check_view3(amressif_ctx *ctx, const amr_array4 *a, const int *lo, const int *hi, int ncomp, const char *what)
{
    if (!a || !a->ptr) return gfail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": null view");
    if (a->dtype != 0) return gfail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": double views only");
    for (int d = 0; d < 3; ++d)
        if (a->lo[d] > lo[d] || a->hi[d] < hi[d])
            return gfail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": view does not cover the box");
    if (a->ncomp < ncomp) return gfail(ctx, AMRESSIF_ERR_ARG, std::string(what) + ": too few components");
    return 0;
}

void face_ext(const G::Geom &g, int d, int *ext)
{
    for (int e = 0; e < 3; ++e) ext[e] = e < g.dim ? g.n + (e == d) : 1;
}

}  // namespace

// ================================================================== set-up
bool amr_gen_wanted(const amressif_inputs *in)
{
    if (in->spacedim == 3) return true;
    for (int d = 0; d < 2; ++d)
        if (in->phy_bc_lo[d] == 111 || in->phy_bc_hi[d] == 111) return true;
    return false;
}

int amr_gen_create(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    const amressif_inputs &in = c.in;
    const int dim = in.spacedim, n = in.n_cell;
    if (dim != 2 && dim != 3) return gfail(ctx, AMRESSIF_ERR_ARG, "spacedim must be 2 or 3");
    if (n < 4 || (dim == 3 && n > 1024)) return gfail(ctx, AMRESSIF_ERR_ARG, "n_cell out of range for the general path");
    if (in.precision != AMRESSIF_FP64)
        return gfail(ctx, AMRESSIF_ERR_UNSUPPORTED, "3D and periodic runs are FP64 only (precision = 0)");
    if (in.poisson_solver != 0)
        return gfail(ctx, AMRESSIF_ERR_UNSUPPORTED, "3D and periodic runs use the multigrid Poisson solve");
    auto *gp = new G::Ctx();
    if (const char *e = std::getenv("AMRESSIF_GEN_FLUXZ")) gp->flux_z = std::atoi(e) ? 1 : 0;
    if (const char *e = std::getenv("AMRESSIF_GEN_FZC")) gp->flux_zc = std::max(1, std::min(1024, std::atoi(e)));
    c.gen = gp;
    G::Ctx &gc = *gp;
    G::Geom &g = gc.g;
    g.dim = dim;
    g.n = n;
    g.nk = dim == 3 ? n : 1;
    // BCRec mapping (main.cpp:810-835) and the cont2cart branches per side
    for (int d = 0; d < dim; ++d) {
        const int lo = in.phy_bc_lo[d], hi = in.phy_bc_hi[d];
        if ((lo == 111) != (hi == 111))
            return gfail(ctx, AMRESSIF_ERR_UNSUPPORTED, "a periodic direction needs 111 on both sides");
        if (!(lo == 111 || std::abs(lo) == 135 || std::abs(lo) == 165 || std::abs(lo) == 195))
            return gfail(ctx, AMRESSIF_ERR_BC, "Invalid bc_lo");
        // the reference's hi-side test reads phy_bc_lo for the 135 case (main.cpp:827)
        if (!(hi == 111 || std::abs(lo) == 135 || std::abs(hi) == 165 || std::abs(hi) == 195))
            return gfail(ctx, AMRESSIF_ERR_BC, "Invalid bc_hi");
        g.per[d] = lo == 111;
        if (!g.per[d])
            for (int side = 0; side < 2; ++side) {
                const int k = side_kind(d, side, side ? hi : lo);
                if (k == -1) return gfail(ctx, AMRESSIF_ERR_BC, "boundary code: Not yet implemented");
                if (k == -3)
                    return gfail(ctx, AMRESSIF_ERR_UNSUPPORTED,
                                 "boundary code leaves the reference's ghost cells unfilled (utilities.cpp:87-346)");
                (side ? g.kind_hi : g.kind_lo)[d] = k;
            }
        g.inflow[d] = in.inflow_waveform[d];
        const double dx = (in.hi_phy_dim[d] - in.lo_phy_dim[d]) / n;
        if (!(dx > 0)) return gfail(ctx, AMRESSIF_ERR_ARG, "hi_phy_dim must exceed lo_phy_dim");
        g.idx[d] = 1.0 / dx;
        g.idx2[d] = 1.0 / (dx * dx);
    }
    g.dt = amressif_dt(&in);
    g.c15 = 1.5 / g.dt;
    g.iren = 1.0 / in.ren;
    g.pseudo = in.PSEUDO_TIMESTEPPING;
    g.ncell = (int64_t)n * n * g.nk;
    g.xs = (int64_t)(n + 4) * (n + 4) * (dim == 3 ? n + 4 : 1);
    const double dtau = 0.4 * g.dt;  // main.cpp:861-869
    gc.rk[0] = dtau * 0.25;
    gc.rk[1] = dtau * (1.0 / 3.0);
    gc.rk[2] = dtau * 0.5;
    gc.rk[3] = dtau * 1.0;
    for (int d = 0; d < dim; ++d) {
        int ext[3];
        face_ext(g, d, ext);
        g.nf[d] = (int64_t)ext[0] * ext[1] * ext[2];
        gc.nb_face[d] = (int)((g.nf[d] + G::kBlock - 1) / G::kBlock);
    }
    // multigrid hierarchy: halve while the side is even and the level holds
    // more than 512 cells; the coarsest level's dense pseudo-inverse
    {
        int nl = n, l = 0;
        for (;;) {
            G::Level &L = gc.lev[l];
            L.n = nl;
            L.ncell = (int64_t)nl * nl * (dim == 3 ? nl : 1);
            for (int d = 0; d < 3; ++d) {
                const double h = d < dim ? (in.hi_phy_dim[d] - in.lo_phy_dim[d]) / nl : 1.0;
                L.a[d] = 1.0 / (h * h);
            }
            ++l;
            if ((nl & 1) || L.ncell <= 512 || l == G::kMaxGenLevels) break;
            nl /= 2;
        }
        gc.nlev = l;
        gc.mb = (int)gc.lev[l - 1].ncell;
        if (gc.lev[l - 1].ncell > 1024)
            return gfail(ctx, AMRESSIF_ERR_UNSUPPORTED,
                         "coarsest multigrid level above 1024 cells (n_cell = m 2^k with m^dim > 1024, m odd)");
    }
    const size_t cellsz = (size_t)g.ncell;
#define GALLOC(p, cnt)                                                                     \
    do {                                                                                   \
        hipError_t e_ = alloc(c, p, cnt);                                                  \
        if (e_ != hipSuccess) return gfail(ctx, AMRESSIF_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)
    for (int d = 0; d < dim; ++d) {
        GALLOC(&gc.U[d], g.nf[d]); GALLOC(&gc.Up[d], g.nf[d]); GALLOC(&gc.C[d], g.nf[d]);
        GALLOC(&gc.B[d], g.nf[d]); GALLOC(&gc.S[d], g.nf[d]);
    }
    GALLOC(&gc.X, (size_t)g.xs * dim);
    GALLOC(&gc.T, cellsz * dim);
    GALLOC(&gc.p, cellsz); GALLOC(&gc.phi, cellsz); GALLOC(&gc.b, cellsz);
    for (int l = 0; l < gc.nlev; ++l) {
        GALLOC(&gc.lev[l].e, (size_t)gc.lev[l].ncell);
        GALLOC(&gc.lev[l].r, (size_t)gc.lev[l].ncell);
    }
    gc.npartials = std::max<int64_t>(2 * ((g.ncell + G::kBlock - 1) / G::kBlock),
                                     (int64_t)gc.nb_face[0] + gc.nb_face[1] + gc.nb_face[2]) + 64;
    GALLOC(&gc.partials, (size_t)gc.npartials);
    {
        std::vector<double> pinv;
        if (!bottom_pinv(g, gc.lev[gc.nlev - 1], pinv))
            return gfail(ctx, AMRESSIF_ERR_ARG, "bottom factorisation failed");
        GALLOC(&gc.pinv, pinv.size());
        GCHK(hipMemcpy(gc.pinv, pinv.data(), pinv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    GALLOC(&c.ctl, 1);
    {
        hipError_t e = hipHostMalloc((void **)&c.ctl_host, sizeof(amr::Ctl), hipHostMallocDefault);
        if (e != hipSuccess) return gfail(ctx, AMRESSIF_ERR_HIP, hipGetErrorString(e));
        std::memset(c.ctl_host, 0, sizeof(amr::Ctl));
        GCHK(hipMemcpy(c.ctl, c.ctl_host, sizeof(amr::Ctl), hipMemcpyHostToDevice));
    }
#undef GALLOC
    // a cavity at rest until a state is loaded
    for (int d = 0; d < dim; ++d) {
        GCHK(G::fill(c.stream, gc.U[d], g.nf[d], 0.0));
        GCHK(G::fill(c.stream, gc.Up[d], g.nf[d], 0.0));
    }
    GCHK(G::fill(c.stream, gc.p, g.ncell, 0.0));
    GCHK(G::fill(c.stream, gc.phi, g.ncell, 0.0));
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

// ================================================================== state
int amr_gen_load_state(amressif_ctx *ctx, const amr_array4 *vel, const amr_array4 *velPrev, const amr_array4 *userCtx)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    if (!vel || !velPrev) return gfail(ctx, AMRESSIF_ERR_ARG, "null argument");
    for (int d = 0; d < g.dim; ++d) {
        int ext[3], lo[3] = {0, 0, 0}, hi[3];
        face_ext(g, d, ext);
        for (int e = 0; e < 3; ++e) hi[e] = ext[e] - 1;
        if (int rc = check_view3(ctx, &vel[d], lo, hi, 1, "velCont")) return rc;
        if (int rc = check_view3(ctx, &velPrev[d], lo, hi, 1, "velContPrev")) return rc;
        GCHK(G::pack(c.stream, vel[d], 0, gc.U[d], ext[0], ext[1], ext[2]));
        GCHK(G::pack(c.stream, velPrev[d], 0, gc.Up[d], ext[0], ext[1], ext[2]));
    }
    if (userCtx) {
        const int lo[3] = {0, 0, 0}, hi[3] = {g.n - 1, g.n - 1, g.nk - 1};
        if (int rc = check_view3(ctx, userCtx, lo, hi, 2, "userCtx")) return rc;
        GCHK(G::pack(c.stream, *userCtx, 0, gc.p, g.n, g.n, g.nk));
        GCHK(G::pack(c.stream, *userCtx, 1, gc.phi, g.n, g.n, g.nk));
    } else {
        GCHK(G::fill(c.stream, gc.p, g.ncell, 0.0));
        GCHK(G::fill(c.stream, gc.phi, g.ncell, 0.0));
    }
    c.pred_valid = 0;
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amr_gen_store_state(amressif_ctx *ctx, amr_array4 *vel, amr_array4 *velPrev, amr_array4 *userCtx,
                        amr_array4 *velCart)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    for (int d = 0; d < g.dim; ++d) {
        int ext[3], lo[3] = {0, 0, 0}, hi[3];
        face_ext(g, d, ext);
        for (int e = 0; e < 3; ++e) hi[e] = ext[e] - 1;
        if (vel) {
            if (int rc = check_view3(ctx, &vel[d], lo, hi, 1, "velCont")) return rc;
            GCHK(G::unpack(c.stream, gc.U[d], vel[d], 0, ext[0], ext[1], ext[2]));
        }
        if (velPrev) {
            if (int rc = check_view3(ctx, &velPrev[d], lo, hi, 1, "velContPrev")) return rc;
            GCHK(G::unpack(c.stream, gc.Up[d], velPrev[d], 0, ext[0], ext[1], ext[2]));
        }
    }
    if (userCtx) {
        const int lo[3] = {0, 0, 0}, hi[3] = {g.n - 1, g.n - 1, g.nk - 1};
        if (int rc = check_view3(ctx, userCtx, lo, hi, 2, "userCtx")) return rc;
        for (int d = 0; d < 3; ++d)
            if ((d < g.dim && (userCtx->lo[d] < -1 || userCtx->hi[d] > g.n)) ||
                (d >= g.dim && (userCtx->lo[d] != 0 || userCtx->hi[d] != 0)))
                return gfail(ctx, AMRESSIF_ERR_ARG, "userCtx: at most one ghost layer");
        GCHK(G::phi_mean(gc, c.stream, c.ctl));
        GCHK(G::unpack_cells_ghosted(gc, c.stream, gc.p, *userCtx, 0, nullptr));
        GCHK(G::unpack_cells_ghosted(gc, c.stream, gc.phi, *userCtx, 1, &c.ctl->phi_mean));
    }
    if (velCart) {
        const int lo[3] = {0, 0, 0}, hi[3] = {g.n - 1, g.n - 1, g.nk - 1};
        if (int rc = check_view3(ctx, velCart, lo, hi, g.dim, "velCart")) return rc;
        for (int d = 0; d < 3; ++d)
            if ((d < g.dim && (velCart->lo[d] < -2 || velCart->hi[d] > g.n + 1)) ||
                (d >= g.dim && (velCart->lo[d] != 0 || velCart->hi[d] != 0)))
                return gfail(ctx, AMRESSIF_ERR_ARG, "velCart: at most two ghost layers");
        GCHK(G::cells(gc, c.stream, gc.U, gc.X, nullptr));
        GCHK(G::unpack_velcart(gc, c.stream, gc.X, *velCart));
    }
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

// ---- MultiFabs of several boxes (amressif_load_state_mf / _store_state_mf)
namespace {

struct Box3 {
    int lo[3], hi[3];
};

// The valid boxes of a MultiFab (nodal: the face direction, or -1 for cells):
// inside the domain, pairwise disjoint and covering it; faces also returns
// the boxes' face windows (the cell box with hi + 1 along the nodal direction)
int check_mf3(amressif_ctx *ctx, const amr_multifab *mf, int nodal, int ncomp, int max_grow, const char *what,
              std::vector<Box3> &win)
{
    const G::Geom &g = ctx->c.gen->g;
    const std::string w(what);
    if (!mf || mf->nbox < 1 || !mf->fab) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": empty MultiFab");
    if (mf->ngrow < 0 || mf->ngrow > max_grow)
        return gfail(ctx, AMRESSIF_ERR_ARG, w + ": ngrow must be in [0, " + std::to_string(max_grow) + "]");
    const int ext[3] = {g.n, g.n, g.nk};
    std::vector<Box3> cells;
    win.clear();
    int64_t vol = 0;
    for (int b = 0; b < mf->nbox; ++b) {
        const amr_array4 &a = mf->fab[b];
        if (!a.ptr) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": null fab");
        if (a.dtype != 0) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": double views only");
        if (a.ncomp < ncomp) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": too few components");
        Box3 c{};
        for (int d = 0; d < 3; ++d) {
            const int gr = d < g.dim ? mf->ngrow : 0;
            c.lo[d] = a.lo[d] + gr;
            c.hi[d] = a.hi[d] - gr - (d == nodal ? 1 : 0);
            if (c.lo[d] < 0 || c.hi[d] >= ext[d] || c.hi[d] < c.lo[d])
                return gfail(ctx, AMRESSIF_ERR_ARG, w + ": a valid box lies outside the domain");
        }
        for (const Box3 &o : cells) {
            bool disjoint = false;
            for (int d = 0; d < 3; ++d) disjoint = disjoint || c.hi[d] < o.lo[d] || o.hi[d] < c.lo[d];
            if (!disjoint) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": valid boxes overlap");
        }
        vol += (int64_t)(c.hi[0] - c.lo[0] + 1) * (c.hi[1] - c.lo[1] + 1) * (c.hi[2] - c.lo[2] + 1);
        cells.push_back(c);
        Box3 f = c;
        if (nodal >= 0) f.hi[nodal] += 1;
        win.push_back(f);
    }
    if (vol != g.ncell) return gfail(ctx, AMRESSIF_ERR_ARG, w + ": valid boxes do not cover the domain");
    return 0;
}

}  // namespace

int amr_gen_load_state_mf(amressif_ctx *ctx, const amr_multifab *vel, const amr_multifab *velPrev,
                          const amr_multifab *userCtx)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    if (!vel || !velPrev) return gfail(ctx, AMRESSIF_ERR_ARG, "null argument");
    // validate everything before the first copy: a failed load leaves the state as it was
    std::vector<Box3> bx[7];
    for (int d = 0; d < g.dim; ++d) {
        if (int rc = check_mf3(ctx, &vel[d], d, 1, 0, "velCont", bx[2 * d])) return rc;
        if (int rc = check_mf3(ctx, &velPrev[d], d, 1, 0, "velContPrev", bx[2 * d + 1])) return rc;
    }
    if (userCtx)
        if (int rc = check_mf3(ctx, userCtx, -1, 2, 1 << 20, "userCtx", bx[6])) return rc;
    for (int d = 0; d < g.dim; ++d) {
        int ext[3];
        face_ext(g, d, ext);
        for (int k = 0; k < 2; ++k) {
            const amr_multifab &mf = k ? velPrev[d] : vel[d];
            double *dst = k ? gc.Up[d] : gc.U[d];
            for (int b = 0; b < mf.nbox; ++b) {
                const Box3 &q = bx[2 * d + k][b];
                GCHK(G::pack_window(c.stream, mf.fab[b], 0, dst, ext[0], ext[1], q.lo, q.hi, 0));
            }
        }
    }
    if (userCtx) {
        for (int b = 0; b < userCtx->nbox; ++b) {
            const Box3 &q = bx[6][b];
            GCHK(G::pack_window(c.stream, userCtx->fab[b], 0, gc.p, g.n, g.n, q.lo, q.hi, 0));
            GCHK(G::pack_window(c.stream, userCtx->fab[b], 1, gc.phi, g.n, g.n, q.lo, q.hi, 0));
        }
    } else {
        GCHK(G::fill(c.stream, gc.p, g.ncell, 0.0));
        GCHK(G::fill(c.stream, gc.phi, g.ncell, 0.0));
    }
    c.pred_valid = 0;
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amr_gen_store_state_mf(amressif_ctx *ctx, amr_multifab *vel, amr_multifab *velPrev, amr_multifab *userCtx,
                           amr_multifab *velCart)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    std::vector<Box3> bx[6], cb;
    for (int d = 0; d < g.dim; ++d) {
        if (vel)
            if (int rc = check_mf3(ctx, &vel[d], d, 1, 0, "velCont", bx[2 * d])) return rc;
        if (velPrev)
            if (int rc = check_mf3(ctx, &velPrev[d], d, 1, 0, "velContPrev", bx[2 * d + 1])) return rc;
    }
    if (userCtx)
        if (int rc = check_mf3(ctx, userCtx, -1, 2, 1, "userCtx", cb)) return rc;
    if (velCart)
        if (int rc = check_mf3(ctx, velCart, -1, g.dim, 2, "velCart", cb)) return rc;
    for (int d = 0; d < g.dim; ++d) {
        int ext[3];
        face_ext(g, d, ext);
        for (int k = 0; k < 2; ++k) {
            amr_multifab *mf = k ? velPrev : vel;
            if (!mf) continue;
            const double *src = k ? gc.Up[d] : gc.U[d];
            for (int b = 0; b < mf[d].nbox; ++b) {
                const Box3 &q = bx[2 * d + k][b];
                GCHK(G::pack_window(c.stream, mf[d].fab[b], 0, const_cast<double *>(src), ext[0], ext[1], q.lo, q.hi,
                                    1));
            }
        }
    }
    // cell views are written whole, ghosts by global index: the neighbouring
    // box's values inside the domain, the mirror or periodic image outside it
    if (userCtx) {
        GCHK(G::phi_mean(gc, c.stream, c.ctl));
        for (int b = 0; b < userCtx->nbox; ++b) {
            GCHK(G::unpack_cells_ghosted(gc, c.stream, gc.p, userCtx->fab[b], 0, nullptr));
            GCHK(G::unpack_cells_ghosted(gc, c.stream, gc.phi, userCtx->fab[b], 1, &c.ctl->phi_mean));
        }
    }
    if (velCart) {
        GCHK(G::cells(gc, c.stream, gc.U, gc.X, nullptr));
        for (int b = 0; b < velCart->nbox; ++b) GCHK(G::unpack_velcart(gc, c.stream, gc.X, velCart->fab[b]));
    }
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

int amr_gen_init_state(amressif_ctx *ctx)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    for (int d = 0; d < g.dim; ++d) {
        GCHK(G::fill(c.stream, gc.U[d], g.nf[d], 0.0));
        GCHK(G::fill(c.stream, gc.Up[d], g.nf[d], 0.0));
    }
    GCHK(G::fill(c.stream, gc.p, g.ncell, 0.0));
    GCHK(G::fill(c.stream, gc.phi, g.ncell, 0.0));
    c.pred_valid = 0;
    GCHK(hipStreamSynchronize(c.stream));
    return AMRESSIF_OK;
}

// ================================================================== step
int amr_gen_advance(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    if (nsteps < 0) return gfail(ctx, AMRESSIF_ERR_ARG, "bad argument");
    if (nsteps == 0) return AMRESSIF_OK;
    if (c.log_cap < nsteps) {
        if (c.log) GCHK(hipFree(c.log));
        c.log = nullptr;
        GCHK(hipMalloc((void **)&c.log, sizeof(amressif_step_stats) * (size_t)nsteps));
        c.log_cap = nsteps;
    }
    GCHK(amr::launch_ctl_set(c, c.in.momentum_tolerance, c.in.poisson_tol_rel, c.in.poisson_tol_abs, c.in.IterNum,
                             c.in.mg_max_iter));
    int status = AMRESSIF_OK;
    for (int s = 0; s < nsteps; ++s) {
        GCHK(G::bdf2(gc, c.stream));
        GCHK(amr::launch_ctl_reset_step(c));
        // ---- momentum predictor
        if (!g.pseudo) {
            GCHK(G::cells(gc, c.stream, gc.S, gc.X, c.ctl));
            GCHK(G::flux(gc, c.stream, gc.S, gc.X, gc.T, c.ctl));
            GCHK(G::update(gc, c.stream, g.dt / 1.5, 1, c.ctl));
            GCHK(G::fin_pseudo(gc, c.stream, c.ctl));
        } else {
            int done = 0, batch = std::max(1, gc.last_P);
            while (!done) {
                for (int it = 0; it < batch; ++it) {
                    for (int sub = 0; sub < 4; ++sub) {
                        GCHK(G::cells(gc, c.stream, gc.S, gc.X, c.ctl));
                        GCHK(G::flux(gc, c.stream, gc.S, gc.X, gc.T, c.ctl));
                        GCHK(G::update(gc, c.stream, gc.rk[sub], sub == 3, c.ctl));
                    }
                    GCHK(G::fin_pseudo(gc, c.stream, c.ctl));
                }
                if (int rc = poll(ctx)) return rc;
                done = c.ctl_host->mom_done;
                batch = std::max(2, batch / 2);
            }
            gc.last_P = std::max(1, c.ctl_host->mom_count);
        }
        // ---- Poisson: b = (1.5/dt) D(u*), its mean removed in the residual
        int nb = 0;
        GCHK(G::divergence(gc, c.stream, gc.S, gc.b, &nb));
        GCHK(G::fin_div(gc, c.stream, nb, c.ctl, 0));
        GCHK(G::mg_start(gc, c.stream, c.ctl));
        GCHK(G::fill(c.stream, gc.phi, g.ncell, 0.0));  // phi0 = 0 (poisson.cpp:63)
        {
            int done = 0, batch = std::max(1, gc.last_V);
            while (!done) {
                for (int it = 0; it < batch; ++it) {
                    int nr = 0;
                    GCHK(G::resid0(gc, c.stream, c.ctl, &nr));
                    GCHK(G::fin_resid(gc, c.stream, nr, c.ctl));
                    GCHK(G::vcycle(gc, c.stream, c.ctl));
                }
                if (int rc = poll(ctx)) return rc;
                done = c.ctl_host->mg_done;
                batch = 2;
            }
            gc.last_V = std::max(1, c.ctl_host->mg_iters + 1);
        }
        // ---- projection, p += phi - mean, divergence of u^{n+1}
        GCHK(G::project(gc, c.stream, c.ctl));
        GCHK(G::divergence(gc, c.stream, gc.U, nullptr, &nb));
        GCHK(G::fin_div(gc, c.stream, nb, c.ctl, 1));
        GCHK(G::fin_step(gc, c.stream, c.ctl, c.log + s));
        if (c.ctl_host->flags & AMRESSIF_FLAG_NONFINITE) {
            status = gfail(ctx, AMRESSIF_ERR_NONFINITE, "non-finite pseudo-iteration norm at step " + std::to_string(s));
            nsteps = s + 1;
            break;
        }
        c.pred_valid = 1;
    }
    if (stats)
        GCHK(hipMemcpyAsync(stats, c.log, sizeof(amressif_step_stats) * (size_t)nsteps, hipMemcpyDeviceToHost,
                            c.stream));
    GCHK(hipStreamSynchronize(c.stream));
    return status;
}

// ================================================================== timing
// which: 0 cells, 1 flux, 2 stage update (the three passes of one stage)
int amr_gen_time_kernel(amressif_ctx *ctx, int which, int reps, double *avg_ms, double *bytes_per_launch)
{
    Ctx &c = ctx->c;
    G::Ctx &gc = *c.gen;
    const G::Geom &g = gc.g;
    if (which < 0 || which > 2) return gfail(ctx, AMRESSIF_ERR_ARG, "unknown general-path kernel");
    hipEvent_t e0, e1;
    GCHK(hipEventCreate(&e0));
    GCHK(hipEventCreate(&e1));
    // a realistic state: S = U, X and T of it
    for (int d = 0; d < g.dim; ++d) {
        GCHK(hipMemcpyAsync(gc.S[d], gc.U[d], g.nf[d] * 8, hipMemcpyDeviceToDevice, c.stream));
        GCHK(hipMemcpyAsync(gc.B[d], gc.U[d], g.nf[d] * 8, hipMemcpyDeviceToDevice, c.stream));
    }
    GCHK(G::bdf2(gc, c.stream));
    GCHK(G::cells(gc, c.stream, gc.S, gc.X, nullptr));
    GCHK(G::flux(gc, c.stream, gc.S, gc.X, gc.T, nullptr));
    auto launch = [&]() -> hipError_t {
        if (which == 0) return G::cells(gc, c.stream, gc.S, gc.X, nullptr);
        if (which == 1) return G::flux(gc, c.stream, gc.S, gc.X, gc.T, nullptr);
        return G::update(gc, c.stream, 0.0, 0, nullptr);  // rk = 0: the faces keep their values
    };
    GCHK(launch());
    GCHK(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; ++r) GCHK(launch());
    GCHK(hipEventRecord(e1, c.stream));
    GCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    GCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    // algorithmic bytes (FP64), each array read or written once per launch:
    // cells: dim face arrays in, dim ghosted cell comps out; flux: dim face
    // arrays and X in, T out; update: T, u*, u^k, C in, u* out
    const double faces = (double)(g.nf[0] + g.nf[1] + (g.dim == 3 ? g.nf[2] : 0));
    const double cellsd = (double)g.ncell * g.dim;
    double bytes = 0.0;
    if (which == 0) bytes = 8.0 * (faces + (double)g.xs * g.dim);
    else if (which == 1) bytes = 8.0 * (faces + (double)g.xs * g.dim + cellsd);
    else bytes = 8.0 * (cellsd + 4.0 * faces);
    if (bytes_per_launch) *bytes_per_launch = bytes;
    c.pred_valid = 0;
    return AMRESSIF_OK;
}
