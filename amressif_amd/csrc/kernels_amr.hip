// Two-level composite step (coarse level + one refined patch, ratio 2), gfx950.
//
// The reference is single-level; this is the two-level forward step that its
// AMReX layer would host.  Geometry: the coarse level is the whole n0 x n0
// domain; the fine patch covers the coarse rows [jc0, n0) at twice the
// resolution (nx = 2 n0, ny = 2 (n0 - jc0)), so its west, east and north sides
// are walls and its south side is the coarse/fine (C/F) interface y = y_I.
//
// Fine-level momentum predictor (FP64): the single-level stage of
// fn_flux_calc.cpp / fn_rhs.cpp / momentum.cpp on the patch, with
//   * wall ghosts by the cont2cart rules (utilities.cpp:87-269) on W, E, N;
//   * south ghosts (rows -1, -2) from the converged coarse predictor:
//     linear in y through coarse rows jc0-1 and jc0-2, and in x through the
//     coarse slope (one-sided at the walls);
//   * interface faces v(i, 0) fixed to the coarse faces v0(., jc0),
//     interpolated linearly in x so that every coarse interface face equals
//     the mean of the two fine faces on it.
// Composite projection: L_c phi = (1.5/dt) D_c u* on the uncovered coarse
// cells and the fine cells, Neumann walls.  Fine cells see a south ghost
//   phi(i, -1) = 8/15 phi_c(x_i) + 2/3 phi(i, 0) - 1/5 phi(i, 1)
// (quadratic through the coarse centre and the first two fine centres,
// phi_c interpolated in x as above), and the coarse cell under the interface
// takes the mean of the two fine interface gradients as its north gradient:
// the flux matches across the interface, so the composite divergence of
// u - (dt/1.5) G_c phi is D_c u - (dt/1.5) L_c phi exactly.
#include "device_common.hpp"
#include "internal.hpp"

namespace amr {


// ------------------------------------------------ coarse-to-fine values
__device__ __forceinline__ double cell_u(const double *u, int P, int i, int j)
{
    const size_t k = i + (size_t)P * j;
    return 0.5 * (u[k] + u[k + 1]);
}
__device__ __forceinline__ double cell_v(const double *v, int P, int i, int j)
{
    const size_t k = i + (size_t)P * j;
    return 0.5 * (v[k] + v[k + P]);
}

// value of a coarse row-jc cell field f(ic, jc) at the centre x of fine column i
template <class F>
__device__ __forceinline__ double tangential(const F &f, int n0, int i, int jc)
{
    const int ic = i >> 1;
    const double sgn = (i & 1) ? 1.0 : -1.0;
    double d;
    if (ic == 0) d = f(1, jc) - f(0, jc);
    else if (ic == n0 - 1) d = f(n0 - 1, jc) - f(n0 - 2, jc);
    else d = 0.5 * (f(ic + 1, jc) - f(ic - 1, jc));
    return f(ic, jc) + sgn * 0.25 * d;
}

// ---------------------------------------------- fine momentum predictor
// cell velocities of the patch with two ghost layers
__global__ __launch_bounds__(256) void k_fine_cells(AmrGeom a, const double *u, const double *v, const double *u0,
                                                    const double *v0, double *XU, double *XV)
{
    const int gi = blockIdx.x * 64 + (threadIdx.x & 63), gj = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    if (gi >= nx + 4 || gj >= ny + 4) return;
    const int i = gi - 2, j = gj - 2;
    const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
    double U = 0.0, V = 0.0;
    if (xin && yin) {
        U = cell_u(u, a.P1, i, j);
        V = cell_v(v, a.P1, i, j);
    } else if (yin) {  // west / east wall ghosts
        const int mi = i < 0 ? -i - 1 : 2 * nx - 1 - i;
        ghost_rule<double, false>(i < 0 ? 0 : 1, i < 0 ? a.bc_w : a.bc_e, cell_u(u, a.P1, mi, j),
                                  cell_v(v, a.P1, mi, j), 0.0, 0.0, U, V);
    } else if (xin && j >= ny) {  // north wall (lid) ghosts
        const int mj = 2 * ny - 1 - j;
        ghost_rule<double, false>(3, a.bc_n, cell_u(u, a.P1, i, mj), cell_v(v, a.P1, i, mj), a.inflow0, a.inflow1,
                                  U, V);
    } else if (xin) {  // south: C/F ghosts from the coarse predictor
        auto fu = [&](int ic, int jc) { return cell_u(u0, a.P0, ic, jc); };
        auto fv = [&](int ic, int jc) { return cell_v(v0, a.P0, ic, jc); };
        const double U1 = tangential(fu, a.n0, i, a.jc0 - 1), U2 = tangential(fu, a.n0, i, a.jc0 - 2);
        const double V1 = tangential(fv, a.n0, i, a.jc0 - 1), V2 = tangential(fv, a.n0, i, a.jc0 - 2);
        const double w = j == -1 ? 0.25 : -0.25;
        U = U1 + w * (U1 - U2);
        V = V1 + w * (V1 - V2);
    }
    const size_t o = gi + (size_t)a.Pg1 * gj;
    XU[o] = U;
    XV[o] = V;
}

// interface faces v(i, 0) from the coarse faces v0(., jc0): linear in x
// through the coarse slope, so the two fine faces on a coarse face average to it
__global__ void k_fine_interface(AmrGeom a, double *v, const double *v0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nx) return;
    auto f = [&](int ic, int jc) { return v0[ic + (size_t)a.P0 * jc]; };
    v[i] = tangential(f, a.n0, i, a.jc0);
}

__device__ __forceinline__ double xg(const double *X, int Pg, int i, int j) { return X[(i + 2) + (size_t)Pg * (j + 2)]; }

// T = -div(QUICK flux) + viscous flux of one component at interior cell (i, j)
__device__ double fine_T(const AmrGeom &a, const double *X, const double *u, const double *v, int i, int j)
{
    const int Pg = a.Pg1, P = a.P1;
    auto fx = [&](int f) {  // x-face (f, j)
        return quick_flux<double>(u[f + (size_t)P * j], xg(X, Pg, f - 2, j), xg(X, Pg, f - 1, j), xg(X, Pg, f, j),
                                  xg(X, Pg, f + 1, j));
    };
    auto fy = [&](int f) {  // y-face (i, f)
        return quick_flux<double>(v[i + (size_t)P * f], xg(X, Pg, i, f - 2), xg(X, Pg, i, f - 1), xg(X, Pg, i, f),
                                  xg(X, Pg, i, f + 1));
    };
    const double ih = 1.0 / a.h1, ih2 = ih * ih;
    const double C = (fx(i + 1) - fx(i)) * ih + (fy(j + 1) - fy(j)) * ih;
    const double c = xg(X, Pg, i, j);
    const double Vs = ((xg(X, Pg, i + 1, j) - 2.0 * c + xg(X, Pg, i - 1, j)) * ih2 +
                       (xg(X, Pg, i, j + 1) - 2.0 * c + xg(X, Pg, i, j - 1)) * ih2) *
                      a.iren;
    return (0.0 - C) + Vs;
}

// one RK4 dual-time stage on the patch (mode 1: last stage, norm partials);
// wall and interface faces keep their base value
__global__ __launch_bounds__(256) void k_fine_stage(AmrGeom a, const double *us, const double *vs, const double *ub,
                                                    const double *vb, const double *cu, const double *cv,
                                                    const double *p, const double *XU, const double *XV, double *uo,
                                                    double *vo, double rk, int mode, double *partials, const Ctl *ctl)
{
    __shared__ double sh[16];
    if (ctl->mom_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny, P = a.P1;
    const double ih = 1.0 / a.h1;
    double sx = 0.0, sy = 0.0;
    if (i <= nx && j <= ny) {
        const size_t k = i + (size_t)P * j;
        // the last stage writes the base in place: read everything before the stores
        double un = 0.0, vn = 0.0, ubk = 0.0, vbk = 0.0;
        if (j < ny) {
            ubk = ub[k];
            un = ubk;
            if (i > 0 && i < nx) {
                const double R = -(p[k] - p[k - 1]) * ih +
                                 0.5 * (fine_T(a, XU, us, vs, i - 1, j) + fine_T(a, XU, us, vs, i, j));
                un = ubk + rk * (R - a.c15 * us[k] + cu[k]);
            }
            const double d = ubk - un;
            sx = fabs(d * d * ((i == 0 || i == nx) ? 0.5 : 1.0));
        }
        if (i < nx) {
            vbk = vb[k];
            vn = vbk;
            if (j > 0 && j < ny) {
                const double R = -(p[k] - p[k - P]) * ih +
                                 0.5 * (fine_T(a, XV, us, vs, i, j - 1) + fine_T(a, XV, us, vs, i, j));
                vn = vbk + rk * (R - a.c15 * vs[k] + cv[k]);
            }
            const double d = vbk - vn;
            sy = fabs(d * d * ((j == 0 || j == ny) ? 0.5 : 1.0));
        }
        if (j < ny) uo[k] = un;
        if (i < nx) vo[k] = vn;
    }
    if (mode == 1) {
        sx = block_reduce<false>(sx, sh);
        sy = block_reduce<false>(sy, sh + 8);
        if (threadIdx.x == 0) {
            const size_t b = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
            partials[2 * b] = sx;
            partials[2 * b + 1] = sy;
        }
    }
}

__global__ __launch_bounds__(1024) void k_fine_fin(const double *partials, int nb, Ctl *ctl, double npts)
{
    __shared__ double sh[32];
    if (ctl->mom_done) return;
    double sx = 0.0, sy = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { sx += partials[2 * b]; sy += partials[2 * b + 1]; }
    sx = block_reduce<false>(sx, sh);
    sy = block_reduce<false>(sy, sh + 16);
    if (threadIdx.x == 0) {
        const double ex = sqrt(sx) / sqrt(npts), ey = sqrt(sy) / sqrt(npts);
        const double ne = ex < ey ? ey : ex;
        ctl->norm_error = ne;
        ctl->mom_count += 1;
        if (!(ne == ne) || isinf(ne)) { ctl->flags |= AMRESSIF_FLAG_NONFINITE; ctl->mom_done = 1; }
        else if (ctl->mom_count > ctl->iter_num) { ctl->flags |= AMRESSIF_FLAG_CAP; ctl->mom_done = 1; }
        else if (ne > 1.e2) { ctl->flags |= AMRESSIF_FLAG_DIVERGED; ctl->mom_done = 1; }
        else if (!(ne > ctl->tol_mom)) ctl->mom_done = 1;
    }
}

__global__ void k_fine_ctl_reset(Ctl *ctl, double tol, int iter_num)
{
    ctl->mom_done = 0;
    ctl->mom_count = 0;
    ctl->flags = 0;
    ctl->norm_error = 0.0;
    ctl->tol_mom = tol;
    ctl->iter_num = iter_num;
}

// BDF2 constant and history on a face array pair
__global__ __launch_bounds__(256) void k_fine_bdf2(double *u, double *up, double *c, int nx, int ny, int P, double dt)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)P * j;
    const double un = u[k], um = up[k];
    c[k] = (1.5 / dt) * un + (0.5 / dt) * (un - um);
    up[k] = un;
}

// ------------------------------- fine momentum predictor, refinement form
// The mixed step's momentum iteration on the patch (FP64): the residual
// r = R(u^k) - (1.5/dt) u^k + c of the implicit BDF2 system, then
// u^{k+1} = u^k + cr r + cj R'(u^k) r with R' the tangent-linear flux
// operator (QUICK with the upwind choice frozen at u^k, viscous flux of r).
// Wall and interface faces keep their value (r = 0 there), and the C/F south
// ghosts depend on the coarse predictor only, so their increment is 0.

// r on every face (0 on the wall and interface faces)
__global__ __launch_bounds__(256) void k_fine_resid(AmrGeom a, const double *u, const double *v, const double *cu,
                                                    const double *cv, const double *p, const double *XU,
                                                    const double *XV, double *ru, double *rv, const Ctl *ctl)
{
    if (ctl->mom_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny, P = a.P1;
    if (i > nx || j > ny) return;
    const double ih = 1.0 / a.h1;
    const size_t k = i + (size_t)P * j;
    if (j < ny) {
        double r = 0.0;
        if (i > 0 && i < nx) {
            const double R = -(p[k] - p[k - 1]) * ih + 0.5 * (fine_T(a, XU, u, v, i - 1, j) + fine_T(a, XU, u, v, i, j));
            r = R - a.c15 * u[k] + cu[k];
        }
        ru[k] = r;
    }
    if (i < nx) {
        double r = 0.0;
        if (j > 0 && j < ny) {
            const double R = -(p[k] - p[k - P]) * ih + 0.5 * (fine_T(a, XV, u, v, i, j - 1) + fine_T(a, XV, u, v, i, j));
            r = R - a.c15 * v[k] + cv[k];
        }
        rv[k] = r;
    }
}

// cell values of r with two ghost layers: the wall rules without their
// constants, 0 below the interface
__global__ __launch_bounds__(256) void k_fine_dcells(AmrGeom a, const double *ru, const double *rv, double *DU,
                                                     double *DV, const Ctl *ctl)
{
    if (ctl->mom_done) return;
    const int gi = blockIdx.x * 64 + (threadIdx.x & 63), gj = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    if (gi >= nx + 4 || gj >= ny + 4) return;
    const int i = gi - 2, j = gj - 2;
    const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
    double U = 0.0, V = 0.0;
    if (xin && yin) {
        U = cell_u(ru, a.P1, i, j);
        V = cell_v(rv, a.P1, i, j);
    } else if (yin) {  // west / east wall ghosts
        const int mi = i < 0 ? -i - 1 : 2 * nx - 1 - i;
        ghost_rule<double, true>(i < 0 ? 0 : 1, i < 0 ? a.bc_w : a.bc_e, cell_u(ru, a.P1, mi, j),
                                 cell_v(rv, a.P1, mi, j), 0.0, 0.0, U, V);
    } else if (xin && j >= ny) {  // north wall ghosts
        const int mj = 2 * ny - 1 - j;
        ghost_rule<double, true>(3, a.bc_n, cell_u(ru, a.P1, i, mj), cell_v(rv, a.P1, i, mj), 0.0, a.inflow1, U, V);
    }
    const size_t o = gi + (size_t)a.Pg1 * gj;
    DU[o] = U;
    DV[o] = V;
}

// derivative of QUICK(vel, X) in direction (dvel, dX), the upwind choice frozen at vel
__device__ __forceinline__ double quick_tl64(double vel, double dvel, double a2, double a1, double b0, double b1,
                                             double da2, double da1, double db0, double db1)
{
    const bool pos = vel >= 0.0;
    // phi = 3/8 downwind + 6/8 upwind - 1/8 far-upwind
    const double xd = pos ? b0 : a1, xu = pos ? a1 : b0, xf = pos ? a2 : b1;
    const double dd = pos ? db0 : da1, du = pos ? da1 : db0, df = pos ? da2 : db1;
    const double phid = 0.375 * dd + 0.75 * du - 0.125 * df;
    const double phix = 0.375 * xd + 0.75 * xu - 0.125 * xf;
    return vel * phid + dvel * phix;
}

// R'(u) r of one component at interior cell (i, j): T' = -div(QUICK') + viscous flux of r
__device__ double fine_TL(const AmrGeom &a, const double *X, const double *D, const double *u, const double *v,
                          const double *ru, const double *rv, int i, int j)
{
    const int Pg = a.Pg1, P = a.P1;
    auto fx = [&](int f) {  // x-face (f, j)
        const size_t k = f + (size_t)P * j;
        return quick_tl64(u[k], ru[k], xg(X, Pg, f - 2, j), xg(X, Pg, f - 1, j), xg(X, Pg, f, j), xg(X, Pg, f + 1, j),
                          xg(D, Pg, f - 2, j), xg(D, Pg, f - 1, j), xg(D, Pg, f, j), xg(D, Pg, f + 1, j));
    };
    auto fy = [&](int f) {  // y-face (i, f)
        const size_t k = i + (size_t)P * f;
        return quick_tl64(v[k], rv[k], xg(X, Pg, i, f - 2), xg(X, Pg, i, f - 1), xg(X, Pg, i, f), xg(X, Pg, i, f + 1),
                          xg(D, Pg, i, f - 2), xg(D, Pg, i, f - 1), xg(D, Pg, i, f), xg(D, Pg, i, f + 1));
    };
    const double ih = 1.0 / a.h1, ih2 = ih * ih;
    const double C = (fx(i + 1) - fx(i)) * ih + (fy(j + 1) - fy(j)) * ih;
    const double c = xg(D, Pg, i, j);
    const double Vs = ((xg(D, Pg, i + 1, j) - 2.0 * c + xg(D, Pg, i - 1, j)) * ih2 +
                       (xg(D, Pg, i, j + 1) - 2.0 * c + xg(D, Pg, i, j - 1)) * ih2) *
                      a.iren;
    return (0.0 - C) + Vs;
}

// u^{k+1} = u^k + cr r + cj R'(u^k) r into (uo, vo) on every face, norm
// partials of the change; after convergence a copy of u^k (the caller's
// buffer exchange then keeps the converged state)
__global__ __launch_bounds__(256) void k_fine_tangent(AmrGeom a, const double *u, const double *v, const double *XU,
                                                      const double *XV, const double *ru, const double *rv,
                                                      const double *DU, const double *DV, double *uo, double *vo,
                                                      double cr, double cj, double *partials, const Ctl *ctl)
{
    __shared__ double sh[16];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny, P = a.P1;
    const size_t k = i + (size_t)P * j;
    if (ctl->mom_done) {  // the same for every thread of the launch
        if (i <= nx && j < ny) uo[k] = u[k];
        if (i < nx && j <= ny) vo[k] = v[k];
        return;
    }
    double sx = 0.0, sy = 0.0;
    if (i <= nx && j <= ny) {
        if (j < ny) {
            double du = 0.0;
            if (i > 0 && i < nx) {
                const double J = 0.5 * (fine_TL(a, XU, DU, u, v, ru, rv, i - 1, j) + fine_TL(a, XU, DU, u, v, ru, rv, i, j));
                du = cr * ru[k] + cj * J;
            }
            uo[k] = u[k] + du;
            sx = fabs(du * du * ((i == 0 || i == nx) ? 0.5 : 1.0));
        }
        if (i < nx) {
            double dv = 0.0;
            if (j > 0 && j < ny) {
                const double J = 0.5 * (fine_TL(a, XV, DV, u, v, ru, rv, i, j - 1) + fine_TL(a, XV, DV, u, v, ru, rv, i, j));
                dv = cr * rv[k] + cj * J;
            }
            vo[k] = v[k] + dv;
            sy = fabs(dv * dv * ((j == 0 || j == ny) ? 0.5 : 1.0));
        }
    }
    sx = block_reduce<false>(sx, sh);
    sy = block_reduce<false>(sy, sh + 8);
    if (threadIdx.x == 0) {
        const size_t b = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * b] = sx;
        partials[2 * b + 1] = sy;
    }
}

// -------------------------------------------------- composite operator
// x0: coarse cells (pitch P0, active rows < jc0), x1: fine cells (pitch P1)
__device__ __forceinline__ double cf_ghost(const AmrGeom &a, const double *x0, const double *x1, int i)
{
    auto f = [&](int ic, int jc) { return x0[ic + (size_t)a.P0 * jc]; };
    const double pc = tangential(f, a.n0, i, a.jc0 - 1);
    return (8.0 / 15.0) * pc + (2.0 / 3.0) * x1[i] - 0.2 * x1[i + a.P1];
}

__device__ double coarse_L(const AmrGeom &a, const double *x0, const double *x1, int i, int j)
{
    const int P = a.P0, n = a.n0;
    const size_t k = i + (size_t)P * j;
    const double c = x0[k], ih = 1.0 / a.h0;
    const double fw = i > 0 ? (c - x0[k - 1]) * ih : 0.0;
    const double fe = i < n - 1 ? (x0[k + 1] - c) * ih : 0.0;
    const double fs = j > 0 ? (c - x0[k - P]) * ih : 0.0;
    double fn;
    if (j < a.jc0 - 1) fn = (x0[k + P] - c) * ih;
    else {  // mean of the two fine interface gradients above
        const int f0 = 2 * i;
        const double g0 = (x1[f0] - cf_ghost(a, x0, x1, f0)) / a.h1;
        const double g1 = (x1[f0 + 1] - cf_ghost(a, x0, x1, f0 + 1)) / a.h1;
        fn = 0.5 * (g0 + g1);
    }
    return (fe - fw) * ih + (fn - fs) * ih;
}

__device__ double fine_L(const AmrGeom &a, const double *x0, const double *x1, int i, int j)
{
    const int P = a.P1;
    const size_t k = i + (size_t)P * j;
    const double c = x1[k], ih = 1.0 / a.h1;
    const double fw = i > 0 ? (c - x1[k - 1]) * ih : 0.0;
    const double fe = i < a.nx - 1 ? (x1[k + 1] - c) * ih : 0.0;
    const double fs = j > 0 ? (c - x1[k - P]) * ih : (c - cf_ghost(a, x0, x1, i)) * ih;
    const double fn = j < a.ny - 1 ? (x1[k + P] - c) * ih : 0.0;
    return (fe - fw) * ih + (fn - fs) * ih;
}

// composite cells as one flat index: k < n0 jc0 coarse (row-major), then fine
__device__ __forceinline__ bool flat_cell(const AmrGeom &a, long long k, int &lev, int &i, int &j)
{
    const long long nc = (long long)a.n0 * a.jc0;
    if (k < nc) { lev = 0; i = (int)(k % a.n0); j = (int)(k / a.n0); return true; }
    k -= nc;
    if (k < (long long)a.nx * a.ny) { lev = 1; i = (int)(k % a.nx); j = (int)(k / a.nx); return true; }
    return false;
}

// y = b - L_c x (or y = L_c x with b null); partials: ||y||inf, ||y||_2^2
__global__ __launch_bounds__(256) void k_amr_apply(AmrGeom a, const double *x0, const double *x1, const double *b0,
                                                   const double *b1, double *y0, double *y1, double *partials)
{
    __shared__ double sh[16];
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    double m = 0.0, s2 = 0.0;
    if (flat_cell(a, k, lev, i, j)) {
        double r;
        if (lev == 0) {
            const size_t q = i + (size_t)a.P0 * j;
            r = coarse_L(a, x0, x1, i, j);
            if (b0) r = b0[q] - r;
            y0[q] = r;
        } else {
            const size_t q = i + (size_t)a.P1 * j;
            r = fine_L(a, x0, x1, i, j);
            if (b1) r = b1[q] - r;
            y1[q] = r;
        }
        m = fabs(r);
        s2 = r * r;
    }
    m = block_reduce<true>(m, sh);
    s2 = block_reduce<false>(s2, sh + 8);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = m;
        partials[2 * blockIdx.x + 1] = s2;
    }
}

// partial dot products <w, V_t>, t < nv, over the composite cells (vectors
// are stored as [coarse segment | fine segment], each segment off0 / off1)
__global__ __launch_bounds__(256) void k_amr_dots(AmrGeom a, const double *w, const double *V, size_t vs, int nv,
                                                  size_t off1, double *partials)
{
    __shared__ double sh[16];
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    size_t q = 0;
    const bool ok = flat_cell(a, k, lev, i, j);
    if (ok) q = lev == 0 ? i + (size_t)a.P0 * j : off1 + i + (size_t)a.P1 * j;
    for (int t = 0; t < nv; ++t) {
        double s = ok ? w[q] * V[t * vs + q] : 0.0;
        s = block_reduce<false>(s, sh);
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * nv + t] = s;
    }
}

// w = alpha w + sum_t c_t V_t over the composite cells (c on the device)
__global__ __launch_bounds__(256) void k_amr_combine(AmrGeom a, double *w, double alpha, const double *V, size_t vs,
                                                     int nv, const double *coef, size_t off1)
{
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    const size_t q = lev == 0 ? i + (size_t)a.P0 * j : off1 + i + (size_t)a.P1 * j;
    double s = alpha * w[q];
    for (int t = 0; t < nv; ++t) s += coef[t] * V[t * vs + q];
    w[q] = s;
}

// partial sums for the composite mean (area weighted) and the plain sum
__global__ __launch_bounds__(256) void k_amr_wsum(AmrGeom a, const double *x0, const double *x1, double *partials)
{
    __shared__ double sh[16];
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    double s = 0.0;
    if (flat_cell(a, k, lev, i, j))
        s = lev == 0 ? x0[i + (size_t)a.P0 * j] * a.h0 * a.h0 : x1[i + (size_t)a.P1 * j] * a.h1 * a.h1;
    s = block_reduce<false>(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_amr_shift(AmrGeom a, double *x0, double *x1, const double *mean)
{
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    if (lev == 0) x0[i + (size_t)a.P0 * j] -= *mean;
    else x1[i + (size_t)a.P1 * j] -= *mean;
}

// ------------------------------------------------------ preconditioner
// coarse right-hand side: r on uncovered cells, the mean of the four fine
// residuals on covered cells
// (guard: the device composite FGMRES's control block; its Arnoldi-step
// kernels do nothing once the cycle or the solve has ended)
__device__ __forceinline__ bool cgm_iter_skip(const Ctl *g) { return g && (g->mg_done | g->gm_stop); }

template <class T>
__global__ __launch_bounds__(256) void k_amr_restrict(AmrGeom a, const double *r0, const double *r1, T *rc,
                                                      const Ctl *guard)
{
    if (cgm_iter_skip(guard)) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= a.n0 || j >= a.n0) return;
    double v;
    if (j < a.jc0) v = r0[i + (size_t)a.P0 * j];
    else {
        const int fi = 2 * i, fj = 2 * (j - a.jc0);
        const size_t q = fi + (size_t)a.P1 * fj;
        v = 0.25 * (r1[q] + r1[q + 1] + r1[q + a.P1] + r1[q + a.P1 + 1]);
    }
    rc[i + (size_t)a.P0 * j] = (T)v;
}

// coarse correction onto the composite vector: z0 = e, z1 = e of the covering
// coarse cell
template <class T>
__global__ __launch_bounds__(256) void k_amr_prolong(AmrGeom a, const T *e, double *z0, double *z1, const Ctl *guard)
{
    if (cgm_iter_skip(guard)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    if (lev == 0) z0[i + (size_t)a.P0 * j] = (double)e[i + (size_t)a.P0 * j];
    else z1[i + (size_t)a.P1 * j] = (double)e[(i >> 1) + (size_t)a.P0 * (a.jc0 + (j >> 1))];
}

// one red or black Gauss-Seidel half-sweep of L_1 z1 = r1 on the fine patch
// (coarse values z0 fixed in the C/F ghost)
__global__ __launch_bounds__(256) void k_amr_fine_gs(AmrGeom a, const double *z0, double *z1, const double *r1,
                                                     int color, const Ctl *guard)
{
    if (cgm_iter_skip(guard)) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= a.nx || j >= a.ny || ((i + j) & 1) != color) return;
    const int P = a.P1;
    const size_t k = i + (size_t)P * j;
    const double ih2 = 1.0 / (a.h1 * a.h1);
    double off = 0.0, diag = 0.0;
    if (i > 0) { off += z1[k - 1]; diag += 1.0; }
    if (i < a.nx - 1) { off += z1[k + 1]; diag += 1.0; }
    if (j < a.ny - 1) { off += z1[k + P]; diag += 1.0; }
    if (j > 0) { off += z1[k - P]; diag += 1.0; }
    else {  // ghost = 8/15 pc + 2/3 z(i,0) - 1/5 z(i,1): fold its z(i,0) part into the diagonal
        auto f = [&](int ic, int jc) { return z0[ic + (size_t)a.P0 * jc]; };
        const double pc = tangential(f, a.n0, i, a.jc0 - 1);
        off += (8.0 / 15.0) * pc - 0.2 * z1[k + P];
        diag += 1.0 - 2.0 / 3.0;
    }
    z1[k] = (off - r1[k] / ih2) / diag;
}

// ------------------------------------------------------------ projection
// b = (1.5/dt) D u* on the composite cells (covered coarse cells: 0)
__global__ __launch_bounds__(256) void k_amr_rhs(AmrGeom a, double dt, const double *u0, const double *v0,
                                                 const double *u1, const double *v1, double *b0, double *b1,
                                                 double *partials)
{
    __shared__ double sh[16];
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    double m = 0.0;
    if (flat_cell(a, k, lev, i, j)) {
        if (lev == 0) {
            const size_t q = i + (size_t)a.P0 * j;
            const double d = (u0[q + 1] - u0[q]) / a.h0 + (v0[q + a.P0] - v0[q]) / a.h0;
            b0[q] = (1.5 / dt) * d;
            m = fabs(d);
        } else {
            const size_t q = i + (size_t)a.P1 * j;
            const double d = (u1[q + 1] - u1[q]) / a.h1 + (v1[q + a.P1] - v1[q]) / a.h1;
            b1[q] = (1.5 / dt) * d;
            m = fabs(d);
        }
    }
    m = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// u = u* - (dt/1.5) G_c phi on the fine faces and the uncovered coarse faces
// below the interface; p += phi
__global__ __launch_bounds__(256) void k_amr_project(AmrGeom a, double dt, const double *phi0, const double *phi1,
                                                     double *u0, double *v0, double *u1, double *v1, double *p0,
                                                     double *p1)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const double kk = dt / 1.5;
    // fine faces and cells (i < nx + 1, j < ny + 1)
    if (i <= a.nx && j <= a.ny) {
        const int P = a.P1;
        const size_t q = i + (size_t)P * j;
        if (j < a.ny && i > 0 && i < a.nx) u1[q] -= kk * (phi1[q] - phi1[q - 1]) / a.h1;
        if (i < a.nx && j < a.ny) {
            const double g = j > 0 ? (phi1[q] - phi1[q - P]) / a.h1 : (phi1[q] - cf_ghost(a, phi0, phi1, i)) / a.h1;
            v1[q] -= kk * g;
            p1[q] += phi1[q];
        }
    }
    // uncovered coarse faces and cells
    if (i <= a.n0 && j < a.jc0) {
        const int P = a.P0;
        const size_t q = i + (size_t)P * j;
        if (i > 0 && i < a.n0) u0[q] -= kk * (phi0[q] - phi0[q - 1]) / a.h0;
        if (i < a.n0) {
            if (j > 0) v0[q] -= kk * (phi0[q] - phi0[q - P]) / a.h0;
            p0[q] += phi0[q];
        }
    }
}

// coarse data on and above the interface from the fine level: faces as the
// mean of the two fine faces they cover, cells as the mean of four
__global__ __launch_bounds__(256) void k_amr_average_down(AmrGeom a, const double *u1, const double *v1,
                                                          const double *p1, double *u0, double *v0, double *p0)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), jr = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int j = a.jc0 + jr;
    if (i > a.n0 || j > a.n0) return;
    const int fi = 2 * i, fj = 2 * jr, P1 = a.P1;
    const size_t q = i + (size_t)a.P0 * j;
    if (j < a.n0) {
        const size_t f = fi + (size_t)P1 * fj;
        u0[q] = 0.5 * (u1[f] + u1[f + P1]);
    }
    if (i < a.n0) {
        const size_t f = fi + (size_t)P1 * fj;
        v0[q] = 0.5 * (v1[f] + v1[f + 1]);
        if (j < a.n0) p0[q] = 0.25 * (p1[f] + p1[f + 1] + p1[f + P1] + p1[f + P1 + 1]);
    }
}

// column reduction of per-block partials laid out [block][ncol]: column 0 is
// a max when is_max, every other column a sum; one workgroup per column
__global__ __launch_bounds__(1024) void k_reduce_cols(const double *partials, int nblk, int ncol, int is_max,
                                                      double *out)
{
    __shared__ double sh[32];
    const int t = blockIdx.x;
    const bool mx = is_max && t == 0;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double p = partials[(size_t)b * ncol + t];
        v = mx ? ((p > v || p != p) ? p : v) : v + p;
    }
    v = mx ? block_reduce<true>(v, sh) : block_reduce<false>(v, sh);
    if (threadIdx.x == 0) out[t] = v;
}

// Array4 view (component 0, k = lo[2]) <-> pitched patch array of nx x ny
__global__ __launch_bounds__(256) void k_pack_rect(amr_array4 v, double *dst, int P, int nx, int ny, int unpack)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const long long js = (long long)v.hi[0] - v.lo[0] + 1;
    double *a = (double *)v.ptr + (i - v.lo[0]) + js * (j - v.lo[1]);
    if (unpack) *a = dst[i + (size_t)P * j];
    else dst[i + (size_t)P * j] = *a;
}

// ------------------------------------------------------------ launchers
static dim3 g2(int nx, int ny) { return dim3((nx + 63) / 64, (ny + 3) / 4); }
static int flat_blocks(const AmrGeom &a) { return (int)(((long long)a.n0 * a.jc0 + (long long)a.nx * a.ny + 255) / 256); }

int amr_flat_blocks(const AmrGeom &a) { return flat_blocks(a); }

hipError_t amr_fine_cells(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *u0,
                          const double *v0, double *XU, double *XV)
{
    hipLaunchKernelGGL(k_fine_cells, g2(a.nx + 4, a.ny + 4), dim3(256), 0, s, a, u, v, u0, v0, XU, XV);
    return hipGetLastError();
}
hipError_t amr_fine_interface(const AmrGeom &a, hipStream_t s, double *v, const double *v0)
{
    hipLaunchKernelGGL(k_fine_interface, dim3((a.nx + 255) / 256), dim3(256), 0, s, a, v, v0);
    return hipGetLastError();
}
hipError_t amr_fine_stage(const AmrGeom &a, hipStream_t s, const double *us, const double *vs, const double *ub,
                          const double *vb, const double *cu, const double *cv, const double *p, const double *XU,
                          const double *XV, double *uo, double *vo, double rk, int mode, double *partials,
                          const Ctl *ctl, int *nblocks)
{
    const dim3 g = g2(a.nx + 1, a.ny + 1);
    *nblocks = (int)(g.x * g.y);
    hipLaunchKernelGGL(k_fine_stage, g, dim3(256), 0, s, a, us, vs, ub, vb, cu, cv, p, XU, XV, uo, vo, rk, mode,
                       partials, ctl);
    return hipGetLastError();
}
hipError_t amr_fine_resid(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *cu,
                          const double *cv, const double *p, const double *XU, const double *XV, double *ru,
                          double *rv, const Ctl *ctl)
{
    hipLaunchKernelGGL(k_fine_resid, g2(a.nx + 1, a.ny + 1), dim3(256), 0, s, a, u, v, cu, cv, p, XU, XV, ru, rv, ctl);
    return hipGetLastError();
}
hipError_t amr_fine_dcells(const AmrGeom &a, hipStream_t s, const double *ru, const double *rv, double *DU, double *DV,
                           const Ctl *ctl)
{
    hipLaunchKernelGGL(k_fine_dcells, g2(a.nx + 4, a.ny + 4), dim3(256), 0, s, a, ru, rv, DU, DV, ctl);
    return hipGetLastError();
}
hipError_t amr_fine_tangent(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *XU,
                            const double *XV, const double *ru, const double *rv, const double *DU, const double *DV,
                            double *uo, double *vo, double cr, double cj, double *partials, const Ctl *ctl,
                            int *nblocks)
{
    const dim3 g = g2(a.nx + 1, a.ny + 1);
    *nblocks = (int)(g.x * g.y);
    hipLaunchKernelGGL(k_fine_tangent, g, dim3(256), 0, s, a, u, v, XU, XV, ru, rv, DU, DV, uo, vo, cr, cj, partials,
                       ctl);
    return hipGetLastError();
}
hipError_t amr_fine_fin(const AmrGeom &a, hipStream_t s, const double *partials, int nb, Ctl *ctl)
{
    hipLaunchKernelGGL(k_fine_fin, dim3(1), dim3(1024), 0, s, partials, nb, ctl, (double)a.nx * (double)a.ny);
    return hipGetLastError();
}
hipError_t amr_fine_ctl_reset(hipStream_t s, Ctl *ctl, double tol, int iter_num)
{
    hipLaunchKernelGGL(k_fine_ctl_reset, dim3(1), dim3(1), 0, s, ctl, tol, iter_num);
    return hipGetLastError();
}
hipError_t amr_fine_bdf2(const AmrGeom &a, hipStream_t s, double *u, double *up, double *c, int dir, double dt)
{
    const int nx = dir == 0 ? a.nx + 1 : a.nx, ny = dir == 0 ? a.ny : a.ny + 1;
    hipLaunchKernelGGL(k_fine_bdf2, g2(nx, ny), dim3(256), 0, s, u, up, c, nx, ny, a.P1, dt);
    return hipGetLastError();
}
hipError_t amr_apply(const AmrGeom &a, hipStream_t s, const double *x0, const double *x1, const double *b0,
                     const double *b1, double *y0, double *y1, double *partials)
{
    hipLaunchKernelGGL(k_amr_apply, dim3(flat_blocks(a)), dim3(256), 0, s, a, x0, x1, b0, b1, y0, y1, partials);
    return hipGetLastError();
}
hipError_t amr_dots(const AmrGeom &a, hipStream_t s, const double *w, const double *V, size_t vs, int nv,
                    size_t off1, double *partials)
{
    hipLaunchKernelGGL(k_amr_dots, dim3(flat_blocks(a)), dim3(256), 0, s, a, w, V, vs, nv, off1, partials);
    return hipGetLastError();
}
hipError_t amr_combine(const AmrGeom &a, hipStream_t s, double *w, double alpha, const double *V, size_t vs, int nv,
                       const double *coef, size_t off1)
{
    hipLaunchKernelGGL(k_amr_combine, dim3(flat_blocks(a)), dim3(256), 0, s, a, w, alpha, V, vs, nv, coef, off1);
    return hipGetLastError();
}
hipError_t amr_wsum(const AmrGeom &a, hipStream_t s, const double *x0, const double *x1, double *partials)
{
    hipLaunchKernelGGL(k_amr_wsum, dim3(flat_blocks(a)), dim3(256), 0, s, a, x0, x1, partials);
    return hipGetLastError();
}
hipError_t amr_shift(const AmrGeom &a, hipStream_t s, double *x0, double *x1, const double *mean)
{
    hipLaunchKernelGGL(k_amr_shift, dim3(flat_blocks(a)), dim3(256), 0, s, a, x0, x1, mean);
    return hipGetLastError();
}
hipError_t amr_restrict(const AmrGeom &a, hipStream_t s, const double *r0, const double *r1, void *rc, int mixed,
                        const Ctl *guard)
{
    if (mixed)
        hipLaunchKernelGGL(k_amr_restrict<float>, g2(a.n0, a.n0), dim3(256), 0, s, a, r0, r1, (float *)rc, guard);
    else
        hipLaunchKernelGGL(k_amr_restrict<double>, g2(a.n0, a.n0), dim3(256), 0, s, a, r0, r1, (double *)rc, guard);
    return hipGetLastError();
}
hipError_t amr_prolong(const AmrGeom &a, hipStream_t s, const void *e, double *z0, double *z1, int mixed,
                       const Ctl *guard)
{
    if (mixed)
        hipLaunchKernelGGL(k_amr_prolong<float>, dim3(flat_blocks(a)), dim3(256), 0, s, a, (const float *)e, z0, z1,
                           guard);
    else
        hipLaunchKernelGGL(k_amr_prolong<double>, dim3(flat_blocks(a)), dim3(256), 0, s, a, (const double *)e, z0, z1,
                           guard);
    return hipGetLastError();
}
hipError_t amr_fine_gs(const AmrGeom &a, hipStream_t s, const double *z0, double *z1, const double *r1, int color,
                       const Ctl *guard)
{
    hipLaunchKernelGGL(k_amr_fine_gs, g2(a.nx, a.ny), dim3(256), 0, s, a, z0, z1, r1, color, guard);
    return hipGetLastError();
}

// ---------------------------------------------- device composite FGMRES
// FGMRES(m) on the composite operator with the whole Arnoldi process on the
// device, as the single-level solver (kernels_gmres.hip) runs it: the
// control block's gm_stop ends a cycle, mg_done the solve.  The Arnoldi-step
// kernels run while neither is set; the cycle-edge kernels (y = R^-1 g,
// x += Z y, the true residual and the next cycle's start) run only once a
// cycle has ended, so the host can enqueue batches of steps followed by a
// cycle edge and wait once per batch.
__device__ __forceinline__ bool cgm_edge_skip(const Ctl *g) { return g->mg_done || !g->gm_stop; }

__device__ __forceinline__ size_t comp_index(const AmrGeom &a, int lev, int i, int j, size_t off1)
{
    return lev == 0 ? i + (size_t)a.P0 * j : off1 + i + (size_t)a.P1 * j;
}

__global__ void k_cgm_begin(Ctl *ctl, GmState *st)
{
    ctl->mg_done = 0;
    ctl->gm_stop = 1;  // "between cycles": the first edge starts the first cycle from x = 0
    ctl->mg_iters = 0;
    ctl->mg_started = 0;
    st->j = 0;
    st->total = 0;
}

// y = R^-1 g over the ended cycle's columns (single thread), then x += Z y
__global__ void k_cgm_solve(Ctl *ctl, GmState *st)
{
    if (cgm_edge_skip(ctl)) return;
    const int jc = st->j;
    for (int r = jc - 1; r >= 0; --r) {
        double acc = st->gv[r];
        for (int q = r + 1; q < jc; ++q) acc -= st->H[r * kGmMax + q] * st->y[q];
        st->y[r] = acc / st->H[r * kGmMax + r];
    }
}

__global__ __launch_bounds__(256) void k_cgm_update(AmrGeom a, double *x, const double *Z, size_t vs, const GmState *st,
                                                    const Ctl *ctl, size_t off1)
{
    if (cgm_edge_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    const size_t q = comp_index(a, lev, i, j, off1);
    const int jc = st->j;
    double acc = x[q];
    for (int t = 0; t < jc; ++t) acc += st->y[t] * Z[t * vs + q];
    x[q] = acc;
}

// V_0 = b - L_c x; partials ||V_0||inf, ||V_0||_2^2 (two per block)
__global__ __launch_bounds__(256) void k_cgm_residual(AmrGeom a, const double *x, const double *b, double *V,
                                                      const Ctl *ctl, size_t off1, double *partials)
{
    __shared__ double sh[16];
    if (cgm_edge_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    double m = 0.0, s2 = 0.0;
    if (flat_cell(a, k, lev, i, j)) {
        const size_t q = comp_index(a, lev, i, j, off1);
        const double r = b[q] - (lev == 0 ? coarse_L(a, x, x + off1, i, j) : fine_L(a, x, x + off1, i, j));
        V[q] = r;
        m = (fabs(r) > m || r != r) ? fabs(r) : m;
        s2 = r * r;
    }
    m = block_reduce<true>(m, sh);
    s2 = block_reduce<false>(s2, sh + 8);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = m;
        partials[2 * blockIdx.x + 1] = s2;
    }
}

// stopping test on the true residual (||r||inf <= max(tol_abs, tol_rel ||b||inf),
// ||b||inf from the first residual, x = 0); otherwise start a cycle
__global__ __launch_bounds__(1024) void k_cgm_fin_residual(const double *partials, int nblk, Ctl *ctl, GmState *st,
                                                          int m)
{
    __shared__ double sh[32];
    if (cgm_edge_skip(ctl)) return;
    double mx = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double pm = partials[2 * b];
        mx = (pm > mx || pm != pm) ? pm : mx;
        s2 += partials[2 * b + 1];
    }
    mx = block_reduce<true>(mx, sh);
    s2 = block_reduce<false>(s2, sh);
    if (threadIdx.x != 0) return;
    ctl->res_norm = mx;
    if (!ctl->mg_started) {
        ctl->b_norm = mx;
        ctl->mg_started = 1;
    }
    const double thr = fmax(ctl->tol_abs, ctl->tol_rel * ctl->b_norm);
    const double beta = sqrt(s2);
    st->j = 0;
    if (mx <= thr || !(mx == mx) || beta == 0.0) {
        ctl->mg_done = 1;
        if (!(mx == mx)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
        return;
    }
    if (ctl->mg_iters >= ctl->mg_max_iter) {
        ctl->mg_done = 1;
        ctl->flags |= AMRESSIF_FLAG_MG_CAP;
        return;
    }
    st->beta = beta;
    st->target = thr * beta / mx;  // the 2-norm target of the Givens estimate
    st->scale = 1.0 / beta;
    for (int q = 0; q <= m; ++q) st->gv[q] = 0.0;
    st->gv[0] = beta;
    ctl->gm_stop = 0;
}

// V_j <- V_j scale (normalised); the preconditioner's input vcur = V_j
__global__ __launch_bounds__(256) void k_cgm_load(AmrGeom a, double *V, size_t vs, double *vcur, const GmState *st,
                                                  const Ctl *ctl, size_t off1)
{
    if (cgm_iter_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    const size_t q = comp_index(a, lev, i, j, off1);
    double *v = V + (size_t)st->j * vs;
    const double x = v[q] * st->scale;
    v[q] = x;
    vcur[q] = x;
}

// Z_j = zcur; V_{j+1} = L_c zcur
__global__ __launch_bounds__(256) void k_cgm_apply(AmrGeom a, const double *zcur, double *Z, double *V, size_t vs,
                                                   const GmState *st, const Ctl *ctl, size_t off1)
{
    if (cgm_iter_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    if (!flat_cell(a, k, lev, i, j)) return;
    const size_t q = comp_index(a, lev, i, j, off1);
    const int jj = st->j;
    Z[(size_t)jj * vs + q] = zcur[q];
    V[(size_t)(jj + 1) * vs + q] = lev == 0 ? coarse_L(a, zcur, zcur + off1, i, j) : fine_L(a, zcur, zcur + off1, i, j);
}

// partials of <V_{j+1}, V_t>, t <= j, into slot t of each block's m + 2
__global__ __launch_bounds__(256) void k_cgm_dots(AmrGeom a, const double *V, size_t vs, const GmState *st,
                                                  const Ctl *ctl, size_t off1, double *partials, int m)
{
    __shared__ double sh[16];
    if (cgm_iter_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    const bool ok = flat_cell(a, k, lev, i, j);
    const size_t q = ok ? comp_index(a, lev, i, j, off1) : 0;
    const int jj = st->j;
    const double wq = ok ? V[(size_t)(jj + 1) * vs + q] : 0.0;
    for (int t = 0; t <= jj; ++t) {
        double acc = ok ? wq * V[(size_t)t * vs + q] : 0.0;
        acc = block_reduce<false>(acc, sh);
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * (m + 2) + t] = acc;
    }
}

// V_{j+1} -= sum_t c_t V_t; partial ||V_{j+1}||_2^2 into slot m + 1
__global__ __launch_bounds__(256) void k_cgm_orth(AmrGeom a, double *V, size_t vs, const GmState *st, const Ctl *ctl,
                                                  size_t off1, double *partials, int m)
{
    __shared__ double sh[16];
    if (cgm_iter_skip(ctl)) return;
    const long long k = blockIdx.x * 256LL + threadIdx.x;
    int lev, i, j;
    double s2 = 0.0;
    if (flat_cell(a, k, lev, i, j)) {
        const size_t q = comp_index(a, lev, i, j, off1);
        const int jj = st->j;
        double w = V[(size_t)(jj + 1) * vs + q];
        for (int t = 0; t <= jj; ++t) w -= st->c[t] * V[(size_t)t * vs + q];
        V[(size_t)(jj + 1) * vs + q] = w;
        s2 = w * w;
    }
    s2 = block_reduce<false>(s2, sh);
    if (threadIdx.x == 0) partials[(size_t)blockIdx.x * (m + 2) + m + 1] = s2;
}

hipError_t amr_cgm_begin(hipStream_t s, Ctl *ctl, GmState *st)
{
    hipLaunchKernelGGL(k_cgm_begin, dim3(1), dim3(1), 0, s, ctl, st);
    return hipGetLastError();
}

hipError_t amr_cgm_edge(const AmrGeom &a, hipStream_t s, const CgmBufs &g)
{
    hipLaunchKernelGGL(k_cgm_solve, dim3(1), dim3(1), 0, s, g.ctl, g.st);
    hipLaunchKernelGGL(k_cgm_update, dim3(flat_blocks(a)), dim3(256), 0, s, a, g.x, (const double *)g.Z, g.vs,
                       (const GmState *)g.st, (const Ctl *)g.ctl, g.off1);
    hipLaunchKernelGGL(k_cgm_residual, dim3(flat_blocks(a)), dim3(256), 0, s, a, (const double *)g.x,
                       (const double *)g.b, g.V, (const Ctl *)g.ctl, g.off1, g.part);
    hipLaunchKernelGGL(k_cgm_fin_residual, dim3(1), dim3(1024), 0, s, (const double *)g.part, flat_blocks(a), g.ctl,
                       g.st, g.m);
    hipLaunchKernelGGL(k_cgm_load, dim3(flat_blocks(a)), dim3(256), 0, s, a, g.V, g.vs, g.vcur, (const GmState *)g.st,
                       (const Ctl *)g.ctl, g.off1);
    return hipGetLastError();
}

hipError_t amr_cgm_apply(const AmrGeom &a, hipStream_t s, const CgmBufs &g)
{
    hipLaunchKernelGGL(k_cgm_apply, dim3(flat_blocks(a)), dim3(256), 0, s, a, (const double *)g.zcur, g.Z, g.V, g.vs,
                       (const GmState *)g.st, (const Ctl *)g.ctl, g.off1);
    return hipGetLastError();
}

// classical Gram-Schmidt twice, then ||w||, the Givens rotations and the
// stopping estimate; then the next V_j is normalised into vcur
hipError_t amr_cgm_orthonormalise(const AmrGeom &a, hipStream_t s, const CgmBufs &g)
{
    const int nb = flat_blocks(a);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_cgm_dots, dim3(nb), dim3(256), 0, s, a, (const double *)g.V, g.vs, (const GmState *)g.st,
                           (const Ctl *)g.ctl, g.off1, g.part, g.m);
        if (hipError_t e = launch_gm_fin_dots_raw(s, g.part, nb, g.m, g.st, g.ctl, pass)) return e;
        hipLaunchKernelGGL(k_cgm_orth, dim3(nb), dim3(256), 0, s, a, g.V, g.vs, (const GmState *)g.st,
                           (const Ctl *)g.ctl, g.off1, g.part, g.m);
    }
    if (hipError_t e = launch_gm_fin_arnoldi_raw(s, g.part, nb, g.m, g.st, g.ctl)) return e;
    hipLaunchKernelGGL(k_cgm_load, dim3(nb), dim3(256), 0, s, a, g.V, g.vs, g.vcur, (const GmState *)g.st,
                       (const Ctl *)g.ctl, g.off1);
    return hipGetLastError();
}
hipError_t amr_rhs(const AmrGeom &a, hipStream_t s, double dt, const double *u0, const double *v0, const double *u1,
                   const double *v1, double *b0, double *b1, double *partials)
{
    hipLaunchKernelGGL(k_amr_rhs, dim3(flat_blocks(a)), dim3(256), 0, s, a, dt, u0, v0, u1, v1, b0, b1, partials);
    return hipGetLastError();
}
hipError_t amr_project(const AmrGeom &a, hipStream_t s, double dt, const double *phi0, const double *phi1, double *u0,
                       double *v0, double *u1, double *v1, double *p0, double *p1)
{
    const int nx = std::max(a.nx + 1, a.n0 + 1), ny = std::max(a.ny + 1, a.jc0);
    hipLaunchKernelGGL(k_amr_project, g2(nx, ny), dim3(256), 0, s, a, dt, phi0, phi1, u0, v0, u1, v1, p0, p1);
    return hipGetLastError();
}
hipError_t launch_reduce_cols(hipStream_t s, const double *partials, int nblk, int ncol, int is_max, double *out)
{
    hipLaunchKernelGGL(k_reduce_cols, dim3(ncol), dim3(1024), 0, s, partials, nblk, ncol, is_max, out);
    return hipGetLastError();
}
hipError_t launch_pack_rect(hipStream_t s, const amr_array4 &v, double *dst, int P, int nx, int ny)
{
    hipLaunchKernelGGL(k_pack_rect, g2(nx, ny), dim3(256), 0, s, v, dst, P, nx, ny, 0);
    return hipGetLastError();
}
hipError_t launch_unpack_rect(hipStream_t s, const double *src, const amr_array4 &v, int P, int nx, int ny)
{
    hipLaunchKernelGGL(k_pack_rect, g2(nx, ny), dim3(256), 0, s, v, (double *)src, P, nx, ny, 1);
    return hipGetLastError();
}
hipError_t amr_average_down(const AmrGeom &a, hipStream_t s, const double *u1, const double *v1, const double *p1,
                            double *u0, double *v0, double *p0)
{
    hipLaunchKernelGGL(k_amr_average_down, g2(a.n0 + 1, a.n0 - a.jc0 + 1), dim3(256), 0, s, a, u1, v1, p1, u0, v0, p0);
    return hipGetLastError();
}

}  // namespace amr
