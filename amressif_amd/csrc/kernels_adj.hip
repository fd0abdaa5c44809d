// Tangent-linear and adjoint of the time step (FP64, gfx950).
//
// The forward step (RK4 dual-time momentum solve to convergence, projection)
// is a map (u^n, u^{n-1}, p^n; lid g) -> (u^{n+1}, u^n, p^{n+1}).  Its
// converged momentum predictor u* solves
//     N(u*) = A(u*; g) - G p - (1.5/dt) u* + c(u^n, u^{n-1}) = 0
// on the interior faces (A = face average of -div(QUICK flux) + viscous flux,
// fn_flux_calc.cpp:11-202, 695-766, fn_rhs.cpp:8-68; c is the BDF2 constant,
// momentum.cpp:68-87), and the projection (fn_rhs.cpp:71-125,
// poisson.cpp:192-392) is linear.  Linearised at the recorded u* (upwind
// switch frozen, QUICK with the face velocity sign of u*):
//   tangent:  (A_u - 1.5/dt) du* = G dp - dc - A_g dg
//   adjoint:  (A_u - 1.5/dt)^T s = u*_bar, then p_bar += G^T s,
//             u^n_bar += -(2/dt) s, u^{n-1}_bar += (0.5/dt) s, g_bar -= A_g^T s.
// Both linear systems are solved by the Richardson iteration
// x <- (A x + rhs) / (1.5/dt), whose contraction is ||A|| dt / 1.5 (about 1e-3
// at 512^2, Re = 1000): a few sweeps reach FP64 round-off.
//
// Operator application, three kernels each (no atomics, every output gathered):
//   tangent  A x:   k_cells (cell values of x with the homogeneous ghost
//                   rules, plus the lid perturbation) -> k_tl_flux (QUICK
//                   tangent fluxes on every face) -> k_tl_out (cell
//                   divergence + viscous, face average, Richardson update)
//   adjoint A^T w:  k_adj_flux (flux adjoints from the face-average and
//                   divergence transposes) -> k_adj_cells (cell adjoints
//                   gathered over the QUICK stencils and the viscous stencil,
//                   ghost cells folded onto their mirrors) -> k_adj_out
//                   (cell-to-face transpose + face-velocity term, update)
// Cell arrays with the two ghost layers have pitch Pg, index
// (i + 2) + Pg (j + 2); face and plain cell arrays use the step's pitch P.
#include "device_common.hpp"
#include "internal.hpp"

namespace amr {

// QUICK weights of the four cells a2, a1, b0, b1 of one face, upwind switch
// frozen by the sign of the face velocity (fn_flux_calc.cpp:66-73):
// vel >= 0: -a2/8 + 6 a1/8 + 3 b0/8; vel < 0: 3 a1/8 + 6 b0/8 - b1/8.
__device__ __forceinline__ double qw(bool pos, int role)
{
    // role 0 = a2, 1 = a1, 2 = b0, 3 = b1
    if (pos) return role == 0 ? -0.125 : role == 1 ? 0.75 : role == 2 ? 0.375 : 0.0;
    return role == 0 ? 0.0 : role == 1 ? 0.375 : role == 2 ? 0.75 : -0.125;
}

__device__ __forceinline__ double qsel(bool pos, double a2, double a1, double b0, double b1)
{
    return pos ? (-0.125 * a2 + 0.75 * a1 + 0.375 * b0) : (0.375 * a1 + 0.75 * b0 - 0.125 * b1);
}

// sign of the homogeneous ghost rule (ghost = sign * mirror) for one side and
// component (utilities.cpp:87-269; constants dropped)
__device__ __forceinline__ double ghost_sign(int side, int code, int comp)
{
    if (code == 135) return -1.0;
    const bool slip = side == 1 ? code == 1 : code == -135;
    if (!slip) return -1.0;
    // slip: the normal component is mirrored with a sign change, the tangential one kept
    const int normal = side < 2 ? 0 : 1;
    return comp == normal ? -1.0 : 1.0;
}

__device__ __forceinline__ bool lid_side_uses_g(int code) { return code != 135 && code != -135; }

struct AdjGeom {
    int n, P, Pg;
    double dx, dy, idx2, idy2, iren, c15;
    int bc_lo0, bc_hi0, bc_lo1, bc_hi1;
    double inflow0, inflow1;
    const double *lid;
};

static AdjGeom adj_geom(const Ctx &c)
{
    const Geometry &g = c.g;
    AdjGeom a{};
    a.n = g.n;
    a.P = g.P;
    a.Pg = c.Pg;
    a.dx = g.dx;
    a.dy = g.dy;
    a.idx2 = 1.0 / (g.dx * g.dx);
    a.idy2 = 1.0 / (g.dy * g.dy);
    a.iren = 1.0 / g.ren;
    a.c15 = 1.5 / g.dt;
    a.bc_lo0 = g.bc_lo0; a.bc_hi0 = g.bc_hi0; a.bc_lo1 = g.bc_lo1; a.bc_hi1 = g.bc_hi1;
    a.inflow0 = g.inflow0; a.inflow1 = g.inflow1;
    a.lid = c.lid;
    return a;
}

// ------------------------------------------------------------ cell values
// Cell values (2 ghost layers) of faces (fu, fv): the base velocity
// (homog = 0: the full cont2cart rules with the lid) or a perturbation
// (homog = 1: constants dropped, plus 2 dlid on the north ghosts of U where
// the north rule carries the lid).  Corners are 0 (no stencil reads them).
__global__ __launch_bounds__(256) void k_cells(AdjGeom a, const double *fu, const double *fv, double *XU, double *XV,
                                               int homog, const double *dlid, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int gi = blockIdx.x * 64 + (threadIdx.x & 63), gj = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n;
    if (gi >= n + 4 || gj >= n + 4) return;
    const int i = gi - 2, j = gj - 2;
    const bool xin = i >= 0 && i < n, yin = j >= 0 && j < n;
    auto interior = [&](int ci, int cj, double &U, double &V) {
        const size_t k = ci + (size_t)a.P * cj;
        U = 0.5 * (fu[k] + fu[k + 1]);
        V = 0.5 * (fv[k] + fv[k + a.P]);
    };
    double U = 0.0, V = 0.0;
    if (xin && yin) {
        interior(i, j, U, V);
    } else if (yin) {  // x ghost
        const int mi = i < 0 ? -i - 1 : 2 * n - 1 - i;
        double um, vm;
        interior(mi, j, um, vm);
        const int side = i < 0 ? 0 : 1, code = i < 0 ? a.bc_lo0 : a.bc_hi0;
        if (homog) ghost_rule<double, true>(side, code, um, vm, 0.0, 0.0, U, V);
        else ghost_rule<double, false>(side, code, um, vm, 0.0, 0.0, U, V);
    } else if (xin) {  // y ghost
        const int mj = j < 0 ? -j - 1 : 2 * n - 1 - j;
        double um, vm;
        interior(i, mj, um, vm);
        const int side = j < 0 ? 2 : 3, code = j < 0 ? a.bc_lo1 : a.bc_hi1;
        if (homog) {
            ghost_rule<double, true>(side, code, um, vm, 0.0, 0.0, U, V);
            if (side == 3 && dlid && lid_side_uses_g(code)) U += 2.0 * dlid[i];
        } else {
            const double gx = a.lid ? a.lid[i] : a.inflow0;
            ghost_rule<double, false>(side, code, um, vm, gx, a.inflow1, U, V);
        }
    }
    const size_t o = gi + (size_t)a.Pg * gj;
    XU[o] = U;
    XV[o] = V;
}

// ---------------------------------------------------------------- tangent
// QUICK tangent fluxes: F = vel phi_sel(dX) + dvel phi_sel(X) on x-faces
// (transport u, stencil along i) and y-faces (transport v, along j).
__global__ __launch_bounds__(256) void k_tl_flux(AdjGeom a, const double *ub, const double *vb, const double *XU,
                                                 const double *XV, const double *du, const double *dv,
                                                 const double *DU, const double *DV, double *Fxu, double *Fxv,
                                                 double *Fyu, double *Fyv, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, Pg = a.Pg;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)a.P * j;
    if (j < n) {  // x-face (i, j): cells (i-2 .. i+1, j)
        const double vel = ub[k], dvel = du[k];
        const bool pos = vel >= 0.0;
        const size_t c0 = (i - 2 + 2) + (size_t)Pg * (j + 2);  // cell i-2
        const double fu = vel * qsel(pos, DU[c0], DU[c0 + 1], DU[c0 + 2], DU[c0 + 3]) +
                          dvel * qsel(pos, XU[c0], XU[c0 + 1], XU[c0 + 2], XU[c0 + 3]);
        const double fv = vel * qsel(pos, DV[c0], DV[c0 + 1], DV[c0 + 2], DV[c0 + 3]) +
                          dvel * qsel(pos, XV[c0], XV[c0 + 1], XV[c0 + 2], XV[c0 + 3]);
        Fxu[k] = fu;
        Fxv[k] = fv;
    }
    if (i < n) {  // y-face (i, j): cells (i, j-2 .. j+1)
        const double vel = vb[k], dvel = dv[k];
        const bool pos = vel >= 0.0;
        const size_t c0 = (i + 2) + (size_t)Pg * (j - 2 + 2);
        const size_t s = Pg;
        const double fu = vel * qsel(pos, DU[c0], DU[c0 + s], DU[c0 + 2 * s], DU[c0 + 3 * s]) +
                          dvel * qsel(pos, XU[c0], XU[c0 + s], XU[c0 + 2 * s], XU[c0 + 3 * s]);
        const double fv = vel * qsel(pos, DV[c0], DV[c0 + s], DV[c0 + 2 * s], DV[c0 + 3 * s]) +
                          dvel * qsel(pos, XV[c0], XV[c0 + s], XV[c0 + 2 * s], XV[c0 + 3 * s]);
        Fyu[k] = fu;
        Fyv[k] = fv;
    }
}

// block partials of the weighted squared change and of the squared iterate
// (weights 1/2 on wall faces, as the pseudo-iteration norm)
__device__ __forceinline__ void lin_partials(double d2, double x2, double *partials)
{
    __shared__ double sh[16];
    d2 = block_reduce<false>(d2, sh);
    x2 = block_reduce<false>(x2, sh + 8);
    if (threadIdx.x == 0) {
        const size_t b = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * b] = d2;
        partials[2 * b + 1] = x2;
    }
}

// x_new = (A x + q) / c15 on interior faces, 0 on wall-normal faces
__global__ __launch_bounds__(256) void k_tl_out(AdjGeom a, const double *Fxu, const double *Fxv, const double *Fyu,
                                                const double *Fyv, const double *DU, const double *DV,
                                                const double *qu, const double *qv, const double *xu_old,
                                                const double *xv_old, double *xu_new, double *xv_new,
                                                double *partials, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P, Pg = a.Pg;
    double d2 = 0.0, x2 = 0.0;
    auto T = [&](const double *Fx, const double *Fy, const double *D, int ci, int cj) {
        const size_t k = ci + (size_t)P * cj;
        const double C = (Fx[k + 1] - Fx[k]) / a.dx + (Fy[k + P] - Fy[k]) / a.dy;
        const size_t g = (ci + 2) + (size_t)Pg * (cj + 2);
        const double Vs = ((D[g + 1] - 2.0 * D[g] + D[g - 1]) * a.idx2 + (D[g + Pg] - 2.0 * D[g] + D[g - Pg]) * a.idy2) *
                          a.iren;
        return -C + Vs;
    };
    if (i <= n && j <= n) {
        const size_t k = i + (size_t)P * j;
        if (j < n) {
            double xn = 0.0;
            if (i > 0 && i < n) xn = (0.5 * (T(Fxu, Fyu, DU, i - 1, j) + T(Fxu, Fyu, DU, i, j)) + qu[k]) / a.c15;
            const double w = (i == 0 || i == n) ? 0.5 : 1.0, d = xn - xu_old[k];
            d2 += w * d * d;
            x2 += w * xn * xn;
            xu_new[k] = xn;
        }
        if (i < n) {
            double xn = 0.0;
            if (j > 0 && j < n) xn = (0.5 * (T(Fxv, Fyv, DV, i, j - 1) + T(Fxv, Fyv, DV, i, j)) + qv[k]) / a.c15;
            const double w = (j == 0 || j == n) ? 0.5 : 1.0, d = xn - xv_old[k];
            d2 += w * d * d;
            x2 += w * xn * xn;
            xv_new[k] = xn;
        }
    }
    lin_partials(d2, x2, partials);
}

// ---------------------------------------------------------------- adjoint
// T_bar of interior cells from the face adjoint w (wall faces read as 0):
// the transpose of the face average R(i) = (T(i-1) + T(i)) / 2.
__device__ __forceinline__ double tbar_u(const AdjGeom &a, const double *wu, int ci, int cj)
{
    if (ci < 0 || ci >= a.n || cj < 0 || cj >= a.n) return 0.0;
    const size_t k = ci + (size_t)a.P * cj;
    const double w0 = ci == 0 ? 0.0 : wu[k], w1 = ci + 1 == a.n ? 0.0 : wu[k + 1];
    return 0.5 * (w0 + w1);
}
__device__ __forceinline__ double tbar_v(const AdjGeom &a, const double *wv, int ci, int cj)
{
    if (ci < 0 || ci >= a.n || cj < 0 || cj >= a.n) return 0.0;
    const size_t k = ci + (size_t)a.P * cj;
    const double w0 = cj == 0 ? 0.0 : wv[k], w1 = cj + 1 == a.n ? 0.0 : wv[k + a.P];
    return 0.5 * (w0 + w1);
}

// flux adjoints: T = -div F + ..., so F_bar(face i) = (T_bar(i) - T_bar(i-1)) / h
__global__ __launch_bounds__(256) void k_adj_flux(AdjGeom a, const double *wu, const double *wv, double *Bxu,
                                                  double *Bxv, double *Byu, double *Byv, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)a.P * j;
    if (j < n) {
        Bxu[k] = (tbar_u(a, wu, i, j) - tbar_u(a, wu, i - 1, j)) / a.dx;
        Bxv[k] = (tbar_v(a, wv, i, j) - tbar_v(a, wv, i - 1, j)) / a.dx;
    }
    if (i < n) {
        Byu[k] = (tbar_u(a, wu, i, j) - tbar_u(a, wu, i, j - 1)) / a.dy;
        Byv[k] = (tbar_v(a, wv, i, j) - tbar_v(a, wv, i, j - 1)) / a.dy;
    }
}

// Adjoint of the cell value (gi, gj), ghosts included, before ghost folding:
// the QUICK stencils of the up-to-four x-faces and four y-faces that read it,
// and the viscous stencils of its interior neighbours.
__device__ void xbar_at(const AdjGeom &a, const double *ub, const double *vb, const double *Bxu, const double *Bxv,
                        const double *Byu, const double *Byv, const double *wu, const double *wv, int gi, int gj,
                        double &XU, double &XV)
{
    const int n = a.n, P = a.P;
    double su = 0.0, sv = 0.0;
    if (gj >= 0 && gj < n) {  // x-faces f = gi + m read cell gi in role 3 - (m + 1) ... (a2 for m = 2)
#pragma unroll
        for (int m = -1; m <= 2; ++m) {
            const int f = gi + m;
            if (f < 0 || f > n) continue;
            const size_t k = f + (size_t)P * gj;
            const double vel = ub[k];
            const int role = 2 - m;  // m = 2 -> a2 (0), 1 -> a1, 0 -> b0, -1 -> b1
            const double w = vel * qw(vel >= 0.0, role);
            su += w * Bxu[k];
            sv += w * Bxv[k];
        }
    }
    if (gi >= 0 && gi < n) {  // y-faces f = gj + m
#pragma unroll
        for (int m = -1; m <= 2; ++m) {
            const int f = gj + m;
            if (f < 0 || f > n) continue;
            const size_t k = gi + (size_t)P * f;
            const double vel = vb[k];
            const int role = 2 - m;
            const double w = vel * qw(vel >= 0.0, role);
            su += w * Byu[k];
            sv += w * Byv[k];
        }
    }
    // viscous: V(c) = ((X(c+1) - 2 X(c) + X(c-1)) / dx^2 + (...) / dy^2) / ren on interior c
    double vu = 0.0, vvv = 0.0;
    const bool in = gi >= 0 && gi < n && gj >= 0 && gj < n;
    if (in) {
        vu -= 2.0 * (a.idx2 + a.idy2) * tbar_u(a, wu, gi, gj);
        vvv -= 2.0 * (a.idx2 + a.idy2) * tbar_v(a, wv, gi, gj);
    }
    vu += a.idx2 * (tbar_u(a, wu, gi - 1, gj) + tbar_u(a, wu, gi + 1, gj)) +
          a.idy2 * (tbar_u(a, wu, gi, gj - 1) + tbar_u(a, wu, gi, gj + 1));
    vvv += a.idx2 * (tbar_v(a, wv, gi - 1, gj) + tbar_v(a, wv, gi + 1, gj)) +
           a.idy2 * (tbar_v(a, wv, gi, gj - 1) + tbar_v(a, wv, gi, gj + 1));
    XU = su + a.iren * vu;
    XV = sv + a.iren * vvv;
}

// interior cell adjoints with the ghost cells that mirror onto them folded in;
// with lidg, the north-ghost adjoints of U (rows n and n+1) for the lid
__global__ __launch_bounds__(256) void k_adj_cells(AdjGeom a, const double *ub, const double *vb, const double *Bxu,
                                                   const double *Bxv, const double *Byu, const double *Byv,
                                                   const double *wu, const double *wv, double *XbU, double *XbV,
                                                   double *lidg, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n;
    if (i >= n || j >= n) return;
    double U, V;
    xbar_at(a, ub, vb, Bxu, Bxv, Byu, Byv, wu, wv, i, j, U, V);
    auto fold = [&](int gi, int gj, int side, int code) {
        double gu, gv;
        xbar_at(a, ub, vb, Bxu, Bxv, Byu, Byv, wu, wv, gi, gj, gu, gv);
        U += ghost_sign(side, code, 0) * gu;
        V += ghost_sign(side, code, 1) * gv;
        return gu;
    };
    // ghost (-1 - i) mirrors cell i on the low side, (2n - 1 - i) on the high side
    if (i <= 1) fold(-1 - i, j, 0, a.bc_lo0);
    if (i >= n - 2) fold(2 * n - 1 - i, j, 1, a.bc_hi0);
    if (j <= 1) fold(i, -1 - j, 2, a.bc_lo1);
    if (j >= n - 2) {
        const double gu = fold(i, 2 * n - 1 - j, 3, a.bc_hi1);
        if (lidg) lidg[(j == n - 1 ? 0 : n) + i] = gu;
    }
    const size_t k = i + (size_t)a.P * j;
    XbU[k] = U;
    XbV[k] = V;
}

// s_new = (A^T s + face-velocity term - rhs) / c15 on interior faces
__global__ __launch_bounds__(256) void k_adj_out(AdjGeom a, const double *XbU, const double *XbV, const double *Bxu,
                                                 const double *Bxv, const double *Byu, const double *Byv,
                                                 const double *ub, const double *vb, const double *XU,
                                                 const double *XV, const double *ru, const double *rv,
                                                 const double *su_old, const double *sv_old, double *su_new,
                                                 double *sv_new, double *partials, const Ctl *ctl)
{
    if (ctl && ctl->lin_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P, Pg = a.Pg;
    double d2 = 0.0, x2 = 0.0;
    if (i <= n && j <= n) {
        const size_t k = i + (size_t)P * j;
        if (j < n) {
            double sn = 0.0;
            if (i > 0 && i < n) {
                const bool pos = ub[k] >= 0.0;
                const size_t c0 = i + (size_t)Pg * (j + 2);  // cell i-2 in the ghosted array
                const double pu = qsel(pos, XU[c0], XU[c0 + 1], XU[c0 + 2], XU[c0 + 3]);
                const double pv = qsel(pos, XV[c0], XV[c0 + 1], XV[c0 + 2], XV[c0 + 3]);
                const double at = 0.5 * (XbU[k - 1] + XbU[k]) + Bxu[k] * pu + Bxv[k] * pv;
                sn = (at - ru[k]) / a.c15;
            }
            const double w = (i == 0 || i == n) ? 0.5 : 1.0, d = sn - su_old[k];
            d2 += w * d * d;
            x2 += w * sn * sn;
            su_new[k] = sn;
        }
        if (i < n) {
            double sn = 0.0;
            if (j > 0 && j < n) {
                const bool pos = vb[k] >= 0.0;
                const size_t c0 = (i + 2) + (size_t)Pg * j;  // cell (i, j-2)
                const double pu = qsel(pos, XU[c0], XU[c0 + Pg], XU[c0 + 2 * Pg], XU[c0 + 3 * Pg]);
                const double pv = qsel(pos, XV[c0], XV[c0 + Pg], XV[c0 + 2 * Pg], XV[c0 + 3 * Pg]);
                const double at = 0.5 * (XbV[k - P] + XbV[k]) + Byu[k] * pu + Byv[k] * pv;
                sn = (at - rv[k]) / a.c15;
            }
            const double w = (j == 0 || j == n) ? 0.5 : 1.0, d = sn - sv_old[k];
            d2 += w * d * d;
            x2 += w * sn * sn;
            sv_new[k] = sn;
        }
    }
    lin_partials(d2, x2, partials);
}

// -------------------------------------------------------- loop control
__global__ void k_lin_start(Ctl *ctl, double tol, int cap)
{
    ctl->lin_done = 0;
    ctl->lin_count = 0;
    ctl->lin_cap = cap;
    ctl->lin_tol = tol;
    ctl->lin_rel = 0.0;
}

// stop when the weighted RMS change is at most tol times the iterate's
__global__ __launch_bounds__(1024) void k_fin_lin(const double *partials, int nb, Ctl *ctl)
{
    __shared__ double sh[32];
    if (ctl->lin_done) return;
    double d2 = 0.0, x2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { d2 += partials[2 * b]; x2 += partials[2 * b + 1]; }
    d2 = block_reduce<false>(d2, sh);
    x2 = block_reduce<false>(x2, sh + 16);
    if (threadIdx.x == 0) {
        ctl->lin_count += 1;
        const double rel = x2 > 0.0 ? sqrt(d2 / x2) : (d2 > 0.0 ? 1.0 : 0.0);
        ctl->lin_rel = rel;
        if (!(rel == rel)) { ctl->flags |= AMRESSIF_FLAG_NONFINITE; ctl->lin_done = 1; }
        else if (rel <= ctl->lin_tol) ctl->lin_done = 1;
        else if (ctl->lin_count >= ctl->lin_cap) { ctl->flags |= AMRESSIF_FLAG_LIN_CAP; ctl->lin_done = 1; }
    }
}

// ------------------------------------------------------- step pieces
// tangent: q = dc - G dp = (2/dt) du - (0.5/dt) du_prev - G dp on interior
// faces (0 on walls); x0 = q / c15
__global__ __launch_bounds__(256) void k_tl_rhs(AdjGeom a, double dt, const double *du, const double *dv,
                                                const double *dup, const double *dvp, const double *dp, double *qu,
                                                double *qv, double *xu, double *xv)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)P * j;
    if (j < n) {
        double q = 0.0;
        if (i > 0 && i < n) q = (2.0 / dt) * du[k] - (0.5 / dt) * dup[k] - (dp[k] - dp[k - 1]) / a.dx;
        qu[k] = q;
        xu[k] = q / a.c15;
    }
    if (i < n) {
        double q = 0.0;
        if (j > 0 && j < n) q = (2.0 / dt) * dv[k] - (0.5 / dt) * dvp[k] - (dp[k] - dp[k - P]) / a.dy;
        qv[k] = q;
        xv[k] = q / a.c15;
    }
}

// adjoint, projection part: b = p_bar + (dt/1.5) D(u_bar) (wall faces read as 0)
__global__ __launch_bounds__(256) void k_adj_phibar(AdjGeom a, double dt, const double *au, const double *av,
                                                    const double *pb, double *b)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P;
    if (i >= n || j >= n) return;
    const size_t k = i + (size_t)P * j;
    const double uw = i == 0 ? 0.0 : au[k], ue = i + 1 == n ? 0.0 : au[k + 1];
    const double vs = j == 0 ? 0.0 : av[k], vn = j + 1 == n ? 0.0 : av[k + P];
    b[k] = pb[k] + (dt / 1.5) * ((ue - uw) / a.dx + (vn - vs) / a.dy);
}

// adjoint: rhs of the momentum system r = u_bar - (1.5/dt) G b_bar on interior
// faces; s0 = -r / c15
__global__ __launch_bounds__(256) void k_adj_rhs(AdjGeom a, double dt, const double *au, const double *av,
                                                 const double *bb, double *ru, double *rv, double *su, double *sv)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)P * j;
    if (j < n) {
        double r = 0.0;
        if (i > 0 && i < n) r = au[k] - (1.5 / dt) * (bb[k] - bb[k - 1]) / a.dx;
        ru[k] = r;
        su[k] = -r / a.c15;
    }
    if (i < n) {
        double r = 0.0;
        if (j > 0 && j < n) r = av[k] - (1.5 / dt) * (bb[k] - bb[k - P]) / a.dy;
        rv[k] = r;
        sv[k] = -r / a.c15;
    }
}

// adjoint, last part: p_bar += G^T s = -D(s); u^n_bar = u_prev_bar - (2/dt) s;
// u^{n-1}_bar = (0.5/dt) s
__global__ __launch_bounds__(256) void k_adj_finish(AdjGeom a, double dt, const double *su, const double *sv,
                                                    double *au, double *av, double *aup, double *avp, double *pb)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n, P = a.P;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)P * j;
    if (i < n && j < n) {
        const double sw = i == 0 ? 0.0 : su[k], se = i + 1 == n ? 0.0 : su[k + 1];
        const double ss = j == 0 ? 0.0 : sv[k], sn = j + 1 == n ? 0.0 : sv[k + P];
        pb[k] -= (se - sw) / a.dx + (sn - ss) / a.dy;
    }
    if (j < n) {
        const double s = (i > 0 && i < n) ? su[k] : 0.0;
        au[k] = (i > 0 && i < n) ? aup[k] - (2.0 / dt) * s : 0.0;
        aup[k] = (0.5 / dt) * s;
    }
    if (i < n) {
        const double s = (j > 0 && j < n) ? sv[k] : 0.0;
        av[k] = (j > 0 && j < n) ? avp[k] - (2.0 / dt) * s : 0.0;
        avp[k] = (0.5 / dt) * s;
    }
}

// g_bar -= A_g^T s: the north ghosts of U carry 2 g
__global__ void k_adj_lid(const double *lidg, double *glid, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) glid[i] -= 2.0 * (lidg[i] + lidg[n + i]);
}

// ------------------------------------------------------------ launchers
static dim3 face_grid(int n) { return dim3((n + 1 + 63) / 64, (n + 1 + 3) / 4); }
static dim3 cell_grid(int n) { return dim3((n + 63) / 64, (n + 3) / 4); }

int lin_partials_count(const Ctx &c)
{
    const dim3 g = face_grid(c.g.n);
    return (int)(g.x * g.y);
}

hipError_t launch_lin_start(const Ctx &c, double tol, int cap)
{
    hipLaunchKernelGGL(k_lin_start, dim3(1), dim3(1), 0, c.stream, c.ctl, tol, cap);
    return hipGetLastError();
}

hipError_t launch_cells(const Ctx &c, const double *fu, const double *fv, double *XU, double *XV, int homog,
                        const double *dlid, int guarded)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_cells, dim3((n + 4 + 63) / 64, (n + 4 + 3) / 4), dim3(256), 0, c.stream, adj_geom(c), fu, fv,
                       XU, XV, homog, dlid, guarded ? c.ctl : nullptr);
    return hipGetLastError();
}

// one guarded Richardson sweep of the tangent system: x_new = (A x + q)/c15
hipError_t launch_tl_sweep(const Ctx &c, const AdjWork &w, const double *xu, const double *xv, double *xu_new,
                           double *xv_new, const double *dlid)
{
    const AdjGeom a = adj_geom(c);
    const int n = c.g.n;
    if (lin_partials_count(c) > kMaxPartials) return hipErrorInvalidValue;
    hipError_t e = launch_cells(c, xu, xv, w.DU, w.DV, 1, dlid, 1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tl_flux, face_grid(n), dim3(256), 0, c.stream, a, w.base_u, w.base_v, w.XU, w.XV, xu, xv,
                       w.DU, w.DV, w.F[0], w.F[1], w.F[2], w.F[3], c.ctl);
    hipLaunchKernelGGL(k_tl_out, face_grid(n), dim3(256), 0, c.stream, a, w.F[0], w.F[1], w.F[2], w.F[3], w.DU, w.DV,
                       w.Q[0], w.Q[1], xu, xv, xu_new, xv_new, c.partials, c.ctl);
    hipLaunchKernelGGL(k_fin_lin, dim3(1), dim3(1024), 0, c.stream, c.partials, lin_partials_count(c), c.ctl);
    return hipGetLastError();
}

// one guarded Richardson sweep of the adjoint system: s_new = (A^T s - r)/c15
hipError_t launch_adj_sweep(const Ctx &c, const AdjWork &w, const double *su, const double *sv, double *su_new,
                            double *sv_new)
{
    const AdjGeom a = adj_geom(c);
    const int n = c.g.n;
    if (lin_partials_count(c) > kMaxPartials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_adj_flux, face_grid(n), dim3(256), 0, c.stream, a, su, sv, w.F[0], w.F[1], w.F[2], w.F[3],
                       c.ctl);
    hipLaunchKernelGGL(k_adj_cells, cell_grid(n), dim3(256), 0, c.stream, a, w.base_u, w.base_v, w.F[0], w.F[1],
                       w.F[2], w.F[3], su, sv, w.XbU, w.XbV, (double *)nullptr, c.ctl);
    hipLaunchKernelGGL(k_adj_out, face_grid(n), dim3(256), 0, c.stream, a, w.XbU, w.XbV, w.F[0], w.F[1], w.F[2],
                       w.F[3], w.base_u, w.base_v, w.XU, w.XV, w.Q[0], w.Q[1], su, sv, su_new, sv_new, c.partials,
                       c.ctl);
    hipLaunchKernelGGL(k_fin_lin, dim3(1), dim3(1024), 0, c.stream, c.partials, lin_partials_count(c), c.ctl);
    return hipGetLastError();
}

// g_bar -= A_g^T s (unguarded; after the adjoint solve)
hipError_t launch_adj_lid(const Ctx &c, const AdjWork &w, const double *su, const double *sv, double *glid)
{
    const AdjGeom a = adj_geom(c);
    const int n = c.g.n;
    hipLaunchKernelGGL(k_adj_flux, face_grid(n), dim3(256), 0, c.stream, a, su, sv, w.F[0], w.F[1], w.F[2], w.F[3],
                       (const Ctl *)nullptr);
    hipLaunchKernelGGL(k_adj_cells, cell_grid(n), dim3(256), 0, c.stream, a, w.base_u, w.base_v, w.F[0], w.F[1],
                       w.F[2], w.F[3], su, sv, w.XbU, w.XbV, w.lidg, (const Ctl *)nullptr);
    hipLaunchKernelGGL(k_adj_lid, dim3((n + 255) / 256), dim3(256), 0, c.stream, w.lidg, glid, n);
    return hipGetLastError();
}

hipError_t launch_tl_rhs(const Ctx &c, const AdjWork &w, const double *du, const double *dv, const double *dup,
                         const double *dvp, const double *dp, double *xu, double *xv)
{
    hipLaunchKernelGGL(k_tl_rhs, face_grid(c.g.n), dim3(256), 0, c.stream, adj_geom(c), c.g.dt, du, dv, dup, dvp, dp,
                       w.Q[0], w.Q[1], xu, xv);
    return hipGetLastError();
}

hipError_t launch_adj_phibar(const Ctx &c, const double *au, const double *av, const double *pb, double *b)
{
    hipLaunchKernelGGL(k_adj_phibar, cell_grid(c.g.n), dim3(256), 0, c.stream, adj_geom(c), c.g.dt, au, av, pb, b);
    return hipGetLastError();
}

hipError_t launch_adj_rhs(const Ctx &c, const AdjWork &w, const double *au, const double *av, const double *bb,
                          double *su, double *sv)
{
    hipLaunchKernelGGL(k_adj_rhs, face_grid(c.g.n), dim3(256), 0, c.stream, adj_geom(c), c.g.dt, au, av, bb, w.Q[0],
                       w.Q[1], su, sv);
    return hipGetLastError();
}

hipError_t launch_adj_finish(const Ctx &c, const double *su, const double *sv, double *au, double *av, double *aup,
                             double *avp, double *pb)
{
    hipLaunchKernelGGL(k_adj_finish, face_grid(c.g.n), dim3(256), 0, c.stream, adj_geom(c), c.g.dt, su, sv, au, av,
                       aup, avp, pb);
    return hipGetLastError();
}

}  // namespace amr
