// The fused RK4 dual-time stage -- the dominant kernel of the step (gfx950).
//
// One launch = one stage over the whole grid: cell velocities from faces with
// the wall ghost rules (cont2cart, utilities.cpp:16-350), QUICK half-node
// fluxes (fn_flux_calc.cpp:11-202), viscous flux (:695-766), flux wall ghosts
// (main.cpp:945-1014), face pressure gradient (fn_flux_calc.cpp:825-911),
// momentum RHS (fn_rhs.cpp:8-68), dual-time residual and stage update
// (momentum.cpp:68-87), and on the last stage the pseudo-iteration norm
// partials (utilities.cpp:421-493).
//
// Tiling: a block of 256 threads owns TX x TY = 64 x 16 cells.  Phase 1 builds
// the cell velocities of the tile plus a 3-cell halo in LDS (ghost cells from
// their mirror cell, so no ghost storage in HBM); phase 2 gives each thread a
// 4-row strip of cells and computes the total flux T = -C + V there, with
// the QUICK face fluxes evaluated from the LDS row/column windows; phase 3
// reuses the LDS buffer for T and updates the faces.  Tiles are handed to the
// 8 XCDs in contiguous horizontal bands so halo rows stay in one L2.
//
// Two arithmetic forms:
//  * FP64 (the reference's): u* <- u^k + rk (R(u*) - (1.5/dt) u* + c) with the
//    BDF2 constant c = (1.5/dt) u^n + (0.5/dt)(u^n - u^{n-1}) folded once per step.
//  * mixed: the stage works on the increment d = u* - u^k in FP32,
//    d <- rk (R_T(u^k + d) - (1.5/dt) d + w), w = c - G p - (1.5/dt) u^k,
//    which is the same update rewritten; u^k is carried in FP32 for the stencil
//    and the step's accumulated change S = u^k - u^n in FP32 (u^n stays FP64).
#include <type_traits>

#include "device_common.hpp"
#include "internal.hpp"

namespace amr {

constexpr int TX = 64, TY = 16, NT = 256, RPT = TY / 4;
constexpr int CW = TX + 5, CH = TY + 5;  // cells: cols i0-3..i0+TX+1, rows j0-3..j0+TY+1

struct StageArgs {
    // FP64 form
    const double *us, *vs, *ub, *vb, *cu, *cv, *p;
    double *uo, *vo;
    // mixed form
    const float *uk, *vk, *ud, *vd;
    float *dou, *dov, *wu, *wv, *su, *sv, *ukn, *vkn;
    const double *lid;
    double *partials;
    const Ctl *ctl;
    int n, P;
    double inv_dx, inv_dy, inv_dx2, inv_dy2, inv_ren, c15, rk;
    int bc_lo0, bc_hi0, bc_lo1, bc_hi1;
    double inflow0, inflow1;
    int mode;   // 0 stage, 1 last stage (+ norm partials), 2 RHS only (FP64)
    int remap;  // XCD-banded tile order
};

template <bool MIXED>
__global__ __launch_bounds__(NT) void k_stage(StageArgs a)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    if (a.ctl && a.ctl->mom_done) return;
    __shared__ R sbuf[2 * CH * CW];
    __shared__ double sred[16];
    R(*sU)[CW] = reinterpret_cast<R(*)[CW]>(sbuf);
    R(*sV)[CW] = reinterpret_cast<R(*)[CW]>(sbuf + CH * CW);

    const int n = a.n, P = a.P, tid = threadIdx.x;
    int bx = blockIdx.x, by = blockIdx.y;
    if (a.remap) {  // block b runs on XCD b % 8: give each XCD a contiguous band of tiles
        const int B = gridDim.x * gridDim.y, b = bx + gridDim.x * by;
        const int t = (b & 7) * (B >> 3) + (b >> 3);
        bx = t % gridDim.x;
        by = t / gridDim.x;
    }
    const int i0 = bx * TX, j0 = by * TY;

    auto fu = [&](int i, int j) -> R {  // stencil-input x-face velocity
        const size_t k = i + (size_t)P * j;
        if constexpr (MIXED) return a.ud ? a.uk[k] + a.ud[k] : a.uk[k];
        else return a.us[k];
    };
    auto fv = [&](int i, int j) -> R {
        const size_t k = i + (size_t)P * j;
        if constexpr (MIXED) return a.vd ? a.vk[k] + a.vd[k] : a.vk[k];
        else return a.vs[k];
    };

    // ---- phase 1: cell velocities of tile + halo
    for (int e = tid; e < CW * CH; e += NT) {
        const int li = e % CW, lj = e / CW, ci = i0 - 3 + li, cj = j0 - 3 + lj;
        const bool inx = ci >= 0 && ci < n, iny = cj >= 0 && cj < n;
        R U = 0, V = 0;
        if (inx && iny) {
            U = R(0.5) * (fu(ci, cj) + fu(ci + 1, cj));
            V = R(0.5) * (fv(ci, cj) + fv(ci, cj + 1));
        } else if (!inx && iny && ci >= -2 && ci <= n + 1) {
            const int mi = ci < 0 ? -ci - 1 : 2 * n - 1 - ci;
            const R um = R(0.5) * (fu(mi, cj) + fu(mi + 1, cj));
            const R vm = R(0.5) * (fv(mi, cj) + fv(mi, cj + 1));
            ghost_rule<R>(ci < 0 ? 0 : 1, ci < 0 ? a.bc_lo0 : a.bc_hi0, um, vm, R(0), R(0), U, V);
        } else if (inx && !iny && cj >= -2 && cj <= n + 1) {
            const int mj = cj < 0 ? -cj - 1 : 2 * n - 1 - cj;
            const R um = R(0.5) * (fu(ci, mj) + fu(ci + 1, mj));
            const R vm = R(0.5) * (fv(ci, mj) + fv(ci, mj + 1));
            const R gx = (R)(a.lid ? a.lid[ci] : a.inflow0);
            ghost_rule<R>(cj < 0 ? 2 : 3, cj < 0 ? a.bc_lo1 : a.bc_hi1, um, vm, gx, (R)a.inflow1, U, V);
        }
        sU[lj][li] = U;
        sV[lj][li] = V;
    }
    __syncthreads();

    // ---- phase 2: total cell flux T = -C + V on a 4-row strip per thread
    const R idx = (R)a.inv_dx, idy = (R)a.inv_dy, idx2 = (R)a.inv_dx2, idy2 = (R)a.inv_dy2, iren = (R)a.inv_ren;
    auto cell_T = [&](int ci, int cj, R &TU, R &TV) {
        const int li = ci - (i0 - 3), lj = cj - (j0 - 3);
        const R uw = fu(ci, cj), ue = fu(ci + 1, cj), vsf = fv(ci, cj), vnf = fv(ci, cj + 1);
        {
            const R m2 = sU[lj][li - 2], m1 = sU[lj][li - 1], c0 = sU[lj][li], p1 = sU[lj][li + 1], p2 = sU[lj][li + 2];
            const R s2 = sU[lj - 2][li], s1 = sU[lj - 1][li], n1 = sU[lj + 1][li], n2 = sU[lj + 2][li];
            const R C = (quick_flux<R>(ue, m1, c0, p1, p2) - quick_flux<R>(uw, m2, m1, c0, p1)) * idx +
                        (quick_flux<R>(vnf, s1, c0, n1, n2) - quick_flux<R>(vsf, s2, s1, c0, n1)) * idy;
            const R Vf = ((p1 - R(2) * c0 + m1) * idx2 + (n1 - R(2) * c0 + s1) * idy2) * iren;
            TU = (R(0) - C) + Vf;
        }
        {
            const R m2 = sV[lj][li - 2], m1 = sV[lj][li - 1], c0 = sV[lj][li], p1 = sV[lj][li + 1], p2 = sV[lj][li + 2];
            const R s2 = sV[lj - 2][li], s1 = sV[lj - 1][li], n1 = sV[lj + 1][li], n2 = sV[lj + 2][li];
            const R C = (quick_flux<R>(ue, m1, c0, p1, p2) - quick_flux<R>(uw, m2, m1, c0, p1)) * idx +
                        (quick_flux<R>(vnf, s1, c0, n1, n2) - quick_flux<R>(vsf, s2, s1, c0, n1)) * idy;
            const R Vf = ((p1 - R(2) * c0 + m1) * idx2 + (n1 - R(2) * c0 + s1) * idy2) * iren;
            TV = (R(0) - C) + Vf;
        }
    };
    const int tx = tid & 63, ty = tid >> 6;
    const int ci = i0 + tx;
    R tU[RPT], tV[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int cj = j0 + ty * RPT + k;
        tU[k] = 0;
        tV[k] = 0;
        if (ci < n && cj < n) cell_T(ci, cj, tU[k], tV[k]);
    }
    // extra strip: T_U of column i0-1 (threads 0..TY-1), T_V of row j0-1 (threads TY..TY+TX-1)
    R xT = 0;
    if (tid < TY) {
        const int cj = j0 + tid;
        R tu = 0, tv = 0;
        if (i0 >= 1 && cj < n) cell_T(i0 - 1, cj, tu, tv);
        xT = tu;
    } else if (tid < TY + TX) {
        const int cc = i0 + tid - TY;
        R tu = 0, tv = 0;
        if (j0 >= 1 && cc < n) cell_T(cc, j0 - 1, tu, tv);
        xT = tv;
    }
    __syncthreads();  // every read of sU/sV is done: reuse the buffer for T
    R(*sTU)[TX + 1] = reinterpret_cast<R(*)[TX + 1]>(sbuf);            // [TY][TX+1], col = ci - (i0-1)
    R(*sTV)[TX] = reinterpret_cast<R(*)[TX]>(sbuf + TY * (TX + 1));    // [TY+1][TX], row = cj - (j0-1)
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        sTU[ty * RPT + k][tx + 1] = tU[k];
        sTV[ty * RPT + k + 1][tx] = tV[k];
    }
    if (tid < TY) sTU[tid][0] = xT;
    else if (tid < TY + TX) sTV[0][tid - TY] = xT;
    __syncthreads();

    // ---- phase 3: face RHS and update.  Wall-normal faces get R = 0 exactly
    //      (flux ghost = -mirror, pressure ghost = mirror).
    double nx_ = 0.0, ny_ = 0.0;
    const int i = ci;
    auto x_face = [&](int fi, int fj, R RT) {
        const size_t k = fi + (size_t)P * fj;
        const bool wall = fi == 0 || fi == n;
        if constexpr (!MIXED) {
            double R_ = 0.0;
            if (!wall) R_ = -(a.p[k] - a.p[k - 1]) * a.inv_dx + RT;
            if (a.mode == 2) { a.uo[k] = R_; return; }
            const double base = a.ub[k];
            const double unew = base + a.rk * (R_ - a.c15 * a.us[k] + a.cu[k]);
            a.uo[k] = unew;
            if (a.mode == 1) { const double d = base - unew; nx_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
        } else {
            const float dold = a.ud ? a.ud[k] : 0.f;
            const float w = a.wu[k];
            const float dn = (float)a.rk * ((wall ? 0.f : RT) - (float)a.c15 * dold + w);
            a.dou[k] = dn;
            if (a.mode == 1) {
                a.ukn[k] = a.uk[k] + dn;
                a.wu[k] = w - (float)a.c15 * dn;
                a.su[k] += dn;
                const double d = (double)dn;
                nx_ += d * d * (wall ? 0.5 : 1.0);
            }
        }
    };
    auto y_face = [&](int fi, int fj, R RT) {
        const size_t k = fi + (size_t)P * fj;
        const bool wall = fj == 0 || fj == n;
        if constexpr (!MIXED) {
            double R_ = 0.0;
            if (!wall) R_ = -(a.p[k] - a.p[k - P]) * a.inv_dy + RT;
            if (a.mode == 2) { a.vo[k] = R_; return; }
            const double base = a.vb[k];
            const double vnew = base + a.rk * (R_ - a.c15 * a.vs[k] + a.cv[k]);
            a.vo[k] = vnew;
            if (a.mode == 1) { const double d = base - vnew; ny_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
        } else {
            const float dold = a.vd ? a.vd[k] : 0.f;
            const float w = a.wv[k];
            const float dn = (float)a.rk * ((wall ? 0.f : RT) - (float)a.c15 * dold + w);
            a.dov[k] = dn;
            if (a.mode == 1) {
                a.vkn[k] = a.vk[k] + dn;
                a.wv[k] = w - (float)a.c15 * dn;
                a.sv[k] += dn;
                const double d = (double)dn;
                ny_ += d * d * (wall ? 0.5 : 1.0);
            }
        }
    };
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = ty * RPT + k, j = j0 + r;
        if (i <= n && j < n) {
            const R RT = (i == 0 || i == n) ? R(0) : R(0.5) * (sTU[r][tx] + sTU[r][tx + 1]);
            x_face(i, j, RT);
            if (i == n - 1 && i0 + TX == n) x_face(n, j, R(0));  // east wall face
        }
        if (i < n && j <= n) {
            const R RT = (j == 0 || j == n) ? R(0) : R(0.5) * (sTV[r][tx] + sTV[r + 1][tx]);
            y_face(i, j, RT);
            if (j == n - 1 && j0 + TY == n) y_face(i, n, R(0));  // north wall face
        }
    }
    if (a.mode == 1) {
        const double bx_ = block_reduce<false>(nx_, sred);
        const double by_ = block_reduce<false>(ny_, sred + 8);
        if (tid == 0) {
            const size_t b = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
            a.partials[2 * b] = bx_;
            a.partials[2 * b + 1] = by_;
        }
    }
}

static StageArgs base_args(const Ctx &c, double rk, int mode)
{
    const Geometry &g = c.g;
    StageArgs a{};
    a.lid = c.lid;
    a.partials = c.partials;
    a.ctl = c.ctl;
    a.n = g.n;
    a.P = g.P;
    a.inv_dx = 1.0 / g.dx;
    a.inv_dy = 1.0 / g.dy;
    a.inv_dx2 = 1.0 / (g.dx * g.dx);
    a.inv_dy2 = 1.0 / (g.dy * g.dy);
    a.inv_ren = 1.0 / g.ren;
    a.c15 = 1.5 / g.dt;
    a.rk = rk;
    a.bc_lo0 = g.bc_lo0; a.bc_hi0 = g.bc_hi0; a.bc_lo1 = g.bc_lo1; a.bc_hi1 = g.bc_hi1;
    a.inflow0 = g.inflow0; a.inflow1 = g.inflow1;
    a.mode = mode;
    return a;
}

static dim3 stage_grid(int n, int *remap)
{
    dim3 grid((n + TX - 1) / TX, (n + TY - 1) / TY);
    *remap = ((grid.x * grid.y) % 8 == 0) ? 1 : 0;
    return grid;
}

hipError_t launch_stage(const Ctx &c, const StageLaunch &s, int *nblocks)
{
    StageArgs a = base_args(c, s.rk, s.mode);
    a.us = s.us; a.vs = s.vs; a.ub = s.ub; a.vb = s.vb; a.cu = s.cu; a.cv = s.cv; a.p = s.p;
    a.uo = s.uo; a.vo = s.vo;
    if (s.mode == 2 || s.unguarded) a.ctl = nullptr;
    const dim3 grid = stage_grid(c.g.n, &a.remap);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    if (nblocks) *nblocks = grid.x * grid.y;
    hipLaunchKernelGGL(k_stage<false>, grid, dim3(NT), 0, c.stream, a);
    return hipGetLastError();
}

hipError_t launch_stage_mixed(const Ctx &c, const MixedStageLaunch &s, int *nblocks)
{
    StageArgs a = base_args(c, s.rk, s.mode);
    a.uk = s.uk; a.vk = s.vk; a.ud = s.ud; a.vd = s.vd;
    a.dou = s.dou; a.dov = s.dov; a.wu = s.wu; a.wv = s.wv; a.su = s.su; a.sv = s.sv;
    a.ukn = s.ukn; a.vkn = s.vkn;
    if (s.unguarded) a.ctl = nullptr;
    const dim3 grid = stage_grid(c.g.n, &a.remap);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    if (nblocks) *nblocks = grid.x * grid.y;
    hipLaunchKernelGGL(k_stage<true>, grid, dim3(NT), 0, c.stream, a);
    return hipGetLastError();
}

// Start of a mixed step: u^k = u^n in FP32, S = 0, w = (0.5/dt)(u^n - u^{n-1}) - G p
// (= c - G p - (1.5/dt) u^n without the cancellation), velContPrev <- velCont.
__global__ __launch_bounds__(256) void k_bdf2_mixed(const double *u, double *up, const double *p, float *uk,
                                                    float *w, float *s, int nx, int ny, int P, int n, int dir,
                                                    double dt, double inv_h)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)P * j;
    const double un = u[k], um = up[k];
    const int f = dir == 0 ? i : j;
    const double gp = (f == 0 || f == n) ? 0.0 : (p[k] - p[k - (dir == 0 ? 1 : P)]) * inv_h;
    up[k] = un;
    uk[k] = (float)un;
    w[k] = (float)((0.5 / dt) * (un - um) - gp);
    s[k] = 0.f;
}

hipError_t launch_bdf2_mixed(const Ctx &c)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_bdf2_mixed, dim3((n + 1 + 63) / 64, (n + 3) / 4), dim3(256), 0, c.stream, c.U[0],
                       c.Uprev[0], c.p, c.uk32[0][0], c.w32[0], c.s32[0], n + 1, n, c.g.P, n, 0, c.g.dt, 1.0 / c.g.dx);
    hipLaunchKernelGGL(k_bdf2_mixed, dim3((n + 63) / 64, (n + 1 + 3) / 4), dim3(256), 0, c.stream, c.U[1],
                       c.Uprev[1], c.p, c.uk32[0][1], c.w32[1], c.s32[1], n, n + 1, c.g.P, n, 1, c.g.dt, 1.0 / c.g.dy);
    return hipGetLastError();
}

}  // namespace amr
