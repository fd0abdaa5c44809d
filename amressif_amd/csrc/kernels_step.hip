// HIP kernels of the fractional-step time step, gfx950.
//
// (The fused RK stage itself lives in kernels_stage.hip.)
//
// Data layout (fine level, one box): pitched row-major arrays with pitch P
// (a multiple of 64 doubles, rows 512 B aligned): x-face u(i,j) at
// u[i + P*j], i in [0,n], j in [0,n); y-face v(i,j), i in [0,n), j in [0,n];
// cells q(i,j), i,j in [0,n).  Ghost cells are never stored: wall ghost
// values are produced in registers/LDS from their mirror cells.
#include "device_common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace amr {

// ------------------------------------------------------- norm finalize
// Error_Computation (utilities.cpp:421-493): e_c = sqrt(s_c)/sqrt(npts),
// normError = max(e_x, e_y); then the loop control of main.cpp:1023-1039.
__global__ __launch_bounds__(1024) void k_fin_pseudo(const double *partials, int nb, Ctl *ctl, int n)
{
    __shared__ double sh[32];
    // the partials' loads go out before the flag's value is needed (one memory
    // round trip instead of two); they stay within the partials buffer
    const int done = __hip_atomic_load(&ctl->mom_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double sx = 0.0, sy = 0.0;
    for_records<2, 2>(partials, nb, [&](const double *o) {
        sx += o[0];
        sy += o[1];
    });
    if (done) return;
    double red[2] = {sx, sy};
    block_reduce_n<0u>(red, sh);
    sx = red[0];
    sy = red[1];
    if (threadIdx.x == 0) {
        const double npts = (double)n * (double)n;
        const double ex = sqrt(sx) / sqrt(npts), ey = sqrt(sy) / sqrt(npts);
        const double ne = (ex < ey) ? ey : ex;
        ctl->norm_error = ne;
        ctl->mom_count += 1;
        if (!(ne == ne) || isinf(ne)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
        if (ctl->mom_count > ctl->iter_num) { ctl->flags |= AMRESSIF_FLAG_CAP; ctl->mom_done = 1; }
        else if (ne > 1.e2) { ctl->flags |= AMRESSIF_FLAG_DIVERGED; ctl->mom_done = 1; }
        else if (!(ne > ctl->tol_mom)) ctl->mom_done = 1;
    }
}

// The same for k_mixed_pair's records of three.  start_mg (the last
// pseudo-iteration of a batch that the step's Poisson phase is enqueued
// behind): also k_mg_start's work, from the final flag of this launch: skip
// the phase unless the loop has ended, reset the solve's control, and the mean
// of b = (1.5/dt) D(u*) from the net outflow in the partials of the last
// launch that did work (a launch past convergence writes none).
__global__ __launch_bounds__(1024) void k_fin_pseudo3(const double *partials, int nb, Ctl *ctl, int n, int start_mg,
                                                      double dt)
{
    __shared__ double sh[48];
    const int done = __hip_atomic_load(&ctl->mom_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double sx = 0.0, sy = 0.0, sb = 0.0;
    for_records<3, 3>(partials, nb, [&](const double *o) {
        sx += o[0];
        sy += o[1];
        sb += o[2];
    });
    if (done && !start_mg) return;
    double red[3] = {sx, sy, sb};
    block_reduce_n<0u>(red, sh);
    if (threadIdx.x == 0) {
        if (!done) {
            const double npts = (double)n * (double)n;
            const double ex = sqrt(red[0]) / sqrt(npts), ey = sqrt(red[1]) / sqrt(npts);
            const double ne = (ex < ey) ? ey : ex;
            ctl->norm_error = ne;
            ctl->mom_count += 1;
            if (!(ne == ne) || isinf(ne)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
            if (ctl->mom_count > ctl->iter_num) { ctl->flags |= AMRESSIF_FLAG_CAP; ctl->mom_done = 1; }
            else if (ne > 1.e2) { ctl->flags |= AMRESSIF_FLAG_DIVERGED; ctl->mom_done = 1; }
            else if (!(ne > ctl->tol_mom)) ctl->mom_done = 1;
        }
        if (start_mg) {  // k_mg_start with require_mom
            const int skip = !ctl->mom_done;
            ctl->mg_done = skip;
            ctl->mg_skip = skip;
            ctl->gm_stop = 0;
            ctl->mg_iters = 0;
            ctl->mg_started = 0;
            ctl->b_norm_fixed = 0;
            ctl->res_norm = 0.0;
            ctl->b_norm = 0.0;
            ctl->phi_sum = 0.0;
            ctl->phi_mean = 0.0;
            if (!skip) ctl->b_mean = (1.5 / dt) * red[2] / ((double)n * (double)n);
        }
    }
}

hipError_t launch_fin_pseudo(const Ctx &c, int nblocks, int start_mg)
{
    if (c.mp_on)
        hipLaunchKernelGGL(k_fin_pseudo3, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n,
                           start_mg, c.g.dt);
    else
        hipLaunchKernelGGL(k_fin_pseudo, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n);
    return hipGetLastError();
}

// PSEUDO_TIMESTEPPING = 0 (main.cpp:900-926): Error_Computation of the one
// explicit update, then the branch leaves the loop before countIter is
// incremented (no cap or divergence test); a non-finite norm is flagged.
__global__ __launch_bounds__(1024) void k_fin_explicit(const double *partials, int nb, Ctl *ctl, int n)
{
    __shared__ double sh[32];
    double sx = 0.0, sy = 0.0;
    for_records<2, 2>(partials, nb, [&](const double *o) {
        sx += o[0];
        sy += o[1];
    });
    double red[2] = {sx, sy};
    block_reduce_n<0u>(red, sh);
    if (threadIdx.x == 0) {
        const double npts = (double)n * (double)n;
        const double ex = sqrt(red[0]) / sqrt(npts), ey = sqrt(red[1]) / sqrt(npts);
        const double ne = (ex < ey) ? ey : ex;
        ctl->norm_error = ne;
        if (!(ne == ne) || isinf(ne)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
        ctl->mom_done = 1;
    }
}

hipError_t launch_fin_explicit(const Ctx &c, int nblocks)
{
    hipLaunchKernelGGL(k_fin_explicit, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n);
    return hipGetLastError();
}

// plain weighted norm of (ub - uo) for the module entry point
__global__ __launch_bounds__(256) void k_norm_partials(const double *ub, const double *vb, const double *uo,
                                                       const double *vo, int n, int P, double *partials)
{
    __shared__ double sh[16];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double sx = 0.0, sy = 0.0;
    if (j < n && i <= n) {
        const double d = ub[i + (size_t)P * j] - uo[i + (size_t)P * j];
        sx = fabs(d * d * ((i > 0 && i < n) ? 1.0 : 0.5));
    }
    if (j <= n && i < n) {
        const double d = vb[i + (size_t)P * j] - vo[i + (size_t)P * j];
        sy = fabs(d * d * ((j > 0 && j < n) ? 1.0 : 0.5));
    }
    sx = block_reduce<false>(sx, sh);
    sy = block_reduce<false>(sy, sh + 8);
    if (threadIdx.x == 0) {
        const size_t b = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * b] = sx;
        partials[2 * b + 1] = sy;
    }
}

__global__ __launch_bounds__(1024) void k_fin_norm_only(const double *partials, int nb, int n, double *out)
{
    __shared__ double sh[32];
    double sx = 0.0, sy = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { sx += partials[2 * b]; sy += partials[2 * b + 1]; }
    sx = block_reduce<false>(sx, sh);
    sy = block_reduce<false>(sy, sh + 16);
    if (threadIdx.x == 0) {
        const double npts = (double)n * (double)n;
        const double ex = sqrt(sx) / sqrt(npts), ey = sqrt(sy) / sqrt(npts);
        *out = (ex < ey) ? ey : ex;
    }
}

hipError_t launch_norm(const Ctx &c, const double *ub, const double *vb, const double *uo, const double *vo,
                       double *out_dev)
{
    const int n = c.g.n;
    dim3 grid((n + 1 + 63) / 64, (n + 1 + 3) / 4);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_norm_partials, grid, dim3(256), 0, c.stream, ub, vb, uo, vo, n, c.g.P, c.partials);
    hipLaunchKernelGGL(k_fin_norm_only, dim3(1), dim3(1024), 0, c.stream, c.partials, (int)(grid.x * grid.y), n,
                       out_dev);
    return hipGetLastError();
}

// ------------------------------------------------------------ control
__global__ void k_ctl_reset_step(Ctl *ctl)
{
    ctl->mom_done = 0;
    ctl->mom_count = 0;
    ctl->flags = 0;
    ctl->norm_error = 1.e9;
}

hipError_t launch_ctl_reset_step(const Ctx &c)
{
    hipLaunchKernelGGL(k_ctl_reset_step, dim3(1), dim3(1), 0, c.stream, c.ctl);
    return hipGetLastError();
}

__global__ void k_ctl_set(Ctl *ctl, double tol_mom, double tol_rel, double tol_abs, int iter_num, int mg_max_iter)
{
    ctl->tol_mom = tol_mom;
    ctl->tol_rel = tol_rel;
    ctl->tol_abs = tol_abs;
    ctl->iter_num = iter_num;
    ctl->mg_max_iter = mg_max_iter;
}

hipError_t launch_ctl_set(const Ctx &c, double tol_mom, double tol_rel, double tol_abs, int iter_num,
                          int mg_max_iter)
{
    hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(1), 0, c.stream, c.ctl, tol_mom, tol_rel, tol_abs, iter_num,
                       mg_max_iter);
    return hipGetLastError();
}

// ---------------------------------------------------------------- BDF2
// main.cpp:879-884 with the dual-time constant folded once per step:
// c = (1.5/dt) u^n + (0.5/dt)(u^n - u^{n-1}); velContPrev <- velCont.
__global__ __launch_bounds__(256) void k_bdf2(double *u, double *up, double *c, int nx, int ny, int P, double dt)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)P * j;
    const double un = u[k], um = up[k];
    c[k] = (1.5 / dt) * un + (0.5 / dt) * (un - um);
    up[k] = un;
}

hipError_t launch_bdf2(const Ctx &c)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_bdf2, dim3((n + 1 + 63) / 64, (n + 3) / 4), dim3(256), 0, c.stream, c.U[0], c.Uprev[0],
                       c.C[0], n + 1, n, c.g.P, c.g.dt);
    hipLaunchKernelGGL(k_bdf2, dim3((n + 63) / 64, (n + 1 + 3) / 4), dim3(256), 0, c.stream, c.U[1], c.Uprev[1],
                       c.C[1], n, n + 1, c.g.P, c.g.dt);
    return hipGetLastError();
}

// The pressure-gradient term of the momentum RHS (fn_flux_calc.cpp:825-911,
// fn_rhs.cpp:8-68) folded into the BDF2 constant: p is the same in every
// stage of a step, so c += -(p(i) - p(i - 1)) / dx on the interior x-faces and
// likewise on the interior y-faces (the stage kernel's expression), and the
// FP64 stage then reads no p.  Wall faces keep c (their RHS is 0).
__global__ __launch_bounds__(256) void k_fold_gradp(double *cu, double *cv, const double *p, int n, int P,
                                                    double inv_dx, double inv_dy)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i > n || j > n) return;
    const size_t k = i + (size_t)P * j;
    if (i > 0 && i < n && j < n) cu[k] += -(p[k] - p[k - 1]) * inv_dx;
    if (j > 0 && j < n && i < n) cv[k] += -(p[k] - p[k - P]) * inv_dy;
}

hipError_t launch_fold_gradp(const Ctx &c)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_fold_gradp, dim3((n + 1 + 63) / 64, (n + 1 + 3) / 4), dim3(256), 0, c.stream, c.C[0], c.C[1],
                       (const double *)c.p, n, c.g.P, 1.0 / c.g.dx, 1.0 / c.g.dy);
    return hipGetLastError();
}

// ------------------------------------------------------------ Poisson RHS
// fn_rhs.cpp:71-125: D = (u_{i+1}-u_i)/dx + (v_{j+1}-v_j)/dy, b = (1.5/dt) D,
// block partials of ||D||inf and sum b.
// With au/av non-null the velocity is u + au (the mixed step's u^n + S).
__global__ __launch_bounds__(256) void k_poisson_rhs(const double *u, const double *v, const float *au,
                                                     const float *av, double *b, int n, int P, double dx, double dy,
                                                     double dt, double *partials)
{
    __shared__ double sh[16];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double m = 0.0, s = 0.0;
    if (i < n && j < n) {
        const size_t k = i + (size_t)P * j;
        double uw = u[k], ue = u[k + 1], vs = v[k], vn = v[k + P];
        if (au) { uw += (double)au[k]; ue += (double)au[k + 1]; vs += (double)av[k]; vn += (double)av[k + P]; }
        const double d = (ue - uw) / dx + (vn - vs) / dy;
        const double bb = (1.5 / dt) * d;
        b[k] = bb;
        m = fabs(d);
        s = bb;
    }
    m = block_reduce<true>(m, sh);
    s = block_reduce<false>(s, sh + 8);
    if (threadIdx.x == 0) {
        const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * bid] = m;
        partials[2 * bid + 1] = s;
    }
}

hipError_t launch_poisson_rhs(const Ctx &c, const double *u, const double *v, double *b, int *nblocks,
                              const float *au, const float *av)
{
    const int n = c.g.n;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    *nblocks = grid.x * grid.y;
    hipLaunchKernelGGL(k_poisson_rhs, grid, dim3(256), 0, c.stream, u, v, au, av, b, n, c.g.P, c.g.dx, c.g.dy, c.g.dt,
                       c.partials);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_fin_prhs(const double *partials, int nb, Ctl *ctl, int n)
{
    __shared__ double sh[32];
    double m = 0.0, s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double pm = partials[2 * b];
        m = (pm > m || pm != pm) ? pm : m;
        s += partials[2 * b + 1];
    }
    m = block_reduce<true>(m, sh);
    s = block_reduce<false>(s, sh + 16);
    if (threadIdx.x == 0) {
        ctl->div_star = m;
        ctl->rhs_sum = s;
        ctl->b_mean = s / ((double)n * (double)n);
    }
}

hipError_t launch_fin_prhs(const Ctx &c, int nblocks)
{
    hipLaunchKernelGGL(k_fin_prhs, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n);
    return hipGetLastError();
}

// mean of a cell array into ctl->b_mean (module Poisson entry point)
__global__ __launch_bounds__(256) void k_cell_sum(const double *b, int n, int P, double *partials)
{
    __shared__ double sh[16];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double s = (i < n && j < n) ? b[i + (size_t)P * j] : 0.0;
    s = block_reduce<false>(s, sh);
    if (threadIdx.x == 0) {
        const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * bid] = 0.0;
        partials[2 * bid + 1] = s;
    }
}

hipError_t launch_mean_b(const Ctx &c, const double *b)
{
    const int n = c.g.n;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cell_sum, grid, dim3(256), 0, c.stream, b, n, c.g.P, c.partials);
    hipLaunchKernelGGL(k_fin_prhs, dim3(1), dim3(1024), 0, c.stream, c.partials, (int)(grid.x * grid.y), c.ctl, n);
    return hipGetLastError();
}

// make the singular Neumann problem solvable: b <- b - mean(b)
__global__ __launch_bounds__(256) void k_center_b(double *b, int n, int P, const Ctl *ctl)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    b[i + (size_t)P * j] -= ctl->b_mean;
}

hipError_t launch_center_b(const Ctx &c, double *b)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_center_b, dim3((n + 63) / 64, (n + 3) / 4), dim3(256), 0, c.stream, b, n, c.g.P, c.ctl);
    return hipGetLastError();
}

// ------------------------------------------------------------ projection
// update_solution (poisson.cpp:192-392): u = u* - (G_x phi) dt / 1.5 on
// interior faces (wall faces keep u*: the Neumann ghost makes G phi = 0),
// p <- p + (phi - mean phi), and the partial ||D(u^{n+1})||inf.
// With the next step's BDF2 folded in (nx.cu non-null): on every face the
// constant of the next step, c = (1.5/dt) u^{n+1} + (0.5/dt)(u^{n+1} - u^n)
// with u^n from nx.up (as k_bdf2 forms it), and a copy of u^{n+1} to nx.w,
// which becomes that step's u^n; nx.up itself still holds u^n afterwards.
// nx.w null: no copy (the caller keeps u^n and u^{n+1} in separate arrays).
struct NextBdf {
    const double *up_u, *up_v;
    double *cu, *cv, *wu, *wv;
    int gradp;  // k_project_rows: also fold -G p^{n+1} into c (interior faces)
};
// A projection enqueued behind a batch of V-cycles before the host has seen
// the solve end: it does nothing unless the solve has ended (mg_done, not a
// skipped phase), and takes phi from the iterate buffer the cycle count left
// it in (phi after an even count, phi_alt after an odd one).
struct ProjGate {
    const double *phi_alt;
    int gated;
};
__device__ __forceinline__ bool proj_gate(const ProjGate &g, const Ctl *ctl, const double *&phi)
{
    if (g.gated && !(ctl->mg_done && !ctl->mg_skip)) return false;
    if (g.phi_alt && (ctl->mg_iters & 1)) phi = g.phi_alt;
    return true;
}

__global__ __launch_bounds__(256) void k_project(const double *us, const double *vs, const float *au,
                                                 const float *av, double *uo, double *vo, const double *p,
                                                 double *p_out, const double *phi, int n, int P, double dx,
                                                 double dy, double dt,
                                                 const Ctl *ctl, int use_mean, double *partials, int ntx, int nty,
                                                 NextBdf nx, ProjGate gate)
{
    __shared__ double sh[16];
    if (!proj_gate(gate, ctl, phi)) return;
    const CellTile t = cell_tile(ntx, nty);
    const double pm = use_mean ? ctl->phi_mean : 0.0;
    double m = 0.0;
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
        const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + 4 * pass + (threadIdx.x >> 6);
        if (!t.live || i >= n || j >= n) continue;
        const size_t k = i + (size_t)P * j;
        const double ph = phi[k];
        auto ux = [&](int fi) -> double {  // new x-face velocity at (fi, j)
            const size_t f = fi + (size_t)P * j;
            const double ustar = au ? us[f] + (double)au[f] : us[f];
            if (fi == 0 || fi == n) return ustar;
            const double g = (phi[f] - phi[f - 1]) / dx;
            return ustar - (g * dt / 1.5);
        };
        auto vy = [&](int fj) -> double {
            const size_t f = i + (size_t)P * fj;
            const double vstar = av ? vs[f] + (double)av[f] : vs[f];
            if (fj == 0 || fj == n) return vstar;
            const double g = (phi[f] - phi[f - P]) / dy;
            return vstar - (g * dt / 1.5);
        };
        const double uw = ux(i), ue = ux(i + 1), vsth = vy(j), vn = vy(j + 1);
        auto put = [&](double *o, const double *up, double *cc, double *wc, size_t f, double un) {
            o[f] = un;
            if (cc) {
                const double um = up[f];
                cc[f] = (1.5 / dt) * un + (0.5 / dt) * (un - um);
                if (wc) wc[f] = un;
            }
        };
        put(uo, nx.up_u, nx.cu, nx.wu, k, uw);
        put(vo, nx.up_v, nx.cv, nx.wv, k, vsth);
        if (i == n - 1) put(uo, nx.up_u, nx.cu, nx.wu, k + 1, ue);
        if (j == n - 1) put(vo, nx.up_v, nx.cv, nx.wv, k + P, vn);
        p_out[k] = p[k] + (ph - pm);
        const double dv = fabs((ue - uw) / dx + (vn - vsth) / dy);
        m = (dv > m || dv != dv) ? dv : m;
    }
    m = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// The same update in the row-march layout of the stage kernel: a wave owns a
// band of 62 columns (64 lanes: one halo lane on each side) and a chunk of
// rows.  phi of the rows j and j + 1 and the new y-face j stay in registers,
// phi(i - 1) and the new x-face i + 1 come by DPP, so each face is formed once
// and a cell costs one load of each input and one store of each output.  The
// next row's loads go out one iteration ahead, a row's stores one iteration
// behind (one wait for memory per row).  x-face n is written by the lane with
// i = n, y-face n by the row n - 1.  Per-face arithmetic is that of k_project;
// with nx.gradp (p_out separate from p) the next step's constant also takes
// -G p^{n+1} on the interior faces, as k_fold_gradp adds it: p^{n+1} of the
// west cell by DPP, of the south cell from the row before.
constexpr int kProjWB = 62;
__global__ __launch_bounds__(256) void k_project_rows(const double *us, const double *vs, const float *au,
                                                      const float *av, double *uo, double *vo, const double *p,
                                                      double *p_out, const double *phi, int n, int P, double dx,
                                                      double dy,
                                                      double dt, const Ctl *ctl, int use_mean, double *partials,
                                                      int nbands, int nwaves, int rows, NextBdf nx, ProjGate gate)
{
    __shared__ double sh[4];
    if (!proj_gate(gate, ctl, phi)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int w = blk * 4 + wv;
    const bool live = w < nwaves;  // a wave past the last one only joins the reduction
    const int band = live ? w % nbands : 0, chunk = live ? w / nbands : 0;
    const int i = band * kProjWB + lane - 1;
    const int j0 = chunk * rows, jend = live ? min(j0 + rows, n) : j0;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool own = lane >= 1 && lane <= kProjWB;
    const bool outc = xin && own, outx = xface && own;
    const bool fold = nx.cu != nullptr;
    const double pm = use_mean ? ctl->phi_mean : 0.0;
    // with nx.gradp, p_out is a separate array: p (p^n) is read-only here, so
    // the halo lane and the row before the chunk can form p^{n+1} of cells
    // that other waves own
    const bool gradp = fold && nx.gradp;
    const double inv_dx = 1.0 / dx, inv_dy = 1.0 / dy;
    struct Raw {
        double phn;       // phi (i, j + 1)
        double us, vsn;   // u* at x-face (i, j), v* at y-face (i, j + 1)
        double pc;        // p (i, j)
        double upu, upv;  // u^n at the faces (i, j) (fold)
        double upvn;      // u^n at y-face (i, n) (fold, row n - 1)
    };
    auto load = [&](int j, Raw &q) {
        q.phn = q.us = q.vsn = q.pc = q.upu = q.upv = q.upvn = 0.0;
        const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * j;
        if (xin && j + 1 < n) q.phn = phi[k + P];
        if (xface) q.us = au ? us[k] + (double)au[k] : us[k];
        if (xin) q.vsn = av ? vs[k + P] + (double)av[k + P] : vs[k + P];
        if (outc || (gradp && xin)) q.pc = p[k];
        if (fold) {
            if (outx) q.upu = nx.up_u[k];
            if (outc) {
                q.upv = nx.up_v[k];
                if (j == n - 1) q.upvn = nx.up_v[k + P];
            }
        }
    };
    auto bdf = [&](double un, double um) { return (1.5 / dt) * un + (0.5 / dt) * (un - um); };
    // prologue: phi (i, j0), the new y-face (i, j0) and, folding -G p, p^{n+1} (i, j0 - 1)
    double phC = 0.0, vprev = 0.0, pnS = 0.0;
    if (jend > j0 && xin) {
        const size_t k = i + (size_t)P * j0;
        phC = phi[k];
        const double vstar = av ? vs[k] + (double)av[k] : vs[k];
        vprev = j0 == 0 ? vstar : vstar - (((phC - phi[k - P]) / dy) * dt / 1.5);
        if (gradp && j0 > 0) pnS = p[k - P] + (phi[k - P] - pm);
    }
    double m = 0.0;
    // deferred stores of row j - 1
    size_t sk = 0;
    bool sx = false, sy = false, syn = false;
    double su = 0.0, scu = 0.0, sv = 0.0, scv = 0.0, sp = 0.0, svn = 0.0, scvn = 0.0;
    auto flush = [&]() {
        if (sx) {
            uo[sk] = su;
            if (fold) { nx.cu[sk] = scu; if (nx.wu) nx.wu[sk] = su; }
        }
        if (sy) {
            vo[sk] = sv;
            p_out[sk] = sp;
            if (fold) { nx.cv[sk] = scv; if (nx.wv) nx.wv[sk] = sv; }
        }
        if (syn) {
            vo[sk + P] = svn;
            if (fold) { nx.cv[sk + P] = scvn; if (nx.wv) nx.wv[sk + P] = svn; }
        }
    };
    Raw cur, nxt;
    if (jend > j0) load(j0, cur);
    for (int j = j0; j < jend; ++j) {
        if (j + 1 < jend) load(j + 1, nxt);  // prefetch
        flush();
        sx = sy = syn = false;
        const double phW = lane_prev(phC);  // (every lane shifts: no divergence around DPP)
        const double uw = (i == 0 || i == n) ? cur.us : cur.us - (((phC - phW) / dx) * dt / 1.5);
        const double ue = lane_next(uw);
        const double vn = j + 1 == n ? cur.vsn : cur.vsn - (((cur.phn - phC) / dy) * dt / 1.5);
        const double pn = cur.pc + (phC - pm);  // p^{n+1} (i, j)
        const double pnW = lane_prev(pn);
        sk = (size_t)(i < 0 ? 0 : i) + (size_t)P * j;
        if (outx) {
            sx = true;
            su = uw;
            if (fold) scu = bdf(uw, cur.upu);
            if (gradp && i > 0 && i < n) scu += -(pn - pnW) * inv_dx;  // as k_fold_gradp adds it
        }
        if (outc) {
            sy = true;
            sv = vprev;
            sp = pn;
            if (fold) scv = bdf(vprev, cur.upv);
            if (gradp && j > 0) scv += -(pn - pnS) * inv_dy;
            if (j == n - 1) {
                syn = true;
                svn = vn;
                if (fold) scvn = bdf(vn, cur.upvn);
            }
            const double dv = fabs((ue - uw) / dx + (vn - vprev) / dy);
            m = (dv > m || dv != dv) ? dv : m;
        }
        phC = cur.phn;
        vprev = vn;
        pnS = pn;
        cur = nxt;
    }
    flush();
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_down(m, o, 64);
        m = (a > m || a != a) ? a : m;
    }
    if (lane == 0) sh[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = sh[0];
        for (int x = 1; x < 4; ++x) q = (sh[x] > q || sh[x] != sh[x]) ? sh[x] : q;
        partials[blockIdx.x] = q;
    }
}

hipError_t launch_project_ex(const Ctx &c, const double *us, const double *vs, double *uo, double *vo, double *p,
                             const double *phi, int use_mean, int *nblocks, const float *au, const float *av,
                             int next_bdf, double *p_out, const double *phi_alt, int gated, double *const *un,
                             int no_copy)
{
    if (!p_out) p_out = p;
    const ProjGate gate{phi_alt, gated};
    const int gradp = next_bdf && p_out != p && c.project_rows > 0;
    double *const *up = un ? un : c.Uprev;
    const NextBdf nx = next_bdf ? NextBdf{up[0], up[1], c.C[0], c.C[1], no_copy ? nullptr : c.S2[0],
                                          no_copy ? nullptr : c.S2[1], gradp}
                                : NextBdf{};
    const int n = c.g.n;
    if (c.project_rows > 0) {  // row-march kernel
        const int rows = c.project_rows;
        const int nbands = (n + 1 + kProjWB - 1) / kProjWB;  // x-face n needs a lane inside a band
        const int nwaves = nbands * ((n + rows - 1) / rows);
        const int nb = ((nwaves + 3) / 4 + 7) / 8 * 8;  // four waves per workgroup, a multiple of 8
        if (nb > kMaxPartials) return hipErrorInvalidValue;
        *nblocks = nb;
        hipLaunchKernelGGL(k_project_rows, dim3(nb), dim3(256), 0, c.stream, us, vs, au, av, uo, vo, (const double *)p,
                           p_out, phi, n, c.g.P, c.g.dx, c.g.dy, c.g.dt, (const Ctl *)c.ctl, use_mean, c.partials,
                           nbands, nwaves, rows, nx, gate);
        return hipGetLastError();
    }
    const int ntx = (n + 63) / 64, nty = (n + 15) / 16;
    const int nb = (ntx * nty + 7) / 8 * 8;  // 64 x 16 tiles, a multiple of 8 workgroups
    if (nb > kMaxPartials) return hipErrorInvalidValue;
    *nblocks = nb;
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), 0, c.stream, us, vs, au, av, uo, vo, (const double *)p, p_out, phi, n, c.g.P, c.g.dx,
                       c.g.dy, c.g.dt, c.ctl, use_mean, c.partials, ntx, nty, nx, gate);
    return hipGetLastError();
}

// reset_next: also clear the pseudo-iteration control for the next step (as
// k_ctl_reset_step does), after this step's statistics are logged
__global__ __launch_bounds__(1024) void k_fin_project(const double *partials, int nb, Ctl *ctl,
                                                      amressif_step_stats *log, int step_index, int reset_next,
                                                      int gated)
{
    __shared__ double sh[16];
    const bool skip = gated && !(ctl->mg_done && !ctl->mg_skip);  // its projection did nothing (ProjGate)
    double m = 0.0;
    for_records<1, 1>(partials, nb, [&](const double *o) {  // (loads issued before the flag is waited for)
        m = (o[0] > m || o[0] != o[0]) ? o[0] : m;
    });
    if (skip) return;
    m = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) {
        ctl->div_new = m;
        if (log && step_index >= 0) {
            amressif_step_stats s;
            s.pseudo_iters = ctl->mom_count;
            s.vcycles = ctl->mg_iters;
            s.flags = ctl->flags;
            s.pad = 0;
            s.norm_error = ctl->norm_error;
            s.div_star = ctl->div_star;
            s.rhs_sum = ctl->rhs_sum;
            s.poisson_res = ctl->b_norm > 0.0 ? ctl->res_norm / ctl->b_norm : ctl->res_norm;
            s.div_new = m;
            log[step_index] = s;
        }
        if (reset_next) {
            ctl->mom_done = 0;
            ctl->mom_count = 0;
            ctl->flags = 0;
            ctl->norm_error = 1.e9;
        }
    }
}

hipError_t launch_fin_project(const Ctx &c, int nblocks, int step_index, int reset_next, int gated)
{
    hipLaunchKernelGGL(k_fin_project, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.log,
                       step_index, reset_next, gated);
    return hipGetLastError();
}

// ----------------------------------------------------- layout transforms
// Array4 views: element (i, j, comp) at ptr[(i-lo0) + js*(j-lo1) + ns*comp].
struct A4 {
    double *ptr;
    int lo0, lo1, hi0, hi1;
    int64_t js, ns;
};

static A4 make_a4(const amr_array4 &a)
{
    A4 r;
    r.ptr = (double *)a.ptr;
    r.lo0 = a.lo[0]; r.lo1 = a.lo[1]; r.hi0 = a.hi[0]; r.hi1 = a.hi[1];
    r.js = (int64_t)(a.hi[0] - a.lo[0] + 1);
    r.ns = r.js * (int64_t)(a.hi[1] - a.lo[1] + 1) * (int64_t)(a.hi[2] - a.lo[2] + 1);
    return r;
}

__global__ __launch_bounds__(256) void k_pack(A4 a, int comp, double *dst, int nx, int ny, int P)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    dst[i + (size_t)P * j] = a.ptr[(i - a.lo0) + a.js * (j - a.lo1) + a.ns * comp];
}

__global__ __launch_bounds__(256) void k_unpack(A4 a, int comp, const double *src, int nx, int ny, int P,
                                                const double *shift)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const double s = shift ? *shift : 0.0;
    a.ptr[(i - a.lo0) + a.js * (j - a.lo1) + a.ns * comp] = src[i + (size_t)P * j] - s;
}

// cells including the ghosts the view holds: Neumann mirror in both
// directions (main.cpp:1078-1158; corners are mirror-of-mirror)
__global__ __launch_bounds__(256) void k_unpack_cell_ghosts(A4 a, int comp, const double *src, int n, int P,
                                                            const double *shift)
{
    const int i = a.lo0 + blockIdx.x * 64 + (threadIdx.x & 63), j = a.lo1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i > a.hi0 || j > a.hi1) return;
    const int mi = i < 0 ? -i - 1 : (i >= n ? 2 * n - 1 - i : i);
    const int mj = j < 0 ? -j - 1 : (j >= n ? 2 * n - 1 - j : j);
    if (mi < 0 || mi >= n || mj < 0 || mj >= n) return;
    const double s = shift ? *shift : 0.0;
    a.ptr[(i - a.lo0) + a.js * (j - a.lo1) + a.ns * comp] = src[mi + (size_t)P * mj] - s;
}

// one box of a multi-box MultiFab: the window [i0, i0 + wx) x [j0, j0 + wy)
// (global indices, inside the view) between the view and the pitched array
__global__ __launch_bounds__(256) void k_pack_win(A4 a, int comp, double *dst, int i0, int j0, int wx, int wy,
                                                  int P)
{
    const int di = blockIdx.x * 64 + (threadIdx.x & 63), dj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (di >= wx || dj >= wy) return;
    const int i = i0 + di, j = j0 + dj;
    dst[i + (size_t)P * j] = a.ptr[(i - a.lo0) + a.js * (j - a.lo1) + a.ns * comp];
}

__global__ __launch_bounds__(256) void k_unpack_win(A4 a, int comp, const double *src, int i0, int j0, int wx,
                                                    int wy, int P)
{
    const int di = blockIdx.x * 64 + (threadIdx.x & 63), dj = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (di >= wx || dj >= wy) return;
    const int i = i0 + di, j = j0 + dj;
    a.ptr[(i - a.lo0) + a.js * (j - a.lo1) + a.ns * comp] = src[i + (size_t)P * j];
}

hipError_t launch_pack_window(const Ctx &c, const amr_array4 &a, int comp, double *dst, int i0, int j0, int i1,
                              int j1)
{
    const int wx = i1 - i0 + 1, wy = j1 - j0 + 1;
    if (wx <= 0 || wy <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_win, dim3((wx + 63) / 64, (wy + 3) / 4), dim3(256), 0, c.stream, make_a4(a), comp,
                       dst, i0, j0, wx, wy, c.g.P);
    return hipGetLastError();
}

hipError_t launch_unpack_window(const Ctx &c, const double *src, const amr_array4 &a, int comp, int i0, int j0,
                                int i1, int j1)
{
    const int wx = i1 - i0 + 1, wy = j1 - j0 + 1;
    if (wx <= 0 || wy <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack_win, dim3((wx + 63) / 64, (wy + 3) / 4), dim3(256), 0, c.stream, make_a4(a), comp,
                       src, i0, j0, wx, wy, c.g.P);
    return hipGetLastError();
}

hipError_t launch_pack_face(const Ctx &c, const amr_array4 &a, double *dst, int dir)
{
    const int n = c.g.n, nx = dir == 0 ? n + 1 : n, ny = dir == 0 ? n : n + 1;
    hipLaunchKernelGGL(k_pack, dim3((nx + 63) / 64, (ny + 3) / 4), dim3(256), 0, c.stream, make_a4(a), 0, dst, nx,
                       ny, c.g.P);
    return hipGetLastError();
}

hipError_t launch_unpack_face(const Ctx &c, const double *src, const amr_array4 &a, int dir)
{
    const int n = c.g.n, nx = dir == 0 ? n + 1 : n, ny = dir == 0 ? n : n + 1;
    hipLaunchKernelGGL(k_unpack, dim3((nx + 63) / 64, (ny + 3) / 4), dim3(256), 0, c.stream, make_a4(a), 0, src,
                       nx, ny, c.g.P, (const double *)nullptr);
    return hipGetLastError();
}

hipError_t launch_pack_cell(const Ctx &c, const amr_array4 &a, int comp, double *dst)
{
    const int n = c.g.n;
    hipLaunchKernelGGL(k_pack, dim3((n + 63) / 64, (n + 3) / 4), dim3(256), 0, c.stream, make_a4(a), comp, dst, n,
                       n, c.g.P);
    return hipGetLastError();
}

hipError_t launch_unpack_cell_shift(const Ctx &c, const double *src, const amr_array4 &a, int comp,
                                    int fill_ghosts, const double *shift)
{
    const int n = c.g.n;
    if (fill_ghosts) {
        const int wx = a.hi[0] - a.lo[0] + 1, wy = a.hi[1] - a.lo[1] + 1;
        hipLaunchKernelGGL(k_unpack_cell_ghosts, dim3((wx + 63) / 64, (wy + 3) / 4), dim3(256), 0, c.stream,
                           make_a4(a), comp, src, n, c.g.P, shift);
    } else {
        hipLaunchKernelGGL(k_unpack, dim3((n + 63) / 64, (n + 3) / 4), dim3(256), 0, c.stream, make_a4(a), comp,
                           src, n, n, c.g.P, shift);
    }
    return hipGetLastError();
}

hipError_t launch_unpack_cell(const Ctx &c, const double *src, const amr_array4 &a, int comp, int fill_ghosts)
{
    return launch_unpack_cell_shift(c, src, a, comp, fill_ghosts, nullptr);
}

// cont2cart into an Array4 velCart (2 comps, 2 ghosts), including the
// corner values the reference's W, E, S, N order leaves (rule_y(rule_x)).
__global__ __launch_bounds__(256) void k_cont2cart_out(A4 a, const double *u, const double *v, int n, int P,
                                                      int lo0, int hi0, int lo1, int hi1, double in0, double in1,
                                                      const double *lid)
{
    const int i = a.lo0 + blockIdx.x * 64 + (threadIdx.x & 63), j = a.lo1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i > a.hi0 || j > a.hi1) return;
    if (i < -2 || i > n + 1 || j < -2 || j > n + 1) return;
    const int mi = i < 0 ? -i - 1 : (i >= n ? 2 * n - 1 - i : i);
    const int mj = j < 0 ? -j - 1 : (j >= n ? 2 * n - 1 - j : j);
    const size_t k = mi + (size_t)P * mj;
    double U = 0.5 * (u[k] + u[k + 1]);
    double V = 0.5 * (v[k] + v[k + P]);
    if (i != mi) {
        double gu, gv;
        ghost_rule(i < 0 ? 0 : 1, i < 0 ? lo0 : hi0, U, V, 0.0, 0.0, gu, gv);
        U = gu; V = gv;
    }
    if (j != mj) {
        double gu, gv;
        const double gx = (lid && i >= 0 && i < n) ? lid[i] : in0;
        ghost_rule(j < 0 ? 2 : 3, j < 0 ? lo1 : hi1, U, V, gx, in1, gu, gv);
        U = gu; V = gv;
    }
    const int64_t o = (i - a.lo0) + a.js * (j - a.lo1);
    a.ptr[o] = U;
    a.ptr[o + a.ns] = V;
}

hipError_t launch_cont2cart_out(const Ctx &c, const double *u, const double *v, const amr_array4 &vc)
{
    const int wx = vc.hi[0] - vc.lo[0] + 1, wy = vc.hi[1] - vc.lo[1] + 1;
    const Geometry &g = c.g;
    hipLaunchKernelGGL(k_cont2cart_out, dim3((wx + 63) / 64, (wy + 3) / 4), dim3(256), 0, c.stream, make_a4(vc), u,
                       v, g.n, g.P, g.bc_lo0, g.bc_hi0, g.bc_lo1, g.bc_hi1, g.inflow0, g.inflow1, c.lid);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_copy2d(const double *src, double *dst, int nx, int ny, int P)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    dst[i + (size_t)P * j] = src[i + (size_t)P * j];
}

hipError_t launch_copy_faces(const Ctx &c, const double *src, double *dst, int dir)
{
    const int n = c.g.n, nx = dir == 0 ? n + 1 : n, ny = dir == 0 ? n : n + 1;
    hipLaunchKernelGGL(k_copy2d, dim3((nx + 63) / 64, (ny + 3) / 4), dim3(256), 0, c.stream, src, dst, nx, ny,
                       c.g.P);
    return hipGetLastError();
}

// plain 16-byte-vector copy (bandwidth calibration)
__global__ __launch_bounds__(256) void k_copy_bytes(const double2 *src, double2 *dst, int64_t n16)
{
    for (int64_t k = blockIdx.x * (int64_t)256 + threadIdx.x; k < n16; k += (int64_t)gridDim.x * 256) dst[k] = src[k];
}

hipError_t launch_copy_bytes(const Ctx &c, const void *src, void *dst, int64_t bytes)
{
    const int64_t n16 = bytes / 16;
    const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_copy_bytes, dim3((unsigned)blocks), dim3(256), 0, c.stream, (const double2 *)src,
                       (double2 *)dst, n16);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill(double *dst, int64_t count, double value)
{
    const int64_t k = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (k < count) dst[k] = value;
}

hipError_t launch_fill(const Ctx &c, double *dst, int64_t count, double value)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c.stream, dst, count, value);
    return hipGetLastError();
}

// build the combined BDF2 constant from two previous states (module entry)
__global__ __launch_bounds__(256) void k_bdf2_const(const double *un, const double *um, double *c, int nx, int ny,
                                                    int P, double dt)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)P * j;
    c[k] = (1.5 / dt) * un[k] + (0.5 / dt) * (un[k] - um[k]);
}

hipError_t launch_bdf2_const(const Ctx &c, const double *un, const double *um, double *cc, int dir)
{
    const int n = c.g.n, nx = dir == 0 ? n + 1 : n, ny = dir == 0 ? n : n + 1;
    hipLaunchKernelGGL(k_bdf2_const, dim3((nx + 63) / 64, (ny + 3) / 4), dim3(256), 0, c.stream, un, um, cc, nx, ny,
                       c.g.P, c.g.dt);
    return hipGetLastError();
}

}  // namespace amr
