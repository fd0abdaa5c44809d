// Geometric multigrid for the pressure-Poisson projection, gfx950.
//
// Operator: L phi = (E+W-2P)/dx^2 + (N+S-2P)/dy^2 with homogeneous Neumann
// walls (the reference's MLABecLaplacian with alpha = 0, beta = -1, B = 1 and
// foextrap -> Neumann, Source/src1_rever/poisson.cpp:85-153).  The Neumann
// mirror drops the out-of-domain neighbour from the stencil and from the
// diagonal, so no ghost cells are stored.
//
// Solve (poisson.cpp:181-185): iterate phi += V(b - L phi) until
// ||b - L phi||inf <= max(tol_abs, tol_rel ||b||inf) with b made solvable by
// removing its mean.  V is one V-cycle on the correction from zero: two
// red-black Gauss-Seidel smooths before and after, 2x2 average restriction,
// piecewise-constant prolongation, and a dense pseudo-inverse solve on the
// coarsest level.  The V-cycle is templated on its working precision: FP64,
// or FP32 inside FP64 iterative refinement (FP64 residual, correction and
// reductions).
#include "internal.hpp"

#include <cmath>

namespace amr {

template <bool IsMax>
__device__ double block_reduce_mg(double v, double *sh)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_down(v, o, 64);
        if (IsMax) v = (w > v || w != w) ? w : v;
        else v += w;
    }
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sh[0];
        for (int w = 1; w < nw; ++w) {
            if (IsMax) r = (sh[w] > r || sh[w] != sh[w]) ? sh[w] : r;
            else r += sh[w];
        }
    }
    return r;
}

// ----------------------------------------------------------- smoothing
// one colour of red-black Gauss-Seidel: colour c updates (i+j)%2 == c; a
// thread owns one cell of that colour.
template <class T>
__global__ __launch_bounds__(256) void k_gsrb(T *e, const T *r, int nx, int ny, int pitch, T ax, T ay, int color,
                                              const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int t = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = 2 * t + ((j + color) & 1);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)pitch * j;
    T s = 0, d = 0;
    if (i > 0) { s += ax * e[k - 1]; d += ax; }
    if (i < nx - 1) { s += ax * e[k + 1]; d += ax; }
    if (j > 0) { s += ay * e[k - pitch]; d += ay; }
    if (j < ny - 1) { s += ay * e[k + pitch]; d += ay; }
    e[k] = (s - r[k]) / d;
}

template <class T>
__device__ __forceinline__ T lap_at(const T *e, int i, int j, int nx, int ny, int pitch, T ax, T ay)
{
    const size_t k = i + (size_t)pitch * j;
    const T P = e[k];
    const T W = i > 0 ? e[k - 1] : P;
    const T E = i < nx - 1 ? e[k + 1] : P;
    const T S = j > 0 ? e[k - pitch] : P;
    const T N = j < ny - 1 ? e[k + pitch] : P;
    return (E + W - T(2) * P) * ax + (N + S - T(2) * P) * ay;
}

// residual of the correction equation on the fine level, averaged over each
// 2x2 block into the coarse right-hand side
template <class T>
__global__ __launch_bounds__(256) void k_resid_restrict(const T *e, const T *r, int nx, int ny, int pf, T ax, T ay,
                                                        T *rc, int ncx, int ncy, int pc, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int I = blockIdx.x * 64 + (threadIdx.x & 63), J = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (I >= ncx || J >= ncy) return;
    const int i = 2 * I, j = 2 * J;
    const T t00 = r[i + (size_t)pf * j] - lap_at(e, i, j, nx, ny, pf, ax, ay);
    const T t10 = r[i + 1 + (size_t)pf * j] - lap_at(e, i + 1, j, nx, ny, pf, ax, ay);
    const T t01 = r[i + (size_t)pf * (j + 1)] - lap_at(e, i, j + 1, nx, ny, pf, ax, ay);
    const T t11 = r[i + 1 + (size_t)pf * (j + 1)] - lap_at(e, i + 1, j + 1, nx, ny, pf, ax, ay);
    rc[I + (size_t)pc * J] = T(0.25) * (((t00 + t10) + t01) + t11);
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(T *ef, int nx, int ny, int pf, const T *ec, int pc,
                                                     const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    ef[i + (size_t)pf * j] += ec[(i >> 1) + (size_t)pc * (j >> 1)];
}

template <class T>
__global__ __launch_bounds__(256) void k_zero_level(T *e, int nx, int ny, int pitch, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    e[i + (size_t)pitch * j] = T(0);
}

// coarsest level: e = pinv(L) r, pinv precomputed on the host (mean-zero
// least-squares solution of the singular Neumann problem)
template <class T>
__global__ __launch_bounds__(256) void k_bottom(const T *pinv, const T *r, int nx, int ny, int pitch, T *e,
                                                const Ctl *ctl)
{
    if (ctl->mg_done) return;
    __shared__ T sr[1024];
    const int m = nx * ny;
    for (int k = threadIdx.x; k < m; k += blockDim.x) sr[k] = r[(k % nx) + (size_t)pitch * (k / nx)];
    __syncthreads();
    for (int row = threadIdx.x; row < m; row += blockDim.x) {
        T s = 0;
        const T *pr = pinv + (size_t)row * m;
        for (int k = 0; k < m; ++k) s += pr[k] * sr[k];
        e[(row % nx) + (size_t)pitch * (row / nx)] = s;
    }
}

template <class T>
static hipError_t vcycle_impl(const Ctx &c, T *const *E, T *const *R, const T *pinv)
{
    const Ctl *ctl = c.ctl;
    for (int l = 0; l < c.nlev - 1; ++l) {
        const Level &L = c.lev[l], &Cl = c.lev[l + 1];
        const T ax = (T)L.ax, ay = (T)L.ay;
        dim3 gc((L.nx + 63) / 64, (L.ny + 3) / 4), gh(((L.nx + 1) / 2 + 63) / 64, (L.ny + 3) / 4);
        hipLaunchKernelGGL(k_zero_level<T>, gc, dim3(256), 0, c.stream, E[l], L.nx, L.ny, L.pitch, ctl);
        for (int s = 0; s < 2; ++s)
            for (int col = 0; col < 2; ++col)
                hipLaunchKernelGGL(k_gsrb<T>, gh, dim3(256), 0, c.stream, E[l], (const T *)R[l], L.nx, L.ny, L.pitch,
                                   ax, ay, col, ctl);
        hipLaunchKernelGGL(k_resid_restrict<T>, dim3((Cl.nx + 63) / 64, (Cl.ny + 3) / 4), dim3(256), 0, c.stream,
                           (const T *)E[l], (const T *)R[l], L.nx, L.ny, L.pitch, ax, ay, R[l + 1], Cl.nx, Cl.ny,
                           Cl.pitch, ctl);
    }
    {
        const Level &B = c.lev[c.nlev - 1];
        hipLaunchKernelGGL(k_bottom<T>, dim3(1), dim3(256), 0, c.stream, pinv, (const T *)R[c.nlev - 1], B.nx, B.ny,
                           B.pitch, E[c.nlev - 1], ctl);
    }
    for (int l = c.nlev - 2; l >= 0; --l) {
        const Level &L = c.lev[l], &Cl = c.lev[l + 1];
        const T ax = (T)L.ax, ay = (T)L.ay;
        dim3 gc((L.nx + 63) / 64, (L.ny + 3) / 4), gh(((L.nx + 1) / 2 + 63) / 64, (L.ny + 3) / 4);
        hipLaunchKernelGGL(k_prolong_add<T>, gc, dim3(256), 0, c.stream, E[l], L.nx, L.ny, L.pitch,
                           (const T *)E[l + 1], Cl.pitch, ctl);
        for (int s = 0; s < 2; ++s)
            for (int col = 0; col < 2; ++col)
                hipLaunchKernelGGL(k_gsrb<T>, gh, dim3(256), 0, c.stream, E[l], (const T *)R[l], L.nx, L.ny, L.pitch,
                                   ax, ay, col, ctl);
    }
    return hipGetLastError();
}

// diagnostic: one unguarded colour sweep on the finest level (FP64 arrays)
__global__ __launch_bounds__(256) void k_gsrb_free(double *e, const double *r, int nx, int ny, int pitch, double ax,
                                                   double ay, int color)
{
    const int t = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = 2 * t + ((j + color) & 1);
    if (i >= nx || j >= ny) return;
    const size_t k = i + (size_t)pitch * j;
    double s = 0, d = 0;
    if (i > 0) { s += ax * e[k - 1]; d += ax; }
    if (i < nx - 1) { s += ax * e[k + 1]; d += ax; }
    if (j > 0) { s += ay * e[k - pitch]; d += ay; }
    if (j < ny - 1) { s += ay * e[k + pitch]; d += ay; }
    e[k] = (s - r[k]) / d;
}

hipError_t launch_gsrb_fine(const Ctx &c, int color)
{
    const Level &L = c.lev[0];
    dim3 gh(((L.nx + 1) / 2 + 63) / 64, (L.ny + 3) / 4);
    hipLaunchKernelGGL(k_gsrb_free, gh, dim3(256), 0, c.stream, L.e, (const double *)L.r, L.nx, L.ny, L.pitch, L.ax,
                       L.ay, color);
    return hipGetLastError();
}

hipError_t launch_vcycle(const Ctx &c, int mixed)
{
    if (mixed) {
        float *E[kMaxLevels], *R[kMaxLevels];
        for (int l = 0; l < c.nlev; ++l) { E[l] = c.lev[l].e32; R[l] = c.lev[l].r32; }
        return vcycle_impl<float>(c, E, R, c.pinv32);
    }
    double *E[kMaxLevels], *R[kMaxLevels];
    for (int l = 0; l < c.nlev; ++l) { E[l] = c.lev[l].e; R[l] = c.lev[l].r; }
    return vcycle_impl<double>(c, E, R, c.pinv);
}

// ----------------------------------------------- outer FP64 iteration
__global__ void k_mg_start(Ctl *ctl)
{
    ctl->mg_done = 0;
    ctl->mg_iters = 0;
    ctl->mg_started = 0;
    ctl->res_norm = 0.0;
    ctl->b_norm = 0.0;
    ctl->phi_sum = 0.0;
    ctl->phi_mean = 0.0;
}

hipError_t launch_mg_start(const Ctx &c)
{
    hipLaunchKernelGGL(k_mg_start, dim3(1), dim3(1), 0, c.stream, c.ctl);
    return hipGetLastError();
}

// r = b - L phi on the fine level (FP64), optionally rounded to FP32 for the
// mixed V-cycle; block partials of ||r||inf and sum phi
__global__ __launch_bounds__(256) void k_residual0(const double *phi, const double *b, double *r, float *r32, int n,
                                                   int P, double ax, double ay, const Ctl *ctl, double *partials)
{
    __shared__ double sh[16];
    if (ctl->mg_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double m = 0.0, s = 0.0;
    if (i < n && j < n) {
        const size_t k = i + (size_t)P * j;
        const double rr = b[k] - lap_at(phi, i, j, n, n, P, ax, ay);
        if (r) r[k] = rr;
        if (r32) r32[k] = (float)rr;
        m = fabs(rr);
        s = phi[k];
    }
    m = block_reduce_mg<true>(m, sh);
    s = block_reduce_mg<false>(s, sh + 8);
    if (threadIdx.x == 0) {
        const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * bid] = m;
        partials[2 * bid + 1] = s;
    }
}

hipError_t launch_residual(const Ctx &c, const double *phi, const double *b, double *r, int *nblocks)
{
    const int n = c.g.n;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    *nblocks = grid.x * grid.y;
    const bool mixed = c.in.precision == AMRESSIF_MIXED;
    hipLaunchKernelGGL(k_residual0, grid, dim3(256), 0, c.stream, phi, b, mixed ? nullptr : r,
                       mixed ? c.lev[0].r32 : nullptr, n, c.g.P, c.lev[0].ax, c.lev[0].ay, c.ctl, c.partials);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_fin_resid(const double *partials, int nb, Ctl *ctl, int n)
{
    __shared__ double sh[32];
    if (ctl->mg_done) return;
    double m = 0.0, s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double pm = partials[2 * b];
        m = (pm > m || pm != pm) ? pm : m;
        s += partials[2 * b + 1];
    }
    m = block_reduce_mg<true>(m, sh);
    s = block_reduce_mg<false>(s, sh + 16);
    if (threadIdx.x == 0) {
        ctl->res_norm = m;
        ctl->phi_sum = s;
        ctl->phi_mean = s / ((double)n * (double)n);
        if (!ctl->mg_started) {
            ctl->b_norm = m;  // phi = 0 on the first evaluation
            ctl->mg_started = 1;
        } else {
            ctl->mg_iters += 1;
        }
        const double thr = fmax(ctl->tol_abs, ctl->tol_rel * ctl->b_norm);
        if (m <= thr || !(m == m)) ctl->mg_done = 1;
        else if (ctl->mg_iters >= ctl->mg_max_iter) { ctl->mg_done = 1; ctl->flags |= AMRESSIF_FLAG_MG_CAP; }
    }
}

hipError_t launch_fin_resid(const Ctx &c, int nblocks)
{
    hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n);
    return hipGetLastError();
}

template <class T>
__global__ __launch_bounds__(256) void k_add_correction(double *phi, const T *e, int n, int P, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const size_t k = i + (size_t)P * j;
    phi[k] += (double)e[k];
}

hipError_t launch_add_correction(const Ctx &c, double *phi, int mixed)
{
    const int n = c.g.n;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if (mixed)
        hipLaunchKernelGGL(k_add_correction<float>, grid, dim3(256), 0, c.stream, phi, (const float *)c.lev[0].e32, n,
                           c.g.P, c.ctl);
    else
        hipLaunchKernelGGL(k_add_correction<double>, grid, dim3(256), 0, c.stream, phi, (const double *)c.lev[0].e, n,
                           c.g.P, c.ctl);
    return hipGetLastError();
}

}  // namespace amr
