// Restarted flexible GMRES for the pressure-Poisson equation, gfx950.
//
// The reference's GMRESPOISSON (GMRES_Poisson.H:14-76, GMRES_Poisson.cpp:5-136)
// plugs a matrix-free Poisson operator into amrex::GMRES, but its apply()
// fills only periodic ghosts (not the Neumann operator of the MLMG solve) and
// its preconditioner is 8 weighted-Jacobi sweeps; the step never calls it.
// This is the solver that path calls for: FGMRES(m) on the same Neumann
// 5-point operator as the multigrid solve, right-preconditioned by one
// V-cycle (FP64, or FP32 in the mixed mode), all vectors and the Arnoldi
// process in FP64, orthogonalisation by classical Gram-Schmidt applied twice,
// dot products as deterministic block partials reduced in one workgroup.
// Stopping test as the multigrid solve: ||b - L x||inf <= max(tol_abs,
// tol_rel ||b||inf), checked with the true residual at every restart; inside
// a cycle the Givens residual estimate (2-norm) is compared with the target
// scaled by the cycle's ||r||_2 / ||r||inf.
//
// Device-side control (Ctl): mg_done ends the solve (every V-cycle kernel
// already honours it), gm_stop ends the current Arnoldi cycle (the V-cycle
// kernels honour it too).
#include "device_common.hpp"
#include "internal.hpp"

namespace amr {

constexpr int GX = 64, GY = 16;  // cells per GMRES block (256 threads, 4 rows each)

struct GmArgs {
    int n, P;
    double ax, ay;
    int m;            // restart length
    int nblk;         // blocks of the vector kernels
    double *V;        // m+1 Arnoldi vectors, stride vs
    double *Z;        // m preconditioned vectors, stride vs
    size_t vs;
    GmState *st;
    double *part;     // nblk x (m + 2) partials
    Ctl *ctl;
};

__device__ __forceinline__ bool gm_skip(const GmArgs &g) { return g.ctl->mg_done | g.ctl->gm_stop; }

__device__ __forceinline__ double lap(const double *x, int i, int j, int n, int P, double ax, double ay)
{
    const size_t k = i + (size_t)P * j;
    const double c = x[k];
    const double w = i > 0 ? x[k - 1] : c, e = i < n - 1 ? x[k + 1] : c;
    const double s = j > 0 ? x[k - P] : c, nn = j < n - 1 ? x[k + P] : c;
    return (e + w - 2.0 * c) * ax + (nn + s - 2.0 * c) * ay;
}

// r = b - L x (into V_0, unnormalised); partials: ||r||inf, ||r||_2^2, sum x
__global__ __launch_bounds__(256) void k_gm_residual(GmArgs g, const double *x, const double *b)
{
    __shared__ double sh[16];
    if (g.ctl->mg_done) return;
    double m = 0.0, s2 = 0.0, sx = 0.0;
    for (int q = 0; q < GY / 4; ++q) {
        const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
        if (i < g.n && j < g.n) {
            const size_t k = i + (size_t)g.P * j;
            const double r = b[k] - lap(x, i, j, g.n, g.P, g.ax, g.ay);
            g.V[k] = r;
            m = (fabs(r) > m || r != r) ? fabs(r) : m;
            s2 += r * r;
            sx += x[k];
        }
    }
    m = block_reduce<true>(m, sh);
    s2 = block_reduce<false>(s2, sh);
    sx = block_reduce<false>(sx, sh);
    if (threadIdx.x == 0) {
        double *p = g.part + (size_t)(blockIdx.x + gridDim.x * blockIdx.y) * 3;
        p[0] = m;
        p[1] = s2;
        p[2] = sx;
    }
}

// true-residual test; start a new cycle: beta = ||r||_2, g = beta e_1
__global__ __launch_bounds__(1024) void k_gm_fin_residual(GmArgs g)
{
    __shared__ double sh[32];
    Ctl *ctl = g.ctl;
    if (ctl->mg_done) return;
    double m = 0.0, s2 = 0.0, sx = 0.0;
    for (int b = threadIdx.x; b < g.nblk; b += blockDim.x) {
        const double *p = g.part + (size_t)b * 3;
        m = (p[0] > m || p[0] != p[0]) ? p[0] : m;
        s2 += p[1];
        sx += p[2];
    }
    m = block_reduce<true>(m, sh);
    s2 = block_reduce<false>(s2, sh);
    sx = block_reduce<false>(sx, sh);
    if (threadIdx.x == 0) {
        GmState *st = g.st;
        ctl->res_norm = m;
        ctl->phi_sum = sx;
        ctl->phi_mean = sx / ((double)g.n * (double)g.n);
        if (!ctl->mg_started) {
            ctl->b_norm = m;
            ctl->mg_started = 1;
        }
        const double thr = fmax(ctl->tol_abs, ctl->tol_rel * ctl->b_norm);
        const double beta = sqrt(s2);
        st->j = 0;
        ctl->gm_stop = 0;
        if (m <= thr || !(m == m) || beta == 0.0) {
            ctl->mg_done = 1;
            if (!(m == m)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
        } else if (ctl->mg_iters >= ctl->mg_max_iter) {
            ctl->mg_done = 1;
            ctl->flags |= AMRESSIF_FLAG_MG_CAP;
        } else {
            st->beta = beta;
            st->target = thr * beta / m;  // 2-norm target for the Givens estimate
            st->scale = 1.0 / beta;
            for (int q = 0; q <= g.m; ++q) st->gv[q] = 0.0;
            st->gv[0] = beta;
        }
    }
}

// V_j <- V_j * scale (normalise); preconditioner input r_0 = (T) V_j
template <class T>
__global__ __launch_bounds__(256) void k_gm_load(GmArgs g, T *r0)
{
    if (gm_skip(g)) return;
    const int jj = g.st->j;
    const double sc = g.st->scale;
    double *v = g.V + (size_t)jj * g.vs;
    for (int q = 0; q < GY / 4; ++q) {
        const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
        if (i < g.n && j < g.n) {
            const size_t k = i + (size_t)g.P * j;
            const double x = v[k] * sc;
            v[k] = x;
            r0[k] = (T)x;
        }
    }
}

// Z_j = (double) e (the V-cycle's correction); w = L Z_j into V_{j+1}
template <class T>
__global__ __launch_bounds__(256) void k_gm_apply(GmArgs g, const T *e)
{
    if (gm_skip(g)) return;
    const int jj = g.st->j;
    double *z = g.Z + (size_t)jj * g.vs, *w = g.V + (size_t)(jj + 1) * g.vs;
    const int n = g.n, P = g.P;
    for (int q = 0; q < GY / 4; ++q) {
        const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
        if (i < n && j < n) {
            const size_t k = i + (size_t)P * j;
            const double c = (double)e[k];
            const double wv = i > 0 ? (double)e[k - 1] : c, ev = i < n - 1 ? (double)e[k + 1] : c;
            const double sv = j > 0 ? (double)e[k - P] : c, nv = j < n - 1 ? (double)e[k + P] : c;
            z[k] = c;
            w[k] = (ev + wv - 2.0 * c) * g.ax + (nv + sv - 2.0 * c) * g.ay;
        }
    }
}

// partial dot products <V_{j+1}, V_i>, i = 0..j
__global__ __launch_bounds__(256) void k_gm_dots(GmArgs g)
{
    __shared__ double sh[16];
    if (gm_skip(g)) return;
    const int jj = g.st->j;
    const double *w = g.V + (size_t)(jj + 1) * g.vs;
    const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
    for (int iv = 0; iv <= jj; ++iv) {
        const double *v = g.V + (size_t)iv * g.vs;
        double s = 0.0;
        for (int q = 0; q < GY / 4; ++q) {
            const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
            if (i < g.n && j < g.n) {
                const size_t k = i + (size_t)g.P * j;
                s += w[k] * v[k];
            }
        }
        s = block_reduce<false>(s, sh);
        if (threadIdx.x == 0) g.part[bid * (g.m + 2) + iv] = s;
    }
}

// coefficients c_i = sum of the partials; pass 0 sets H(:, j), pass 1 adds
__global__ __launch_bounds__(1024) void k_gm_fin_dots(GmArgs g, int pass)
{
    __shared__ double sh[32];
    if (gm_skip(g)) return;
    GmState *st = g.st;
    const int jj = st->j;
    for (int iv = 0; iv <= jj; ++iv) {
        double s = 0.0;
        for (int b = threadIdx.x; b < g.nblk; b += blockDim.x) s += g.part[(size_t)b * (g.m + 2) + iv];
        s = block_reduce<false>(s, sh);
        if (threadIdx.x == 0) {
            st->c[iv] = s;
            double &h = st->H[iv * kGmMax + jj];
            h = pass ? h + s : s;
        }
        __syncthreads();
    }
}

// w -= sum_i c_i V_i; partials of ||w||^2 (slot m+1)
__global__ __launch_bounds__(256) void k_gm_orth(GmArgs g)
{
    __shared__ double sh[16];
    if (gm_skip(g)) return;
    const GmState *st = g.st;
    const int jj = st->j;
    double *w = g.V + (size_t)(jj + 1) * g.vs;
    double s2 = 0.0;
    for (int q = 0; q < GY / 4; ++q) {
        const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
        if (i < g.n && j < g.n) {
            const size_t k = i + (size_t)g.P * j;
            double x = w[k];
            for (int iv = 0; iv <= jj; ++iv) x -= st->c[iv] * g.V[(size_t)iv * g.vs + k];
            w[k] = x;
            s2 += x * x;
        }
    }
    s2 = block_reduce<false>(s2, sh);
    if (threadIdx.x == 0) g.part[(blockIdx.x + (size_t)gridDim.x * blockIdx.y) * (g.m + 2) + g.m + 1] = s2;
}

// h_{j+1,j} = ||w||, Givens rotations, residual estimate, next j
__global__ __launch_bounds__(1024) void k_gm_fin_arnoldi(GmArgs g)
{
    __shared__ double sh[32];
    if (gm_skip(g)) return;
    double s2 = 0.0;
    for (int b = threadIdx.x; b < g.nblk; b += blockDim.x) s2 += g.part[(size_t)b * (g.m + 2) + g.m + 1];
    s2 = block_reduce<false>(s2, sh);
    if (threadIdx.x != 0) return;
    GmState *st = g.st;
    Ctl *ctl = g.ctl;
    const int jj = st->j;
    const double hn = sqrt(s2);
    double *H = st->H;
    for (int iv = 0; iv < jj; ++iv) {  // earlier rotations on column jj
        const double a = H[iv * kGmMax + jj], b = H[(iv + 1) * kGmMax + jj];
        H[iv * kGmMax + jj] = st->cs[iv] * a + st->sn[iv] * b;
        H[(iv + 1) * kGmMax + jj] = -st->sn[iv] * a + st->cs[iv] * b;
    }
    const double a = H[jj * kGmMax + jj];
    const double rr = sqrt(a * a + hn * hn);
    const double cs = rr > 0.0 ? a / rr : 1.0, sn = rr > 0.0 ? hn / rr : 0.0;
    st->cs[jj] = cs;
    st->sn[jj] = sn;
    H[jj * kGmMax + jj] = rr;
    st->gv[jj + 1] = -sn * st->gv[jj];
    st->gv[jj] = cs * st->gv[jj];
    st->scale = hn > 0.0 ? 1.0 / hn : 0.0;
    st->j = jj + 1;
    st->total += 1;
    ctl->mg_iters += 1;
    const double est = fabs(st->gv[jj + 1]);
    if (!(est == est)) {
        ctl->flags |= AMRESSIF_FLAG_NONFINITE;
        ctl->mg_done = 1;
    }
    if (est <= st->target || hn == 0.0 || st->j >= g.m || ctl->mg_iters >= ctl->mg_max_iter) ctl->gm_stop = 1;
}

// y = R^{-1} g (back substitution over the cycle's j columns); single thread
__global__ void k_gm_solve(GmArgs g)
{
    if (g.ctl->mg_done) return;
    GmState *st = g.st;
    const int jc = st->j;
    for (int r = jc - 1; r >= 0; --r) {
        double s = st->gv[r];
        for (int q = r + 1; q < jc; ++q) s -= st->H[r * kGmMax + q] * st->y[q];
        st->y[r] = s / st->H[r * kGmMax + r];
    }
}

// x += Z y
__global__ __launch_bounds__(256) void k_gm_update(GmArgs g, double *x)
{
    if (g.ctl->mg_done) return;
    const GmState *st = g.st;
    const int jc = st->j;
    for (int q = 0; q < GY / 4; ++q) {
        const int i = blockIdx.x * GX + (threadIdx.x & 63), j = blockIdx.y * GY + q * 4 + (threadIdx.x >> 6);
        if (i < g.n && j < g.n) {
            const size_t k = i + (size_t)g.P * j;
            double s = x[k];
            for (int iv = 0; iv < jc; ++iv) s += st->y[iv] * g.Z[(size_t)iv * g.vs + k];
            x[k] = s;
        }
    }
}

__global__ void k_gm_reset(GmState *st, Ctl *ctl)
{
    st->j = 0;
    st->total = 0;
    ctl->gm_stop = 0;
}

// ------------------------------------------------------------ launchers
static GmArgs gm_args(const Ctx &c)
{
    GmArgs g{};
    g.n = c.g.n;
    g.P = c.g.P;
    g.ax = c.lev[0].ax;
    g.ay = c.lev[0].ay;
    g.m = c.gm_m;
    g.nblk = ((c.g.n + GX - 1) / GX) * ((c.g.n + GY - 1) / GY);
    g.V = c.gmV;
    g.Z = c.gmZ;
    g.vs = c.gm_vs;
    g.st = c.gm;
    g.part = c.gm_part;
    g.ctl = c.ctl;
    return g;
}

static dim3 gm_grid(const Ctx &c) { return dim3((c.g.n + GX - 1) / GX, (c.g.n + GY - 1) / GY); }

size_t gm_partials_doubles(const Ctx &c, int m)
{
    return (size_t)((c.g.n + GX - 1) / GX) * ((c.g.n + GY - 1) / GY) * (m + 2);
}

hipError_t launch_gm_reset(const Ctx &c)
{
    hipLaunchKernelGGL(k_gm_reset, dim3(1), dim3(1), 0, c.stream, c.gm, c.ctl);
    return hipGetLastError();
}

// true residual of x, convergence test, start of a cycle (V_0 normalised and
// handed to the preconditioner)
hipError_t launch_gm_cycle_start(const Ctx &c, const double *x, const double *b, int mixed)
{
    const GmArgs g = gm_args(c);
    hipLaunchKernelGGL(k_gm_residual, gm_grid(c), dim3(256), 0, c.stream, g, x, b);
    hipLaunchKernelGGL(k_gm_fin_residual, dim3(1), dim3(1024), 0, c.stream, g);
    if (mixed) hipLaunchKernelGGL(k_gm_load<float>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].r32);
    else hipLaunchKernelGGL(k_gm_load<double>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].r);
    return hipGetLastError();
}

// one Arnoldi step: V-cycle on V_j, Z_j, w = L Z_j, CGS2, Givens, V_{j+1}
hipError_t launch_gm_iteration(const Ctx &c, int mixed)
{
    const GmArgs g = gm_args(c);
    hipError_t e = launch_vcycle(c, mixed);
    if (e != hipSuccess) return e;
    if (mixed) hipLaunchKernelGGL(k_gm_apply<float>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].e32b);
    else hipLaunchKernelGGL(k_gm_apply<double>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].e2);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_gm_dots, gm_grid(c), dim3(256), 0, c.stream, g);
        hipLaunchKernelGGL(k_gm_fin_dots, dim3(1), dim3(1024), 0, c.stream, g, pass);
        hipLaunchKernelGGL(k_gm_orth, gm_grid(c), dim3(256), 0, c.stream, g);
    }
    hipLaunchKernelGGL(k_gm_fin_arnoldi, dim3(1), dim3(1024), 0, c.stream, g);
    if (mixed) hipLaunchKernelGGL(k_gm_load<float>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].r32);
    else hipLaunchKernelGGL(k_gm_load<double>, gm_grid(c), dim3(256), 0, c.stream, g, c.lev[0].r);
    return hipGetLastError();
}

hipError_t launch_gm_cycle_end(const Ctx &c, double *x)
{
    const GmArgs g = gm_args(c);
    hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(1), 0, c.stream, g);
    hipLaunchKernelGGL(k_gm_update, gm_grid(c), dim3(256), 0, c.stream, g, x);
    return hipGetLastError();
}

// The Arnoldi finalizers on another vector layout (the two-level composite
// solve): they read only the block partials (nblk x (m + 2)), the state and
// the control block.
static GmArgs gm_raw(double *part, int nblk, int m, GmState *st, Ctl *ctl)
{
    GmArgs g{};
    g.m = m;
    g.nblk = nblk;
    g.part = part;
    g.st = st;
    g.ctl = ctl;
    return g;
}

hipError_t launch_gm_fin_dots_raw(hipStream_t s, double *part, int nblk, int m, GmState *st, Ctl *ctl, int pass)
{
    hipLaunchKernelGGL(k_gm_fin_dots, dim3(1), dim3(1024), 0, s, gm_raw(part, nblk, m, st, ctl), pass);
    return hipGetLastError();
}

hipError_t launch_gm_fin_arnoldi_raw(hipStream_t s, double *part, int nblk, int m, GmState *st, Ctl *ctl)
{
    hipLaunchKernelGGL(k_gm_fin_arnoldi, dim3(1), dim3(1024), 0, s, gm_raw(part, nblk, m, st, ctl));
    return hipGetLastError();
}

}  // namespace amr
