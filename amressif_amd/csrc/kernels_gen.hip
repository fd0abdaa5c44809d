// Kernels of the general path (gen.hpp): the fractional step on n^dim cells,
// dim 2 or 3, with periodic directions beside the wall codes, FP64.
//
// One RK4 dual-time stage is three streaming passes:
//   k_gen_cells  : cell velocities X from the faces with every ghost layer the
//                  stencils read (cont2cart, utilities.cpp:16-350; periodic
//                  images and the wall rules, corners in the reference's pass
//                  order);
//   k_gen_flux   : fluxTotal T = -C + V per cell: the QUICK half-node fluxes
//                  of both faces per direction (fn_flux_calc.cpp:11-202) and
//                  the viscous Laplacian (:695-766), X read through L2;
//   k_gen_update : every face: R = (T(left) + T(right))/2 with the flux ghosts
//                  of main.cpp:939-1014 (wall faces R = 0, periodic images),
//                  u* <- u^k + rk (R - (1.5/dt) u* + C), the BDF2 constant C
//                  holding -G p (momentum.cpp:68-87, fn_rhs.cpp:8-68); on the
//                  last stage the Error_Computation partials (utilities.cpp:
//                  421-493) and u^k <- u*.
// Then the Poisson RHS (fn_rhs.cpp:71-125), a geometric V-cycle iteration with
// red-black Gauss-Seidel on Neumann / periodic neighbours and a dense
// pseudo-inverse bottom (poisson.cpp:41-186), and the projection
// (poisson.cpp:192-392).  Blocks are handed to the XCDs in contiguous ranges
// (xcd_block), so the rows and planes a stencil shares sit in one L2.
#include "device_common.hpp"
#include "gen.hpp"

namespace amr {
namespace gen {

namespace {

constexpr int B = kBlock;

__device__ __forceinline__ int64_t gid() { return (int64_t)xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x; }

// (i, j, k) of the linear index q over an nx x ny x .. box (x fastest) in
// 32-bit unsigned arithmetic: every box here has fewer than 2^31 points
// (n <= 1024 in 3D, n <= 16384 in 2D), and two 32-bit divisions cost a
// fraction of the 64-bit div/mod sequence.
__device__ __forceinline__ void split3(int64_t q, int nx, int ny, int *x)
{
    const unsigned u = (unsigned)q, a = u / (unsigned)nx, b = a / (unsigned)ny;
    x[0] = (int)(u - a * (unsigned)nx);
    x[1] = (int)(a - b * (unsigned)ny);
    x[2] = (int)b;
}

__device__ __forceinline__ int64_t fidx(const Geom &g, int d, int i, int j, int k)
{
    const int64_t nx = g.n + (d == 0), ny = g.n + (d == 1);
    return i + nx * (j + ny * (int64_t)k);
}
__device__ __forceinline__ int64_t cidx(const Geom &g, int i, int j, int k)
{
    return i + (int64_t)g.n * (j + (int64_t)g.n * k);
}
__device__ __forceinline__ int64_t cidx(const Geom &g, const int *x) { return cidx(g, x[0], x[1], x[2]); }
__device__ __forceinline__ int64_t xidx(const Geom &g, int i, int j, int k)
{
    const int64_t s = g.n + 4;
    const int gk = g.dim == 3 ? 2 : 0;
    return (i + 2) + s * ((j + 2) + s * (int64_t)(k + gk));
}

// apply a wall rule to the mirror values v (one ghost layer crossing of dim d)
__device__ __forceinline__ void wall_rule(const Geom &g, int kind, int d, double *v)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= g.dim) break;
        if (kind == 0) v[c] = -v[c];
        else if (kind == 1) v[c] = c == d ? -v[c] : v[c];
        else if (kind == 2) v[c] = 2.0 - v[c];
        else v[c] = 2.0 * g.inflow[c] - v[c];
    }
}

// ------------------------------------------------------------ cell velocities
__global__ __launch_bounds__(B) void k_gen_cells(Geom g, const double *f0, const double *f1, const double *f2,
                                                 double *X, const Ctl *guard)
{
    if (guard && guard->mom_done) return;
    const int64_t s = g.n + 4, sk = g.dim == 3 ? s : 1;
    const int64_t q = gid();
    if (q >= s * s * sk) return;
    const int n = g.n;
    int x[3];
    split3(q, (int)s, (int)s, x);
    x[0] -= 2;
    x[1] -= 2;
    x[2] = g.dim == 3 ? x[2] - 2 : 0;
    int b[3] = {x[0], x[1], x[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        if (x[d] < 0 || x[d] >= n) b[d] = g.per[d] ? (x[d] + n) % n : (x[d] < 0 ? -x[d] - 1 : 2 * n - 1 - x[d]);
    }
    const double *f[3] = {f0, f1, f2};
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= g.dim) break;
        int e[3] = {b[0], b[1], b[2]};
        e[c] += 1;
        v[c] = 0.5 * (f[c][fidx(g, c, b[0], b[1], b[2])] + f[c][fidx(g, c, e[0], e[1], e[2])]);
    }
    // the wall passes in the reference's order W, E, S, N, B, F: the rule of
    // the lowest direction applies first (corners are the later passes'
    // rules of the earlier passes' ghosts)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        if ((x[d] < 0 || x[d] >= n) && !g.per[d]) wall_rule(g, x[d] < 0 ? g.kind_lo[d] : g.kind_hi[d], d, v);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (c < g.dim) X[q + (int64_t)c * g.xs] = v[c];
}

// ------------------------------------------------------------ fluxTotal
__device__ __forceinline__ double quick(double vel, double a2, double a1, double b0, double b1)
{
    return quick_flux<double>(vel, a2, a1, b0, b1);
}

__global__ __launch_bounds__(B) void k_gen_flux(Geom g, const double *f0, const double *f1, const double *f2,
                                                const double *X, double *T, const Ctl *guard)
{
    if (guard && guard->mom_done) return;
    const int64_t q = gid();
    if (q >= g.ncell) return;
    const int n = g.n;
    int x[3];
    split3(q, n, n, x);
    const double *f[3] = {f0, f1, f2};
    const int64_t xc = xidx(g, x[0], x[1], x[2]);
    int64_t st[3];
    st[0] = 1;
    st[1] = n + 4;
    st[2] = (int64_t)(n + 4) * (n + 4);
    // face velocities of both faces per direction
    double fw[3], fe[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        int e[3] = {x[0], x[1], x[2]};
        e[d] += 1;
        fw[d] = f[d][fidx(g, d, x[0], x[1], x[2])];
        fe[d] = f[d][fidx(g, d, e[0], e[1], e[2])];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= g.dim) break;
        const double *Xc = X + (int64_t)c * g.xs + xc;
        double conv = 0.0, lap[3] = {0.0, 0.0, 0.0};
        const double C = Xc[0];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= g.dim) break;
            const int64_t o = st[d];
            const double m2 = Xc[-2 * o], m1 = Xc[-o], p1 = Xc[o], p2 = Xc[2 * o];
            const double Fw = quick(fw[d], m2, m1, C, p1), Fe = quick(fe[d], m1, C, p1, p2);
            const double t = (Fe - Fw) * g.idx[d];
            conv = d == 0 ? t : conv + t;
            lap[d] = (p1 - 2.0 * C + m1) * g.idx2[d];
        }
        // ((E - 2C + W)/dx^2 + (N - 2C + S)/dy^2)/ren [+ ((B - 2C + F)/dz^2)/ren]
        double visc = (lap[0] + lap[1]) * g.iren;
        if (g.dim == 3) visc = visc + lap[2] * g.iren;
        T[q + (int64_t)c * g.ncell] = (0.0 - conv) + visc;
    }
}

// The same fluxTotal for a 3D box, marched along z: a workgroup owns a 64 x 4
// column tile and zc planes, and each thread keeps its column's X(z - 2 ..
// z + 2) of every component in registers (the next plane loaded one step
// ahead), so the z-stencil reads X from memory once per cell instead of five
// times, and the x and y neighbours of the current plane come from the
// tile's own and its halo's loads of that plane (L1 and L2).  k_gen_flux's
// one-thread-per-cell order reads the five planes of a cell from five
// distant parts of X, which at 256^3 (540 KB per plane and component) do not
// stay in an XCD's L2.  Per cell the arithmetic is k_gen_flux's, term by term.
constexpr int kFzTX = 64, kFzTY = 4;
__global__ __launch_bounds__(256) void k_gen_flux_z(Geom g, const double *f0, const double *f1, const double *f2,
                                                    const double *X, double *T, const Ctl *guard, int ntx, int nty,
                                                    int zc)
{
    if (guard && guard->mom_done) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tx = blk % ntx, ty = (blk / ntx) % nty, tz = blk / (ntx * nty);
    const int n = g.n;
    const int x = tx * kFzTX + (int)(threadIdx.x & 63), y = ty * kFzTY + (int)(threadIdx.x >> 6);
    const int z0 = tz * zc, z1 = min(z0 + zc, n);
    if (x >= n || y >= n || z0 >= n) return;  // no barrier below: a thread past the edge just ends
    const int64_t s = n + 4, sz = s * s, xs = g.xs;
    int64_t xc = xidx(g, x, y, z0);
    double w[3][5];  // X(x, y, z - 2 .. z + 2) per component
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 5; ++k) w[c][k] = X[c * xs + xc + (k - 2) * sz];
    double fzl = f2[fidx(g, 2, x, y, z0)], fzh = f2[fidx(g, 2, x, y, z0 + 1)];
    for (int z = z0; z < z1; ++z) {
        double nw[3] = {0.0, 0.0, 0.0}, nfz = 0.0;
        if (z + 1 < z1) {  // the next step's new plane z + 3 (a ghost plane at most) and z-face z + 2
#pragma unroll
            for (int c = 0; c < 3; ++c) nw[c] = X[c * xs + xc + 3 * sz];
            nfz = f2[fidx(g, 2, x, y, z + 2)];
        }
        const double fw0 = f0[fidx(g, 0, x, y, z)], fe0 = f0[fidx(g, 0, x + 1, y, z)];
        const double fw1 = f1[fidx(g, 1, x, y, z)], fe1 = f1[fidx(g, 1, x, y + 1, z)];
        const int64_t q = cidx(g, x, y, z);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double *Xc = X + c * xs + xc;
            const double C = w[c][2];
            double m2 = Xc[-2], m1 = Xc[-1], p1 = Xc[1], p2 = Xc[2];
            double Fw = quick(fw0, m2, m1, C, p1), Fe = quick(fe0, m1, C, p1, p2);
            double conv = (Fe - Fw) * g.idx[0];
            const double l0 = (p1 - 2.0 * C + m1) * g.idx2[0];
            m2 = Xc[-2 * s]; m1 = Xc[-s]; p1 = Xc[s]; p2 = Xc[2 * s];
            Fw = quick(fw1, m2, m1, C, p1);
            Fe = quick(fe1, m1, C, p1, p2);
            conv = conv + (Fe - Fw) * g.idx[1];
            const double l1 = (p1 - 2.0 * C + m1) * g.idx2[1];
            m2 = w[c][0]; m1 = w[c][1]; p1 = w[c][3]; p2 = w[c][4];
            Fw = quick(fzl, m2, m1, C, p1);
            Fe = quick(fzh, m1, C, p1, p2);
            conv = conv + (Fe - Fw) * g.idx[2];
            const double l2 = (p1 - 2.0 * C + m1) * g.idx2[2];
            double visc = (l0 + l1) * g.iren;
            visc = visc + l2 * g.iren;
            T[q + (int64_t)c * g.ncell] = (0.0 - conv) + visc;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            w[c][0] = w[c][1]; w[c][1] = w[c][2]; w[c][2] = w[c][3]; w[c][3] = w[c][4]; w[c][4] = nw[c];
        }
        fzl = fzh;
        fzh = nfz;
        xc += sz;
    }
}

// block range of face direction d in a grid over every face (nb[d] blocks each)
__device__ __forceinline__ int face_dir(const int *nb, int blk, int &lb)
{
    if (blk < nb[0]) { lb = blk; return 0; }
    if (blk < nb[0] + nb[1]) { lb = blk - nb[0]; return 1; }
    lb = blk - nb[0] - nb[1];
    return 2;
}
__device__ __forceinline__ void face_coords(const Geom &g, int d, int64_t q, int *x)
{
    split3(q, g.n + (d == 0), g.n + (d == 1), x);
}

struct FaceArrays {
    double *s[3], *b[3];
    const double *c[3];
    int nb[3];
};

// ------------------------------------------------------------ stage update
__global__ __launch_bounds__(B) void k_gen_update(Geom g, FaceArrays a, const double *T, double rk, int last,
                                                  double *partials, const Ctl *guard)
{
    __shared__ double sh[B / 64];
    if (guard && guard->mom_done) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    int lb;
    const int d = face_dir(a.nb, blk, lb);
    const int64_t q = (int64_t)lb * B + threadIdx.x;
    const int n = g.n;
    double sq = 0.0;
    if (q < g.nf[d]) {
        int x[3];
        face_coords(g, d, q, x);
        const int xd = x[d];
        const bool wlo = xd == 0 && !g.per[d], whi = xd == n && !g.per[d];
        double R;
        if (g.pseudo && (wlo || whi)) {
            R = 0.0;  // the wall ghost -T(mirror) cancels T (main.cpp:953-1012)
        } else {
            int l[3] = {x[0], x[1], x[2]}, r[3] = {x[0], x[1], x[2]};
            l[d] = xd == 0 ? n - 1 : xd - 1;
            r[d] = xd == n ? 0 : xd;
            const double *Td = T + (int64_t)d * g.ncell;
            // explicit form: fluxTotal's wall ghosts keep setVal's 0
            const double tl = wlo ? 0.0 : Td[cidx(g, l)], tr = whi ? 0.0 : Td[cidx(g, r)];
            R = 0.5 * (tl + tr);
        }
        const double us = a.s[d][q], ub = a.b[d][q];
        const double unew = ub + rk * (R - g.c15 * us + a.c[d][q]);
        a.s[d][q] = unew;
        if (last) {
            const double w = (xd == 0 || xd == n) ? 0.5 : 1.0;
            const double df = ub - unew;
            sq = w * df * df;
            a.b[d][q] = unew;
        }
    }
    if (last) {
        const double r = block_reduce<false>(sq, sh);
        if (threadIdx.x == 0) partials[blk] = r;
    }
}

// Error_Computation from the per-direction block partials; the loop control
// of main.cpp:898-1039 (countIter, IterNum cap, divergence guard), or for
// PSEUDO_TIMESTEPPING = 0 the single update's norm with countIter 0
__global__ __launch_bounds__(B) void k_gen_fin_pseudo(Geom g, const double *partials, int nb0, int nb1, int nb2,
                                                      Ctl *ctl)
{
    __shared__ double sh[B / 64];
    if (ctl->mom_done) return;
    const int nb[3] = {nb0, nb1, nb2};
    double e[3] = {0.0, 0.0, 0.0};
    int off = 0;
    const double npts = (double)g.ncell;
    for (int d = 0; d < g.dim; ++d) {
        double s = 0.0;
        for (int k = threadIdx.x; k < nb[d]; k += B) s += partials[off + k];
        s = block_reduce<false>(s, sh);
        if (threadIdx.x == 0) e[d] = sqrt(s) / sqrt(npts);
        off += nb[d];
    }
    if (threadIdx.x != 0) return;
    double ne = (e[0] < e[1]) ? e[1] : e[0];
    if (g.dim == 3) ne = (ne < e[2]) ? e[2] : ne;
    ctl->norm_error = ne;
    if (!(ne == ne) || isinf(ne)) ctl->flags |= AMRESSIF_FLAG_NONFINITE;
    if (!g.pseudo) {
        ctl->mom_done = 1;
        return;
    }
    ctl->mom_count += 1;
    if (ctl->mom_count > ctl->iter_num) { ctl->flags |= AMRESSIF_FLAG_CAP; ctl->mom_done = 1; }
    else if (ne > 1.e2) { ctl->flags |= AMRESSIF_FLAG_DIVERGED; ctl->mom_done = 1; }
    else if (!(ne > ctl->tol_mom)) ctl->mom_done = 1;
}

// ------------------------------------------------------------ BDF2 constant
// main.cpp:879-884 with the pressure term folded: C = (1.5/dt) u^n + (0.5/dt)
// (u^n - u^{n-1}) - G p (p's ghosts: Neumann mirror, so G p = 0 on wall
// faces; periodic images); then u^{n-1} <- u^n and u^k = u* = u^n.
struct Bdf2Arrays {
    double *u[3], *up[3], *c[3], *b[3], *s[3];
    const double *p;
    int nb[3];
};
__global__ __launch_bounds__(B) void k_gen_bdf2(Geom g, Bdf2Arrays a)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    int lb;
    const int d = face_dir(a.nb, blk, lb);
    const int64_t q = (int64_t)lb * B + threadIdx.x;
    if (q >= g.nf[d]) return;
    const int n = g.n;
    int x[3];
    face_coords(g, d, q, x);
    const int xd = x[d];
    double gp = 0.0;
    if (g.per[d] || (xd > 0 && xd < n)) {
        int l[3] = {x[0], x[1], x[2]}, r[3] = {x[0], x[1], x[2]};
        l[d] = xd == 0 ? n - 1 : xd - 1;
        r[d] = xd == n ? 0 : xd;
        gp = (a.p[cidx(g, r)] - a.p[cidx(g, l)]) * g.idx[d];
    }
    const double un = a.u[d][q], um = a.up[d][q];
    a.c[d][q] = g.c15 * un + (0.5 / g.dt) * (un - um) - gp;
    a.up[d][q] = un;
    a.b[d][q] = un;
    a.s[d][q] = un;
}

// ------------------------------------------------------------ divergence
__global__ __launch_bounds__(B) void k_gen_div(Geom g, const double *f0, const double *f1, const double *f2,
                                               double *b, double *partials)
{
    __shared__ double sh[B / 64];
    const int64_t q = gid();
    double m = 0.0, sum = 0.0;
    if (q < g.ncell) {
        const int n = g.n;
        int x[3];
        split3(q, n, n, x);
        const double *f[3] = {f0, f1, f2};
        double D = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= g.dim) break;
            int e[3] = {x[0], x[1], x[2]};
            e[d] += 1;
            const double t = (f[d][fidx(g, d, e[0], e[1], e[2])] - f[d][fidx(g, d, x[0], x[1], x[2])]) * g.idx[d];
            D = d == 0 ? t : D + t;
        }
        m = fabs(D);
        if (D != D) m = D;
        if (b) {
            const double bv = g.c15 * D;
            b[q] = bv;
            sum = bv;
        }
    }
    const double rm = block_reduce<true>(m, sh);
    const double rs = block_reduce<false>(sum, sh);
    if (threadIdx.x == 0) {
        partials[2 * (int64_t)blockIdx.x] = rm;
        partials[2 * (int64_t)blockIdx.x + 1] = rs;
    }
}

__global__ __launch_bounds__(B) void k_gen_fin_div(const double *partials, int nb, Ctl *ctl, double ncell, int which)
{
    __shared__ double sh[B / 64];
    double m = 0.0, s = 0.0;
    for (int k = threadIdx.x; k < nb; k += B) {
        const double v = partials[2 * k];
        m = (v > m || v != v) ? v : m;
        s += partials[2 * k + 1];
    }
    m = block_reduce<true>(m, sh);
    s = block_reduce<false>(s, sh);
    if (threadIdx.x != 0) return;
    if (which == 0) {
        ctl->div_star = m;
        ctl->rhs_sum = s;
        ctl->b_mean = s / ncell;
    } else {
        ctl->div_new = m;
    }
}

// ------------------------------------------------------------ multigrid
// Laplacian of e at cell x of a level (n cells per side, a = 1/h^2): Neumann
// wall neighbours are the cell itself (their term drops), periodic ones wrap
__device__ __forceinline__ double lap(const Geom &g, const Level &L, const double *e, const int *x, int64_t q)
{
    const int n = L.n;
    const double ec = e[q];
    double s = 0.0;
    int64_t st = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        const int64_t lo = x[d] > 0 ? q - st : (g.per[d] ? q + st * (n - 1) : q);
        const int64_t hi = x[d] < n - 1 ? q + st : (g.per[d] ? q - st * (n - 1) : q);
        s += (e[hi] - 2.0 * ec + e[lo]) * L.a[d];
        st *= n;
    }
    return s;
}

__device__ __forceinline__ void lcoords(const Geom &g, int n, int64_t q, int *x)
{
    split3(q, n, n, x);
    if (g.dim != 3) x[2] = 0;
}

__global__ void k_gen_mg_start(Ctl *ctl)
{
    ctl->mg_done = 0;
    ctl->mg_iters = 0;
    ctl->mg_started = 0;
}

// r0 = (b - mean b) - L phi and its max-norm partials
__global__ __launch_bounds__(B) void k_gen_resid0(Geom g, Level L, const double *phi, const double *b,
                                                  double *partials, const Ctl *ctl)
{
    __shared__ double sh[B / 64];
    if (ctl->mg_done) return;
    const int64_t q = gid();
    double m = 0.0;
    if (q < L.ncell) {
        int x[3];
        lcoords(g, L.n, q, x);
        const double r = (b[q] - ctl->b_mean) - lap(g, L, phi, x, q);
        L.r[q] = r;
        m = fabs(r);
        if (r != r) m = r;
    }
    m = block_reduce<true>(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// the stopping test ||r||inf <= max(tol_abs, tol_rel ||b - mean||inf)
// (poisson.cpp:181-182); the first residual (phi = 0) is b - mean itself
__global__ __launch_bounds__(B) void k_gen_fin_resid(const double *partials, int nb, Ctl *ctl)
{
    __shared__ double sh[B / 64];
    if (ctl->mg_done) return;
    double m = 0.0;
    for (int k = threadIdx.x; k < nb; k += B) {
        const double v = partials[k];
        m = (v > m || v != v) ? v : m;
    }
    m = block_reduce<true>(m, sh);
    if (threadIdx.x != 0) return;
    if (!ctl->mg_started) {
        ctl->b_norm = m;
        ctl->mg_started = 1;
    }
    ctl->res_norm = m;
    const double thr = fmax(ctl->tol_abs, ctl->tol_rel * ctl->b_norm);
    if (m <= thr || !(m == m)) ctl->mg_done = 1;
    else if (ctl->mg_iters >= ctl->mg_max_iter) {
        ctl->flags |= AMRESSIF_FLAG_MG_CAP;
        ctl->mg_done = 1;
    } else ctl->mg_iters += 1;
}

// red-black Gauss-Seidel sweep of L e = r; first: e = 0 before the sweep (the
// first red sweep reads no neighbour)
__global__ __launch_bounds__(B) void k_gen_gsrb(Geom g, Level L, int color, int first, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int64_t q = gid();
    if (q >= L.ncell) return;
    int x[3];
    lcoords(g, L.n, q, x);
    if (((x[0] + x[1] + x[2]) & 1) != color) return;
    const int n = L.n;
    double off = 0.0, diag = 0.0;
    int64_t st = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        const bool haslo = x[d] > 0 || g.per[d], hashi = x[d] < n - 1 || g.per[d];
        if (haslo) {
            if (!first) off += L.a[d] * L.e[x[d] > 0 ? q - st : q + st * (n - 1)];
            diag += L.a[d];
        }
        if (hashi) {
            if (!first) off += L.a[d] * L.e[x[d] < n - 1 ? q + st : q - st * (n - 1)];
            diag += L.a[d];
        }
        st *= n;
    }
    L.e[q] = (off - L.r[q]) / diag;
}

// r_c = 2^-dim sum over the children of (r - L e)
__global__ __launch_bounds__(B) void k_gen_restrict(Geom g, Level F, Level Cl, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int64_t q = gid();
    if (q >= Cl.ncell) return;
    int c[3];
    lcoords(g, Cl.n, q, c);
    double s = 0.0;
    const int nz = g.dim == 3 ? 2 : 1;
    for (int dz = 0; dz < nz; ++dz)
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                int f[3] = {2 * c[0] + dx, 2 * c[1] + dy, g.dim == 3 ? 2 * c[2] + dz : 0};
                const int64_t k = f[0] + (int64_t)F.n * (f[1] + (int64_t)F.n * f[2]);
                s += F.r[k] - lap(g, F, F.e, f, k);
            }
    Cl.r[q] = (g.dim == 3 ? 0.125 : 0.25) * s;
}

// cell-centred (bi/tri)linear prolongation: per direction 3/4 of the coarse
// cell and 1/4 of its neighbour on the fine cell's side (the periodic image,
// or the cell itself past a wall: the Neumann mirror)
__global__ __launch_bounds__(B) void k_gen_prolong(Geom g, Level F, Level Cl, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int64_t q = gid();
    if (q >= F.ncell) return;
    int f[3];
    lcoords(g, F.n, q, f);
    int c[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
    const int m = Cl.n;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.dim) break;
        c[d] = f[d] >> 1;
        int t = (f[d] & 1) ? c[d] + 1 : c[d] - 1;
        if (t < 0 || t >= m) t = g.per[d] ? (t + m) % m : c[d];
        nb[d] = t;
    }
    double s = 0.0;
    const int corners = 1 << g.dim;
    for (int k = 0; k < corners; ++k) {
        const int x = (k & 1) ? nb[0] : c[0], y = (k & 2) ? nb[1] : c[1], z = (k & 4) ? nb[2] : c[2];
        const double w = ((k & 1) ? 0.25 : 0.75) * ((k & 2) ? 0.25 : 0.75) * (g.dim == 3 ? ((k & 4) ? 0.25 : 0.75) : 1.0);
        s += w * Cl.e[x + (int64_t)m * (y + (int64_t)m * z)];
    }
    F.e[q] += s;
}

// e = pinv r on the coarsest level (pinv symmetric, read by columns)
__global__ __launch_bounds__(B) void k_gen_bottom(Level L, const double *pinv, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int64_t i = (int64_t)blockIdx.x * B + threadIdx.x;
    const int64_t m = L.ncell;
    if (i >= m) return;
    double s = 0.0;
    for (int64_t j = 0; j < m; ++j) s += pinv[j * m + i] * L.r[j];
    L.e[i] = s;
}

__global__ __launch_bounds__(B) void k_gen_correct(Level L, double *phi, const Ctl *ctl)
{
    if (ctl->mg_done) return;
    const int64_t q = gid();
    if (q < L.ncell) phi[q] += L.e[q];
}

// ------------------------------------------------------------ projection
__global__ __launch_bounds__(B) void k_gen_sum(int64_t m, const double *x, double *partials)
{
    __shared__ double sh[B / 64];
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    const double v = q < m ? x[q] : 0.0;
    const double s = block_reduce<false>(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(B) void k_gen_fin_mean(const double *partials, int nb, double m, Ctl *ctl)
{
    __shared__ double sh[B / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += B) s += partials[k];
    s = block_reduce<false>(s, sh);
    if (threadIdx.x == 0) ctl->phi_mean = s / m;
}

// update_solution: u_d = u*_d - (G_d phi) dt/1.5 on the faces (G phi = 0 on
// wall faces: phi's ghosts mirror), then p += phi - mean(phi) on the cells
struct ProjArrays {
    double *u[3];
    const double *s[3];
    double *p;
    const double *phi;
    int nb[4];  // face blocks per direction, then cell blocks
};
__global__ __launch_bounds__(B) void k_gen_project(Geom g, ProjArrays a, const Ctl *ctl)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int n = g.n;
    if (blk >= a.nb[0] + a.nb[1] + a.nb[2]) {
        const int64_t q = (int64_t)(blk - a.nb[0] - a.nb[1] - a.nb[2]) * B + threadIdx.x;
        if (q < g.ncell) a.p[q] = a.p[q] + (a.phi[q] - ctl->phi_mean);
        return;
    }
    int lb;
    const int d = face_dir(a.nb, blk, lb);
    const int64_t q = (int64_t)lb * B + threadIdx.x;
    if (q >= g.nf[d]) return;
    int x[3];
    face_coords(g, d, q, x);
    const int xd = x[d];
    double gphi = 0.0;
    if (g.per[d] || (xd > 0 && xd < n)) {
        int l[3] = {x[0], x[1], x[2]}, r[3] = {x[0], x[1], x[2]};
        l[d] = xd == 0 ? n - 1 : xd - 1;
        r[d] = xd == n ? 0 : xd;
        gphi = (a.phi[cidx(g, r)] - a.phi[cidx(g, l)]) * g.idx[d];
    }
    a.u[d][q] = a.s[d][q] - (gphi * g.dt / 1.5);
}

__global__ void k_gen_fin_step(Ctl *ctl, amressif_step_stats *log)
{
    amressif_step_stats st;
    st.pseudo_iters = ctl->mom_count;
    st.vcycles = ctl->mg_iters;
    st.flags = ctl->flags;
    st.pad = 0;
    st.norm_error = ctl->norm_error;
    st.div_star = ctl->div_star;
    st.rhs_sum = ctl->rhs_sum;
    st.poisson_res = ctl->b_norm > 0 ? ctl->res_norm / ctl->b_norm : ctl->res_norm;
    st.div_new = ctl->div_new;
    *log = st;
}

__global__ __launch_bounds__(B) void k_gen_fill(double *dst, int64_t count, double value)
{
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    if (q < count) dst[q] = value;
}

// ------------------------------------------------------------ Array4 views
struct V4 {
    double *ptr;
    int lo[3], hi[3];
    int64_t js, ks, ns;
};
V4 view(const amr_array4 &a)
{
    V4 v;
    v.ptr = (double *)a.ptr;
    for (int d = 0; d < 3; ++d) {
        v.lo[d] = a.lo[d];
        v.hi[d] = a.hi[d];
    }
    v.js = a.hi[0] - a.lo[0] + 1;
    v.ks = v.js * (a.hi[1] - a.lo[1] + 1);
    v.ns = v.ks * (a.hi[2] - a.lo[2] + 1);
    return v;
}
__device__ __forceinline__ int64_t vidx(const V4 &v, int i, int j, int k, int c)
{
    return (i - v.lo[0]) + v.js * (j - v.lo[1]) + v.ks * (k - v.lo[2]) + v.ns * c;
}

__global__ __launch_bounds__(B) void k_gen_pack(V4 v, int comp, double *dst, int ex, int ey, int ez, int unpack)
{
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    if (q >= (int64_t)ex * ey * ez) return;
    const int i = (int)(q % ex), j = (int)((q / ex) % ey), k = (int)(q / ((int64_t)ex * ey));
    const int64_t e = vidx(v, i, j, k, comp);
    if (unpack) v.ptr[e] = dst[q];
    else dst[q] = v.ptr[e];
}

// every cell of the view: ghosts by the Neumann mirror or the periodic image
__global__ __launch_bounds__(B) void k_gen_unpack_ghosted(Geom g, const double *src, V4 v, int comp,
                                                          const double *shift)
{
    const int64_t wx = v.hi[0] - v.lo[0] + 1, wy = v.hi[1] - v.lo[1] + 1, wz = v.hi[2] - v.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    if (q >= wx * wy * wz) return;
    int x[3] = {v.lo[0] + (int)(q % wx), v.lo[1] + (int)((q / wx) % wy), v.lo[2] + (int)(q / (wx * wy))};
    int m[3] = {x[0], x[1], x[2]};
    const int n = g.n;
    for (int d = 0; d < g.dim; ++d)
        if (x[d] < 0 || x[d] >= n) {
            if (g.per[d]) m[d] = ((x[d] % n) + n) % n;
            else m[d] = x[d] < 0 ? -x[d] - 1 : 2 * n - 1 - x[d];
        }
    const double s = shift ? *shift : 0.0;
    v.ptr[vidx(v, x[0], x[1], x[2], comp)] = src[cidx(g, m)] - s;
}

__global__ __launch_bounds__(B) void k_gen_unpack_velcart(Geom g, const double *X, V4 v)
{
    const int64_t wx = v.hi[0] - v.lo[0] + 1, wy = v.hi[1] - v.lo[1] + 1, wz = v.hi[2] - v.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    if (q >= wx * wy * wz) return;
    const int x[3] = {v.lo[0] + (int)(q % wx), v.lo[1] + (int)((q / wx) % wy), v.lo[2] + (int)(q / (wx * wy))};
    for (int c = 0; c < g.dim; ++c) v.ptr[vidx(v, x[0], x[1], x[2], c)] = X[xidx(g, x[0], x[1], x[2]) + (int64_t)c * g.xs];
}

inline int nblk(int64_t m) { return (int)((m + B - 1) / B); }

}  // namespace

// ============================================================ launchers
hipError_t cells(const Ctx &c, hipStream_t s, double *const *f, double *X, const Ctl *guard)
{
    hipLaunchKernelGGL(k_gen_cells, dim3(nblk(c.g.xs)), dim3(B), 0, s, c.g, f[0], f[1], c.g.dim == 3 ? f[2] : f[0],
                       X, guard);
    return hipGetLastError();
}

hipError_t flux(const Ctx &c, hipStream_t s, double *const *f, const double *X, double *T, const Ctl *guard)
{
    if (c.g.dim == 3 && c.flux_z) {
        const int n = c.g.n, ntx = (n + kFzTX - 1) / kFzTX, nty = (n + kFzTY - 1) / kFzTY;
        const int zc = c.flux_zc, nz = (n + zc - 1) / zc;
        hipLaunchKernelGGL(k_gen_flux_z, dim3(ntx * nty * nz), dim3(256), 0, s, c.g, f[0], f[1], f[2], X, T, guard,
                           ntx, nty, zc);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_gen_flux, dim3(nblk(c.g.ncell)), dim3(B), 0, s, c.g, f[0], f[1],
                       c.g.dim == 3 ? f[2] : f[0], X, T, guard);
    return hipGetLastError();
}

hipError_t update(const Ctx &c, hipStream_t s, double rk, int last, const Ctl *guard)
{
    FaceArrays a{};
    for (int d = 0; d < 3; ++d) {
        a.s[d] = d < c.g.dim ? c.S[d] : nullptr;
        a.b[d] = d < c.g.dim ? c.B[d] : nullptr;
        a.c[d] = d < c.g.dim ? c.C[d] : nullptr;
        a.nb[d] = d < c.g.dim ? c.nb_face[d] : 0;
    }
    const int nb = a.nb[0] + a.nb[1] + a.nb[2];
    hipLaunchKernelGGL(k_gen_update, dim3(nb), dim3(B), 0, s, c.g, a, c.T, rk, last, c.partials, guard);
    return hipGetLastError();
}

hipError_t fin_pseudo(const Ctx &c, hipStream_t s, Ctl *ctl)
{
    hipLaunchKernelGGL(k_gen_fin_pseudo, dim3(1), dim3(B), 0, s, c.g, c.partials, c.nb_face[0], c.nb_face[1],
                       c.g.dim == 3 ? c.nb_face[2] : 0, ctl);
    return hipGetLastError();
}

hipError_t bdf2(const Ctx &c, hipStream_t s)
{
    Bdf2Arrays a{};
    for (int d = 0; d < 3; ++d) {
        const bool on = d < c.g.dim;
        a.u[d] = on ? c.U[d] : nullptr;
        a.up[d] = on ? c.Up[d] : nullptr;
        a.c[d] = on ? c.C[d] : nullptr;
        a.b[d] = on ? c.B[d] : nullptr;
        a.s[d] = on ? c.S[d] : nullptr;
        a.nb[d] = on ? c.nb_face[d] : 0;
    }
    a.p = c.p;
    hipLaunchKernelGGL(k_gen_bdf2, dim3(a.nb[0] + a.nb[1] + a.nb[2]), dim3(B), 0, s, c.g, a);
    return hipGetLastError();
}

hipError_t divergence(const Ctx &c, hipStream_t s, double *const *f, double *b, int *nblocks)
{
    *nblocks = nblk(c.g.ncell);
    hipLaunchKernelGGL(k_gen_div, dim3(*nblocks), dim3(B), 0, s, c.g, f[0], f[1], c.g.dim == 3 ? f[2] : f[0], b,
                       c.partials);
    return hipGetLastError();
}

hipError_t fin_div(const Ctx &c, hipStream_t s, int nblocks, Ctl *ctl, int which)
{
    hipLaunchKernelGGL(k_gen_fin_div, dim3(1), dim3(B), 0, s, c.partials, nblocks, ctl, (double)c.g.ncell, which);
    return hipGetLastError();
}

hipError_t mg_start(const Ctx &c, hipStream_t s, Ctl *ctl)
{
    hipLaunchKernelGGL(k_gen_mg_start, dim3(1), dim3(1), 0, s, ctl);
    return hipGetLastError();
}

hipError_t resid0(const Ctx &c, hipStream_t s, const Ctl *ctl, int *nblocks)
{
    *nblocks = nblk(c.lev[0].ncell);
    hipLaunchKernelGGL(k_gen_resid0, dim3(*nblocks), dim3(B), 0, s, c.g, c.lev[0], c.phi, c.b, c.partials, ctl);
    return hipGetLastError();
}

hipError_t fin_resid(const Ctx &c, hipStream_t s, int nblocks, Ctl *ctl)
{
    hipLaunchKernelGGL(k_gen_fin_resid, dim3(1), dim3(B), 0, s, c.partials, nblocks, ctl);
    return hipGetLastError();
}

hipError_t vcycle(const Ctx &c, hipStream_t s, const Ctl *ctl)
{
    const int L = c.nlev;
    for (int l = 0; l + 1 < L; ++l) {
        const Level &F = c.lev[l];
        const dim3 gr(nblk(F.ncell));
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 0, 1, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 1, 0, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 0, 0, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 1, 0, ctl);
        hipLaunchKernelGGL(k_gen_restrict, dim3(nblk(c.lev[l + 1].ncell)), dim3(B), 0, s, c.g, F, c.lev[l + 1], ctl);
    }
    hipLaunchKernelGGL(k_gen_bottom, dim3(nblk(c.lev[L - 1].ncell)), dim3(B), 0, s, c.lev[L - 1], c.pinv, ctl);
    for (int l = L - 2; l >= 0; --l) {
        const Level &F = c.lev[l];
        const dim3 gr(nblk(F.ncell));
        hipLaunchKernelGGL(k_gen_prolong, gr, dim3(B), 0, s, c.g, F, c.lev[l + 1], ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 0, 0, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 1, 0, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 0, 0, ctl);
        hipLaunchKernelGGL(k_gen_gsrb, gr, dim3(B), 0, s, c.g, F, 1, 0, ctl);
    }
    hipLaunchKernelGGL(k_gen_correct, dim3(nblk(c.lev[0].ncell)), dim3(B), 0, s, c.lev[0], c.phi, ctl);
    return hipGetLastError();
}

hipError_t phi_mean(const Ctx &c, hipStream_t s, Ctl *ctl)
{
    const int nb = nblk(c.g.ncell);
    hipLaunchKernelGGL(k_gen_sum, dim3(nb), dim3(B), 0, s, c.g.ncell, c.phi, c.partials);
    hipLaunchKernelGGL(k_gen_fin_mean, dim3(1), dim3(B), 0, s, c.partials, nb, (double)c.g.ncell, ctl);
    return hipGetLastError();
}

hipError_t project(const Ctx &c, hipStream_t s, Ctl *ctl)
{
    if (hipError_t e = phi_mean(c, s, ctl)) return e;
    ProjArrays a{};
    for (int d = 0; d < 3; ++d) {
        const bool on = d < c.g.dim;
        a.u[d] = on ? c.U[d] : nullptr;
        a.s[d] = on ? c.S[d] : nullptr;
        a.nb[d] = on ? c.nb_face[d] : 0;
    }
    a.nb[3] = nblk(c.g.ncell);
    a.p = c.p;
    a.phi = c.phi;
    hipLaunchKernelGGL(k_gen_project, dim3(a.nb[0] + a.nb[1] + a.nb[2] + a.nb[3]), dim3(B), 0, s, c.g, a, ctl);
    return hipGetLastError();
}

hipError_t fin_step(const Ctx &c, hipStream_t s, Ctl *ctl, amressif_step_stats *log)
{
    hipLaunchKernelGGL(k_gen_fin_step, dim3(1), dim3(1), 0, s, ctl, log);
    return hipGetLastError();
}

hipError_t fill(hipStream_t s, double *dst, int64_t count, double value)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_fill, dim3(nblk(count)), dim3(B), 0, s, dst, count, value);
    return hipGetLastError();
}

hipError_t pack(hipStream_t s, const amr_array4 &a, int comp, double *dst, int ex, int ey, int ez)
{
    hipLaunchKernelGGL(k_gen_pack, dim3(nblk((int64_t)ex * ey * ez)), dim3(B), 0, s, view(a), comp, dst, ex, ey, ez,
                       0);
    return hipGetLastError();
}

hipError_t unpack(hipStream_t s, const double *src, const amr_array4 &a, int comp, int ex, int ey, int ez)
{
    hipLaunchKernelGGL(k_gen_pack, dim3(nblk((int64_t)ex * ey * ez)), dim3(B), 0, s, view(a), comp,
                       const_cast<double *>(src), ex, ey, ez, 1);
    return hipGetLastError();
}

// the window [lo, hi] (inclusive, inside the view) between a view and an
// internal array of extents ex x ey x ..: view -> array (unpack 0) or back
__global__ __launch_bounds__(B) void k_gen_pack_window(V4 v, int comp, double *dst, int ex, int ey, int lo0, int lo1,
                                                       int lo2, int wx, int wy, int wz, int unpack)
{
    const int64_t q = (int64_t)blockIdx.x * B + threadIdx.x;
    if (q >= (int64_t)wx * wy * wz) return;
    const int i = lo0 + (int)(q % wx), j = lo1 + (int)((q / wx) % wy), k = lo2 + (int)(q / ((int64_t)wx * wy));
    const int64_t e = vidx(v, i, j, k, comp), d = i + (int64_t)ex * (j + (int64_t)ey * k);
    if (unpack) v.ptr[e] = dst[d];
    else dst[d] = v.ptr[e];
}

hipError_t pack_window(hipStream_t s, const amr_array4 &a, int comp, double *dst, int ex, int ey, const int *lo,
                       const int *hi, int unpack)
{
    const int wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = hi[2] - lo[2] + 1;
    hipLaunchKernelGGL(k_gen_pack_window, dim3(nblk((int64_t)wx * wy * wz)), dim3(B), 0, s, view(a), comp, dst, ex, ey,
                       lo[0], lo[1], lo[2], wx, wy, wz, unpack);
    return hipGetLastError();
}

hipError_t unpack_cells_ghosted(const Ctx &c, hipStream_t s, const double *src, const amr_array4 &a, int comp,
                                const double *shift)
{
    const V4 v = view(a);
    const int64_t m = (int64_t)(v.hi[0] - v.lo[0] + 1) * (v.hi[1] - v.lo[1] + 1) * (v.hi[2] - v.lo[2] + 1);
    hipLaunchKernelGGL(k_gen_unpack_ghosted, dim3(nblk(m)), dim3(B), 0, s, c.g, src, v, comp, shift);
    return hipGetLastError();
}

hipError_t unpack_velcart(const Ctx &c, hipStream_t s, const double *X, const amr_array4 &a)
{
    const V4 v = view(a);
    const int64_t m = (int64_t)(v.hi[0] - v.lo[0] + 1) * (v.hi[1] - v.lo[1] + 1) * (v.hi[2] - v.lo[2] + 1);
    hipLaunchKernelGGL(k_gen_unpack_velcart, dim3(nblk(m)), dim3(B), 0, s, c.g, X, v);
    return hipGetLastError();
}

}  // namespace gen
}  // namespace amr
