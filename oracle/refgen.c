/*
 * refgen.c -- CPU FP64 restatement of AMRESSIF's fractional-step time step
 * (Source/src1_rever) in 2D and 3D, with periodic directions beside the wall
 * codes: the oracle of the library's general path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker, never by the
 * product library.
 *
 * Each function follows one reference function, with the reference's
 * AMREX_SPACEDIM = 3 branches for the third direction (cited per function).
 * The periodic semantics are this oracle's own definition (refgen.h): the
 * reference's periodic path is not a usable target, because its flux and
 * phi/p ghost passes overwrite the periodic ghosts on every side.
 *
 * Parity pins (tests/test_oracle_gen.py): equal to the 2D oracle (refcpu.c,
 * itself pinned) on every 2D wall configuration; a z-independent 3D state
 * with z periodic steps as the 2D state does; a cyclic permutation of the
 * axes commutes with the 3D step; translation along a periodic direction
 * commutes with the step; the Poisson solve against the exact FFT/DCT
 * solution; Taylor-Green (2D) and Arnold-Beltrami-Childress (3D) decay at
 * second order in the cell size.
 */
#include "refgen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(const rg_params *p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* ---------------------------------------------------------------- layout */
static inline int nk_of(const rg_params *p) { return p->dim == 3 ? p->n : 1; }

long rg_face_size(const rg_params *p, int d)
{
    const long n = p->n;
    return (n + (d == 0)) * (n + (d == 1)) * (p->dim == 3 ? n + (d == 2) : 1);
}

long rg_cell_size(const rg_params *p, int G)
{
    const long s = p->n + 2 * G;
    return s * s * (p->dim == 3 ? s : 1);
}

/* face (x[0], x[1], x[2]) of direction d */
static inline long FI(const rg_params *p, int d, const int *x)
{
    const long n = p->n, nx = n + (d == 0), ny = n + (d == 1);
    return x[0] + nx * (x[1] + ny * (long)x[2]);
}

/* cell (x[0], x[1], x[2]) comp c of an array with G ghosts */
static inline long CI(const rg_params *p, int G, const int *x, int c)
{
    const long s = p->n + 2 * G, sk = p->dim == 3 ? s : 1, gk = p->dim == 3 ? G : 0;
    return (x[0] + G) + s * ((x[1] + G) + s * ((x[2] + gk) + sk * (long)c));
}

static inline int periodic(const rg_params *p, int d) { return p->bc_lo[d] == 111 && p->bc_hi[d] == 111; }

/* flattened loop over the valid cells: x from the flat index q */
static inline void unflat(const rg_params *p, long q, int *x)
{
    const long n = p->n;
    x[0] = (int)(q % n);
    x[1] = (int)((q / n) % n);
    x[2] = (int)(q / (n * n));
}
static inline long ncells(const rg_params *p) { return (long)p->n * p->n * nk_of(p); }

/* flattened loop over the faces of direction d */
static inline void unflat_face(const rg_params *p, int d, long q, int *x)
{
    const long n = p->n, nx = n + (d == 0), ny = n + (d == 1);
    x[0] = (int)(q % nx);
    x[1] = (int)((q / nx) % ny);
    x[2] = (int)(q / (nx * ny));
}

/* FillBoundary with the domain's periodicity for an array of G ghosts and
 * ncomp comps: each ghost cell whose out-of-domain coordinates are all in
 * periodic directions takes the periodic image */
static void fill_periodic(const rg_params *p, double *a, int G, int ncomp)
{
    const int n = p->n, dim = p->dim;
    int any = 0;
    for (int d = 0; d < dim; ++d) any |= periodic(p, d);
    if (!any) return;
    const int klo = dim == 3 ? -G : 0, khi = dim == 3 ? n + G : 1;
    for (int c = 0; c < ncomp; ++c)
        for (int k = klo; k < khi; ++k)
            for (int j = -G; j < n + G; ++j)
                for (int i = -G; i < n + G; ++i) {
                    int x[3] = {i, j, k}, m[3] = {i, j, k}, out = 0, ok = 1;
                    for (int d = 0; d < dim; ++d)
                        if (x[d] < 0 || x[d] >= n) {
                            out = 1;
                            if (!periodic(p, d)) ok = 0;
                            m[d] = (x[d] + n) % n;
                        }
                    if (out && ok) a[CI(p, G, x, c)] = a[CI(p, G, m, c)];
                }
}

/* --------------------------------------------------------------- cont2cart
 * Source/src1_rever/utilities.cpp:16-350.  Face -> cell average (:47-54,
 * the third component with the z-faces), FillBoundary (:60), then the wall
 * ghost layers in the order W, E, S, N (:87-269), B, F (:271-346), each over
 * the grown box, skipping periodic sides.  Mirror of x_d is -x_d-1 on the lo
 * side and (n-x_d)+(n-1) on the hi side.  Rules: 135 every comp -mirror;
 * -135 (east: the code 1) the normal comp -mirror, the others +mirror; 165
 * 2 - mirror, except north: 2 inflow - mirror.  Returns 0, -1 for a code the
 * reference aborts on, -2 for a one-sided 111, -3 for a code whose ghosts the
 * reference leaves unfilled.
 */
static int side_rule(int d, int hi, int code, int *kind)
{
    /* kind: 0 all -mirror, 1 slip (normal -, tangential +), 2 unit inflow (2 - mirror),
     * 3 north inflow (2 inflow - mirror) */
    if (code == 135) { *kind = 0; return 0; }
    if (code == 195 || code == -195) return -1;
    if (d == 0 && !hi) {
        if (code == -135) { *kind = 1; return 0; }
        if (code == 165 || code == -165) { *kind = 2; return 0; }
        return -3;
    }
    if (d == 0 && hi) {
        if (code == -165) return -1;
        if (code == 1) { *kind = 1; return 0; }
        if (code == 165) { *kind = 2; return 0; }
        return -3;
    }
    if (code == -165) return -1;
    if (code == -135) { *kind = 1; return 0; }
    if (code == 165) { *kind = (d == 1 && hi) ? 3 : 2; return 0; }
    return -3;
}

int rg_cont2cart(const rg_params *p, double *vc, double *const *f)
{
    const int n = p->n, dim = p->dim;
    set_threads(p);
    for (int d = 0; d < dim; ++d)
        if ((p->bc_lo[d] == 111) != (p->bc_hi[d] == 111)) return -2;
    const long nc = ncells(p);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nc; ++q) {
        int x[3];
        unflat(p, q, x);
        for (int c = 0; c < dim; ++c) {
            int xp[3] = {x[0], x[1], x[2]};
            xp[c] += 1;
            vc[CI(p, 2, x, c)] = 0.5 * (f[c][FI(p, c, x)] + f[c][FI(p, c, xp)]);
        }
    }
    fill_periodic(p, vc, 2, dim);
    const int klo = dim == 3 ? -2 : 0, khi = dim == 3 ? n + 2 : 1;
    for (int d = 0; d < dim; ++d)
        for (int hi = 0; hi < 2; ++hi) {
            if (periodic(p, d)) continue;
            int kind = 0;
            const int rc = side_rule(d, hi, hi ? p->bc_hi[d] : p->bc_lo[d], &kind);
            if (rc) return rc;
            for (int k = klo; k < khi; ++k)
                for (int j = -2; j < n + 2; ++j)
                    for (int i = -2; i < n + 2; ++i) {
                        int x[3] = {i, j, k};
                        if (hi ? x[d] < n : x[d] >= 0) continue;
                        int m[3] = {i, j, k};
                        m[d] = hi ? (n - x[d]) + (n - 1) : -x[d] - 1;
                        for (int c = 0; c < dim; ++c) {
                            const double mv = vc[CI(p, 2, m, c)];
                            double g;
                            if (kind == 0) g = -mv;
                            else if (kind == 1) g = c == d ? -mv : mv;
                            else if (kind == 2) g = 2 - mv;
                            else g = 2 * p->inflow[c] - mv;
                            vc[CI(p, 2, x, c)] = g;
                        }
                    }
        }
    return 0;
}

/* --------------------------------------------------- QUICK convective flux
 * Source/src1_rever/fn_flux_calc.cpp:11-202: on the faces of direction d,
 * con = 0.5 f_d, ud = con + |con|, up = con - |con| (:66-68, 99-101,
 * 117-119); the half-node flux of each comp (:71-79, 103-110, 121-128); the
 * cell divergence summed x, y, z (:161-180); fluxTotal -= C (:194-199).
 * fh[d * dim + c]: comp c on the faces of direction d (fluxHalfN1/2/3[d]).
 */
static inline double quick_face(double up, double ud, double a2, double a1, double b0, double b1)
{
    return up * ((1.0 / 8.0) * (-b1 - 2 * b0 + 3 * a1) + b0) + ud * ((1.0 / 8.0) * (-a2 - 2 * a1 + 3 * b0) + a1);
}

void rg_convective_flux(const rg_params *p, const double *vc, double *const *f, double *const *fh, double *fconv,
                        double *ftot)
{
    const int dim = p->dim;
    set_threads(p);
    for (int d = 0; d < dim; ++d) {
        const long nf = rg_face_size(p, d);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nf; ++q) {
            int x[3];
            unflat_face(p, d, q, x);
            const double con = 0.5 * f[d][q];
            const double ud = con + fabs(con), up = con - fabs(con);
            int m2[3] = {x[0], x[1], x[2]}, m1[3] = {x[0], x[1], x[2]}, p1[3] = {x[0], x[1], x[2]};
            m2[d] -= 2;
            m1[d] -= 1;
            p1[d] += 1;
            for (int c = 0; c < dim; ++c)
                fh[d * dim + c][q] = quick_face(up, ud, vc[CI(p, 2, m2, c)], vc[CI(p, 2, m1, c)],
                                                vc[CI(p, 2, x, c)], vc[CI(p, 2, p1, c)]);
        }
    }
    const long nc = ncells(p);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nc; ++q) {
        int x[3];
        unflat(p, q, x);
        for (int c = 0; c < dim; ++c) {
            double s = 0.0;
            for (int d = 0; d < dim; ++d) {
                int xp[3] = {x[0], x[1], x[2]};
                xp[d] += 1;
                const double t = (fh[d * dim + c][FI(p, d, xp)] - fh[d * dim + c][FI(p, d, x)]) / p->dx[d];
                s = d == 0 ? t : s + t;
            }
            fconv[CI(p, 0, x, c)] = s;
            ftot[CI(p, 1, x, c)] = ftot[CI(p, 1, x, c)] - s;
        }
    }
}

/* ------------------------------------------------------------ viscous flux
 * Source/src1_rever/fn_flux_calc.cpp:695-766: ((E - 2C + W)/dx^2 +
 * (N - 2C + S)/dy^2)/ren, plus ((B - 2C + F)/dz^2)/ren in 3D (:731-737);
 * fluxTotal += V (:753-765).
 */
void rg_viscous_flux(const rg_params *p, const double *vc, double *fvisc, double *ftot)
{
    const int dim = p->dim;
    set_threads(p);
    const long nc = ncells(p);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nc; ++q) {
        int x[3];
        unflat(p, q, x);
        int nb[3][2][3];
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) {
                nb[d][s][0] = x[0]; nb[d][s][1] = x[1]; nb[d][s][2] = x[2];
                nb[d][s][d] += s ? 1 : -1;
            }
        for (int c = 0; c < dim; ++c) {
            const double C = vc[CI(p, 2, x, c)];
            const double W = vc[CI(p, 2, nb[0][0], c)], E = vc[CI(p, 2, nb[0][1], c)];
            const double S = vc[CI(p, 2, nb[1][0], c)], N = vc[CI(p, 2, nb[1][1], c)];
            double v = ((E - 2 * C + W) / (p->dx[0] * p->dx[0]) + (N - 2 * C + S) / (p->dx[1] * p->dx[1])) / p->ren;
            if (dim == 3) {
                const double F = vc[CI(p, 2, nb[2][0], c)], B = vc[CI(p, 2, nb[2][1], c)];
                v = v + ((B - 2 * C + F) / (p->dx[2] * p->dx[2])) / p->ren;
            }
            fvisc[CI(p, 0, x, c)] = v;
            ftot[CI(p, 1, x, c)] = ftot[CI(p, 1, x, c)] + v;
        }
    }
}

/* ----------------------------------------------------- face gradients
 * Source/src1_rever/fn_flux_calc.cpp:825-911: on every face of direction d,
 * G_d q = (q(x) - q(x - e_d))/dx_d (:884-896, z :905-907), q = userCtx comp.
 */
void rg_gradient(const rg_params *p, const double *ctx, int comp, double *const *g)
{
    const int dim = p->dim;
    set_threads(p);
    for (int d = 0; d < dim; ++d) {
        const long nf = rg_face_size(p, d);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nf; ++q) {
            int x[3], m[3];
            unflat_face(p, d, q, x);
            m[0] = x[0]; m[1] = x[1]; m[2] = x[2];
            m[d] -= 1;
            g[d][q] = (ctx[CI(p, 1, x, comp)] - ctx[CI(p, 1, m, comp)]) / p->dx[d];
        }
    }
}

/* ------------------------------------------------- flux ghosts
 * Source/src1_rever/main.cpp:939-1014: FillBoundary, then on each wall side
 * the normal comp's ghost = -mirror (:953-1012); the other comps' domain
 * ghosts keep the 0 of setVal (:935).  Periodic sides keep FillBoundary's
 * images (this oracle's periodic semantics).
 */
void rg_flux_ghosts(const rg_params *p, double *ft)
{
    const int n = p->n, dim = p->dim;
    fill_periodic(p, ft, 1, dim);
    const int klo = dim == 3 ? -1 : 0, khi = dim == 3 ? n + 1 : 1;
    for (int d = 0; d < dim; ++d)
        for (int hi = 0; hi < 2; ++hi) {
            if (periodic(p, d)) continue;
            for (int k = klo; k < khi; ++k)
                for (int j = -1; j < n + 1; ++j)
                    for (int i = -1; i < n + 1; ++i) {
                        int x[3] = {i, j, k};
                        if (hi ? x[d] < n : x[d] >= 0) continue;
                        int m[3] = {i, j, k};
                        m[d] = hi ? (n - x[d]) + (n - 1) : -x[d] - 1;
                        ft[CI(p, 1, x, d)] = -ft[CI(p, 1, m, d)];
                    }
        }
}

/* ------------------------------------------- momentum_righthand_side_calc
 * Source/src1_rever/fn_rhs.cpp:8-68: R_d = -G_d p + 0.5 (T_d(x - e_d) + T_d(x))
 * (:45, :55, z :60).
 */
void rg_momentum_rhs(const rg_params *p, const double *ft, double *const *gp, double *const *rhs)
{
    const int dim = p->dim;
    set_threads(p);
    for (int d = 0; d < dim; ++d) {
        const long nf = rg_face_size(p, d);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nf; ++q) {
            int x[3], m[3];
            unflat_face(p, d, q, x);
            m[0] = x[0]; m[1] = x[1]; m[2] = x[2];
            m[d] -= 1;
            rhs[d][q] = -gp[d][q] + 0.5 * (ft[CI(p, 1, m, d)] + ft[CI(p, 1, x, d)]);
        }
    }
}

/* ------------------------------------------------------ Error_Computation
 * Source/src1_rever/utilities.cpp:421-493 with StagL2Norm (:669-746): d =
 * u^k - u* on the faces, s_d = sum w d^2 with w = 1/2 on the first and last
 * face along the normal, e_d = sqrt(s_d)/sqrt(npts), normError = max(ex, ey)
 * then max(., ez) (std::max order).
 */
double rg_error_computation(const rg_params *p, double *const *uk, double *const *us)
{
    const int n = p->n, dim = p->dim;
    const double npts = (double)n * (double)n * (dim == 3 ? (double)n : 1.0);
    double e[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        const long nf = rg_face_size(p, d);
        double s = 0.0;
        for (long q = 0; q < nf; ++q) {
            int x[3];
            unflat_face(p, d, q, x);
            const double w = (x[d] > 0 && x[d] < n) ? 1.0 : 0.5;
            const double df = uk[d][q] - us[d][q];
            s += w * df * df;
        }
        e[d] = sqrt(s) / sqrt(npts);
    }
    double ne = (e[0] < e[1]) ? e[1] : e[0];
    if (dim == 3) ne = (ne < e[2]) ? e[2] : ne;
    return ne;
}

/* ------------------------------------------------------- Poisson RHS
 * Source/src1_rever/fn_rhs.cpp:71-125: D = sum_d (f_d(x + e_d) - f_d(x))/dx_d
 * (:93-101), ||D||inf (:103-104), b = (1.5/dt) D (:116), sum b (:123-124).
 */
void rg_poisson_rhs(const rg_params *p, double *const *f, double *rhs, double *div_linf, double *rhs_sum)
{
    const int dim = p->dim;
    const long nc = ncells(p);
    double m = 0.0, s = 0.0;
    for (long q = 0; q < nc; ++q) {
        int x[3];
        unflat(p, q, x);
        double D = 0.0;
        for (int d = 0; d < dim; ++d) {
            int xp[3] = {x[0], x[1], x[2]};
            xp[d] += 1;
            const double t = (f[d][FI(p, d, xp)] - f[d][FI(p, d, x)]) / p->dx[d];
            D = d == 0 ? t : D + t;
        }
        if (fabs(D) > m || D != D) m = fabs(D);
        const double b = (1.5 / p->dt) * D;
        rhs[CI(p, 1, x, 0)] = b;
        s += b;
    }
    if (div_linf) *div_linf = m;
    if (rhs_sum) *rhs_sum = s;
}

/* ------------------------------------------------- userCtx ghost fill
 * Source/src1_rever/main.cpp:1078-1158 (poisson.cpp:312-391): homogeneous
 * Neumann mirror for both comps on the wall sides; periodic images on the
 * periodic ones.
 */
void rg_userctx_ghosts(const rg_params *p, double *ctx)
{
    const int n = p->n, dim = p->dim;
    fill_periodic(p, ctx, 1, 2);
    const int klo = dim == 3 ? -1 : 0, khi = dim == 3 ? n + 1 : 1;
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < dim; ++d)
            for (int hi = 0; hi < 2; ++hi) {
                if (periodic(p, d)) continue;
                for (int k = klo; k < khi; ++k)
                    for (int j = -1; j < n + 1; ++j)
                        for (int i = -1; i < n + 1; ++i) {
                            int x[3] = {i, j, k};
                            if (hi ? x[d] < n : x[d] >= 0) continue;
                            int m[3] = {i, j, k};
                            m[d] = hi ? n - 1 : 0;
                            ctx[CI(p, 1, x, c)] = ctx[CI(p, 1, m, c)];
                        }
            }
}

/* ================================================================ Poisson
 * Source/src1_rever/poisson.cpp:41-186: MLABecLaplacian alpha = 0, beta =
 * -1, B = 1, i.e. L phi = sum_d (phi(x+e_d) - 2 phi + phi(x-e_d))/dx_d^2
 * with homogeneous Neumann on wall sides (foextrap, :85-120) and periodic
 * images on periodic ones; phi0 = 0 (:63).  The RHS mean is removed (the
 * operator is singular in both cases) and phi is returned mean-zero.
 * Geometric multigrid (AMReX's documented MLMG cycle: V-cycle, 2 + 2 GSRB
 * sweeps, 2^dim averaging, piecewise-constant prolongation) to a coarsest
 * level of at most 512 cells or an odd side, solved by its dense
 * pseudo-inverse; stop at ||r||inf <= tol_rel ||b||inf.
 */
typedef struct {
    int n;             /* cells per side on this level */
    double a[3];       /* 1/dx_d^2 */
    double *e, *r, *t; /* correction, right-hand side, scratch */
} gl_level;

typedef struct {
    int nl;
    gl_level lev[24];
    double *pinv; /* bottom pseudo-inverse, m x m */
    int m;
} gl_hier;

static inline long lidx(const rg_params *p, int n, const int *x)
{
    return x[0] + (long)n * (x[1] + (long)n * x[2]);
}

/* neighbour of x along d by s = +-1 on a level of n cells: -1 if the face is
 * a Neumann wall (the term drops), else the flat index (periodic wrap) */
static inline long lnb(const rg_params *p, int n, const int *x, int d, int s)
{
    int y[3] = {x[0], x[1], x[2]};
    y[d] += s;
    if (y[d] < 0 || y[d] >= n) {
        if (!periodic(p, d)) return -1;
        y[d] = (y[d] + n) % n;
    }
    return lidx(p, n, y);
}

static void gl_apply(const rg_params *p, const gl_level *L, const double *x, double *out)
{
    const int n = L->n, dim = p->dim;
    const long m = (long)n * n * (dim == 3 ? n : 1);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < m; ++q) {
        int c[3] = {(int)(q % n), (int)((q / n) % n), (int)(q / ((long)n * n))};
        double s = 0.0;
        for (int d = 0; d < dim; ++d) {
            const long lo = lnb(p, n, c, d, -1), hi = lnb(p, n, c, d, +1);
            const double xl = lo < 0 ? x[q] : x[lo], xh = hi < 0 ? x[q] : x[hi];
            s += (xh - 2 * x[q] + xl) * L->a[d];
        }
        out[q] = s;
    }
}

static void gl_gsrb(const rg_params *p, const gl_level *L, double *x, const double *b, int color)
{
    const int n = L->n, dim = p->dim;
    const long m = (long)n * n * (dim == 3 ? n : 1);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < m; ++q) {
        int c[3] = {(int)(q % n), (int)((q / n) % n), (int)(q / ((long)n * n))};
        if (((c[0] + c[1] + c[2]) & 1) != color) continue;
        double off = 0.0, diag = 0.0;
        for (int d = 0; d < dim; ++d)
            for (int s = -1; s <= 1; s += 2) {
                const long nb = lnb(p, n, c, d, s);
                if (nb < 0) continue;
                off += L->a[d] * x[nb];
                diag += L->a[d];
            }
        x[q] = (off - b[q]) / diag;
    }
}

static int chol(double *A, int m)
{
    for (int j = 0; j < m; ++j) {
        double s = A[(long)j * m + j];
        for (int k = 0; k < j; ++k) s -= A[(long)j * m + k] * A[(long)j * m + k];
        if (!(s > 0)) return -1;
        const double l = sqrt(s);
        A[(long)j * m + j] = l;
        for (int i = j + 1; i < m; ++i) {
            double t = A[(long)i * m + j];
            for (int k = 0; k < j; ++k) t -= A[(long)i * m + k] * A[(long)j * m + k];
            A[(long)i * m + j] = t / l;
        }
    }
    return 0;
}

/* pinv(L) = J - (J - L)^{-1}, J = 11^T/m: J - L is SPD (L's null space is
 * the constants under Neumann and periodic sides alike) */
static int gl_pinv(const rg_params *p, const gl_level *L, double *out)
{
    const int n = L->n, dim = p->dim;
    const int m = n * n * (dim == 3 ? n : 1);
    double *A = (double *)calloc((size_t)m * m, sizeof(double));
    double *col = (double *)calloc((size_t)m, sizeof(double));
    double *img = (double *)calloc((size_t)m, sizeof(double));
    for (int j = 0; j < m; ++j) {
        memset(col, 0, sizeof(double) * (size_t)m);
        col[j] = 1.0;
        gl_apply(p, L, col, img);
        for (int i = 0; i < m; ++i) A[(long)i * m + j] = 1.0 / m - img[i];
    }
    int rc = chol(A, m);
    if (rc == 0)
        for (int j = 0; j < m; ++j) {
            /* solve (J - L) y = e_j */
            memset(col, 0, sizeof(double) * (size_t)m);
            col[j] = 1.0;
            for (int i = 0; i < m; ++i) {
                double t = col[i];
                for (int k = 0; k < i; ++k) t -= A[(long)i * m + k] * col[k];
                col[i] = t / A[(long)i * m + i];
            }
            for (int i = m - 1; i >= 0; --i) {
                double t = col[i];
                for (int k = i + 1; k < m; ++k) t -= A[(long)k * m + i] * col[k];
                col[i] = t / A[(long)i * m + i];
            }
            for (int i = 0; i < m; ++i) out[(long)i * m + j] = 1.0 / m - col[i];
        }
    free(A);
    free(col);
    free(img);
    return rc;
}

static long lcells(const rg_params *p, int n) { return (long)n * n * (p->dim == 3 ? n : 1); }

static int gl_build(const rg_params *p, gl_hier *H)
{
    int n = p->n, l = 0;
    for (;;) {
        gl_level *L = &H->lev[l];
        L->n = n;
        for (int d = 0; d < 3; ++d) {
            const double h = p->dx[d] * (double)(p->n / n);
            L->a[d] = 1.0 / (h * h);
        }
        const long m = lcells(p, n);
        L->e = (double *)calloc((size_t)m, sizeof(double));
        L->r = (double *)calloc((size_t)m, sizeof(double));
        L->t = (double *)calloc((size_t)m, sizeof(double));
        ++l;
        if ((n & 1) || m <= 512 || l == 24) break;
        n /= 2;
    }
    H->nl = l;
    H->m = (int)lcells(p, H->lev[l - 1].n);
    if (H->m > 4096) return -1;
    H->pinv = (double *)calloc((size_t)H->m * H->m, sizeof(double));
    return gl_pinv(p, &H->lev[l - 1], H->pinv);
}

static void gl_free(gl_hier *H)
{
    for (int l = 0; l < H->nl; ++l) {
        free(H->lev[l].e);
        free(H->lev[l].r);
        free(H->lev[l].t);
    }
    free(H->pinv);
}

/* V-cycle on level l for L e = r (e from 0) */
static void gl_vcycle(const rg_params *p, gl_hier *H, int l)
{
    gl_level *L = &H->lev[l];
    const long m = lcells(p, L->n);
    if (l == H->nl - 1) {
        for (long i = 0; i < m; ++i) {
            double s = 0.0;
            for (long j = 0; j < m; ++j) s += H->pinv[i * m + j] * L->r[j];
            L->e[i] = s;
        }
        return;
    }
    memset(L->e, 0, sizeof(double) * (size_t)m);
    for (int s = 0; s < 2; ++s) {
        gl_gsrb(p, L, L->e, L->r, 0);
        gl_gsrb(p, L, L->e, L->r, 1);
    }
    gl_apply(p, L, L->e, L->t);
    gl_level *Cl = &H->lev[l + 1];
    const int nc = Cl->n, dim = p->dim, n = L->n;
    const long mc = lcells(p, nc);
    const double w = dim == 3 ? 0.125 : 0.25;
    for (long q = 0; q < mc; ++q) {
        int c[3] = {(int)(q % nc), (int)((q / nc) % nc), (int)(q / ((long)nc * nc))};
        double s = 0.0;
        for (int dz = 0; dz < (dim == 3 ? 2 : 1); ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dxx = 0; dxx < 2; ++dxx) {
                    int f[3] = {2 * c[0] + dxx, 2 * c[1] + dy, dim == 3 ? 2 * c[2] + dz : 0};
                    const long k = lidx(p, n, f);
                    s += L->r[k] - L->t[k];
                }
        Cl->r[q] = w * s;
    }
    gl_vcycle(p, H, l + 1);
    for (long q = 0; q < m; ++q) {
        int f[3] = {(int)(q % n), (int)((q / n) % n), (int)(q / ((long)n * n))};
        int c[3] = {f[0] / 2, f[1] / 2, f[2] / 2};
        L->e[q] += Cl->e[lidx(p, nc, c)];
    }
    for (int s = 0; s < 2; ++s) {
        gl_gsrb(p, L, L->e, L->r, 0);
        gl_gsrb(p, L, L->e, L->r, 1);
    }
}

void rg_apply_laplacian(const rg_params *p, const double *phi, double *out)
{
    gl_level L;
    L.n = p->n;
    for (int d = 0; d < 3; ++d) L.a[d] = 1.0 / (p->dx[d] * p->dx[d]);
    gl_apply(p, &L, phi, out);
}

int rg_poisson_solve(const rg_params *p, double *sol, const double *rhs, int *iters, double *res_rel)
{
    const long m = ncells(p);
    set_threads(p);
    double *b = (double *)malloc(sizeof(double) * (size_t)m);
    double *x = (double *)calloc((size_t)m, sizeof(double));
    double *t = (double *)malloc(sizeof(double) * (size_t)m);
    double mb = 0.0;
    for (long q = 0; q < m; ++q) {
        int c[3];
        unflat(p, q, c);
        b[q] = rhs[CI(p, 1, c, 0)];
        mb += b[q];
    }
    mb /= (double)m;
    double bnorm = 0.0;
    for (long q = 0; q < m; ++q) {
        b[q] -= mb;
        if (fabs(b[q]) > bnorm) bnorm = fabs(b[q]);
    }
    gl_hier H;
    memset(&H, 0, sizeof(H));
    int rc = gl_build(p, &H);
    int it = 0;
    double rr = 0.0;
    if (rc == 0) {
        gl_level *F = &H.lev[0];
        const int maxit = p->mg_max_iter > 0 ? p->mg_max_iter : 200;
        for (;;) {
            gl_apply(p, F, x, t);
            rr = 0.0;
            for (long q = 0; q < m; ++q) {
                F->r[q] = b[q] - t[q];
                if (fabs(F->r[q]) > rr || F->r[q] != F->r[q]) rr = fabs(F->r[q]);
            }
            if (rr <= p->poisson_tol_rel * bnorm || it >= maxit || rr != rr) break;
            gl_vcycle(p, &H, 0);
            for (long q = 0; q < m; ++q) x[q] += F->e[q];
            ++it;
        }
    }
    gl_free(&H);
    double mx = 0.0;
    for (long q = 0; q < m; ++q) mx += x[q];
    mx /= (double)m;
    for (long q = 0; q < m; ++q) x[q] -= mx;
    rg_apply_laplacian(p, x, t);
    rr = 0.0;
    for (long q = 0; q < m; ++q)
        if (fabs(b[q] - t[q]) > rr) rr = fabs(b[q] - t[q]);
    const long ng = rg_cell_size(p, 1);
    for (long q = 0; q < ng; ++q) sol[q] = 0.0;
    for (long q = 0; q < m; ++q) {
        int c[3];
        unflat(p, q, c);
        sol[CI(p, 1, c, 0)] = x[q];
    }
    if (iters) *iters = it;
    if (res_rel) *res_rel = bnorm > 0 ? rr / bnorm : rr;
    free(b);
    free(x);
    free(t);
    return rc;
}

/* --------------------------------------------------- update_solution
 * Source/src1_rever/poisson.cpp:192-392: face gradients of phi (:216),
 * u_d = u*_d - G_d phi dt/1.5 (:251, :256, z :261), cont2cart (:267),
 * p <- p + phi (:309), userCtx ghosts (:312-391).
 */
int rg_update_solution(const rg_params *p, double *ctx, double *vc, double *const *u, double *const *us)
{
    const int dim = p->dim;
    for (int d = 0; d < dim; ++d) {
        const long nf = rg_face_size(p, d);
        for (long q = 0; q < nf; ++q) {
            int x[3], m[3];
            unflat_face(p, d, q, x);
            m[0] = x[0]; m[1] = x[1]; m[2] = x[2];
            m[d] -= 1;
            const double g = (ctx[CI(p, 1, x, 1)] - ctx[CI(p, 1, m, 1)]) / p->dx[d];
            u[d][q] = us[d][q] - (g * p->dt / 1.5);
        }
    }
    const int rc = rg_cont2cart(p, vc, u);
    const long nc = ncells(p);
    for (long q = 0; q < nc; ++q) {
        int x[3];
        unflat(p, q, x);
        ctx[CI(p, 1, x, 0)] = ctx[CI(p, 1, x, 0)] + ctx[CI(p, 1, x, 1)];
    }
    rg_userctx_ghosts(p, ctx);
    return rc;
}

/* ------------------------------------------------------------ advance
 * Source/src1_rever/main.cpp:875-1202 as the 2D oracle's rc_advance, per
 * direction: BDF2 history (879-884), RK4 dual time with rk = dtau{1/4, 1/3,
 * 1/2, 1}, dtau = 0.4 dt (860-870), the loop while normError > tol (898),
 * the stage sequence (929-1017) with runge_kutta4_pseudo_time_stepping
 * (momentum.cpp:9-94), velCont <- velStar (1023-1026), caps (1030-1039),
 * or PSEUDO_TIMESTEPPING = 0's single explicit update (900-926,
 * momentum.cpp:96-191; fluxTotal's wall ghosts keep 0, the periodic ones
 * FillBoundary's images); Poisson (1058-1060), phi into userCtx
 * (1078-1158), update_solution (1164), divergence (1180).
 */
int rg_advance(const rg_params *p, double *const *u, double *const *up, double *ctx, double *vc, int nsteps,
               rg_step_stats *stats)
{
    const int dim = p->dim;
    double *diff[3] = {0}, *star[3] = {0}, *rhs[3] = {0}, *gp[3] = {0}, *fh[9] = {0};
    long nf[3] = {0};
    for (int d = 0; d < dim; ++d) {
        nf[d] = rg_face_size(p, d);
        diff[d] = (double *)malloc(sizeof(double) * (size_t)nf[d]);
        star[d] = (double *)malloc(sizeof(double) * (size_t)nf[d]);
        rhs[d] = (double *)malloc(sizeof(double) * (size_t)nf[d]);
        gp[d] = (double *)malloc(sizeof(double) * (size_t)nf[d]);
        for (int c = 0; c < dim; ++c) fh[d * dim + c] = (double *)malloc(sizeof(double) * (size_t)nf[d]);
    }
    const long n1 = rg_cell_size(p, 1), n0 = rg_cell_size(p, 0);
    double *ft = (double *)malloc(sizeof(double) * (size_t)n1 * dim);
    double *fconv = (double *)malloc(sizeof(double) * (size_t)n0 * dim);
    double *fvisc = (double *)malloc(sizeof(double) * (size_t)n0 * dim);
    double *prhs = (double *)calloc((size_t)n1, sizeof(double));
    double *psol = (double *)calloc((size_t)n1, sizeof(double));
    const double dtau = 0.4 * p->dt;
    const double rk[4] = {dtau * 0.25, dtau * (1.0 / 3.0), dtau * 0.5, dtau * 1.0};
    int rc = rg_cont2cart(p, vc, u);
    for (int step = 0; step < nsteps && rc == 0; ++step) {
        rg_step_stats st;
        memset(&st, 0, sizeof(st));
        for (int d = 0; d < dim; ++d)
            for (long q = 0; q < nf[d]; ++q) {
                diff[d][q] = u[d][q] - up[d][q];
                up[d][q] = u[d][q];
                star[d][q] = u[d][q];
            }
        int count = 0;
        double normError = 1.e9;
        if (!p->pseudo) {
            rg_gradient(p, ctx, 0, gp);
            memset(ft, 0, sizeof(double) * (size_t)n1 * dim);
            rg_convective_flux(p, vc, star, fh, fconv, ft);
            rg_viscous_flux(p, vc, fvisc, ft);
            fill_periodic(p, ft, 1, dim);
            rg_momentum_rhs(p, ft, gp, rhs);
            for (int d = 0; d < dim; ++d)
                for (long q = 0; q < nf[d]; ++q) {
                    rhs[d][q] = rhs[d][q] + (1.0 / (2.0 * p->dt)) * diff[d][q];
                    star[d][q] = star[d][q] + ((p->dt / 1.5) * rhs[d][q]);
                }
            normError = rg_error_computation(p, u, star);
            for (int d = 0; d < dim; ++d) memcpy(u[d], star[d], sizeof(double) * (size_t)nf[d]);
        }
        while (p->pseudo && normError > p->mom_tol) {
            for (int sub = 0; sub < 4; ++sub) {
                rg_gradient(p, ctx, 0, gp);
                memset(ft, 0, sizeof(double) * (size_t)n1 * dim);
                rg_convective_flux(p, vc, star, fh, fconv, ft);
                rg_viscous_flux(p, vc, fvisc, ft);
                rg_flux_ghosts(p, ft);
                rg_momentum_rhs(p, ft, gp, rhs);
                for (int d = 0; d < dim; ++d)
                    for (long q = 0; q < nf[d]; ++q) {
                        rhs[d][q] = rhs[d][q] - (1.5 / p->dt) * (star[d][q] - up[d][q]) + (0.5 / p->dt) * diff[d][q];
                        star[d][q] = u[d][q] + (rk[sub] * rhs[d][q]);
                    }
                rc = rg_cont2cart(p, vc, star);
                if (rc) break;
            }
            if (rc) break;
            normError = rg_error_computation(p, u, star);
            for (int d = 0; d < dim; ++d) memcpy(u[d], star[d], sizeof(double) * (size_t)nf[d]);
            ++count;
            if (count > p->iter_num) { st.flags |= 1; break; }
            if (normError > 1.e2) { st.flags |= 2; break; }
        }
        if (rc) break;
        if (!(normError == normError) || isinf(normError)) st.flags |= 4;
        st.pseudo_iters = count;
        st.norm_error = normError;
        rg_poisson_rhs(p, star, prhs, &st.div_star, &st.rhs_sum);
        rc = rg_poisson_solve(p, psol, prhs, &st.vcycles, &st.poisson_res);
        if (rc) break;
        const long nc = ncells(p);
        for (long q = 0; q < nc; ++q) {
            int x[3];
            unflat(p, q, x);
            ctx[CI(p, 1, x, 1)] = psol[CI(p, 1, x, 0)];
        }
        rg_userctx_ghosts(p, ctx);
        rc = rg_update_solution(p, ctx, vc, u, star);
        rg_poisson_rhs(p, u, prhs, &st.div_new, NULL);
        if (stats) stats[step] = st;
    }
    for (int d = 0; d < dim; ++d) {
        free(diff[d]); free(star[d]); free(rhs[d]); free(gp[d]);
        for (int c = 0; c < dim; ++c) free(fh[d * dim + c]);
    }
    free(ft); free(fconv); free(fvisc); free(prhs); free(psol);
    return rc;
}
