/*
 * refcpu.c -- CPU FP64 restatement of AMRESSIF's fractional-step time step
 * (the Source/src1_rever variant that Exec/GNUmakefile builds), written from
 * the reference's behaviour, one function per reference function.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle.  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker
 * or as the timed CPU baseline, and never by the product library.
 *
 * Parity pins (tests/test_oracle.py):
 *   cont2cart ghost rules ........ closed forms (ghost+mirror)/2 == wall value
 *   QUICK half-node flux ......... exact on quadratics, either flow direction
 *   viscous flux ................. Laplacian exact on quadratics
 *   gradient / divergence ........ D(G phi) == Neumann 5-point Laplacian, exact
 *   RK4 dual-time stage .......... amplification polynomial 1+z+z^2/2+z^3/6+z^4/24
 *   error norm ................... hand-computed weighted RMS
 *   Poisson (MG and DCT) ......... dense pseudo-inverse (numpy) and scipy dctn
 *   whole step ................... divergence-free, wall-normal faces exactly 0,
 *                                   MG and DCT oracles agree
 * The MG V-cycle restates the behaviour AMReX's MLMG documents for this
 * operator (AMReX is not vendored and is unpinned): its iterate sequence and
 * the additive constant of phi are "parity unpinned"; phi is taken mean-zero.
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- layout */
#define XF(i, j) ((size_t)(i) + (size_t)(n + 1) * (size_t)(j))
#define YF(i, j) ((size_t)(i) + (size_t)(n) * (size_t)(j))
#define VC(i, j, c) ((size_t)((i) + 2) + (size_t)(n + 4) * (size_t)((j) + 2) + (size_t)(c) * (size_t)(n + 4) * (size_t)(n + 4))
#define G1(i, j, c) ((size_t)((i) + 1) + (size_t)(n + 2) * (size_t)((j) + 1) + (size_t)(c) * (size_t)(n + 2) * (size_t)(n + 2))
#define C0(i, j, c) ((size_t)(i) + (size_t)(n) * (size_t)(j) + (size_t)(c) * (size_t)(n) * (size_t)(n))

static void set_threads(const rc_params *p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* --------------------------------------------------------------- cont2cart
 * Source/src1_rever/utilities.cpp:16-350.  Face -> cell average (47-54), then
 * the ghost layers g = 1, 2 in the order W, E, S, N, each over the grown box
 * (so the S/N passes write the corners last).  Mirror of i is -i-1 on the lo
 * side and (n-i)+(n-1) on the hi side (utilities.cpp:99, 148).
 * Returns 0, or -1 for a code the reference aborts on, -2 for periodic (111,
 * not supported by this oracle), -3 for a code the reference leaves unfilled.
 */
int rc_cont2cart(const rc_params *p, double *vc, const double *u, const double *v)
{
    const int n = p->n;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            vc[VC(i, j, 0)] = 0.5 * (u[XF(i, j)] + u[XF(i + 1, j)]);
            vc[VC(i, j, 1)] = 0.5 * (v[YF(i, j)] + v[YF(i, j + 1)]);
        }

    /* West (utilities.cpp:87-135) */
    {
        const int code = p->bc_lo[0];
        if (code == 195 || code == -195) return -1;
        if (code == 111) return -2;
        if (code != 135 && code != -135 && code != 165 && code != -165) return -3;
        for (int j = -2; j < n + 2; ++j)
            for (int i = -2; i < 0; ++i) {
                const int m = -i - 1;
                if (code == 135) {
                    vc[VC(i, j, 0)] = -vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(m, j, 1)];
                } else if (code == -135) {
                    vc[VC(i, j, 0)] = -vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = vc[VC(m, j, 1)];
                } else { /* 165 and -165: hard-coded unit inflow */
                    vc[VC(i, j, 0)] = 2 - vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = 2 - vc[VC(m, j, 1)];
                }
            }
    }
    /* East (utilities.cpp:137-184); the slip branch tests the code 1 */
    {
        const int code = p->bc_hi[0];
        if (code == -165 || code == 195 || code == -195) return -1;
        if (code == 111) return -2;
        if (code != 135 && code != 1 && code != 165) return -3;
        for (int j = -2; j < n + 2; ++j)
            for (int i = n; i < n + 2; ++i) {
                const int m = (n - i) + (n - 1);
                if (code == 135) {
                    vc[VC(i, j, 0)] = -vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(m, j, 1)];
                } else if (code == 1) {
                    vc[VC(i, j, 0)] = -vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = vc[VC(m, j, 1)];
                } else {
                    vc[VC(i, j, 0)] = 2 - vc[VC(m, j, 0)];
                    vc[VC(i, j, 1)] = 2 - vc[VC(m, j, 1)];
                }
            }
    }
    /* South (utilities.cpp:189-232) */
    {
        const int code = p->bc_lo[1];
        if (code == -165 || code == 195 || code == -195) return -1;
        if (code == 111) return -2;
        if (code != 135 && code != -135 && code != 165) return -3;
        for (int j = -2; j < 0; ++j)
            for (int i = -2; i < n + 2; ++i) {
                const int m = -j - 1;
                if (code == 135) {
                    vc[VC(i, j, 0)] = -vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(i, m, 1)];
                } else if (code == -135) {
                    vc[VC(i, j, 0)] = vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(i, m, 1)];
                } else {
                    vc[VC(i, j, 0)] = 2 - vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = 2 - vc[VC(i, m, 1)];
                }
            }
    }
    /* North (utilities.cpp:236-275): 165 is the lid, 2*inflow - mirror */
    {
        const int code = p->bc_hi[1];
        if (code == -165 || code == 195 || code == -195) return -1;
        if (code == 111) return -2;
        if (code != 135 && code != -135 && code != 165) return -3;
        for (int j = n; j < n + 2; ++j)
            for (int i = -2; i < n + 2; ++i) {
                const int m = (n - j) + (n - 1);
                if (code == 135) {
                    vc[VC(i, j, 0)] = -vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(i, m, 1)];
                } else if (code == -135) {
                    vc[VC(i, j, 0)] = vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = -vc[VC(i, m, 1)];
                } else {
                    const double gx = (p->lid && i >= 0 && i < n) ? p->lid[i] : p->inflow[0];
                    vc[VC(i, j, 0)] = 2 * gx - vc[VC(i, m, 0)];
                    vc[VC(i, j, 1)] = 2 * p->inflow[1] - vc[VC(i, m, 1)];
                }
            }
    }
    return 0;
}

/* --------------------------------------------------- QUICK convective flux
 * Source/src1_rever/fn_flux_calc.cpp:11-202 (convective_flux_calc_new_quick).
 * x-faces: ucon = 0.5*ucont, ud = ucon+|ucon| (= max(ucont,0)),
 * up = ucon-|ucon| (= min(ucont,0)) (:66-68); half-node flux (:71-73);
 * y-faces with vcont (:99-105); cell divergence (:161-180); fluxTotal -= C
 * (:194-199).  The velocity argument is velStar (main.cpp:936).
 */
static inline double quick_face(double up, double ud, double a2, double a1, double b0, double b1)
{
    /* a2 = U(i-2), a1 = U(i-1), b0 = U(i), b1 = U(i+1), in the reference's operation order */
    return up * ((1.0 / 8.0) * (-b1 - 2 * b0 + 3 * a1) + b0) +
           ud * ((1.0 / 8.0) * (-a2 - 2 * a1 + 3 * b0) + a1);
}

void rc_convective_flux(const rc_params *p, const double *vc, const double *u, const double *v,
                        double *fhx_u, double *fhx_v, double *fhy_u, double *fhy_v,
                        double *fconv, double *ftot)
{
    const int n = p->n;
    const double dx = p->dx, dy = p->dy;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= n; ++i) {
            const double ucon = 0.5 * u[XF(i, j)];
            const double ud = ucon + fabs(ucon);
            const double up = ucon - fabs(ucon);
            for (int c = 0; c < 2; ++c) {
                const double f = quick_face(up, ud, vc[VC(i - 2, j, c)], vc[VC(i - 1, j, c)],
                                            vc[VC(i, j, c)], vc[VC(i + 1, j, c)]);
                if (c == 0) fhx_u[XF(i, j)] = f; else fhx_v[XF(i, j)] = f;
            }
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i < n; ++i) {
            const double vcon = 0.5 * v[YF(i, j)];
            const double vd = vcon + fabs(vcon);
            const double vp = vcon - fabs(vcon);
            for (int c = 0; c < 2; ++c) {
                const double f = quick_face(vp, vd, vc[VC(i, j - 2, c)], vc[VC(i, j - 1, c)],
                                            vc[VC(i, j, c)], vc[VC(i, j + 1, c)]);
                if (c == 0) fhy_u[YF(i, j)] = f; else fhy_v[YF(i, j)] = f;
            }
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            fconv[C0(i, j, 0)] = (fhx_u[XF(i + 1, j)] - fhx_u[XF(i, j)]) / dx +
                                 (fhy_u[YF(i, j + 1)] - fhy_u[YF(i, j)]) / dy;
            fconv[C0(i, j, 1)] = (fhx_v[XF(i + 1, j)] - fhx_v[XF(i, j)]) / dx +
                                 (fhy_v[YF(i, j + 1)] - fhy_v[YF(i, j)]) / dy;
            ftot[G1(i, j, 0)] = ftot[G1(i, j, 0)] - fconv[C0(i, j, 0)];
            ftot[G1(i, j, 1)] = ftot[G1(i, j, 1)] - fconv[C0(i, j, 1)];
        }
}

/* ------------------------------------------------------------ viscous flux
 * Source/src1_rever/fn_flux_calc.cpp:695-766: 5-point Laplacian of velCart
 * divided by ren (:731), then fluxTotal += V (:753-765).
 */
void rc_viscous_flux(const rc_params *p, const double *vc, double *fvisc, double *ftot)
{
    const int n = p->n;
    const double dx = p->dx, dy = p->dy, ren = p->ren;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 2; ++c) {
                const double C = vc[VC(i, j, c)], S = vc[VC(i, j - 1, c)], N = vc[VC(i, j + 1, c)];
                const double W = vc[VC(i - 1, j, c)], E = vc[VC(i + 1, j, c)];
                const double vf = ((E - 2 * C + W) / (dx * dx) + (N - 2 * C + S) / (dy * dy)) / ren;
                fvisc[C0(i, j, c)] = vf;
                ftot[G1(i, j, c)] = ftot[G1(i, j, c)] + vf;
            }
}

/* ----------------------------------------------------- face gradients
 * Source/src1_rever/fn_flux_calc.cpp:825-911 (gradient_calc_approach2):
 * on every face including walls, G_x q(i,j) = (q(i,j) - q(i-1,j))/dx (:884,886),
 * G_y likewise (:894,896), for q = p (comp 0) and phi (comp 1) of userCtx.
 */
void rc_gradient(const rc_params *p, const double *ctx, double *gpx, double *gpy,
                 double *gphix, double *gphiy)
{
    const int n = p->n;
    const double dx = p->dx, dy = p->dy;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= n; ++i) {
            gpx[XF(i, j)] = (ctx[G1(i, j, 0)] - ctx[G1(i - 1, j, 0)]) / dx;
            gphix[XF(i, j)] = (ctx[G1(i, j, 1)] - ctx[G1(i - 1, j, 1)]) / dx;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i < n; ++i) {
            gpy[YF(i, j)] = (ctx[G1(i, j, 0)] - ctx[G1(i, j - 1, 0)]) / dy;
            gphiy[YF(i, j)] = (ctx[G1(i, j, 1)] - ctx[G1(i, j - 1, 1)]) / dy;
        }
}

/* ---------------------------------------------------- flux wall ghosts
 * Source/src1_rever/main.cpp:945-1014: on every domain side, with no BC-code
 * check, the normal component's ghost is minus its mirror; the other
 * component's ghosts keep the 0 from setVal (main.cpp:935).
 */
void rc_flux_wall_ghosts(const rc_params *p, double *ft)
{
    const int n = p->n;
    for (int j = -1; j <= n; ++j) ft[G1(-1, j, 0)] = -ft[G1(0, j, 0)];
    for (int j = -1; j <= n; ++j) ft[G1(n, j, 0)] = -ft[G1(n - 1, j, 0)];
    for (int i = -1; i <= n; ++i) ft[G1(i, -1, 1)] = -ft[G1(i, 0, 1)];
    for (int i = -1; i <= n; ++i) ft[G1(i, n, 1)] = -ft[G1(i, n - 1, 1)];
}

/* ------------------------------------------------------- momentum RHS
 * Source/src1_rever/fn_rhs.cpp:8-68: R_x(i,j) = -G_x p + 0.5*(T_U(i-1,j) + T_U(i,j))
 * (:45), R_y likewise with T_V (:55).
 */
void rc_momentum_rhs(const rc_params *p, const double *ft, const double *gpx, const double *gpy,
                     double *rhsx, double *rhsy)
{
    const int n = p->n;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= n; ++i)
            rhsx[XF(i, j)] = -gpx[XF(i, j)] + 0.5 * (ft[G1(i - 1, j, 0)] + ft[G1(i, j, 0)]);
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i < n; ++i)
            rhsy[YF(i, j)] = -gpy[YF(i, j)] + 0.5 * (ft[G1(i, j - 1, 1)] + ft[G1(i, j, 1)]);
}

/* ------------------------------------------------ dual-time RK4 stage
 * Source/src1_rever/momentum.cpp:9-94: rhs <- rhs - (1.5/dt)(u* - u^n)
 * + (0.5/dt)(u^n - u^{n-1}) (:70, :77); u* <- u^k + rk[s]*rhs (:72, :79);
 * then cont2cart(velCart, velStar) (:93).
 */
int rc_rk4_stage(const rc_params *p, double rk_s, double *rhsx, double *rhsy,
                 double *ustar, double *vstar, const double *ucont, const double *vcont,
                 const double *udiff, const double *vdiff, const double *uprev, const double *vprev,
                 double *vc)
{
    const int n = p->n;
    const double dt = p->dt;
    const size_t nf = (size_t)(n + 1) * (size_t)n;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)nf; ++k) {
        rhsx[k] = rhsx[k] - (1.5 / dt) * (ustar[k] - uprev[k]) + (0.5 / dt) * udiff[k];
        ustar[k] = ucont[k] + (rk_s * rhsx[k]);
        rhsy[k] = rhsy[k] - (1.5 / dt) * (vstar[k] - vprev[k]) + (0.5 / dt) * vdiff[k];
        vstar[k] = vcont[k] + (rk_s * rhsy[k]);
    }
    return rc_cont2cart(p, vc, ustar, vstar);
}

/* --------------------------------------------------- pseudo error norm
 * Source/src1_rever/utilities.cpp:421-493 (Error_Computation) with
 * StagL2Norm / StagInnerProd / SumAbsStag (:669-778): d = u^k - u*, weight
 * 1/2 on the first and last face of the nodal box along the normal (:693,
 * :714), s = sum |d*d*w|, e_c = sqrt(s)/sqrt(npts), result max_c e_c.
 * Rows are summed in parallel and combined in row order (deterministic).
 */
double rc_error_computation(const rc_params *p, const double *ucont, const double *vcont,
                            const double *ustar, const double *vstar)
{
    const int n = p->n;
    double *rowx = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *rowy = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= n; ++j) {
        double sx = 0.0, sy = 0.0;
        if (j < n)
            for (int i = 0; i <= n; ++i) {
                const double d = ucont[XF(i, j)] - ustar[XF(i, j)];
                const double w = (i > 0 && i < n) ? 1.0 : 0.5;
                sx += fabs(d * d * w);
            }
        {
            const double w = (j > 0 && j < n) ? 1.0 : 0.5;
            for (int i = 0; i < n; ++i) {
                const double d = vcont[YF(i, j)] - vstar[YF(i, j)];
                sy += fabs(d * d * w);
            }
        }
        rowx[j] = sx;
        rowy[j] = sy;
    }
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j <= n; ++j) { sx += rowx[j]; sy += rowy[j]; }
    free(rowx);
    free(rowy);
    const double npts = (double)n * (double)n;
    const double ex = sqrt(sx) / sqrt(npts);
    const double ey = sqrt(sy) / sqrt(npts);
    return (ex < ey) ? ey : ex; /* std::max(ex, ey) */
}

/* ------------------------------------------------------- Poisson RHS
 * Source/src1_rever/fn_rhs.cpp:71-125: D = (u_{i+1}-u_i)/dx + (v_{j+1}-v_j)/dy
 * (:93-101), ||D||inf (:103-104), b = (1.5/dt) D (:116), sum b (:123-124).
 */
void rc_poisson_rhs(const rc_params *p, const double *u, const double *v, double *rhs,
                    double *div_linf, double *rhs_sum)
{
    const int n = p->n;
    const double dx = p->dx, dy = p->dy, dt = p->dt;
    double *rmax = (double *)malloc(sizeof(double) * (size_t)n);
    double *rsum = (double *)malloc(sizeof(double) * (size_t)n);
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double m = 0.0, s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double d = (u[XF(i + 1, j)] - u[XF(i, j)]) / dx + (v[YF(i, j + 1)] - v[YF(i, j)]) / dy;
            if (fabs(d) > m || d != d) m = fabs(d);
            const double b = (1.5 / dt) * d;
            rhs[G1(i, j, 0)] = b;
            s += b;
        }
        rmax[j] = m;
        rsum[j] = s;
    }
    double m = 0.0, s = 0.0;
    for (int j = 0; j < n; ++j) {
        if (rmax[j] > m || rmax[j] != rmax[j]) m = rmax[j];
        s += rsum[j];
    }
    free(rmax);
    free(rsum);
    if (div_linf) *div_linf = m;
    if (rhs_sum) *rhs_sum = s;
}

/* ------------------------------------------------- userCtx ghost fill
 * Source/src1_rever/main.cpp:1078-1158 (also poisson.cpp:312-391): homogeneous
 * Neumann mirror for both components on every side, order W, E, S, N.
 */
void rc_userctx_ghosts(const rc_params *p, double *ctx)
{
    const int n = p->n;
    for (int c = 0; c < 2; ++c) {
        for (int j = -1; j <= n; ++j) ctx[G1(-1, j, c)] = ctx[G1(0, j, c)];
        for (int j = -1; j <= n; ++j) ctx[G1(n, j, c)] = ctx[G1(n - 1, j, c)];
        for (int i = -1; i <= n; ++i) ctx[G1(i, -1, c)] = ctx[G1(i, 0, c)];
        for (int i = -1; i <= n; ++i) ctx[G1(i, n, c)] = ctx[G1(i, n - 1, c)];
    }
}

/* ================================================================ Poisson
 * Operator: MLABecLaplacian with alpha=0, beta=-1, B=1 (poisson.cpp:128-153),
 * i.e. L phi = (E+W-2P)/dx^2 + (N+S-2P)/dy^2 with homogeneous Neumann
 * (foextrap -> Neumann, poisson.cpp:85-120), phi0 = 0 (:63), tol_rel 1e-10,
 * tol_abs 0 (:181-182).  The singular RHS is made solvable by removing its
 * mean; phi is returned mean-zero.
 */
typedef struct {
    int nx, ny;
    double ax, ay; /* 1/dx^2, 1/dy^2 on this level */
    double *e, *r, *t; /* correction, rhs, scratch residual */
} mg_level;

static void mg_apply(const mg_level *L, const double *x, double *out)
{
    const int nx = L->nx, ny = L->ny;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double P = x[i + (size_t)nx * j];
            const double W = i > 0 ? x[i - 1 + (size_t)nx * j] : P;
            const double E = i < nx - 1 ? x[i + 1 + (size_t)nx * j] : P;
            const double S = j > 0 ? x[i + (size_t)nx * (j - 1)] : P;
            const double N = j < ny - 1 ? x[i + (size_t)nx * (j + 1)] : P;
            out[i + (size_t)nx * j] = (E + W - 2 * P) * L->ax + (N + S - 2 * P) * L->ay;
        }
}

void rc_apply_laplacian(const rc_params *p, const double *phi, double *out)
{
    mg_level L = {p->n, p->n, 1.0 / (p->dx * p->dx), 1.0 / (p->dy * p->dy), NULL, NULL, NULL};
    set_threads(p);
    mg_apply(&L, phi, out);
}

/* red-black Gauss-Seidel sweep: colour c updates cells with (i+j)%2 == c;
 * the Neumann mirror drops the out-of-domain neighbour from both the
 * neighbour sum and the diagonal. */
static void mg_gsrb(const mg_level *L, double *x, const double *b, int color)
{
    const int nx = L->nx, ny = L->ny;
    const double ax = L->ax, ay = L->ay;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = (j + color) & 1; i < nx; i += 2) {
            double s = 0.0, d = 0.0;
            if (i > 0) { s += ax * x[i - 1 + (size_t)nx * j]; d += ax; }
            if (i < nx - 1) { s += ax * x[i + 1 + (size_t)nx * j]; d += ax; }
            if (j > 0) { s += ay * x[i + (size_t)nx * (j - 1)]; d += ay; }
            if (j < ny - 1) { s += ay * x[i + (size_t)nx * (j + 1)]; d += ay; }
            x[i + (size_t)nx * j] = (s - b[i + (size_t)nx * j]) / d;
        }
    }
}

/* Cholesky factorisation A = R^T R (upper R in place, row-major m x m) */
static int chol(double *A, int m)
{
    for (int k = 0; k < m; ++k) {
        double d = A[(size_t)k * m + k];
        for (int s = 0; s < k; ++s) d -= A[(size_t)s * m + k] * A[(size_t)s * m + k];
        if (!(d > 0)) return -1;
        d = sqrt(d);
        A[(size_t)k * m + k] = d;
        for (int j = k + 1; j < m; ++j) {
            double v = A[(size_t)k * m + j];
            for (int s = 0; s < k; ++s) v -= A[(size_t)s * m + k] * A[(size_t)s * m + j];
            A[(size_t)k * m + j] = v / d;
        }
    }
    return 0;
}

/* pinv(L) = J - (J - L)^{-1}, J = 11^T/m, for the Neumann 5-point Laplacian
 * (J - L is symmetric positive definite since -L >= 0 with null space 1). */
int rc_pinv_neumann(int nx, int ny, double dx, double dy, double *pinv)
{
    const int m = nx * ny;
    const double ax = 1.0 / (dx * dx), ay = 1.0 / (dy * dy);
    double *A = (double *)calloc((size_t)m * m, sizeof(double));
    if (!A) return -1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int r = i + nx * j;
            double d = 0.0;
            if (i > 0) { A[(size_t)r * m + r - 1] -= ax; d += ax; }
            if (i < nx - 1) { A[(size_t)r * m + r + 1] -= ax; d += ax; }
            if (j > 0) { A[(size_t)r * m + r - nx] -= ay; d += ay; }
            if (j < ny - 1) { A[(size_t)r * m + r + nx] -= ay; d += ay; }
            A[(size_t)r * m + r] += d; /* -L */
        }
    for (size_t k = 0; k < (size_t)m * m; ++k) A[k] += 1.0 / m; /* + J */
    if (chol(A, m)) { free(A); return -2; }
    /* invert via solving R^T R x = e_c column by column */
    double *col = (double *)malloc(sizeof(double) * (size_t)m);
    for (int c = 0; c < m; ++c) {
        for (int i = 0; i < m; ++i) col[i] = (i == c) ? 1.0 : 0.0;
        for (int i = 0; i < m; ++i) { /* R^T y = e */
            double v = col[i];
            for (int s = 0; s < i; ++s) v -= A[(size_t)s * m + i] * col[s];
            col[i] = v / A[(size_t)i * m + i];
        }
        for (int i = m - 1; i >= 0; --i) { /* R x = y */
            double v = col[i];
            for (int s = i + 1; s < m; ++s) v -= A[(size_t)i * m + s] * col[s];
            col[i] = v / A[(size_t)i * m + i];
        }
        for (int i = 0; i < m; ++i) pinv[(size_t)i * m + c] = 1.0 / m - col[i];
    }
    free(col);
    free(A);
    return 0;
}

static double vmean(const double *x, size_t m)
{
    double s = 0.0;
    for (size_t k = 0; k < m; ++k) s += x[k];
    return s / (double)m;
}

static double vlinf(const double *x, size_t m)
{
    double r = 0.0;
    for (size_t k = 0; k < m; ++k) {
        const double a = fabs(x[k]);
        if (a > r || a != a) r = a;
    }
    return r;
}

/* bottom levels of more cells than this are solved by the DCT (the same
 * mean-zero solution as the pseudo-inverse, without its O(m^3) set-up) */
#define MG_PINV_MAX 1024

static void dct_solve_rect(int nx, int ny, double hx, double hy, const double *b, double *x);

typedef struct {
    int nlev;
    mg_level lev[32];
    double *pinv;
    int mb;
} mg_hier;

static void mg_build(const rc_params *p, mg_hier *H)
{
    int nx = p->n, ny = p->n;
    double dx = p->dx, dy = p->dy;
    int l = 0;
    for (;;) {
        mg_level *L = &H->lev[l];
        L->nx = nx; L->ny = ny;
        L->ax = 1.0 / (dx * dx); L->ay = 1.0 / (dy * dy);
        L->e = (double *)calloc((size_t)nx * ny, sizeof(double));
        L->r = (double *)calloc((size_t)nx * ny, sizeof(double));
        L->t = (double *)calloc((size_t)nx * ny, sizeof(double));
        ++l;
        if (nx > p->mg_bottom_max && ny > p->mg_bottom_max && nx % 2 == 0 && ny % 2 == 0 && l < 32) {
            nx /= 2; ny /= 2; dx *= 2; dy *= 2;
        } else break;
    }
    H->nlev = l;
    mg_level *B = &H->lev[l - 1];
    H->mb = B->nx * B->ny;
    H->pinv = NULL;
    if (H->mb > MG_PINV_MAX) return; /* the bottom is solved by the DCT instead */
    H->pinv = (double *)malloc(sizeof(double) * (size_t)H->mb * H->mb);
    rc_pinv_neumann(B->nx, B->ny, 1.0 / sqrt(B->ax), 1.0 / sqrt(B->ay), H->pinv);
}

static void mg_free(mg_hier *H)
{
    for (int l = 0; l < H->nlev; ++l) { free(H->lev[l].e); free(H->lev[l].r); free(H->lev[l].t); }
    free(H->pinv);
}

/* one V-cycle for L e = r on level l, e starting from 0: nu1 = nu2 = 2 GSRB
 * smooths (each red then black), residual, 2x2 average restriction,
 * piecewise-constant prolongation added to e, dense pseudo-inverse at the
 * bottom. */
static void mg_vcycle(mg_hier *H, int l)
{
    mg_level *L = &H->lev[l];
    const size_t m = (size_t)L->nx * L->ny;
    if (l == H->nlev - 1) {
        const int mb = H->mb;
        if (!H->pinv) { /* an odd bottom too large for the dense pseudo-inverse: its exact DCT solve */
            dct_solve_rect(L->nx, L->ny, 1.0 / sqrt(L->ax), 1.0 / sqrt(L->ay), L->r, L->e);
            return;
        }
        for (int i = 0; i < mb; ++i) {
            double s = 0.0;
            for (int k = 0; k < mb; ++k) s += H->pinv[(size_t)i * mb + k] * L->r[k];
            L->e[i] = s;
        }
        return;
    }
    memset(L->e, 0, m * sizeof(double));
    for (int s = 0; s < 2; ++s) { mg_gsrb(L, L->e, L->r, 0); mg_gsrb(L, L->e, L->r, 1); }
    mg_apply(L, L->e, L->t);
    for (size_t k = 0; k < m; ++k) L->t[k] = L->r[k] - L->t[k];
    mg_level *Cl = &H->lev[l + 1];
#pragma omp parallel for schedule(static)
    for (int J = 0; J < Cl->ny; ++J)
        for (int I = 0; I < Cl->nx; ++I) {
            const size_t f = (size_t)(2 * I) + (size_t)L->nx * (size_t)(2 * J);
            Cl->r[I + (size_t)Cl->nx * J] = 0.25 * (L->t[f] + L->t[f + 1] + L->t[f + L->nx] + L->t[f + L->nx + 1]);
        }
    mg_vcycle(H, l + 1);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < L->ny; ++j)
        for (int i = 0; i < L->nx; ++i)
            L->e[i + (size_t)L->nx * j] += Cl->e[(i / 2) + (size_t)Cl->nx * (j / 2)];
    for (int s = 0; s < 2; ++s) { mg_gsrb(L, L->e, L->r, 0); mg_gsrb(L, L->e, L->r, 1); }
}

/* exact mean-zero solve of the nx x ny Neumann 5-point problem (cell sizes
 * hx, hy) by the DCT-II diagonalisation: eigenvalues -(4/h^2) sin^2(pi k / 2n)
 * per direction, the constant mode dropped (the pseudo-inverse's solution) */
static void dct_solve_rect(int nx, int ny, double hx, double hy, const double *b, double *x)
{
    const double pi = 3.14159265358979323846;
    double *Qx = (double *)malloc(sizeof(double) * (size_t)nx * nx);
    double *Qy = (double *)malloc(sizeof(double) * (size_t)ny * ny);
    double *T = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    double *H = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nx; ++k)
            Qx[(size_t)i * nx + k] = (k == 0 ? sqrt(1.0 / nx) : sqrt(2.0 / nx)) * cos(pi * k * (i + 0.5) / nx);
    for (int j = 0; j < ny; ++j)
        for (int l = 0; l < ny; ++l)
            Qy[(size_t)j * ny + l] = (l == 0 ? sqrt(1.0 / ny) : sqrt(2.0 / ny)) * cos(pi * l * (j + 0.5) / ny);
    /* T[j][k] = sum_i b[j][i] Qx[i][k] */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nx; ++k) {
            double s = 0.0;
            for (int i = 0; i < nx; ++i) s += b[(size_t)j * nx + i] * Qx[(size_t)i * nx + k];
            T[(size_t)j * nx + k] = s;
        }
    /* H[l][k] = sum_j Qy[j][l] T[j][k] / (lx[k] + ly[l]) */
#pragma omp parallel for schedule(static)
    for (int l = 0; l < ny; ++l)
        for (int k = 0; k < nx; ++k) {
            double s = 0.0;
            for (int j = 0; j < ny; ++j) s += Qy[(size_t)j * ny + l] * T[(size_t)j * nx + k];
            const double sx = sin(pi * k / (2.0 * nx)), sy = sin(pi * l / (2.0 * ny));
            const double lam = -4.0 * sx * sx / (hx * hx) - 4.0 * sy * sy / (hy * hy);
            H[(size_t)l * nx + k] = (k == 0 && l == 0) ? 0.0 : s / lam;
        }
    /* T[l][i] = sum_k H[l][k] Qx[i][k] */
#pragma omp parallel for schedule(static)
    for (int l = 0; l < ny; ++l)
        for (int i = 0; i < nx; ++i) {
            double s = 0.0;
            for (int k = 0; k < nx; ++k) s += H[(size_t)l * nx + k] * Qx[(size_t)i * nx + k];
            T[(size_t)l * nx + i] = s;
        }
    /* x[j][i] = sum_l Qy[j][l] T[l][i] */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double s = 0.0;
            for (int l = 0; l < ny; ++l) s += Qy[(size_t)j * ny + l] * T[(size_t)l * nx + i];
            x[(size_t)j * nx + i] = s;
        }
    free(Qx); free(Qy); free(T); free(H);
}

static void dct_solve(const rc_params *p, const double *b, double *x)
{
    dct_solve_rect(p->n, p->n, p->dx, p->dy, b, x);
}

/* Source/src1_rever/poisson.cpp:41-186 (poisson_advance): sol and rhs carry
 * one ghost (AMReX layout); the ghosts of sol are left 0 as by setVal (:63). */
int rc_poisson_solve(const rc_params *p, double *sol, double *rhs, int *iters, double *res_rel)
{
    const int n = p->n;
    const size_t m = (size_t)n * n;
    set_threads(p);
    double *b = (double *)malloc(sizeof(double) * m);
    double *x = (double *)calloc(m, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) b[i + (size_t)n * j] = rhs[G1(i, j, 0)];
    const double mb = vmean(b, m);
    for (size_t k = 0; k < m; ++k) b[k] -= mb;
    const double bnorm = vlinf(b, m);
    int it = 0;
    double rr = 0.0;
    if (p->poisson_kind == 1) {
        dct_solve(p, b, x);
    } else {
        mg_hier H;
        memset(&H, 0, sizeof(H));
        mg_build(p, &H);
        mg_level *F = &H.lev[0];
        const int maxit = p->mg_max_iter > 0 ? p->mg_max_iter : 200;
        for (;;) {
            mg_apply(F, x, F->t);
            for (size_t k = 0; k < m; ++k) F->r[k] = b[k] - F->t[k];
            rr = vlinf(F->r, m);
            if (rr <= p->poisson_tol_rel * bnorm || it >= maxit || rr != rr) break;
            mg_vcycle(&H, 0);
            for (size_t k = 0; k < m; ++k) x[k] += F->e[k];
            ++it;
        }
        mg_free(&H);
    }
    const double mx = vmean(x, m);
    for (size_t k = 0; k < m; ++k) x[k] -= mx;
    {
        mg_level F = {n, n, 1.0 / (p->dx * p->dx), 1.0 / (p->dy * p->dy), NULL, NULL, NULL};
        double *t = (double *)malloc(sizeof(double) * m);
        mg_apply(&F, x, t);
        rr = 0.0;
        for (size_t k = 0; k < m; ++k) {
            const double a = fabs(b[k] - t[k]);
            if (a > rr) rr = a;
        }
        free(t);
    }
    for (size_t k = 0; k < (size_t)(n + 2) * (n + 2); ++k) sol[k] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) sol[G1(i, j, 0)] = x[i + (size_t)n * j];
    if (iters) *iters = it;
    if (res_rel) *res_rel = bnorm > 0 ? rr / bnorm : rr;
    free(b);
    free(x);
    return 0;
}

/* --------------------------------------------------- update_solution
 * Source/src1_rever/poisson.cpp:192-392: face gradients (:216), u = u* -
 * G phi * dt / 1.5 (:251, :256), cont2cart (:267), p <- p + phi (:309),
 * Neumann ghosts of userCtx (:312-391).
 */
int rc_update_solution(const rc_params *p, double *ctx, double *vc, double *ucont, double *vcont,
                       const double *ustar, const double *vstar)
{
    const int n = p->n;
    const double dx = p->dx, dy = p->dy, dt = p->dt;
    set_threads(p);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= n; ++i) {
            const double g = (ctx[G1(i, j, 1)] - ctx[G1(i - 1, j, 1)]) / dx;
            ucont[XF(i, j)] = ustar[XF(i, j)] - (g * dt / 1.5);
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i < n; ++i) {
            const double g = (ctx[G1(i, j, 1)] - ctx[G1(i, j - 1, 1)]) / dy;
            vcont[YF(i, j)] = vstar[YF(i, j)] - (g * dt / 1.5);
        }
    const int rc = rc_cont2cart(p, vc, ucont, vcont);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) ctx[G1(i, j, 0)] = ctx[G1(i, j, 0)] + ctx[G1(i, j, 1)];
    rc_userctx_ghosts(p, ctx);
    return rc;
}

/* ------------------------------------------------------------ advance
 * Source/src1_rever/main.cpp:875-1202: BDF2 history (879-884), RK coefficients
 * rk = dtau*{1/4,1/3,1/2,1}, dtau = 0.4 dt (860-870), pseudo loop while
 * normError > momentum_tolerance (898) with the stage sequence of 929-1017,
 * velCont <- velStar (1023-1026), caps (1030-1039), Poisson (1058-1060), phi
 * into userCtx with ghosts (1078-1158), update_solution (1164), divergence
 * check (1180).
 */
int rc_advance(const rc_params *p, double *ucont, double *vcont, double *uprev, double *vprev,
               double *ctx, double *vc, int nsteps, rc_step_stats *stats)
{
    const int n = p->n;
    const size_t nf = (size_t)(n + 1) * n;
    const size_t ng1 = (size_t)(n + 2) * (n + 2);
    double *udiff = malloc(nf * 8), *vdiff = malloc(nf * 8), *ustar = malloc(nf * 8), *vstar = malloc(nf * 8);
    double *rhsx = malloc(nf * 8), *rhsy = malloc(nf * 8);
    double *fxu = malloc(nf * 8), *fxv = malloc(nf * 8), *fyu = malloc(nf * 8), *fyv = malloc(nf * 8);
    double *gpx = malloc(nf * 8), *gpy = malloc(nf * 8), *gphix = malloc(nf * 8), *gphiy = malloc(nf * 8);
    double *fconv = malloc((size_t)n * n * 2 * 8), *fvisc = malloc((size_t)n * n * 2 * 8);
    double *ftot = malloc(ng1 * 2 * 8), *prhs = calloc(ng1, 8), *psol = calloc(ng1, 8);
    const double dtau = 0.4 * p->dt;
    double rk[4];
    rk[0] = dtau * 0.25;
    rk[1] = dtau * (1.0 / 3.0);
    rk[2] = dtau * 0.5;
    rk[3] = dtau * 1.0;
    int rc = rc_cont2cart(p, vc, ucont, vcont);
    for (int step = 0; step < nsteps && rc == 0; ++step) {
        rc_step_stats st;
        memset(&st, 0, sizeof(st));
        for (size_t k = 0; k < nf; ++k) {
            udiff[k] = ucont[k] - uprev[k]; vdiff[k] = vcont[k] - vprev[k];
            uprev[k] = ucont[k]; vprev[k] = vcont[k];
            ustar[k] = ucont[k]; vstar[k] = vcont[k];
        }
        int count = 0;
        double normError = 1.e9;
        if (!p->pseudo) {
            /* PSEUDO_TIMESTEPPING = 0 (main.cpp:900-926): one explicit update
             * from velCart = cont2cart(u^n); fluxTotal keeps the 0 of setVal in
             * its domain ghosts (no wall-ghost pass on this branch), then
             * explicit_time_marching (momentum.cpp:96-191): rhs += diff/(2 dt),
             * u* += (dt/1.5) rhs; countIter stays 0 (the branch breaks out of
             * the loop before it is counted). */
            rc_gradient(p, ctx, gpx, gpy, gphix, gphiy);
            memset(ftot, 0, ng1 * 2 * 8);
            rc_convective_flux(p, vc, ustar, vstar, fxu, fxv, fyu, fyv, fconv, ftot);
            rc_viscous_flux(p, vc, fvisc, ftot);
            rc_momentum_rhs(p, ftot, gpx, gpy, rhsx, rhsy);
            for (size_t k = 0; k < nf; ++k) {
                rhsx[k] = rhsx[k] + (1.0 / (2.0 * p->dt)) * udiff[k];
                rhsy[k] = rhsy[k] + (1.0 / (2.0 * p->dt)) * vdiff[k];
                ustar[k] = ustar[k] + ((p->dt / 1.5) * rhsx[k]);
                vstar[k] = vstar[k] + ((p->dt / 1.5) * rhsy[k]);
            }
            normError = rc_error_computation(p, ucont, vcont, ustar, vstar);
            memcpy(ucont, ustar, nf * 8);
            memcpy(vcont, vstar, nf * 8);
        }
        while (p->pseudo && normError > p->mom_tol) {
            for (int sub = 0; sub < 4; ++sub) {
                rc_gradient(p, ctx, gpx, gpy, gphix, gphiy);
                memset(ftot, 0, ng1 * 2 * 8);
                rc_convective_flux(p, vc, ustar, vstar, fxu, fxv, fyu, fyv, fconv, ftot);
                rc_viscous_flux(p, vc, fvisc, ftot);
                rc_flux_wall_ghosts(p, ftot);
                rc_momentum_rhs(p, ftot, gpx, gpy, rhsx, rhsy);
                rc = rc_rk4_stage(p, rk[sub], rhsx, rhsy, ustar, vstar, ucont, vcont, udiff, vdiff,
                                  uprev, vprev, vc);
                if (rc) break;
            }
            if (rc) break;
            normError = rc_error_computation(p, ucont, vcont, ustar, vstar);
            memcpy(ucont, ustar, nf * 8);
            memcpy(vcont, vstar, nf * 8);
            ++count;
            if (count > p->iter_num) { st.flags |= 1; break; }
            if (normError > 1.e2) { st.flags |= 2; break; }
        }
        if (rc) break;
        if (!(normError == normError) || isinf(normError)) st.flags |= 4;
        st.pseudo_iters = count;
        st.norm_error = normError;
        rc_poisson_rhs(p, ustar, vstar, prhs, &st.div_star, &st.rhs_sum);
        rc_poisson_solve(p, psol, prhs, &st.vcycles, &st.poisson_res);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) ctx[G1(i, j, 1)] = psol[G1(i, j, 0)];
        rc_userctx_ghosts(p, ctx);
        rc = rc_update_solution(p, ctx, vc, ucont, vcont, ustar, vstar);
        rc_poisson_rhs(p, ucont, vcont, prhs, &st.div_new, NULL);
        if (stats) stats[step] = st;
    }
    free(udiff); free(vdiff); free(ustar); free(vstar); free(rhsx); free(rhsy);
    free(fxu); free(fxv); free(fyu); free(fyv); free(gpx); free(gpy); free(gphix); free(gphiy);
    free(fconv); free(fvisc); free(ftot); free(prhs); free(psol);
    return rc;
}
