/*
 * refcpu.h -- CPU FP64 restatement of AMRESSIF's fractional-step time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or as the timed CPU baseline.  The
 * product library (amressif_amd/libamressif_hip.so) never links or calls it.
 *
 * Layout: every field is one AMReX-style box in Array4 order (i fastest,
 * then j, then component), ghosts included, exactly as the reference
 * allocates it for a single-box 2D run (Source/src1_rever/main.cpp:330-383):
 *   x-face array  (velCont[0] ...): (n+1) x n,  0 ghosts, idx = i + (n+1)*j
 *   y-face array  (velCont[1] ...): n x (n+1),  0 ghosts, idx = i + n*j
 *   velCart   : 2 comps, 2 ghosts, (n+4)^2 per comp, idx = (i+2)+(n+4)*(j+2)+c*(n+4)^2
 *   userCtx   : 2 comps (p, phi), 1 ghost, (n+2)^2 per comp
 *   fluxTotal : 2 comps, 1 ghost
 *   poisson rhs / sol : 1 comp, 1 ghost
 */
#ifndef AMRESSIF_REFCPU_H
#define AMRESSIF_REFCPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;                  /* n_cell: cells per side, one box */
    double dx, dy;          /* cell sizes */
    int bc_lo[2], bc_hi[2]; /* phy_bc_lo / phy_bc_hi codes */
    double inflow[2];       /* inflow_waveform */
    const double *lid;      /* optional north-lid profile (n values, replaces inflow[0]); NULL = uniform */
    double ren;             /* Reynolds number */
    double dt;              /* physical time step */
    double mom_tol;         /* momentum_tolerance */
    int iter_num;           /* IterNum */
    int poisson_kind;       /* 0 = geometric MG V-cycles, 1 = exact DCT-II solve */
    double poisson_tol_rel; /* MG stop: ||r||inf <= tol_rel * ||b||inf */
    int mg_max_iter;        /* MG iteration cap */
    int mg_bottom_max;      /* coarsen while n_level > mg_bottom_max and even */
    int nthreads;           /* OpenMP threads (<=0: runtime default) */
    int pseudo;             /* PSEUDO_TIMESTEPPING: 1 = RK4 dual time (default), 0 = one explicit update */
} rc_params;

typedef struct {
    int pseudo_iters;   /* countIter at loop exit */
    int vcycles;        /* MG V-cycles (0 for DCT) */
    int flags;          /* bit0 cap hit, bit1 diverged (>1e2), bit2 non-finite norm */
    int pad;
    double norm_error;  /* last pseudo-iteration norm */
    double div_star;    /* ||D(u*)||inf before projection */
    double rhs_sum;     /* sum of Poisson RHS */
    double poisson_res; /* final ||r||inf / ||b||inf */
    double div_new;     /* ||D(u^{n+1})||inf after projection */
} rc_step_stats;

/* ---- module-level functions, one per reference function ---- */
int  rc_cont2cart(const rc_params *p, double *velCart, const double *u, const double *v);
void rc_convective_flux(const rc_params *p, const double *velCart, const double *u, const double *v,
                        double *fhx_u, double *fhx_v, double *fhy_u, double *fhy_v,
                        double *fluxConvect, double *fluxTotal);
void rc_viscous_flux(const rc_params *p, const double *velCart, double *fluxViscous, double *fluxTotal);
void rc_gradient(const rc_params *p, const double *userCtx, double *gpx, double *gpy,
                 double *gphix, double *gphiy);
void rc_flux_wall_ghosts(const rc_params *p, double *fluxTotal);
void rc_momentum_rhs(const rc_params *p, const double *fluxTotal, const double *gpx, const double *gpy,
                     double *rhsx, double *rhsy);
int  rc_rk4_stage(const rc_params *p, double rk_s, double *rhsx, double *rhsy,
                  double *ustar, double *vstar, const double *ucont, const double *vcont,
                  const double *udiff, const double *vdiff, const double *uprev, const double *vprev,
                  double *velCart);
double rc_error_computation(const rc_params *p, const double *ucont, const double *vcont,
                            const double *ustar, const double *vstar);
void rc_poisson_rhs(const rc_params *p, const double *u, const double *v, double *rhs,
                    double *div_linf, double *rhs_sum);
int  rc_poisson_solve(const rc_params *p, double *sol, double *rhs, int *iters, double *res_rel);
void rc_userctx_ghosts(const rc_params *p, double *userCtx);
int  rc_update_solution(const rc_params *p, double *userCtx, double *velCart,
                        double *ucont, double *vcont, const double *ustar, const double *vstar);

/* the whole per-step advance (main.cpp time loop), nsteps steps */
int  rc_advance(const rc_params *p, double *ucont, double *vcont, double *uprev, double *vprev,
                double *userCtx, double *velCart, int nsteps, rc_step_stats *stats);

/* Poisson helpers used by tests */
void rc_apply_laplacian(const rc_params *p, const double *phi_valid, double *out_valid);
int  rc_pinv_neumann(int nx, int ny, double dx, double dy, double *pinv /* (nx*ny)^2 */);

#ifdef __cplusplus
}
#endif
#endif
