/*
 * refgen.h -- CPU FP64 restatement of AMRESSIF's fractional-step time step in
 * 2D and 3D with periodic directions (phy_bc 111) beside the wall codes.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle of the library's general
 * path (2D with a periodic direction, and 3D).  Only tests/ may load it, as
 * the checker.  The product library never links or calls it.
 *
 * Layout: one box of n^dim cells (the reference's n_cell in every direction,
 * Source/src1_rever/main.cpp:300-312), AMReX Array4 order (i fastest, then j,
 * then k, then component), ghosts included:
 *   face array of direction d: n+1 along d, n along the others, 0 ghosts
 *   velCart  : dim comps, 2 ghosts   ((n+4)^dim per comp)
 *   userCtx  : 2 comps (p, phi), 1 ghost
 *   fluxTotal: dim comps, 1 ghost
 *   Poisson rhs / sol: 1 comp, 1 ghost
 * In 2D the k extent is 1 (k = 0 only).
 *
 * Periodic semantics.  The reference's periodic path overwrites its periodic
 * ghosts: the flux wall-ghost pass and the phi/p ghost pass run on every side
 * whatever the BC code (main.cpp:945-1014, 1078-1158).  This restatement
 * defines the periodic step instead: every ghost of a periodic direction
 * (velCart, fluxTotal, userCtx, the Poisson operator) is the periodic image
 * (FillBoundary with the domain's periodicity), and the wall passes run only
 * on wall sides.  Face n of a periodic direction is face 0: both are stored
 * (AMReX's nodal layout) and both take the same value.
 */
#ifndef AMRESSIF_REFGEN_H
#define AMRESSIF_REFGEN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim;                /* 2 or 3 */
    int n;                  /* cells per side */
    double dx[3];           /* cell sizes */
    int bc_lo[3], bc_hi[3]; /* phy_bc_lo / phy_bc_hi: 111 periodic (both sides), 135, -135, 165; east slip is 1 */
    double inflow[3];       /* inflow_waveform (the north 165 side) */
    double ren;
    double dt;
    double mom_tol;
    int iter_num;
    double poisson_tol_rel; /* ||r||inf <= tol_rel ||b||inf */
    int mg_max_iter;
    int nthreads;           /* OpenMP threads (<= 0: runtime default) */
    int pseudo;             /* PSEUDO_TIMESTEPPING: 1 RK4 dual time, 0 one explicit update */
} rg_params;

typedef struct {
    int pseudo_iters, vcycles, flags, pad;
    double norm_error, div_star, rhs_sum, poisson_res, div_new;
} rg_step_stats;

/* sizes (doubles) of the arrays above */
long rg_face_size(const rg_params *p, int d);
long rg_cell_size(const rg_params *p, int ghosts);

/* ---- module-level functions, one per reference function ---- */
int  rg_cont2cart(const rg_params *p, double *velCart, double *const *faces);
void rg_convective_flux(const rg_params *p, const double *velCart, double *const *faces, double *const *fh,
                        double *fluxConvect, double *fluxTotal);
void rg_viscous_flux(const rg_params *p, const double *velCart, double *fluxViscous, double *fluxTotal);
void rg_gradient(const rg_params *p, const double *userCtx, int comp, double *const *grad);
void rg_flux_ghosts(const rg_params *p, double *fluxTotal);
void rg_momentum_rhs(const rg_params *p, const double *fluxTotal, double *const *gradp, double *const *rhs);
double rg_error_computation(const rg_params *p, double *const *ucont, double *const *ustar);
void rg_poisson_rhs(const rg_params *p, double *const *faces, double *rhs, double *div_linf, double *rhs_sum);
int  rg_poisson_solve(const rg_params *p, double *sol, const double *rhs, int *iters, double *res_rel);
void rg_userctx_ghosts(const rg_params *p, double *userCtx);
void rg_apply_laplacian(const rg_params *p, const double *phi_valid, double *out_valid);
int  rg_update_solution(const rg_params *p, double *userCtx, double *velCart, double *const *ucont,
                        double *const *ustar);

/* the whole per-step advance, nsteps steps */
int rg_advance(const rg_params *p, double *const *ucont, double *const *uprev, double *userCtx, double *velCart,
               int nsteps, rg_step_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
