/*
 * amressif.h -- C-ABI of the MI355X (gfx950) implementation of AMRESSIF's
 * fractional-step time step (reference: milk-white-way/AMRESSIF,
 * Source/src1_rever).  Plain C11, no exceptions cross the ABI, no torch types.
 *
 * The reference has no advance function: its step is inline in main_main
 * (Source/src1_rever/main.cpp:875-1202) and its module functions take AMReX
 * MultiFabs.  This header replaces those module functions and the step loop
 * with entry points on AMReX Array4 views (device pointers, ghosts included,
 * i fastest, then j, then component): one box per view, and several boxes
 * per MultiFab at the state load/store (amr_multifab).
 *
 * Conventions
 *  - The caller owns every field array (device memory).  A context owns its
 *    working set (internal pitched copies, multigrid hierarchy, reductions).
 *  - Entry points are ordered on the context's HIP stream.  Those that return
 *    host scalars or statistics synchronise that stream.
 *  - One context per stream/thread; no global state.
 *  - Status 0 is success.  Negative codes replace the reference's amrex::Abort
 *    and add explicit failure reports (non-finite norm, unsupported input).
 */
#ifndef AMRESSIF_H
#define AMRESSIF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMRESSIF_OK 0
#define AMRESSIF_ERR_ARG (-1)         /* bad argument / shape mismatch */
#define AMRESSIF_ERR_BC (-2)          /* BC code the reference aborts on (main.cpp:822,832; utilities.cpp:128) */
#define AMRESSIF_ERR_UNSUPPORTED (-3) /* valid in the reference but not built here (see the entry point) */
#define AMRESSIF_ERR_HIP (-4)         /* HIP runtime failure */
#define AMRESSIF_ERR_NONFINITE (-5)   /* non-finite pseudo-iteration norm (reference exits silently) */
#define AMRESSIF_ERR_PARSE (-6)       /* inputs file error */

/* precision modes */
#define AMRESSIF_FP64 0
#define AMRESSIF_MIXED 1 /* FP32 stencils + FP32 V-cycle inside FP64 iterative refinement */

/* step flags (amressif_step_stats.flags) */
#define AMRESSIF_FLAG_CAP 1       /* countIter > IterNum (main.cpp:1030-1034) */
#define AMRESSIF_FLAG_DIVERGED 2  /* normError > 1e2 (main.cpp:1035-1039) */
#define AMRESSIF_FLAG_NONFINITE 4 /* NaN/Inf norm: the reference's while test exits silently (main.cpp:898) */
#define AMRESSIF_FLAG_MG_CAP 8    /* Poisson iteration cap reached before tolerance */
#define AMRESSIF_FLAG_LIN_CAP 16  /* tangent/adjoint momentum solve reached its sweep cap */

/* The reference's ParmParse keys (Source/src1_rever/main.cpp:89-157) plus the
 * solver settings it hard-codes (poisson.cpp:159-182) and this library's
 * precision and solver switches.  Defaults are the reference's. */
typedef struct {
    int n_cell;                 /* required */
    int max_grid_size[3];       /* required (x, y, z) */
    double lo_phy_dim[3];       /* default 0 */
    double hi_phy_dim[3];       /* default 1 */
    int IterNum;                /* required */
    int nsteps;                 /* default 1 */
    double cfl;                 /* default 0.9 */
    double fixed_dt;            /* default -1 (use cfl) */
    double ren;                 /* required */
    double vis;                 /* required, unused by the reference's arithmetic */
    int phy_bc_lo[3];           /* default 0 (invalid) */
    int phy_bc_hi[3];
    double inflow_waveform[3];  /* default 0 */
    int target_resolution;      /* default -1 */
    double momentum_tolerance;  /* default 1e-10 */
    int PSEUDO_TIMESTEPPING;    /* default 1 */
    int plot_int;               /* default -1 */
    int chk_int;                /* default -1 */
    int txt_int;                /* default 0 */
    int chk_out;                /* default 0 */
    /* solver settings the reference hard-codes in poisson_advance */
    double poisson_tol_rel;     /* 1e-10 (poisson.cpp:181) */
    double poisson_tol_abs;     /* 0 (poisson.cpp:182) */
    int mg_max_iter;            /* 2000 (poisson.cpp:159) */
    int mg_bottom_max;          /* coarsen while the level is larger than this and even; 4 */
    /* this library's own settings */
    int precision;              /* AMRESSIF_FP64 or AMRESSIF_MIXED */
    int poisson_solver;         /* 0: multigrid V-cycle iteration (the MLMG solve, default);
                                   1: FGMRES(gmres_restart) right-preconditioned by one V-cycle
                                   (the solver GMRES_Poisson.H:14-76 sets up; same operator) */
    int gmres_restart;          /* restart length, 1..64; default 30 */
    int mixed_momentum;         /* precision = 1 only: the momentum solve of the implicit BDF2 system.
                                   1 (default): FP64 iterative refinement, each correction the FP32
                                   two-term Neumann series -(1/c)(I + R'/c) of the inverse Jacobian,
                                   c = 1.5/dt; 0: the RK4 dual-time recursion to first order in
                                   the Jacobian (the reference's pseudo-time structure) */
    int poisson_warm_start;     /* the step's Poisson solve starts from the linear extrapolation
                                   2 phi^{n-1} - phi^{n-2} of the previous steps' solutions (phi^{n-1}
                                   alone right after a state load) instead of 0 (poisson.cpp:63);
                                   the stopping test stays ||r||inf <= tol_rel ||b||inf.
                                   -1 (default): on for precision = 1, off for FP64; 0 off; 1 on */
    int spacedim;               /* AMREX_SPACEDIM (Exec/GNUmakefile:10, a compile-time DIM in the reference):
                                   2 (default) or 3; the domain is n_cell^spacedim cells */
} amressif_inputs;

/* AMReX Array4 view of one box: ptr addresses element (lo[0], lo[1], lo[2], comp 0);
 * strides are those of a contiguous FAB, jstride = hi0-lo0+1, kstride =
 * jstride*(hi1-lo1+1), nstride = kstride*(hi2-lo2+1). */
typedef struct {
    void *ptr;
    int lo[3];
    int hi[3];
    int ncomp;
    int dtype; /* 0 = double, 1 = float */
} amr_array4;

typedef struct {
    int pseudo_iters;   /* RK4 pseudo-iterations (countIter) */
    int vcycles;        /* Poisson V-cycles */
    int flags;          /* AMRESSIF_FLAG_* */
    int pad;
    double norm_error;  /* last pseudo-iteration norm (Error_Computation) */
    double div_star;    /* ||D(u*)||inf (fn_rhs.cpp:103-104) */
    double rhs_sum;     /* sum of the Poisson RHS (fn_rhs.cpp:123-124) */
    double poisson_res; /* final ||r||inf / ||b||inf */
    double div_new;     /* ||D(u^{n+1})||inf (main.cpp:1180) */
} amressif_step_stats;

typedef struct amressif_ctx amressif_ctx;

/* ---------------------------------------------------------------- inputs */
/* Fill the reference's defaults (main.cpp:89-157; required keys set to -1). */
void amressif_inputs_default(amressif_inputs *in);
/* ParmParse-style reader: "key = v1 v2 ..." lines, '#' comments, then
 * "key=value" overrides from argv (replaces amrex::Initialize + ParmParse,
 * main.cpp:89-157).  Returns AMRESSIF_ERR_PARSE with a message in err. */
int amressif_parse_inputs(const char *path, int argc, const char *const *argv, amressif_inputs *out,
                          char *err, int errlen);
/* dt = cfl/(2 sum 1/dx^2) or fixed_dt (main.cpp:284-294) */
double amressif_dt(const amressif_inputs *in);

/* --------------------------------------------------------------- context */
/* Validates the BC codes (main.cpp:810-835) and allocates the working set.
 * stream: a hipStream_t (NULL = default stream). */
int amressif_create(const amressif_inputs *in, void *hip_stream, amressif_ctx **out);
int amressif_destroy(amressif_ctx *ctx);
const char *amressif_last_error(const amressif_ctx *ctx);
/* Device bytes held by the context. */
int64_t amressif_workspace_bytes(const amressif_ctx *ctx);
/* Per-column north-lid velocity (n_cell doubles, device pointer; NULL resets
 * to the uniform inflow_waveform[0] of utilities.cpp:251-261). */
int amressif_set_lid_profile(amressif_ctx *ctx, const double *dev_lid);

/* ------------------------------------------------- state and time step */
/* Copy the caller's state into the context: velCont[2] and velContPrev[2]
 * (faces, 0 ghosts), userCtx (2 comps p, phi; >= 0 ghosts).  On a 3D context
 * (spacedim = 3) every velCont / velContPrev argument of this section is an
 * array of 3 entries (x, y, z faces), here and in the _mf forms. */
int amressif_load_state(amressif_ctx *ctx, const amr_array4 velCont[2], const amr_array4 velContPrev[2],
                        const amr_array4 *userCtx);
/* Copy the state out.  Any pointer may be NULL.  velCart (2 comps, 2 ghosts)
 * is cont2cart(velCont) with the wall ghost rules; userCtx ghosts get the
 * Neumann mirror (main.cpp:1078-1158). */
int amressif_store_state(const amressif_ctx *ctx, amr_array4 velCont[2], amr_array4 velContPrev[2],
                         amr_array4 *userCtx, amr_array4 *velCart);
/* A MultiFab of several boxes: one Array4 per box (what MultiFab::array(mfi)
 * gives for each MFIter box), each grown by ngrow cells around its valid box.
 * The valid cell boxes must tile the domain (BoxArray::maxSize chopping,
 * main.cpp:317-324).  Face MultiFabs are nodal in their direction
 * (surroundingNodes, main.cpp:333-341): a box's valid faces include both
 * bounding faces, so neighbouring boxes share one face column or row. */
typedef struct {
    int nbox;
    const amr_array4 *fab; /* nbox views (for a store the memory is written through them) */
    int ngrow;
} amr_multifab;

/* amressif_load_state over several boxes: each box's valid region is
 * gathered into the context's one-box working set (a shared face is read
 * from the later box; a consistent MultiFab holds the same value in both).
 * userCtx may be NULL (p = phi = 0).  2D and 3D, walls and periodic
 * directions; boxes are validated before the first copy, so a refused load
 * leaves the state as it was. */
int amressif_load_state_mf(amressif_ctx *ctx, const amr_multifab velCont[2], const amr_multifab velContPrev[2],
                           const amr_multifab *userCtx);
/* amressif_store_state over several boxes: every element of every view is
 * written.  Face MultiFabs have ngrow = 0 (main.cpp:337-338); userCtx
 * ngrow <= 1 and velCart (cont2cart with the wall rules) ngrow <= 2: their
 * ghosts inside the domain get the neighbouring box's values
 * (FillBoundary), ghosts outside it the reference's wall rules, as
 * amressif_store_state writes them.  Any pointer may be NULL. */
int amressif_store_state_mf(const amressif_ctx *ctx, amr_multifab velCont[2], amr_multifab velContPrev[2],
                            amr_multifab *userCtx, amr_multifab *velCart);
/* Advance nsteps time steps (main.cpp:875-1202 per step).  stats: host array
 * of nsteps entries or NULL.  Synchronises the stream. */
int amressif_advance(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats);

/* ------------------------------------------------ initial state and files */
/* hybrid_grid_init (fn_init.cpp:30-631, kn_init.H:8-37) for the cavity: all
 * faces 0, velContPrev = velCont, p = phi = 0.  With psi_coeffs (host array
 * a[kx][ky], kmax x kmax, row-major) the faces instead come from the nodal
 * streamfunction psi = sum a[kx][ky] sin(kx pi x) sin(ky pi y) sin^2(pi x)
 * sin^2(pi y) (x = i/n, y = j/n, 0 on the boundary nodes):
 * u = (psi(i,j+1) - psi(i,j))/dy, v = -(psi(i+1,j) - psi(i,j))/dx, which is
 * discretely divergence-free with zero wall-normal faces (no reference
 * counterpart: the reference only starts from rest).  Synchronises. */
int amressif_init_state(amressif_ctx *ctx, const double *psi_coeffs, int kmax);
/* SaveCheckpoint (utilities.cpp:792-900): writes <dir>/checkpointNNNNN/Header
 * (title, time, BoxArray of the inputs' max_grid_size, main.cpp:317-320) and
 * Level_0/{pressure, vel_xCont, vel_yCont, vel_xContPrev, vel_yContPrev} as
 * AMReX VisMF MultiFabs (_H header + _D_00000 data).  NNNNN = step. */
int amressif_write_checkpoint(amressif_ctx *ctx, const char *dir, int step, double time);
/* LoadCheckpoint (utilities.cpp:904-1022) and the restart re-derivation
 * (main.cpp:386-800): reads the checkpoint directory `path` into the state
 * (velCont, velContPrev, p; phi, which is not saved, = 0) and returns its time.
 * Any BoxArray that tiles the domain is accepted. */
int amressif_read_checkpoint(amressif_ctx *ctx, const char *path, double *time);
/* Export_Flow_Field (utilities.cpp:534-570): WriteSingleLevelPlotfile of
 * (pressure, U, V) into <dir>/<name>NNNNN (HyperCLaw-V1.1 Header and
 * Level_0/Cell VisMF data).  U, V are the cell averages of the faces. */
int amressif_write_plotfile(amressif_ctx *ctx, const char *dir, const char *name, int step, double time);
/* Export_Fluxes (utilities.cpp:496-532, called every plot_int steps at
 * main.cpp:1045-1048): <dir>/pltFluxNNNNN with conv_flux_x, conv_flux_y
 * (fluxConvect), visc_flux_x, visc_flux_y (fluxViscous) and press_grad_x,
 * press_grad_y (fluxPrsGrad, which the reference never fills: 0).  The fluxes
 * are those of the last step's converged predictor u*; the reference exports
 * those of its last RK stage's input, which differs from u* by that stage's
 * increment (below momentum_tolerance).  AMRESSIF_ERR_ARG if no step has run
 * since the state was set. */
int amressif_write_flux_plotfile(amressif_ctx *ctx, const char *dir, int step, double time);

/* ------------------------------------------- module-level entry points
 * Each runs the same kernels as amressif_advance on one call's inputs (Array4
 * device views, copied into and out of the context's working layout). */

/* cont2cart (utilities.cpp:16-350): face -> cell average + wall-BC ghost fill */
int amressif_cont2cart(amressif_ctx *ctx, const amr_array4 velCont[2], amr_array4 *velCart);
/* One stage's momentum RHS before the dual-time terms: gradient_calc_approach2
 * (fn_flux_calc.cpp:825), convective_flux_calc_new_quick (:11), viscous_flux_calc
 * (:695), flux wall ghosts (main.cpp:945-1014), momentum_righthand_side_calc
 * (fn_rhs.cpp:8) -- one fused kernel.  rhs: 2 face arrays. */
int amressif_momentum_rhs(amressif_ctx *ctx, const amr_array4 velStar[2], const amr_array4 *userCtx,
                          amr_array4 rhs[2]);
/* One RK4 dual-time stage (momentum.cpp:9-94 together with the above):
 * velStar <- velCont + rk[sub] * (R(velStar) - (1.5/dt)(velStar - velContPrev)
 * + (0.5/dt) velContDiff), where velContDiff = velContPrev - velContPrevPrev. */
int amressif_rk_stage(amressif_ctx *ctx, int sub, amr_array4 velStar[2], const amr_array4 velCont[2],
                      const amr_array4 velContPrev[2], const amr_array4 velContPrevPrev[2],
                      const amr_array4 *userCtx);
/* The flux fields of convective_flux_calc_new_quick (fn_flux_calc.cpp:11-202)
 * and viscous_flux_calc (:695-766) for velStar: fluxConvect and fluxViscous
 * (cells, 2 comps), and the half-node fluxes fluxHalfN1 (of U) and
 * fluxHalfN2 (of V) on the x- and y-faces, from cont2cart's cell velocities
 * with the wall rules.  Any output may be NULL (the half-node pair together).
 * Evaluated in the reference's operation order (bit-exact with it). */
int amressif_fluxes(amressif_ctx *ctx, const amr_array4 velStar[2], amr_array4 *fluxConvect, amr_array4 *fluxViscous,
                    amr_array4 fluxHalfN1[2], amr_array4 fluxHalfN2[2]);
/* Error_Computation (utilities.cpp:421-493): weighted RMS of velCont - velStar */
int amressif_pseudo_norm(amressif_ctx *ctx, const amr_array4 velCont[2], const amr_array4 velStar[2],
                         double *norm);
/* poisson_righthand_side_calc (fn_rhs.cpp:71-125): rhs = (1.5/dt) D(vel) */
int amressif_poisson_rhs(amressif_ctx *ctx, const amr_array4 vel[2], amr_array4 *rhs, double *div_linf,
                         double *rhs_sum);
/* poisson_advance (poisson.cpp:41-186): L sol = rhs, homogeneous Neumann,
 * ||r||inf <= max(tol_abs, tol_rel ||rhs - mean||inf); sol returned mean-zero. */
int amressif_poisson_solve(amressif_ctx *ctx, amr_array4 *sol, const amr_array4 *rhs, double tol_rel,
                           double tol_abs, int *iters, double *res_rel);
/* No counterpart in the reference: the allocation-free variant of
 * amressif_poisson_rhs, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  It computes rhs = (1.5/dt)
 * D(vel) on the caller's views and writes dev_out[0] = ||D||inf and
 * dev_out[1] = sum rhs to the caller's device buffer (2 doubles), without
 * synchronising.  Its block partials live in a static device array of 128 KiB
 * per device (n_cell <= 16384; larger is refused with AMRESSIF_ERR_ARG).
 * The context's device must be current.  Calls that overlap on one device
 * share that array and race. */
int amressif_poisson_rhs_noalloc(amressif_ctx *ctx, const amr_array4 vel[2], amr_array4 *rhs, double *dev_out);
/* update_solution (poisson.cpp:192-392) with phi from userCtx comp 1:
 * velCont = velStar - G phi dt/1.5, p += phi.  Returns ||D(velCont)||inf. */
int amressif_project(amressif_ctx *ctx, const amr_array4 velStar[2], amr_array4 *userCtx,
                     amr_array4 velCont[2], double *div_linf);

/* ------------------------------------------------- discrete adjoint
 * No reference counterpart (RV has no adjoint).  The step is the map
 * (u^n, u^{n-1}, p^n; lid g) -> (u^{n+1}, u^n, p^{n+1}) of amressif_advance.
 * Its linearisation is taken at the converged momentum predictor u* (QUICK
 * upwind switch frozen at u*), so the tangent and adjoint are those of the
 * converged step: the momentum system (A_u - 1.5/dt) du* = G dp - dc - A_g dg
 * and its transpose are solved by Richardson sweeps to rel. change
 * <= adjoint_tol; the projection reuses the step's multigrid solve (the
 * Neumann Laplacian is symmetric).
 *
 * A tape holds the u* of consecutive forward steps (2 face arrays per step,
 * device memory owned by the tape).  Checkpointing is the caller's: restore a
 * checkpoint with amressif_load_state, re-run its segment with
 * amressif_step_fwd into a reset tape, then amressif_step_bwd over it.
 *
 * The dual state (faces du[2], du_prev[2], cells dp, lid dlid) lives in the
 * context: a perturbation for amressif_step_tlm, an adjoint for
 * amressif_step_bwd (which accumulates into dlid). */
typedef struct amressif_tape amressif_tape;

typedef struct {
    int lin_iters;   /* Richardson sweeps of the momentum system */
    int vcycles;     /* V-cycles of the projection solve */
    int flags;       /* AMRESSIF_FLAG_LIN_CAP, AMRESSIF_FLAG_MG_CAP, AMRESSIF_FLAG_NONFINITE */
    int pad;
    double lin_rel;  /* last relative change of the momentum solve */
} amressif_lin_stats;

int amressif_tape_create(amressif_ctx *ctx, int max_steps, amressif_tape **out);
int amressif_tape_destroy(amressif_tape *tape);
int amressif_tape_reset(amressif_tape *tape);
int amressif_tape_steps(const amressif_tape *tape);
/* Forward of the pair: amressif_advance, also appending each step's u* to the
 * tape (ERR_ARG if the tape would overflow). */
int amressif_step_fwd(amressif_ctx *ctx, int nsteps, amressif_tape *tape, amressif_step_stats *stats);
/* Backward of the pair: the adjoint of the tape's last nsteps steps, newest
 * first, applied to the dual state; the steps are removed from the tape.
 * stats: host array of nsteps entries (newest first) or NULL. */
int amressif_step_bwd(amressif_ctx *ctx, amressif_tape *tape, int nsteps, amressif_lin_stats *stats);
/* Tangent-linear of tape steps [first, first + nsteps) in forward order,
 * applied to the dual state (dlid is the constant lid perturbation). */
int amressif_step_tlm(amressif_ctx *ctx, const amressif_tape *tape, int first, int nsteps,
                      amressif_lin_stats *stats);
/* Dual state I/O: faces du[2], du_prev[2] (0 ghosts), cells dp (1 comp,
 * >= 0 ghosts), dlid (n_cell doubles, device).  NULL loads zeros / skips. */
int amressif_load_dual(amressif_ctx *ctx, const amr_array4 du[2], const amr_array4 du_prev[2],
                       const amr_array4 *dp, const double *dlid);
int amressif_store_dual(amressif_ctx *ctx, amr_array4 du[2], amr_array4 du_prev[2], amr_array4 *dp,
                        double *dlid);
/* Richardson tolerance and sweep cap of the tangent/adjoint momentum solve
 * (defaults 1e-14 and 200). */
int amressif_set_adjoint_solver(amressif_ctx *ctx, double tol, int max_sweeps);

/* ------------------------------------------------- two-level composite step
 * No reference counterpart (RV is single-level).  A coarse n_cell^2 level
 * plus one refined patch (ratio 2) over the coarse rows [patch_jlo, n_cell),
 * full width, so its W/E/N sides are the domain walls and its south side the
 * coarse/fine interface.  One dt (from the fine spacing, or fixed_dt) for both
 * levels.  Per step: coarse predictor (the single-level dual-time loop), fine
 * predictor with ghosts and interface faces from the coarse predictor, average
 * down, composite projection (FGMRES on the composite Neumann operator with a
 * coarse V-cycle + patch Gauss-Seidel preconditioner; flux matching across the
 * interface), update, average down.  Fine arrays: x-faces (2n+1) x ny,
 * y-faces 2n x (ny+1), p 2n x ny with ny = 2 (n_cell - patch_jlo). */
typedef struct amressif_amr amressif_amr;

typedef struct {
    int coarse_iters;   /* coarse pseudo-iterations */
    int fine_iters;     /* fine pseudo-iterations */
    int poisson_iters;  /* FGMRES iterations of the composite solve */
    int flags;          /* AMRESSIF_FLAG_* */
    double div_star;    /* composite ||D(u*)||inf */
    double div_new;     /* composite ||D(u^{n+1})||inf */
    double poisson_res; /* final ||r||inf / ||b||inf */
} amressif_amr_stats;

int amressif_amr_create(const amressif_inputs *in, int patch_jlo, void *hip_stream, amressif_amr **out);
int amressif_amr_destroy(amressif_amr *amr);
const char *amressif_amr_last_error(const amressif_amr *amr);
int amressif_amr_dt(const amressif_amr *amr, double *dt);
/* coarse faces + coarse userCtx (p, phi) as amressif_load_state; fine faces and
 * fine p (1 comp, >= 0 ghosts; NULL = 0).  velContPrev = velCont on both. */
int amressif_amr_load_state(amressif_amr *amr, const amr_array4 coarse_vel[2], const amr_array4 *coarse_ctx,
                            const amr_array4 fine_vel[2], const amr_array4 *fine_p);
int amressif_amr_store_state(amressif_amr *amr, amr_array4 coarse_vel[2], amr_array4 *coarse_ctx,
                             amr_array4 fine_vel[2], amr_array4 *fine_p);
int amressif_amr_advance(amressif_amr *amr, int nsteps, amressif_amr_stats *stats);

/* ------------------------------------------------------------ diagnostics
 * No reference counterpart (the reference prints per-step wall time only,
 * main.cpp:1173-1175).  Average device time of one kernel at the context's
 * shape and current state, timed with HIP events on the context's stream over
 * `reps` back-to-back launches:
 *   which = 0  fused FP64 RK stage (the FP64 step's dominant kernel)
 *   which = 1  finest-level down-leg smoother with the fused restriction (V-cycle precision)
 *   which = 2  plain device-to-device copy of copy_bytes bytes (bandwidth calibration)
 *   which = 3  mixed precision: FP64 residual of the pseudo base
 *   which = 4  mixed precision: FP32 tangent-linear update of the pseudo base
 *   which = 5  finest-level up-leg smoother with the fused FP64 outer step (V-cycle precision)
 *   which = 6  two FP64 RK stages in one pass (stages 1-2 of a pseudo-iteration)
 *   which = 10 mixed precision: one mixed iteration in one pass (residual + tangent-linear update)
 * On a 3D or periodic context (the general path) the kernels are those of one
 * RK4 stage, each over the whole grid:
 *   which = 7  cell velocities with their ghost layers (cont2cart)
 *   which = 8  fluxTotal: QUICK convective and viscous fluxes per cell
 *   which = 9  stage update of every face
 * which = 3, 4 and 10 need a context created with precision = 1.
 * bytes_per_launch receives the kernel's algorithmic HBM bytes per launch. */
int amressif_time_kernel(amressif_ctx *ctx, int which, int reps, int64_t copy_bytes, double *avg_ms,
                         double *bytes_per_launch);

/* No reference counterpart (a measurement hook).  In-step kernel timing:
 * with enable != 0, amressif_advance brackets each launch of the step's main
 * kernels with HIP events on the context's stream and keeps, per kernel, the
 * number of launches that did work and their total time (launches past the
 * converged pseudo-iteration or V-cycle, which exit at once, are not counted).
 * Enabling clears the totals; enable = 0 stops recording and keeps them.
 *   which = 0  FP64 RK stage
 *   which = 1  FP64 last RK stage (+ norm partials)
 *   which = 2  mixed FP64 residual
 *   which = 3  k_tangent: mixed FP32 tangent-linear update
 *   which = 4  finest-level down-leg smoother + restriction (outer-step cycles)
 *   which = 5  finest-level up-leg smoother + FP64 outer step
 *   which = 6  k_project
 *   which = 7  k_poisson_start
 *   which = 8  FP64 RK stages 1-2 in one pass
 *   which = 9  FP64 RK stages 3-4 in one pass (+ norm partials)
 *   which = 10 mixed iteration in one pass (residual + tangent-linear update + norm partials)
 * amressif_kernel_time returns the average time per counted launch (0 if none),
 * the number of counted launches and their average algorithmic HBM bytes (the
 * bytes each launch must read and write, counted once per cell). */
int amressif_kernel_timing(amressif_ctx *ctx, int enable);
int amressif_kernel_time(const amressif_ctx *ctx, int which, double *avg_ms, int64_t *launches,
                         double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
