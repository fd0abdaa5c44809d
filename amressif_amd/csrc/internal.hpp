// Internal types shared by the HIP kernels (kernels_*.hip) and the host
// driver (api.cpp).  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/amressif.h"

namespace amr {

namespace gen {
struct Ctx;  // the general path's working set (gen.hpp)
}

// Device-side control block.  Written only by single-block finalize kernels;
// every guarded kernel reads its `done` flag first and returns if it is set,
// so the host can enqueue work ahead and poll it once per batch.
struct Ctl {
    int mom_done;   // pseudo-iteration loop finished for this step
    int mom_count;  // countIter
    int flags;      // AMRESSIF_FLAG_*
    int mg_done;    // Poisson solve converged (or capped)
    int mg_iters;   // V-cycles done
    int mg_started; // first residual evaluated
    int b_norm_fixed;  // ||b||inf measured before the first residual (warm start)
    int pad1;
    double norm_error;
    double div_star, rhs_sum, b_mean;
    double b_norm, res_norm, phi_sum, phi_mean;
    double div_new;
    double tol_mom, tol_rel, tol_abs;
    int iter_num, mg_max_iter;
    // linear (tangent / adjoint) Richardson solves
    int lin_done, lin_count, lin_cap;
    int gm_stop;  // GMRES: the current Arnoldi cycle is finished (V-cycle kernels skip)
    int mg_skip;  // the step's Poisson phase ran before the pseudo-iterations ended: skipped
    double lin_tol, lin_rel;
};

// Pitched 2D device arrays.  Fine level: x-faces rows j in [0,n), cols
// i in [0,n]; y-faces rows [0,n], cols [0,n); cells rows/cols [0,n).  All
// share one pitch (elements), a multiple of 64.
struct Level {
    int nx, ny, pitch;
    double ax, ay;   // 1/dx^2, 1/dy^2
    double *e, *e2, *r;   // FP64 correction (pre- and post-smoothing buffers) and right-hand side
    float *e32, *e32b, *r32; // the same in FP32 for the mixed V-cycle
    double *rb;              // level 0: second right-hand-side buffer (the fused outer step ping-pongs)
    float *r32b;
};

struct Geometry {
    int n, P;
    double dx, dy, dt;
    int bc_lo0, bc_hi0, bc_lo1, bc_hi1;
    double inflow0, inflow1;
    double ren;
};

constexpr int kMaxLevels = 24;
constexpr int kMaxPartials = 1 << 16;

// Restarted FGMRES state (kernels_gmres.hip), device memory
constexpr int kGmMax = 64;  // largest restart length
struct GmState {
    int j, total, pad0, pad1;  // Arnoldi step in the cycle, steps in the solve
    double beta, target, scale;
    double gv[kGmMax + 1], cs[kGmMax], sn[kGmMax], y[kGmMax], c[kGmMax + 1];
    double H[(kGmMax + 1) * kGmMax];  // Hessenberg, row-major, reduced to R by Givens rotations
};

// Working arrays of the tangent-linear and adjoint step (kernels_adj.hip)
struct AdjWork {
    const double *base_u = nullptr, *base_v = nullptr;  // recorded u* of the step (faces)
    double *XU = nullptr, *XV = nullptr;                // base cell values, two ghost layers
    double *DU = nullptr, *DV = nullptr;                // perturbation cell values, two ghost layers
    double *F[4]{};                                     // face fluxes or flux adjoints (xU, xV, yU, yV)
    double *Q[2]{};                                     // right-hand side of the linear system (faces)
    double *XbU = nullptr, *XbV = nullptr;              // cell adjoints (interior)
    double *lidg = nullptr;                             // north-ghost adjoints of U (2 n)
    double *L[2][2]{};                                  // Richardson iterates (ping-pong, 2 face arrays)
    double *phi_lin[2]{};                               // the projections' Poisson iterates (cells)
    double *phi_mean_save = nullptr;                    // the forward phi's mean, kept across them
};

// Per-launch HIP-event timing of the step's main kernels inside the step loop
// (amressif_kernel_timing).  A record is counted once its loop's final count
// shows the launch did work: launches past the converged pseudo-iteration or
// V-cycle, and a speculative Poisson phase that was skipped, exit at once and
// are dropped.
enum KProfId {
    KP_STAGE = 0,    // k_stage64<0, *>: FP64 RK stage
    KP_STAGE_LAST,   // k_stage64<1, *>: FP64 last RK stage + norm partials
    KP_RESID,        // k_stage64<3, *>: mixed FP64 residual
    KP_TANGENT,      // k_tangent
    KP_MG_DOWN,      // finest down-leg smoother + restriction
    KP_MG_UP,        // finest up-leg smoother + FP64 outer step
    KP_PROJECT,      // k_project
    KP_PSTART,       // k_poisson_start
    KP_PAIR,         // k_stage_pair<false>: FP64 RK stages 1-2
    KP_PAIR_LAST,    // k_stage_pair<true>: FP64 RK stages 3-4 + norm partials
    KP_MPAIR,        // k_mixed_pair: mixed residual + tangent-linear update in one pass
    KP_COUNT
};
enum KProfLoop { KL_MOM = 0, KL_MG, KL_PHASE, KL_ONCE, KL_PROJ };
struct KProf {
    bool on = false;
    struct Rec {
        hipEvent_t e0, e1;
        int id, loop, idx, state;  // state: 0 open, 1 counted, 2 dropped
        double bytes;              // algorithmic HBM bytes of the launch
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double sum_ms[KP_COUNT]{}, sum_bytes[KP_COUNT]{};
    int64_t cnt[KP_COUNT]{};
    int mg_idx = 0;  // V-cycle being enqueued
};
struct Ctx;
hipError_t kprof_begin(const Ctx &c, int id, int loop, int idx, double bytes);
hipError_t kprof_end(const Ctx &c);
void kprof_close(const Ctx &c, int loop, int count);  // open records of the loop: counted if idx < count
hipError_t kprof_collect(const Ctx &c);               // accumulate the finished records

struct Ctx {
    amressif_inputs in{};
    hipStream_t stream = nullptr;
    Geometry g{};
    double rk[4]{};
    // faces: [0] = x-faces (u), [1] = y-faces (v)
    double *U[2]{}, *Uprev[2]{}, *C[2]{}, *S1[2]{}, *S2[2]{};
    // mixed-precision momentum state (FP32 faces): copy of the pseudo base and
    // its residual
    float *uk32[2]{}, *r32[2]{};
    double *p = nullptr, *phi = nullptr, *b = nullptr;
    double *p_alt = nullptr;  // the step's projection writes p^{n+1} here; then the two swap
    int c_gradp = 0;          // C holds -G p of the current p (amr_momentum_begin)
    // dual state of the tangent-linear / adjoint step (faces, faces, cells, lid)
    double *Du[2]{}, *Dup[2]{}, *Dp = nullptr, *Dlid = nullptr;
    AdjWork *adj = nullptr;  // their working arrays (allocated on first use)
    int Pg = 0;              // pitch of the cell arrays with two ghost layers
    double adj_tol = 1e-14;  // Richardson tolerance of the tangent/adjoint momentum solve
    int adj_cap = 200;       // and its sweep cap
    int last_L = 4;          // sweeps of the previous linear solve (batch size)
    int stage_rows = 8;      // k_stage64: output rows per wave
    int tangent_rows = 16;   // k_tangent: output rows per wave
    int stage_order = 1;     // stage kernels: 1 = consecutive waves take horizontally adjacent bands
    int stage_alt = 1;       // k_stage64: odd row chunks march down
    int stage_fast = 1;      // k_stage64: unrolled interior path (0: every wave on the general march)
    int stage_pair = 1;      // FP64 pseudo-iteration as two k_stage_pair launches (0: four k_stage64)
    int pair_rows = 16;      // k_stage_pair: output rows per wave
    // mixed iteration as one k_mixed_pair launch (0: k_stage64 mode 3 + k_tangent);
    // used by the step without a tape (advance) with the multigrid solve
    int mixed_pair = 1;
    int mpair_rows = 16;     // k_mixed_pair: output rows per wave
    int stage_edge_first = 1;  // pair kernels: waves on the general march dispatched first
    int mp_on = 0;           // this step's mixed iterations run as k_mixed_pair (advance_impl)
    int mp_count = 0;        // k_mixed_pair launches of the step so far
    int mp_spec = 0;         // the Poisson phase follows the first batch unseen: its last finalizer starts it
    // multigrid smoothers of the levels above the tail: 2 = wave-column kernel
    // with two columns per lane, 1 = one column per lane (both registers +
    // DPP), 0 = LDS tile kernel; output rows per wave of the wave-column
    // kernels by [FP64, FP32][up leg, down leg, finest up leg], and on the
    // levels below the finest
    int mg_smoother = 2;
    int mg_late = 1;  // finest up-leg smoother: phi and b loaded after the sweeps
    int mg_post3 = 1;  // FP32 V-cycle: three post-smoothing sweeps on the finest level (0: two)
    int mg_pre3 = 1;   // FP32 V-cycle: three pre-smoothing sweeps on the finest level (0: two)
    int mg_up64_single = 1;  // FP64 finest up leg on the one-column kernel (with mg_smoother 2)
    int mg_rows[2][3] = {{8, 8, 8}, {8, 8, 8}};
    int mg_rows_coarse = 4;
    // FGMRES Poisson solver (poisson_solver = 1)
    int gm_m = 0;            // restart length
    double *gmV = nullptr, *gmZ = nullptr, *gm_part = nullptr;
    size_t gm_vs = 0;        // doubles between consecutive vectors
    GmState *gm = nullptr;
    double *phiv[2]{};   // Poisson iterates (ping-pong); phi points at the last solution
    double *phih = nullptr;  // warm start: the solution of the step before (allocated on first use)
    int phih_valid = 0;
    int pred_valid = 0;  // S1 holds the last step's converged predictor u* (set by a step, cleared by a load)
    int tail_level = 0;  // first multigrid level handled by the single-workgroup tail kernel
    // exact solve of that level by the DCT (k_tail_dct): cosine and inverse
    // eigenvalue tables (null: the tail runs the V-cycle down to the bottom)
    int tail_dct = 1;
    int dct_mfma64 = 1;  // FP64 DCT products on the matrix cores (0: on the VALU)
    int start_rows = 8;   // Poisson start: rows per wave of the row-march kernel (0: cell-tile kernel)
    int project_rows = 8;   // projection: rows per wave of the row-march kernel (0: cell-tile kernel)
    double *dct_tab = nullptr;
    float *dct_tab32 = nullptr;
    // a coarsest level above 1024 cells (odd side > 32): its exact DCT solve as
    // four dense products; tables C_x (nx^2), C_y (ny^2), 1/lambda (ny nx) and two
    // work arrays (null: the bottom is the dense pseudo-inverse / k_tail_dct)
    double *big_tab = nullptr, *big_work = nullptr;
    float *big_tab32 = nullptr, *big_work32 = nullptr;
    // FP32 V-cycle: the level above it smoothed inside the same kernel
    // (k_tail_dct_ext; 32, 64 or 128 rows, at most 128 columns)
    int tail_ext = 1;
    int tail_waves = 16;  // waves of that kernel and of the DCT tail kernel (4 or 16)
    const double *lid = nullptr;  // caller-owned device lid profile or null
    Level lev[kMaxLevels]{};
    int nlev = 0;
    double *pinv = nullptr;  // bottom pseudo-inverse, mb x mb
    float *pinv32 = nullptr;
    int mb = 0;
    double *partials = nullptr;  // kMaxPartials * 4 doubles
    Ctl *ctl = nullptr;
    Ctl *ctl_host = nullptr;  // pinned
    hipEvent_t poll_ev = nullptr;  // marks the control-block copy behind a V-cycle batch
    amressif_step_stats *log = nullptr;  // device step log
    int log_cap = 0;
    int last_P = 16, last_V = 8;
    int mg_batch = 0;  // > 0: cap on the first batch of V-cycles of a solve (tests of the later batches)
    int64_t bytes = 0;
    // scratch of the module-level entry points (allocated on first use)
    double *M[4][2]{};
    double *Mc[2]{};
    double *dscalar = nullptr;
    std::vector<void *> allocs;
    std::string err;
    KProf *kp = nullptr;  // in-step kernel timing (null: off)
    gen::Ctx *gen = nullptr;  // 3D or periodic inputs: the general path (gen_path.cpp) owns the state
};

// ---- kernel launchers (kernels_*.hip)
struct StageLaunch {
    const double *us, *vs;  // stencil input faces
    const double *ub, *vb;  // base u^k
    const double *cu, *cv;  // BDF2 constant (1.5/dt) u^n + (0.5/dt)(u^n - u^{n-1})
    const double *p;        // pressure (cells)
    double *uo, *vo;        // outputs
    double rk;
    int mode;               // 0 = stage update, 1 = stage update + norm partials, 2 = RHS only
    int unguarded;          // 1: ignore the control block (diagnostic timing)
    int foldp;              // modes 0, 1: cu, cv already hold -G p (launch_fold_gradp); p is not read
    double rk2;             // launch_stage_pair: the second stage's coefficient
    int pair;               // launch_stage_pair: 1 = stages 1-2, 2 = stages 3-4
    int expl;               // mode 1 with foldp: PSEUDO_TIMESTEPPING = 0's explicit update (rk = dt/1.5,
                            // base = stencil input = u^n; fluxTotal's domain ghosts 0, so wall faces
                            // take half the adjacent cell's flux)
};
hipError_t launch_stage(const Ctx &c, const StageLaunch &a, int *nblocks);
// two stages in one launch (k_stage_pair): stage rk from us into registers,
// then stage rk2 from those into uo; pair 1 = stages 1-2, pair 2 = stages 3-4
// (+ norm partials).  uo must differ from us and ub.
hipError_t launch_stage_pair(const Ctx &c, const StageLaunch &a, int *nblocks);
bool stage_pair_fits(const Ctx &c);
// cu, cv += -G p on the interior faces, once per FP64 step (p fixed over its pseudo-iterations)
hipError_t launch_fold_gradp(const Ctx &c);
// mixed (defect-correction) pseudo-iteration: see kernels_stage.hip
struct MixedStageLaunch {
    int stage;                   // 0: FP64 residual of the base; 1: FP32 tangent-linear update
    double *u64, *v64;           // FP64 pseudo base u^k (stage 1 updates it in place)
    const double *cu, *cv, *p;   // stage 0: BDF2 constant and pressure
    float *uk, *vk;              // FP32 copy of the base (written by stage 0)
    float *ru, *rv;              // FP64 residual of the base, stored FP32 (written by stage 0)
    double cr, cj;               // stage 1: u^{k+1} = u^k + cr r + cj R'(u^k) r
    int unguarded;
    int foldp;                   // stage 0: cu, cv already hold -G p; p is not read
};
hipError_t launch_stage_mixed(const Ctx &c, const MixedStageLaunch &a, int *nblocks);
// one mixed iteration in one launch (k_mixed_pair): u^{k+1} of u^k into uo, vo
// (which must not alias u^k); after convergence it copies u^k
struct MixedPairLaunch {
    const double *u64, *v64;  // pseudo base u^k
    const double *cu, *cv;    // BDF2 constant holding -G p
    double *uo, *vo;          // u^{k+1}
    double cr, cj;            // u^{k+1} = u^k + cr r + cj R'(u^k) r
    int unguarded;
};
hipError_t launch_mixed_pair(const Ctx &c, const MixedPairLaunch &a, int *nblocks);
bool mixed_pair_fits(const Ctx &c);
// the pseudo-iteration norm and loop control (k_mixed_pair's records when the
// step runs it); start_mg: also the Poisson phase's start (k_mg_start)
hipError_t launch_fin_pseudo(const Ctx &c, int nblocks, int start_mg = 0);
// the norm of PSEUDO_TIMESTEPPING = 0's single update: ends the loop, countIter stays 0
hipError_t launch_fin_explicit(const Ctx &c, int nblocks);
hipError_t launch_ctl_reset_step(const Ctx &c);
hipError_t launch_bdf2(const Ctx &c);
hipError_t launch_poisson_rhs(const Ctx &c, const double *u, const double *v, double *b, int *nblocks,
                              const float *au = nullptr, const float *av = nullptr);
hipError_t launch_fin_prhs(const Ctx &c, int nblocks);
hipError_t launch_center_b(const Ctx &c, double *b);
// control fields of a Poisson solve; with the predictor faces us, vs also the
// mean of the b they give (to be followed by launch_poisson_start with them)
hipError_t launch_mg_start(const Ctx &c, int require_mom = 0, const double *us = nullptr,
                           const double *vs = nullptr);
// first residual of a Poisson solve (see kernels_mg.hip), with its finalize
hipError_t launch_poisson_start(const Ctx &c, int mode, int center, const double *phi1, const double *phi2,
                                double *phi0, int mixed, const double *us = nullptr, const double *vs = nullptr);
hipError_t launch_fin_resid(const Ctx &c, int nblocks);
hipError_t launch_vcycle(const Ctx &c, int mixed);
hipError_t launch_vcycle_step(const Ctx &c, int mixed, int par, const double *phi_in, double *phi_out, int *nblocks);
hipError_t launch_correct_residual(const Ctx &c, const double *phi_in, double *phi_out, int with_correction,
                                   int mixed, int *nblocks);
size_t tail_lds_bytes(const Ctx &c, int mixed);
hipError_t configure_tail(Ctx &c);
hipError_t launch_smooth_fine(const Ctx &c, int kind, double *scratch_phi);
hipError_t launch_fin_project(const Ctx &c, int nblocks, int step_index, int reset_next = 0, int gated = 0);
hipError_t launch_cont2cart_out(const Ctx &c, const double *u, const double *v, const amr_array4 &vc);
hipError_t launch_pack_face(const Ctx &c, const amr_array4 &a, double *dst, int dir);
hipError_t launch_unpack_face(const Ctx &c, const double *src, const amr_array4 &a, int dir);
hipError_t launch_pack_cell(const Ctx &c, const amr_array4 &a, int comp, double *dst);
hipError_t launch_unpack_cell(const Ctx &c, const double *src, const amr_array4 &a, int comp, int fill_ghosts);
// window [i0, i1] x [j0, j1] (global indices, inclusive) of one box view <-> pitched array
hipError_t launch_pack_window(const Ctx &c, const amr_array4 &a, int comp, double *dst, int i0, int j0, int i1,
                              int j1);
hipError_t launch_unpack_window(const Ctx &c, const double *src, const amr_array4 &a, int comp, int i0, int j0,
                                int i1, int j1);
hipError_t launch_copy_faces(const Ctx &c, const double *src, double *dst, int dir);
hipError_t launch_fill(const Ctx &c, double *dst, int64_t count, double value);
hipError_t launch_norm(const Ctx &c, const double *ub, const double *vb, const double *uo, const double *vo,
                       double *out_dev);
hipError_t launch_ctl_set(const Ctx &c, double tol_mom, double tol_rel, double tol_abs, int iter_num,
                          int mg_max_iter);
// next_bdf: also form the next step's BDF2 constant in C and a copy of the new
// velocity in S2 (that step then takes S2 as u^n and skips launch_bdf2).
// p_out: where p + phi - mean goes (null: in place).  With next_bdf and a
// separate p_out, the row-march kernel also folds -G p^{n+1} into the next
// constant (project_rows > 0).  gated: enqueued before the host saw the solve
// end (does nothing unless it has); phi_alt: the iterate buffer taken after an
// odd cycle count.  un: where next_bdf takes u^n (null: Uprev); no_copy: no
// copy of the new velocity into S2.
hipError_t launch_project_ex(const Ctx &c, const double *us, const double *vs, double *uo, double *vo, double *p,
                             const double *phi, int use_mean, int *nblocks, const float *au = nullptr,
                             const float *av = nullptr, int next_bdf = 0, double *p_out = nullptr,
                             const double *phi_alt = nullptr, int gated = 0, double *const *un = nullptr,
                             int no_copy = 0);
hipError_t launch_unpack_cell_shift(const Ctx &c, const double *src, const amr_array4 &a, int comp,
                                    int fill_ghosts, const double *shift);
hipError_t launch_copy_bytes(const Ctx &c, const void *src, void *dst, int64_t bytes);
hipError_t launch_mean_b(const Ctx &c, const double *b);
hipError_t launch_bdf2_const(const Ctx &c, const double *un, const double *um, double *cc, int dir);
// flux fields of the faces u, v (kernels_export.hip): half-node QUICK fluxes
// on x-faces (fxu, fxv) and y-faces (fyu, fyv), fluxConvect (cu, cv) and
// fluxViscous (vu, vv); any output pair may be null
hipError_t launch_flux_fields(const Ctx &c, const double *u, const double *v, double *fxu, double *fxv, double *fyu,
                              double *fyv, double *cu, double *cv, double *vu, double *vv);

// FGMRES Poisson solver (kernels_gmres.hip)
size_t gm_partials_doubles(const Ctx &c, int m);
hipError_t launch_gm_reset(const Ctx &c);
hipError_t launch_gm_cycle_start(const Ctx &c, const double *x, const double *b, int mixed);
hipError_t launch_gm_iteration(const Ctx &c, int mixed);
hipError_t launch_gm_cycle_end(const Ctx &c, double *x);

// two-level composite step (kernels_amr.hip): coarse level n0 x n0, fine patch
// over the coarse rows [jc0, n0) at ratio 2
struct AmrGeom {
    int n0, P0, jc0;        // coarse cells per side, pitch, first covered row
    int nx, ny, P1, Pg1;    // fine patch cells, pitch, ghosted-cell pitch
    double h0, h1;          // spacings
    double iren, c15;
    int bc_w, bc_e, bc_n;   // wall codes of the patch sides (the coarse domain's)
    double inflow0, inflow1;
};
int amr_flat_blocks(const AmrGeom &a);
hipError_t amr_fine_cells(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *u0,
                          const double *v0, double *XU, double *XV);
hipError_t amr_fine_interface(const AmrGeom &a, hipStream_t s, double *v, const double *v0);
hipError_t amr_fine_stage(const AmrGeom &a, hipStream_t s, const double *us, const double *vs, const double *ub,
                          const double *vb, const double *cu, const double *cv, const double *p, const double *XU,
                          const double *XV, double *uo, double *vo, double rk, int mode, double *partials,
                          const Ctl *ctl, int *nblocks);
hipError_t amr_fine_fin(const AmrGeom &a, hipStream_t s, const double *partials, int nb, Ctl *ctl);
// the patch's momentum iteration in the refinement form (mixed step): the
// residual r, the cells of r, then u^{k+1} = u^k + cr r + cj R'(u^k) r into
// (uo, vo) with the norm partials (a copy of u^k after convergence)
hipError_t amr_fine_resid(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *cu,
                          const double *cv, const double *p, const double *XU, const double *XV, double *ru,
                          double *rv, const Ctl *ctl);
hipError_t amr_fine_dcells(const AmrGeom &a, hipStream_t s, const double *ru, const double *rv, double *DU, double *DV,
                           const Ctl *ctl);
hipError_t amr_fine_tangent(const AmrGeom &a, hipStream_t s, const double *u, const double *v, const double *XU,
                            const double *XV, const double *ru, const double *rv, const double *DU, const double *DV,
                            double *uo, double *vo, double cr, double cj, double *partials, const Ctl *ctl,
                            int *nblocks);
hipError_t amr_fine_ctl_reset(hipStream_t s, Ctl *ctl, double tol, int iter_num);
hipError_t amr_fine_bdf2(const AmrGeom &a, hipStream_t s, double *u, double *up, double *c, int dir, double dt);
hipError_t amr_apply(const AmrGeom &a, hipStream_t s, const double *x0, const double *x1, const double *b0,
                     const double *b1, double *y0, double *y1, double *partials);
hipError_t amr_dots(const AmrGeom &a, hipStream_t s, const double *w, const double *V, size_t vs, int nv,
                    size_t off1, double *partials);
hipError_t amr_combine(const AmrGeom &a, hipStream_t s, double *w, double alpha, const double *V, size_t vs, int nv,
                       const double *coef, size_t off1);
hipError_t amr_wsum(const AmrGeom &a, hipStream_t s, const double *x0, const double *x1, double *partials);
hipError_t amr_shift(const AmrGeom &a, hipStream_t s, double *x0, double *x1, const double *mean);
// (guard: null, or the device composite FGMRES's control block: the kernel
// does nothing once its Arnoldi cycle or solve has ended)
hipError_t amr_restrict(const AmrGeom &a, hipStream_t s, const double *r0, const double *r1, void *rc, int mixed,
                        const Ctl *guard = nullptr);
hipError_t amr_prolong(const AmrGeom &a, hipStream_t s, const void *e, double *z0, double *z1, int mixed,
                       const Ctl *guard = nullptr);
hipError_t amr_fine_gs(const AmrGeom &a, hipStream_t s, const double *z0, double *z1, const double *r1, int color,
                       const Ctl *guard = nullptr);
// device composite FGMRES (kernels_amr.hip): vectors [coarse | fine] of stride vs
struct CgmBufs {
    Ctl *ctl;        // control: mg_done ends the solve, gm_stop the cycle
    GmState *st;
    double *V, *Z;   // m + 1 Arnoldi vectors, m preconditioned vectors
    double *x, *b;   // iterate and right-hand side
    double *vcur;    // the preconditioner's input (V_j normalised)
    double *zcur;    // and its output
    double *part;    // flat blocks x (m + 2) partials
    size_t vs, off1;
    int m;
};
hipError_t amr_cgm_begin(hipStream_t s, Ctl *ctl, GmState *st);
hipError_t amr_cgm_edge(const AmrGeom &a, hipStream_t s, const CgmBufs &g);
hipError_t amr_cgm_apply(const AmrGeom &a, hipStream_t s, const CgmBufs &g);
hipError_t amr_cgm_orthonormalise(const AmrGeom &a, hipStream_t s, const CgmBufs &g);
// the Arnoldi finalizers of kernels_gmres.hip on an nblk x (m + 2) partials layout
hipError_t launch_gm_fin_dots_raw(hipStream_t s, double *part, int nblk, int m, GmState *st, Ctl *ctl, int pass);
hipError_t launch_gm_fin_arnoldi_raw(hipStream_t s, double *part, int nblk, int m, GmState *st, Ctl *ctl);
hipError_t amr_rhs(const AmrGeom &a, hipStream_t s, double dt, const double *u0, const double *v0, const double *u1,
                   const double *v1, double *b0, double *b1, double *partials);
hipError_t amr_project(const AmrGeom &a, hipStream_t s, double dt, const double *phi0, const double *phi1, double *u0,
                       double *v0, double *u1, double *v1, double *p0, double *p1);
hipError_t amr_average_down(const AmrGeom &a, hipStream_t s, const double *u1, const double *v1, const double *p1,
                            double *u0, double *v0, double *p0);
hipError_t launch_reduce_cols(hipStream_t s, const double *partials, int nblk, int ncol, int is_max, double *out);
hipError_t launch_pack_rect(hipStream_t s, const amr_array4 &v, double *dst, int P, int nx, int ny);
hipError_t launch_unpack_rect(hipStream_t s, const double *src, const amr_array4 &v, int P, int nx, int ny);

// tangent-linear / adjoint step (kernels_adj.hip)
int lin_partials_count(const Ctx &c);
hipError_t launch_lin_start(const Ctx &c, double tol, int cap);
hipError_t launch_cells(const Ctx &c, const double *fu, const double *fv, double *XU, double *XV, int homog,
                        const double *dlid, int guarded);
hipError_t launch_tl_sweep(const Ctx &c, const AdjWork &w, const double *xu, const double *xv, double *xu_new,
                           double *xv_new, const double *dlid);
hipError_t launch_adj_sweep(const Ctx &c, const AdjWork &w, const double *su, const double *sv, double *su_new,
                            double *sv_new);
hipError_t launch_adj_lid(const Ctx &c, const AdjWork &w, const double *su, const double *sv, double *glid);
hipError_t launch_tl_rhs(const Ctx &c, const AdjWork &w, const double *du, const double *dv, const double *dup,
                         const double *dvp, const double *dp, double *xu, double *xv);
hipError_t launch_adj_phibar(const Ctx &c, const double *au, const double *av, const double *pb, double *b);
hipError_t launch_adj_rhs(const Ctx &c, const AdjWork &w, const double *au, const double *av, const double *bb,
                          double *su, double *sv);
hipError_t launch_adj_finish(const Ctx &c, const double *su, const double *sv, double *au, double *av, double *aup,
                             double *avp, double *pb);

}  // namespace amr

// the C-ABI context handle (include/amressif.h) and host helpers shared by
// the single-level driver (api.cpp) and the two-level driver (amr_two_level.cpp)
struct amressif_ctx {
    amr::Ctx c;
};
int amr_momentum_predictor(amressif_ctx *ctx, bool bdf_ready = false, bool reset_done = false);
// its two halves: the step's set-up and first batch of pseudo-iterations (no
// host wait), then the polled continuation until the loop ends
int amr_momentum_begin(amressif_ctx *ctx, bool bdf_ready, int *launched, bool reset_done = false);
int amr_momentum_finish(amressif_ctx *ctx, int launched);
int amr_poll(amressif_ctx *ctx);
int amr_set_ctl(amressif_ctx *ctx);

// host copies of the state in the global one-box layout (row-major, i
// fastest): x faces n rows of n + 1, y faces n + 1 rows of n, cells n x n.
// state_to_host: any pointer may be null; phi is returned mean-removed as
// amressif_store_state gives it.  state_from_host: phi may be null (0), then
// finishes as amressif_load_state does.  Both synchronise the stream.
int amr_state_to_host(amressif_ctx *ctx, double *u, double *v, double *up, double *vp, double *p, double *phi);
int amr_state_from_host(amressif_ctx *ctx, const double *u, const double *v, const double *up, const double *vp,
                        const double *p, const double *phi);
int amr_fail(amressif_ctx *ctx, int code, const std::string &msg);
// the module-level entry points' scratch arrays M[4][2], Mc[2] (allocated on first use)
int amr_module_scratch(amressif_ctx *ctx);

// the general path (gen_path.cpp): 3D inputs or a periodic direction
bool amr_gen_wanted(const amressif_inputs *in);
int amr_gen_create(amressif_ctx *ctx);
int amr_gen_load_state(amressif_ctx *ctx, const amr_array4 *vel, const amr_array4 *velPrev, const amr_array4 *userCtx);
int amr_gen_store_state(amressif_ctx *ctx, amr_array4 *vel, amr_array4 *velPrev, amr_array4 *userCtx,
                        amr_array4 *velCart);
int amr_gen_init_state(amressif_ctx *ctx);
int amr_gen_load_state_mf(amressif_ctx *ctx, const amr_multifab *vel, const amr_multifab *velPrev,
                          const amr_multifab *userCtx);
int amr_gen_store_state_mf(amressif_ctx *ctx, amr_multifab *vel, amr_multifab *velPrev, amr_multifab *userCtx,
                           amr_multifab *velCart);
int amr_gen_advance(amressif_ctx *ctx, int nsteps, amressif_step_stats *stats);
int amr_gen_time_kernel(amressif_ctx *ctx, int which, int reps, double *avg_ms, double *bytes_per_launch);
