// The fused RK4 dual-time stage -- the dominant kernel of the step (gfx950).
//
// One launch = one stage over the whole grid: cell velocities from faces with
// the wall ghost rules (cont2cart, utilities.cpp:16-350), QUICK half-node
// fluxes (fn_flux_calc.cpp:11-202), viscous flux (:695-766), flux wall ghosts
// (main.cpp:945-1014), face pressure gradient (fn_flux_calc.cpp:825-911),
// momentum RHS (fn_rhs.cpp:8-68), dual-time residual and stage update
// (momentum.cpp:68-87), and on the last stage the pseudo-iteration norm
// partials (utilities.cpp:421-493).
//
// Work decomposition: one wavefront owns a band of 58 columns (64 lanes: the
// band plus a 3-lane halo) and a chunk of rows, and marches up the rows.  A
// rolling 5-row window of cell velocities stays in registers (the y-direction
// QUICK stencil); x-direction neighbours come from adjacent lanes through DPP
// wave shifts (no LDS, no barriers).  Each face flux is evaluated once; the
// next row's global loads are issued one iteration ahead and (k_stage64) a
// row's stores one iteration behind, so each row waits for memory once.
// Walls: ghost cells come from their mirror lane (x) or mirror row (y) with
// the cont2cart rules.
//
// Precision forms:
//  * FP64 (the reference's): u* <- u^k + rk (R(u*) - (1.5/dt) u* + c), with the
//    BDF2 constant c = (1.5/dt) u^n + (0.5/dt)(u^n - u^{n-1}) folded once per step.
//  * mixed (iterative refinement): per iteration, k_stage64 in mode 3 forms
//    the FP64 residual r = R(u^k) - (1.5/dt) u^k + c of the implicit BDF2
//    system and k_tangent applies the FP32 tangent-linear stencil R'(u^k) r
//    (QUICK with the upwind switch frozen at u^k, viscous flux of r), then
//    u^{k+1} = u^k + cr r + cj R' r in FP64.  With J = R' - c, c = 1.5/dt,
//    the default is the Newton correction with the two-term Neumann series
//    J^{-1} ~ -(1/c)(I + R'/c): cr = 1/c, cj = 1/c^2 (|R'|/c is ~1e-4 at
//    2048^2, so a few iterations converge).  The alternative keeps the RK4
//    dual-time recursion (a1..a4) to first order in J: cr = a4 - a3 a4 c,
//    cj = a3 a4.  A converged state (r = 0) is a fixed point of both, so the
//    answer is fixed by the FP64 residual.
#include "device_common.hpp"
#include "internal.hpp"

namespace amr {

constexpr int WB = 58;  // output columns per wave (64 lanes = 3 halo + 58 + 3 halo)

struct StageArgs {
    // FP64 form: stencil input, pseudo base, BDF2 constant, pressure, outputs
    const double *us, *vs, *ub, *vb, *cu, *cv, *p;
    double *uo, *vo;
    // mixed form
    float *uk, *vk, *ru, *rv;  // FP32 copy of the base and its FP64 residual (stored FP32)
    double *u64, *v64;         // FP64 pseudo base, updated in place by k_tangent
    double cr, cj;             // u^{k+1} = u^k + cr r + cj R'(u^k) r
    const double *lid;
    double *partials;
    const Ctl *ctl;
    int n, P;
    double inv_dx, inv_dy, inv_dx2, inv_dy2, inv_ren, c15, rk;
    int bc_lo0, bc_hi0, bc_lo1, bc_hi1;
    double inflow0, inflow1;
    int mode;  // k_stage64: 0 stage, 1 last stage (+ norm), 2 RHS only, 3 mixed residual
    int nbands, nchunks;
    int rows;   // output rows per wave (chunk height)
    int order;  // wave -> tile: 0 band-major (chunks of one band consecutive), 1 chunk-major
    int alt;    // k_stage64: odd chunks march down (shared halo rows read at the same time)
};

// wave id -> (band, chunk); blocks of 4 waves, handed to the 8 XCDs in
// contiguous ranges so vertically adjacent chunks share an L2
__device__ __forceinline__ int wave_id(const StageArgs &)
{
    return xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
}

// cell velocities of row jr for this lane's column from raw face loads of the
// mirror row, with the x ghosts from the mirror lane and the y ghost rule.
// Homog: the rules without their constant part (an increment of the velocity).
// gx: the lid (or inflow) velocity of this lane's column, loaded once per wave
// so that the row loops hold no loads besides their prefetch.
template <class R, bool Homog>
__device__ __forceinline__ void finish_cells(const StageArgs &a, int lane, int i, int jr, bool edge, R u0, R u1,
                                             R v0, R v1, R gx, R &U, R &V)
{
    const int n = a.n;
    const int jm = jr < 0 ? -jr - 1 : (jr >= n ? 2 * n - 1 - jr : jr);
    const bool xin = i >= 0 && i < n;
    U = R(0.5) * (u0 + u1);
    V = R(0.5) * (v0 + v1);
    if (edge) {
        const int mi = i < 0 ? -i - 1 : 2 * n - 1 - i;
        int ml = lane + (mi - i);
        ml = ml < 0 ? 0 : (ml > 63 ? 63 : ml);
        const R Um = __shfl(U, ml, 64), Vm = __shfl(V, ml, 64);
        if (!xin) {
            R gu = 0, gv = 0;
            if (i >= -2 && i <= n + 1)
                ghost_rule<R, Homog>(i < 0 ? 0 : 1, i < 0 ? a.bc_lo0 : a.bc_hi0, Um, Vm, R(0), R(0), gu, gv);
            U = gu;
            V = gv;
        }
    }
    if (jr != jm && xin) {
        R gu, gv;
        ghost_rule<R, Homog>(jr < 0 ? 2 : 3, jr < 0 ? a.bc_lo1 : a.bc_hi1, U, V, gx, (R)a.inflow1, gu, gv);
        U = gu;
        V = gv;
    }
}

// ============================================================ FP64 stage
// MODE: the launch's a.mode (0 stage, 1 last stage + norm, 2 RHS only, 3 mixed
// residual) as a template argument, so each form compiles to its own kernel.
// FOLDP (modes 0, 1, 3): the constant c_u, c_v already holds -G p on the
// interior faces (k_fold_gradp or the projection), so p is not read.
template <int MODE, bool FOLDP = false, bool EXPL = false>
__global__ __launch_bounds__(256) void k_stage64(StageArgs a)
{
    using R = double;
    a.mode = MODE;
    if (a.ctl && a.ctl->mom_done) return;
    const int lane = threadIdx.x & 63;
    const int wid = wave_id(a);
    if (wid >= a.nbands * a.nchunks) return;  // whole wave
    const int band = a.order ? wid % a.nbands : wid / a.nchunks;
    const int chunk = a.order ? wid / a.nbands : wid % a.nchunks;
    const int n = a.n, P = a.P;
    const int i0 = band * WB, j0 = chunk * a.rows;
    const int i = i0 + lane - 3;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool edge = i0 < 3 || i0 + 61 > n;
    const bool outlane = lane >= 3 && lane < 3 + WB;
    const double *us = a.us, *vs = a.vs;
    const R gx = (xin && a.lid) ? a.lid[i] : a.inflow0;
    struct Raw {
        R u0, v0, v1, vn, uw, pc, ub, cu, vb, cv;  // u(i + 1) = u0 of lane + 1
    };
    auto load = [&](int jc, Raw &r) {  // everything iteration jc consumes
        const int jr = jc + 2;
        const int jm = jr < 0 ? -jr - 1 : (jr >= n ? 2 * n - 1 - jr : jr);
        r.u0 = r.v0 = r.v1 = 0;
        if (jm >= 0 && jm < n && xface) {
            const size_t k = i + (size_t)P * jm;
            r.u0 = us[k];
            if (xin) { r.v0 = vs[k]; r.v1 = vs[k + P]; }
        }
        r.vn = (xin && jc + 1 >= 0 && jc + 1 <= n) ? vs[i + (size_t)P * (jc + 1)] : R(0);
        const bool row = jc >= 0 && jc < n;
        const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * (jc < 0 ? 0 : jc);
        r.uw = (xface && row) ? us[k] : R(0);
        r.pc = (!FOLDP && xin && row) ? a.p[k] : R(0);
        const bool out = row && jc >= j0 && a.mode != 2 && outlane;
        r.ub = (out && xface && a.mode != 3) ? a.ub[k] : R(0);
        r.cu = (out && xface) ? a.cu[k] : R(0);
        r.vb = (out && xin && a.mode != 3) ? a.vb[k] : R(0);
        r.cv = (out && xin) ? a.cv[k] : R(0);
    };
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren;
    R wU[5], wV[5];  // rolling window: w[q] = cell (i, jc - 2 + q)
    const int jend = min(j0 + a.rows, n);
    if (a.alt && (chunk & 1)) {
        // Odd chunks march down (even ones up): the rows a chunk shares with
        // its neighbours as stencil halo are then read by both at about the same
        // time (both at their start or both at their end), so the second read
        // is served by L2.  Same arithmetic per face as the upward march; the
        // y-face (i, jc + 1) is written at row jc from the values carried from
        // row jc + 1.
        struct RawD {
            R u0, v0, v1, vs, uw, pc, ub, cu, vb, cv;
        };
        auto cells = [&](int jr, RawD &r) {  // face values of cell row jr (its mirror row outside)
            const int jm = jr < 0 ? -jr - 1 : (jr >= n ? 2 * n - 1 - jr : jr);
            r.u0 = r.v0 = r.v1 = 0;
            if (jm >= 0 && jm < n && xface) {
                const size_t k = i + (size_t)P * jm;
                r.u0 = us[k];
                if (xin) { r.v0 = vs[k]; r.v1 = vs[k + P]; }
            }
        };
        auto loadd = [&](int jc, RawD &r) {  // everything iteration jc consumes
            cells(jc - 2, r);
            r.vs = (xin && jc >= 0 && jc <= n) ? vs[i + (size_t)P * jc] : R(0);  // south y-face of row jc
            const bool row = jc >= 0 && jc < n;
            const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * (jc < 0 ? 0 : jc);
            r.uw = (xface && row) ? us[k] : R(0);
            r.pc = (!FOLDP && xin && row) ? a.p[k] : R(0);
            const bool out = row && jc >= j0 && MODE != 2 && outlane;
            r.ub = (out && xface && MODE != 3) ? a.ub[k] : R(0);
            r.cu = (out && xface) ? a.cu[k] : R(0);
            r.vb = (out && xin && MODE != 3) ? a.vb[k] : R(0);
            r.cv = (out && xin) ? a.cv[k] : R(0);
        };
        {
            RawD r;
#pragma unroll
            for (int q = 0; q < 4; ++q) {  // rows jend + 1 .. jend - 2 into w[4] .. w[1]
                cells(jend + 1 - q, r);
                finish_cells<R, false>(a, lane, i, jend + 1 - q, edge, r.u0, lane_next(r.u0), r.v0, r.v1, gx, wU[4 - q],
                                       wV[4 - q]);
            }
        }
        const R vtop = (xin && jend <= n) ? vs[i + (size_t)P * jend] : R(0);  // y-face (i, jend)
        R fnU = quick_flux<R>(vtop, wU[1], wU[2], wU[3], wU[4]), fnV = quick_flux<R>(vtop, wV[1], wV[2], wV[3], wV[4]);
        R tVup = 0, pup = 0, vup = vtop, vbup = 0, cvup = 0;  // row jc + 1's TV, p, south-face v, base, constant
        R tVtop = 0;  // TV of the chunk's top row (EXPL: the north wall faces take half of it)
        double nx_ = 0.0, ny_ = 0.0;
        // deferred stores, as in the upward march below: x-face (i, jc) and
        // y-face (i, jc + 1) of row jc go out at the top of iteration jc - 1
        unsigned skx = 0, sky = 0;  // n <= 16384: every index fits 32 bits
        bool sxf = false, syf = false;
        double sx = 0.0, sy = 0.0;
        float fx0 = 0.f, fx1 = 0.f, fy0 = 0.f, fy1 = 0.f;
        auto flush = [&]() {
            if (sxf) {
                if (MODE == 3) { a.uk[skx] = fx0; a.ru[skx] = fx1; }
                else a.uo[skx] = sx;
            }
            if (syf) {
                if (MODE == 3) { a.vk[sky] = fy0; a.rv[sky] = fy1; }
                else a.vo[sky] = sy;
            }
        };
        RawD cur, nxt;
        loadd(jend - 1, cur);
        for (int jc = jend - 1; jc >= j0 - 1; --jc) {
            if (jc - 1 >= j0 - 1) loadd(jc - 1, nxt);  // prefetch
            flush();
            sxf = syf = false;
            finish_cells<R, false>(a, lane, i, jc - 2, edge, cur.u0, lane_next(cur.u0), cur.v0, cur.v1, gx, wU[0], wV[0]);
            const R vsf = cur.vs;
            const R fsU = quick_flux<R>(vsf, wU[0], wU[1], wU[2], wU[3]);
            const R fsV = quick_flux<R>(vsf, wV[0], wV[1], wV[2], wV[3]);
            const R Uc = wU[2], Vc = wV[2];
            const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
            const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
            const R uw = cur.uw;
            const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
            const R feU = lane_next(fwU), feV = lane_next(fwV);
            const R CU = (feU - fwU) * idx + (fnU - fsU) * idy;
            const R CV = (feV - fwV) * idx + (fnV - fsV) * idy;
            const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (wU[3] - R(2) * Uc + wU[1]) * idy2) * iren;
            const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (wV[3] - R(2) * Vc + wV[1]) * idy2) * iren;
            // EXPL: the domain ghosts of fluxTotal keep setVal's 0 (no wall-ghost pass)
            const bool tcell = xin && jc >= 0 && jc < n;
            const R TU = EXPL && !tcell ? R(0) : (R(0) - CU) + VU, TV = EXPL && !tcell ? R(0) : (R(0) - CV) + VV;
            const R TUw = lane_prev(TU);
            const R pc = cur.pc, pw = lane_prev(pc);
            if (jc >= j0 && outlane && xface) {  // x-face (i, jc); wall faces get R = 0 exactly (EXPL: T/2)
                skx = (unsigned)(i < 0 ? 0 : i) + (unsigned)P * jc;
                const bool wall = i == 0 || i == n;
                const double Rx = EXPL   ? R(0.5) * (TUw + TU)
                                 : wall  ? 0.0
                                 : FOLDP ? R(0.5) * (TUw + TU)
                                         : -(pc - pw) * a.inv_dx + R(0.5) * (TUw + TU);
                sxf = true;
                if (MODE == 2) sx = Rx;
                else if (MODE == 3) {
                    fx0 = (float)uw;
                    fx1 = (float)(Rx - a.c15 * uw + cur.cu);
                } else {
                    const double unew = cur.ub + a.rk * (Rx - a.c15 * uw + cur.cu);
                    sx = unew;
                    if (MODE == 1) { const double d = cur.ub - unew; nx_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
                }
            }
            if (jc + 1 < jend && jc + 1 >= j0 && outlane && xin) {  // y-face (i, jc + 1)
                sky = (unsigned)i + (unsigned)P * (jc + 1);
                const bool wall = jc + 1 == 0;
                const double Ry = EXPL   ? R(0.5) * (TV + tVup)
                                 : wall  ? 0.0
                                 : FOLDP ? R(0.5) * (TV + tVup)
                                         : -(pup - pc) * a.inv_dy + R(0.5) * (TV + tVup);
                syf = true;
                if (MODE == 2) sy = Ry;
                else if (MODE == 3) {
                    fy0 = (float)vup;
                    fy1 = (float)(Ry - a.c15 * vup + cvup);
                } else {
                    const double vnew = vbup + a.rk * (Ry - a.c15 * vup + cvup);
                    sy = vnew;
                    if (MODE == 1) { const double d = vbup - vnew; ny_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
                }
            }
            if (jc == jend - 1) tVtop = TV;
            fnU = fsU; fnV = fsV; tVup = TV; pup = pc; vup = vsf; vbup = cur.vb; cvup = cur.cv;
#pragma unroll
            for (int q = 4; q > 0; --q) { wU[q] = wU[q - 1]; wV[q] = wV[q - 1]; }
            cur = nxt;
        }
        flush();
        if (jend == n && outlane && xin) {  // north wall row of y-faces: R = 0 (EXPL: T(n - 1)/2)
            const size_t k = i + (size_t)P * n;
            if (MODE == 2) a.vo[k] = 0.0;
            else if (MODE == 3) {
                a.vk[k] = (float)vtop;
                a.rv[k] = (float)(-a.c15 * vtop + a.cv[k]);
            } else {
                const double base = a.vb[k];
                const double vnew = base + a.rk * ((EXPL ? R(0.5) * tVtop : 0.0) - a.c15 * vtop + a.cv[k]);
                a.vo[k] = vnew;
                if (MODE == 1) { const double d = base - vnew; ny_ += fabs(d * d * 0.5); }
            }
        }
        if (MODE == 1) {
            for (int o = 32; o > 0; o >>= 1) { nx_ += __shfl_down(nx_, o, 64); ny_ += __shfl_down(ny_, o, 64); }
            if (lane == 0) { a.partials[2 * wid] = nx_; a.partials[2 * wid + 1] = ny_; }
        }
        return;
    }
    {
        Raw r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            load(j0 - 5 + q, r);
            finish_cells<R, false>(a, lane, i, j0 - 3 + q, edge, r.u0, lane_next(r.u0), r.v0, r.v1, gx, wU[q], wV[q]);
        }
    }
    R vn = (xin && j0 - 1 >= 0) ? vs[i + (size_t)P * (j0 - 1)] : R(0);  // own y-face of the row
    R fsU = quick_flux<R>(vn, wU[0], wU[1], wU[2], wU[3]), fsV = quick_flux<R>(vn, wV[0], wV[1], wV[2], wV[3]);
    R tVprev = 0, pprev = 0;
    double nx_ = 0.0, ny_ = 0.0;
    // Row jc's stores are issued at the top of iteration jc + 1, right after
    // that iteration's prefetch.  With loads and stores both in flight the
    // compiler waits for all vector memory operations at once, so the loop is
    // laid out to need one such wait per row (before the row's arithmetic
    // reads the prefetched values), and the loads and stores of a row overlap
    // a full row of arithmetic.
    unsigned sk = 0;  // n <= 16384: every index fits 32 bits
    bool sxf = false, syf = false;
    double sx = 0.0, sy = 0.0;                  // MODE 0-2: the face values
    float fx0 = 0.f, fx1 = 0.f, fy0 = 0.f, fy1 = 0.f;  // MODE 3: FP32 base and residual
    auto flush = [&]() {
        if (sxf) {
            if (MODE == 3) { a.uk[sk] = fx0; a.ru[sk] = fx1; }
            else a.uo[sk] = sx;
        }
        if (syf) {
            if (MODE == 3) { a.vk[sk] = fy0; a.rv[sk] = fy1; }
            else a.vo[sk] = sy;
        }
    };
    Raw cur, nxt;
    load(j0 - 1, cur);
    for (int jc = j0 - 1; jc < jend; ++jc) {
        if (jc + 1 < jend) load(jc + 1, nxt);  // prefetch
        flush();
        sxf = syf = false;
        finish_cells<R, false>(a, lane, i, jc + 2, edge, cur.u0, lane_next(cur.u0), cur.v0, cur.v1, gx, wU[4], wV[4]);
        const R vnf = cur.vn;
        const R fnU = quick_flux<R>(vnf, wU[1], wU[2], wU[3], wU[4]);
        const R fnV = quick_flux<R>(vnf, wV[1], wV[2], wV[3], wV[4]);
        const R Uc = wU[2], Vc = wV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R uw = cur.uw;
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - fsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - fsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (wU[3] - R(2) * Uc + wU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (wV[3] - R(2) * Vc + wV[1]) * idy2) * iren;
        // EXPL: the domain ghosts of fluxTotal keep setVal's 0 (no wall-ghost pass)
        const bool tcell = xin && jc >= 0 && jc < n;
        const R TU = EXPL && !tcell ? R(0) : (R(0) - CU) + VU, TV = EXPL && !tcell ? R(0) : (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        const R pc = cur.pc, pw = lane_prev(pc);
        if (jc >= j0 && outlane) {
            sk = (unsigned)(i < 0 ? 0 : i) + (unsigned)P * jc;
            if (xface) {  // x-face (i, jc); wall faces get R = 0 exactly (EXPL: T/2)
                const bool wall = i == 0 || i == n;
                const double Rx = EXPL   ? R(0.5) * (TUw + TU)
                                 : wall  ? 0.0
                                 : FOLDP ? R(0.5) * (TUw + TU)
                                         : -(pc - pw) * a.inv_dx + R(0.5) * (TUw + TU);
                sxf = true;
                if (MODE == 2) sx = Rx;
                else if (MODE == 3) {
                    fx0 = (float)uw;
                    fx1 = (float)(Rx - a.c15 * uw + cur.cu);
                } else {
                    const double unew = cur.ub + a.rk * (Rx - a.c15 * uw + cur.cu);
                    sx = unew;
                    if (MODE == 1) { const double d = cur.ub - unew; nx_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
                }
            }
            if (xin) {  // y-face (i, jc)
                const bool wall = jc == 0;
                const double Ry = EXPL   ? R(0.5) * (tVprev + TV)
                                 : wall  ? 0.0
                                 : FOLDP ? R(0.5) * (tVprev + TV)
                                         : -(pc - pprev) * a.inv_dy + R(0.5) * (tVprev + TV);
                syf = true;
                if (MODE == 2) sy = Ry;
                else if (MODE == 3) {
                    fy0 = (float)vn;
                    fy1 = (float)(Ry - a.c15 * vn + cur.cv);
                } else {
                    const double vnew = cur.vb + a.rk * (Ry - a.c15 * vn + cur.cv);
                    sy = vnew;
                    if (MODE == 1) { const double d = cur.vb - vnew; ny_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
                }
            }
        }
        fsU = fnU; fsV = fnV; tVprev = TV; pprev = pc; vn = vnf;
#pragma unroll
        for (int q = 0; q < 4; ++q) { wU[q] = wU[q + 1]; wV[q] = wV[q + 1]; }
        cur = nxt;
    }
    flush();
    if (jend == n && outlane && xin) {  // north wall row of y-faces: R = 0 (EXPL: T(n - 1)/2)
        const size_t k = i + (size_t)P * n;
        if (a.mode == 2) a.vo[k] = 0.0;
        else if (a.mode == 3) {
            a.vk[k] = (float)vn;
            a.rv[k] = (float)(-a.c15 * vn + a.cv[k]);
        } else {
            const double base = a.vb[k];
            const double vnew = base + a.rk * ((EXPL ? R(0.5) * tVprev : 0.0) - a.c15 * vn + a.cv[k]);
            a.vo[k] = vnew;
            if (a.mode == 1) { const double d = base - vnew; ny_ += fabs(d * d * 0.5); }
        }
    }
    if (a.mode == 1) {
        for (int o = 32; o > 0; o >>= 1) { nx_ += __shfl_down(nx_, o, 64); ny_ += __shfl_down(ny_, o, 64); }
        if (lane == 0) { a.partials[2 * wid] = nx_; a.partials[2 * wid + 1] = ny_; }
    }
}

// ================================================= FP32 tangent-linear step
// Tangent-linear flux: derivative of QUICK(vel, X) in direction (dvel, dX)
// with the upwind choice frozen at vel: vel*phi(dX) + dvel*phi(X).
__device__ __forceinline__ float quick_tl(float vel, float dvel, float a2, float a1, float b0, float b1, float da2,
                                          float da1, float db0, float db1)
{
    const bool pos = vel >= 0.f;
    // phi = 3/8 downwind + 6/8 upwind - 1/8 far-upwind
    const float xd = pos ? b0 : a1, xu = pos ? a1 : b0, xf = pos ? a2 : b1;
    const float dd = pos ? db0 : da1, du = pos ? da1 : db0, df = pos ? da2 : db1;
    const float phid = 0.375f * dd + 0.75f * du - 0.125f * df;
    const float phix = 0.375f * xd + 0.75f * xu - 0.125f * xf;
    return vel * phid + dvel * phix;
}

// u^{k+1} = u^k + cr r + cj R'(u^k) r in FP64 with R' r formed in FP32 at the
// FP32 base; norm partials of the change.
__global__ __launch_bounds__(256) void k_tangent(StageArgs a)
{
    using R = float;
    if (a.ctl && a.ctl->mom_done) return;
    const int lane = threadIdx.x & 63;
    const int wid = wave_id(a);
    if (wid >= a.nbands * a.nchunks) return;
    const int band = a.order ? wid % a.nbands : wid / a.nchunks;
    const int chunk = a.order ? wid / a.nbands : wid % a.nchunks;
    const int n = a.n, P = a.P;
    const int i0 = band * WB, j0 = chunk * a.rows;
    const int i = i0 + lane - 3;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool edge = i0 < 3 || i0 + 61 > n;
    const float *uk = a.uk, *vk = a.vk, *ru = a.ru, *rv = a.rv;
    const R gx = (R)((xin && a.lid) ? a.lid[i] : a.inflow0);
    const bool outlane = lane >= 3 && lane < 3 + WB;
    struct Raw {
        R u0, u1, v0, v1, d0, d1, e0, e1, vn, dvn, uw, duw;
        double ub, vb;  // FP64 base of the row's own faces, updated below
    };
    auto load = [&](int jc, Raw &r) {
        const int jr = jc + 2;
        const int jm = jr < 0 ? -jr - 1 : (jr >= n ? 2 * n - 1 - jr : jr);
        r.u0 = r.u1 = r.v0 = r.v1 = r.d0 = r.d1 = r.e0 = r.e1 = 0;
        if (jm >= 0 && jm < n && xin) {
            const size_t k = i + (size_t)P * jm;
            r.u0 = uk[k]; r.u1 = uk[k + 1]; r.v0 = vk[k]; r.v1 = vk[k + P];
            r.d0 = ru[k]; r.d1 = ru[k + 1]; r.e0 = rv[k]; r.e1 = rv[k + P];
        }
        const bool vok = xin && jc + 1 >= 0 && jc + 1 <= n;
        r.vn = vok ? vk[i + (size_t)P * (jc + 1)] : R(0);
        r.dvn = vok ? rv[i + (size_t)P * (jc + 1)] : R(0);
        const bool row = jc >= 0 && jc < n;
        const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * (jc < 0 ? 0 : jc);
        r.uw = (xface && row) ? uk[k] : R(0);
        r.duw = (xface && row) ? ru[k] : R(0);
    };
    const R idx = (R)a.inv_dx, idy = (R)a.inv_dy, idx2 = (R)a.inv_dx2, idy2 = (R)a.inv_dy2, iren = (R)a.inv_ren;
    R wU[5], wV[5], dU[5], dV[5];
    {
        Raw r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            load(j0 - 5 + q, r);
            finish_cells<R, false>(a, lane, i, j0 - 3 + q, edge, r.u0, r.u1, r.v0, r.v1, gx, wU[q], wV[q]);
            finish_cells<R, true>(a, lane, i, j0 - 3 + q, edge, r.d0, r.d1, r.e0, r.e1, R(0), dU[q], dV[q]);
        }
    }
    const bool v0ok = xin && j0 - 1 >= 0;
    R vn = v0ok ? vk[i + (size_t)P * (j0 - 1)] : R(0), dvn = v0ok ? rv[i + (size_t)P * (j0 - 1)] : R(0);
    R fsU = quick_tl(vn, dvn, wU[0], wU[1], wU[2], wU[3], dU[0], dU[1], dU[2], dU[3]);
    R fsV = quick_tl(vn, dvn, wV[0], wV[1], wV[2], wV[3], dV[0], dV[1], dV[2], dV[3]);
    R tVprev = 0;
    double nx_ = 0.0, ny_ = 0.0;
    const int jend = min(j0 + a.rows, n);
    const double cr = a.cr, cj = a.cj;
    Raw cur, nxt;
    load(j0 - 1, cur);
    for (int jc = j0 - 1; jc < jend; ++jc) {
        if (jc + 1 < jend) load(jc + 1, nxt);
        finish_cells<R, false>(a, lane, i, jc + 2, edge, cur.u0, cur.u1, cur.v0, cur.v1, gx, wU[4], wV[4]);
        finish_cells<R, true>(a, lane, i, jc + 2, edge, cur.d0, cur.d1, cur.e0, cur.e1, R(0), dU[4], dV[4]);
        const R fnU = quick_tl(cur.vn, cur.dvn, wU[1], wU[2], wU[3], wU[4], dU[1], dU[2], dU[3], dU[4]);
        const R fnV = quick_tl(cur.vn, cur.dvn, wV[1], wV[2], wV[3], wV[4], dV[1], dV[2], dV[3], dV[4]);
        const R Uc = wU[2], Vc = wV[2], dUc = dU[2], dVc = dV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R dUm1 = lane_prev(dUc), dUm2 = lane_prev(dUm1), dUp1 = lane_next(dUc);
        const R dVm1 = lane_prev(dVc), dVm2 = lane_prev(dVm1), dVp1 = lane_next(dVc);
        const R fwU = quick_tl(cur.uw, cur.duw, Um2, Um1, Uc, Up1, dUm2, dUm1, dUc, dUp1);
        const R fwV = quick_tl(cur.uw, cur.duw, Vm2, Vm1, Vc, Vp1, dVm2, dVm1, dVc, dVp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - fsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - fsV) * idy;
        const R VU = ((dUp1 - R(2) * dUc + dUm1) * idx2 + (dU[3] - R(2) * dUc + dU[1]) * idy2) * iren;
        const R VV = ((dVp1 - R(2) * dVc + dVm1) * idx2 + (dV[3] - R(2) * dVc + dV[1]) * idy2) * iren;
        const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        if (jc >= j0 && outlane) {
            const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * jc;
            if (xface) {
                const bool wall = i == 0 || i == n;
                const float Jr = wall ? 0.f : 0.5f * (TUw + TU);
                const double du = cr * (double)cur.duw + cj * (double)Jr;
                a.u64[k] += du;
                nx_ += du * du * (wall ? 0.5 : 1.0);
            }
            if (xin) {
                const bool wall = jc == 0;
                const float Jr = wall ? 0.f : 0.5f * (tVprev + TV);
                const double dv = cr * (double)dvn + cj * (double)Jr;
                a.v64[k] += dv;
                ny_ += dv * dv * (wall ? 0.5 : 1.0);
            }
        }
        fsU = fnU; fsV = fnV; tVprev = TV; vn = cur.vn; dvn = cur.dvn;
#pragma unroll
        for (int q = 0; q < 4; ++q) { wU[q] = wU[q + 1]; wV[q] = wV[q + 1]; dU[q] = dU[q + 1]; dV[q] = dV[q + 1]; }
        cur = nxt;
    }
    if (jend == n && outlane && xin) {  // north wall row: R' r = 0 there
        const size_t k = i + (size_t)P * n;
        const double dv = cr * (double)rv[k];
        a.v64[k] += dv;
        ny_ += dv * dv * 0.5;
    }
    for (int o = 32; o > 0; o >>= 1) { nx_ += __shfl_down(nx_, o, 64); ny_ += __shfl_down(ny_, o, 64); }
    if (lane == 0) { a.partials[2 * wid] = nx_; a.partials[2 * wid + 1] = ny_; }
}

static StageArgs base_args(const Ctx &c, double rk, int mode, int rows)
{
    const Geometry &g = c.g;
    StageArgs a{};
    a.lid = c.lid;
    a.partials = c.partials;
    a.ctl = c.ctl;
    a.n = g.n;
    a.P = g.P;
    a.inv_dx = 1.0 / g.dx;
    a.inv_dy = 1.0 / g.dy;
    a.inv_dx2 = 1.0 / (g.dx * g.dx);
    a.inv_dy2 = 1.0 / (g.dy * g.dy);
    a.inv_ren = 1.0 / g.ren;
    a.c15 = 1.5 / g.dt;
    a.rk = rk;
    a.bc_lo0 = g.bc_lo0; a.bc_hi0 = g.bc_hi0; a.bc_lo1 = g.bc_lo1; a.bc_hi1 = g.bc_hi1;
    a.inflow0 = g.inflow0; a.inflow1 = g.inflow1;
    a.mode = mode;
    a.rows = rows;
    a.order = c.stage_order;
    a.alt = c.stage_alt;
    a.nbands = (g.n + 1 + WB - 1) / WB;
    a.nchunks = (g.n + a.rows - 1) / a.rows;
    return a;
}

static dim3 stage_grid(StageArgs &a)
{
    return dim3((a.nbands * a.nchunks + 3) / 4);
}


hipError_t launch_stage(const Ctx &c, const StageLaunch &s, int *nblocks)
{
    StageArgs a = base_args(c, s.rk, s.mode, c.stage_rows);
    a.us = s.us; a.vs = s.vs; a.ub = s.ub; a.vb = s.vb; a.cu = s.cu; a.cv = s.cv; a.p = s.p;
    a.uo = s.uo; a.vo = s.vo;
    if (s.mode == 2 || s.unguarded) a.ctl = nullptr;
    const dim3 grid = stage_grid(a);
    if ((int64_t)a.nbands * a.nchunks > kMaxPartials) return hipErrorInvalidValue;
    if (nblocks) *nblocks = a.nbands * a.nchunks;  // one norm partial per wave
    if (s.foldp && s.mode > 1) return hipErrorInvalidValue;
    switch (s.mode) {
    case 0:
        if (s.foldp) hipLaunchKernelGGL((k_stage64<0, true>), grid, dim3(256), 0, c.stream, a);
        else hipLaunchKernelGGL((k_stage64<0, false>), grid, dim3(256), 0, c.stream, a);
        break;
    case 1:
        if (s.expl) {
            if (!s.foldp) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_stage64<1, true, true>), grid, dim3(256), 0, c.stream, a);
        } else if (s.foldp) hipLaunchKernelGGL((k_stage64<1, true>), grid, dim3(256), 0, c.stream, a);
        else hipLaunchKernelGGL((k_stage64<1, false>), grid, dim3(256), 0, c.stream, a);
        break;
    case 2: hipLaunchKernelGGL((k_stage64<2, false>), grid, dim3(256), 0, c.stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stage_mixed(const Ctx &c, const MixedStageLaunch &s, int *nblocks)
{
    StageArgs a = base_args(c, 0.0, 3, s.stage == 0 ? c.stage_rows : c.tangent_rows);
    if (s.unguarded) a.ctl = nullptr;
    const dim3 grid = stage_grid(a);
    if ((int64_t)a.nbands * a.nchunks > kMaxPartials) return hipErrorInvalidValue;
    if (nblocks) *nblocks = a.nbands * a.nchunks;
    a.uk = s.uk; a.vk = s.vk; a.ru = s.ru; a.rv = s.rv;
    if (s.stage == 0) {  // FP64 residual of the pseudo base
        a.us = s.u64; a.vs = s.v64; a.cu = s.cu; a.cv = s.cv; a.p = s.p;
        if (s.foldp) hipLaunchKernelGGL((k_stage64<3, true>), grid, dim3(256), 0, c.stream, a);
        else hipLaunchKernelGGL((k_stage64<3, false>), grid, dim3(256), 0, c.stream, a);
    } else {  // FP32 tangent-linear correction, base updated in place
        a.u64 = s.u64; a.v64 = s.v64; a.cr = s.cr; a.cj = s.cj;
        hipLaunchKernelGGL(k_tangent, grid, dim3(256), 0, c.stream, a);
    }
    return hipGetLastError();
}

}  // namespace amr
