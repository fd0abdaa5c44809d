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
#include <type_traits>

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
    double rk2;                // k_stage_pair: the second stage's coefficient
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
    int edge_first;  // pair kernels: waves on the general march dispatched first (pair_wave)
    int eb0, eb1, ec0, ec1;  // pair kernels: leading / trailing bands and chunks on the general march
    int fast;   // k_stage64: interior waves take the unrolled path (rows 8 or 16)
};

// wave id -> (band, chunk); blocks of 4 waves, handed to the 8 XCDs in
// contiguous ranges so vertically adjacent chunks share an L2
__device__ __forceinline__ int wave_id(const StageArgs &)
{
    // wave-uniform by construction; readfirstlane lets the compiler keep the
    // tile coordinates, row offsets and wave-level branches in scalar registers
    return __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6));
}

// Wave -> (band, chunk, id) of a pair kernel with the waves on the general
// march first.  Those waves (every band of the leading ec0 and trailing ec1
// chunks, the leading eb0 and trailing eb1 bands of the other chunks) take
// about twice as long as an interior wave on the unrolled march, and alone
// about as long as the interior waves together; dispatched last they would
// set the kernel's tail.  They fill the first ceil(E / 4) blocks, which the
// dispatcher hands round-robin to the 8 XCDs.  The interior waves fill the
// rest in chunk-major order, each XCD a contiguous range of them (its
// vertically adjacent chunks share its L2).  id (the norm partial's slot) is
// e for edge wave e and E + k for interior wave k: dense in [0, nb nc).
// false: a padding wave of the last edge block or past the last interior one.
__device__ __forceinline__ bool pair_wave(const StageArgs &a, int &band, int &chunk, int &id)
{
    const int nb = a.nbands, nc = a.nchunks;
    const int nbi = nb - a.eb0 - a.eb1, nci = nc - a.ec0 - a.ec1;
    const int I = nbi * nci, E = nb * nc - I;
    const int Be = (E + 3) >> 2;  // blocks of edge waves
    const int b = blockIdx.x, w = (int)(threadIdx.x >> 6);
    if (b < Be) {
        const int e = 4 * b + w;
        if (e >= E) return false;
        id = e;
        const int ne = (a.ec0 + a.ec1) * nb;  // waves of the edge chunks
        if (e < ne) {
            const int c = e / nb;
            band = e - c * nb;
            chunk = c < a.ec0 ? c : nc - a.ec1 + (c - a.ec0);
        } else {
            const int f = e - ne, eb = a.eb0 + a.eb1;
            const int bb = f % eb;
            chunk = a.ec0 + f / eb;
            band = bb < a.eb0 ? bb : nb - a.eb1 + (bb - a.eb0);
        }
        return true;
    }
    // interior block: XCD x = b mod 8, its rank among that XCD's interior blocks
    const int B = gridDim.x, q = B >> 3, r = B & 7, x = b & 7;
    const int qe = Be >> 3, re = Be & 7;
    const int rank = (b >> 3) - (qe + (x < re ? 1 : 0));
    const int base = (x * q + min(x, r)) - (x * qe + min(x, re));
    const int k = 4 * (base + rank) + w;
    if (k >= I) return false;
    id = E + k;
    band = a.eb0 + k % nbi;
    chunk = a.ec0 + k / nbi;
    return true;
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

// ============================================ fully unrolled interior march
// The row march of k_stage64 for a wave whose band and rows touch no wall,
// ghost cell or array edge: every lane's column and every row of the chunk's
// stencil (cells j0 - 3 .. j0 + ROWS + 1, y-faces up to j0 + ROWS + 2) lie
// inside the domain.  With ROWS a compile-time constant the loop unrolls
// fully, so the rolling window, the carried fluxes and the face rows are
// register renames rather than copies, every load is unconditional, and each
// face row is loaded once (u(jc + 3) and v(jc + 4) one iteration ahead of
// their use; the x-face of a row is reused as the cell velocity of that row).
// Per face the arithmetic is k_stage64's (FOLDP, no explicit form), so the two
// paths give identical results.  MODE 0 stage, 1 last stage + norm partials,
// 3 mixed residual.
template <int ROWS, int MODE>
__device__ __forceinline__ void stage_interior(const StageArgs &a, int lane, int i, int j0, int wid)
{
    using R = double;
    const size_t P = a.P;
    const bool outlane = lane >= 3 && lane < 3 + WB;
    // i >= 0 here: every access is a uniform row base plus this lane's 32-bit byte offset
    const unsigned ob = (unsigned)i * 8u;
    auto at = [&](const double *base, int row) -> const double & {
        return *(const double *)((const char *)(base + P * row) + ob);
    };
    auto put = [&](auto *base, int row, auto val) {
        using T = decltype(val);
        *(T *)((char *)(base + P * row) + ob * sizeof(T) / 8u) = val;
    };
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren;
    const R c15 = a.c15, rk = a.rk;
    auto U_of = [&](R u) { return R(0.5) * (u + lane_next(u)); };
    // rows j0 - 3 .. j0 + 1 of u, j0 - 3 .. j0 + 2 of v
    const double *us = a.us, *vs = a.vs;
    R u0 = at(us, j0 - 3), u1 = at(us, j0 - 2), u2 = at(us, j0 - 1), u3 = at(us, j0), u4 = at(us, j0 + 1);
    R v0 = at(vs, j0 - 3), v1 = at(vs, j0 - 2), v2 = at(vs, j0 - 1), v3 = at(vs, j0), v4 = at(vs, j0 + 1),
      v5 = at(vs, j0 + 2);
    R wU[5], wV[5];  // cell (i, jc - 2 + q)
    wU[0] = U_of(u0); wV[0] = R(0.5) * (v0 + v1);
    wU[1] = U_of(u1); wV[1] = R(0.5) * (v1 + v2);
    wU[2] = U_of(u2); wV[2] = R(0.5) * (v2 + v3);
    wU[3] = U_of(u3); wV[3] = R(0.5) * (v3 + v4);
    // carried rows at the top of iteration jc: uA = u(jc), uB = u(jc + 1), uC = u(jc + 2);
    // vA = v(jc), vB = v(jc + 1), vC = v(jc + 2), vD = v(jc + 3)
    R uA = u2, uB = u3, uC = u4, vA = v2, vB = v3, vC = v4, vD = v5;
    R fsU = quick_flux<R>(vA, wU[0], wU[1], wU[2], wU[3]), fsV = quick_flux<R>(vA, wV[0], wV[1], wV[2], wV[3]);
    R tVprev = 0;
    double nx_ = 0.0, ny_ = 0.0;
    R bu = 0, bcu = 0, bv = 0, bcv = 0;  // base and constant of the row
#pragma unroll
    for (int t = 0; t <= ROWS; ++t) {
        const int jc = j0 - 1 + t;
        // one iteration ahead: the next iteration's new rows and the next row's base and constant
        R un = 0, vn = 0, nbu = 0, ncu = 0, nbv = 0, ncv = 0;
        if (t < ROWS) {
            un = at(us, jc + 3);
            vn = at(vs, jc + 4);
            if (MODE != 3) { nbu = at(a.ub, jc + 1); nbv = at(a.vb, jc + 1); }
            ncu = at(a.cu, jc + 1); ncv = at(a.cv, jc + 1);
        }
        wU[4] = U_of(uC);
        wV[4] = R(0.5) * (vC + vD);
        const R fnU = quick_flux<R>(vB, wU[1], wU[2], wU[3], wU[4]);
        const R fnV = quick_flux<R>(vB, wV[1], wV[2], wV[3], wV[4]);
        const R Uc = wU[2], Vc = wV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R uw = uA;
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - fsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - fsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (wU[3] - R(2) * Uc + wU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (wV[3] - R(2) * Vc + wV[1]) * idy2) * iren;
        const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        if (t > 0 && outlane) {  // row jc >= j0: no wall face in this wave
            const double Rx = R(0.5) * (TUw + TU);
            const double Ry = R(0.5) * (tVprev + TV);
            if (MODE == 3) {
                put(a.uk, jc, (float)uw);
                put(a.ru, jc, (float)(Rx - c15 * uw + bcu));
                put(a.vk, jc, (float)vA);
                put(a.rv, jc, (float)(Ry - c15 * vA + bcv));
            } else {
                const double unew = bu + rk * (Rx - c15 * uw + bcu);
                const double vnew = bv + rk * (Ry - c15 * vA + bcv);
                put(a.uo, jc, unew);
                put(a.vo, jc, vnew);
                if (MODE == 1) {
                    const double du = bu - unew, dv = bv - vnew;
                    nx_ += fabs(du * du * 1.0);
                    ny_ += fabs(dv * dv * 1.0);
                }
            }
        }
        fsU = fnU; fsV = fnV; tVprev = TV;
#pragma unroll
        for (int q = 0; q < 4; ++q) { wU[q] = wU[q + 1]; wV[q] = wV[q + 1]; }
        uA = uB; uB = uC; uC = un;
        vA = vB; vB = vC; vC = vD; vD = vn;
        bu = nbu; bcu = ncu; bv = nbv; bcv = ncv;
    }
    if (MODE == 1) {
        for (int o = 32; o > 0; o >>= 1) { nx_ += __shfl_down(nx_, o, 64); ny_ += __shfl_down(ny_, o, 64); }
        if (lane == 0) { a.partials[2 * wid] = nx_; a.partials[2 * wid + 1] = ny_; }
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
    if (FOLDP && !EXPL && MODE != 2 && !edge && j0 >= 3 && a.fast) {
        // the whole stencil inside the domain: the unrolled path
        if (a.rows == 8 && j0 + 8 + 2 <= n) { stage_interior<8, MODE>(a, lane, i, j0, wid); return; }
        if (a.rows == 16 && j0 + 16 + 2 <= n) { stage_interior<16, MODE>(a, lane, i, j0, wid); return; }
    }
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

// ================================================= two RK4 stages in one pass
// Stages s and s + 1 of an FP64 pseudo-iteration (the constant holds -G p) in
// one launch.  A wave owns 52 output columns (a 6-lane halo on each side) and a
// chunk of rows, and runs two upward marches in one loop.  March A is
// k_stage64's stage s (stencil input, base and constant from memory) three
// rows ahead of march B, stage s + 1, which takes A's faces from registers (a
// delay line of four x-face rows and three y-face rows) instead of memory.  A
// is exact on lanes 3..60, so B is on lanes 6..57.  The cells B needs outside
// the domain are the wall ghosts of its own window cells (the ghost of cell row
// -1 - j or 2n - 1 - j is the ghost rule applied to cell row j).  Per face the
// arithmetic is k_stage64's; the memory traffic is one stage's: the stencil
// input, base and constant in and stage s + 1 out, 64 B per cell (B's reads
// of the base and constant repeat A's three rows later and hit in cache).
// The output must not alias the stencil input or the base: A reads both in the
// halo rows and columns that neighbouring waves write.
//
// LAST (the iteration's second pair): norm partials of the change, and when
// the loop has already converged (the launch is skipped) the base is copied to
// the output, so that the caller's exchange of the two buffers keeps the
// converged state.
constexpr int WB2 = 52;  // output columns per wave of k_stage_pair

// The pair for a wave whose whole stencil lies inside the domain (no lane or
// row of A's march, cells j0 - 6 .. j0 + ROWS + 4, touches a wall, a ghost
// cell or an array edge), fully unrolled as stage_interior: A runs its rows
// j0 - 3 .. j0 + ROWS + 2 (t = 1 .. ROWS + 6 after a first iteration without
// output), B its rows j0 - 1 .. j0 + ROWS - 1 three rows behind (t >= 6).
template <int ROWS, bool LAST>
__device__ __forceinline__ void pair_interior(const StageArgs &a, int lane, int i, int j0, int wid)
{
    using R = double;
    const size_t P = a.P;
    const bool outB = lane >= 6 && lane <= 57;
    const unsigned ob = (unsigned)i * 8u;
    auto at = [&](const double *base, int row) -> const double & {
        return *(const double *)((const char *)(base + P * row) + ob);
    };
    auto put = [&](double *base, int row, double val) { *(double *)((char *)(base + P * row) + ob) = val; };
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren;
    const R c15 = a.c15, rkA = a.rk, rkB = a.rk2;
    auto U_of = [&](R u) { return R(0.5) * (u + lane_next(u)); };
    const double *us = a.us, *vs = a.vs;
    // the stencil's T = -C + V at cell (i, row) from the row's window (w[1..3] the rows
    // below, at, above; w[0], w[4] two away), its two y-face fluxes and its west x-face velocity
    auto T_of = [&](const R *wU, const R *wV, R fsU, R fsV, R fnU, R fnV, R uw, R &TU, R &TV) {
        const R Uc = wU[2], Vc = wV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - fsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - fsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (wU[3] - R(2) * Uc + wU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (wV[3] - R(2) * Vc + wV[1]) * idy2) * iren;
        TU = (R(0) - CU) + VU;
        TV = (R(0) - CV) + VV;
    };
    // ---- A: rows j0 - 6 .. j0 - 2 of u, j0 - 6 .. j0 - 1 of v
    const int jA = j0 - 3;  // A's first output row
    R aU[5], aV[5];
    {
        const R u0 = at(us, jA - 3), u1 = at(us, jA - 2), u2 = at(us, jA - 1), u3 = at(us, jA);
        const R v0 = at(vs, jA - 3), v1 = at(vs, jA - 2), v2 = at(vs, jA - 1), v3 = at(vs, jA), v4 = at(vs, jA + 1);
        aU[0] = U_of(u0); aV[0] = R(0.5) * (v0 + v1);
        aU[1] = U_of(u1); aV[1] = R(0.5) * (v1 + v2);
        aU[2] = U_of(u2); aV[2] = R(0.5) * (v2 + v3);
        aU[3] = U_of(u3); aV[3] = R(0.5) * (v3 + v4);
    }
    R uA = at(us, jA - 1), uB = at(us, jA), uC = at(us, jA + 1);
    R vA = at(vs, jA - 1), vB = at(vs, jA), vC = at(vs, jA + 1), vD = at(vs, jA + 2);
    R afsU = quick_flux<R>(vA, aU[0], aU[1], aU[2], aU[3]), afsV = quick_flux<R>(vA, aV[0], aV[1], aV[2], aV[3]);
    R atV = 0;
    R pu = 0, pcu = 0, pv = 0, pcv = 0;  // A's base and constant of its row
    // the same of A's three previous rows (the last is B's row): B takes them
    // from here rather than loading them again
    R d0u = 0, d0c = 0, d0v = 0, d0w = 0, d1u = 0, d1c = 0, d1v = 0, d1w = 0, d2u = 0, d2c = 0, d2v = 0, d2w = 0;
    // ---- B
    R bU[5], bV[5];
    R x0 = 0, x1 = 0, x2 = 0, x3 = 0, y0 = 0, y1 = 0, y2 = 0, y3 = 0;  // A's rows j0 - 3 .. j0
    R uq0 = 0, uq1 = 0, uq2 = 0, vq1 = 0, vq2 = 0, bvn = 0;           // as in k_stage_pair
    R bfsU = 0, bfsV = 0, btV = 0;
    double nx_ = 0.0, ny_ = 0.0;
#pragma unroll
    for (int t = 0; t <= ROWS + 6; ++t) {
        // keep each iteration's instructions in it: the prefetch below is the
        // only look-ahead, so the live registers stay those of one iteration
        __builtin_amdgcn_sched_barrier(0);
        const int jc = jA - 1 + t;  // A's row
        const int jb = jc - 3;      // B's row (t >= 6)
        const R bub = d2u, bcu = d2c, bvb = d2v, bcv = d2w;  // B's base and constant (A's row jc - 3)
        R un = 0, vn = 0, nbu = 0, ncu = 0, nbv = 0, ncv = 0;
        if (t < ROWS + 6) {  // A's next iteration
            un = at(us, jc + 3);
            vn = at(vs, jc + 4);
            nbu = at(a.ub, jc + 1); ncu = at(a.cu, jc + 1); nbv = at(a.vb, jc + 1); ncv = at(a.cv, jc + 1);
        }
        // A, row jc
        aU[4] = U_of(uC);
        aV[4] = R(0.5) * (vC + vD);
        const R fnU = quick_flux<R>(vB, aU[1], aU[2], aU[3], aU[4]);
        const R fnV = quick_flux<R>(vB, aV[1], aV[2], aV[3], aV[4]);
        R TU, TV;
        T_of(aU, aV, afsU, afsV, fnU, fnV, uA, TU, TV);
        const R TUw = lane_prev(TU);
        const R eu = pu + rkA * (R(0.5) * (TUw + TU) - c15 * uA + pcu);   // A's x-face (i, jc)
        const R ev = pv + rkA * (R(0.5) * (atV + TV) - c15 * vA + pcv);   // A's y-face (i, jc)
        afsU = fnU; afsV = fnV; atV = TV;
#pragma unroll
        for (int q = 0; q < 4; ++q) { aU[q] = aU[q + 1]; aV[q] = aV[q + 1]; }
        uA = uB; uB = uC; uC = un;
        vA = vB; vB = vC; vC = vD; vD = vn;
        d2u = d1u; d2c = d1c; d2v = d1v; d2w = d1w;
        d1u = d0u; d1c = d0c; d1v = d0v; d1w = d0w;
        d0u = pu; d0c = pcu; d0v = pv; d0w = pcv;
        pu = nbu; pcu = ncu; pv = nbv; pcv = ncv;
        __builtin_amdgcn_sched_barrier(0);  // A's temporaries end here
        // B
        if (t == 1) { x0 = eu; y0 = ev; }
        if (t == 2) { x1 = eu; y1 = ev; bU[0] = U_of(x0); bV[0] = R(0.5) * (y0 + y1); }
        if (t == 3) { x2 = eu; y2 = ev; bU[1] = U_of(x1); bV[1] = R(0.5) * (y1 + y2); }
        if (t == 4) { x3 = eu; y3 = ev; bU[2] = U_of(x2); bV[2] = R(0.5) * (y2 + y3); }
        if (t == 5) {  // row j0 + 1: B's prologue is complete
            bU[3] = U_of(x3); bV[3] = R(0.5) * (y3 + ev);
            bvn = y2;
            uq0 = x2; uq1 = x3; uq2 = eu; vq1 = y3; vq2 = ev;
            bfsU = quick_flux<R>(bvn, bU[0], bU[1], bU[2], bU[3]);
            bfsV = quick_flux<R>(bvn, bV[0], bV[1], bV[2], bV[3]);
        }
        if (t >= 6) {  // row jb + 3 from A; B's row jb
            const R uq3 = eu, vq3 = ev;
            bU[4] = U_of(uq2);
            bV[4] = R(0.5) * (vq2 + vq3);
            const R gnU = quick_flux<R>(vq1, bU[1], bU[2], bU[3], bU[4]);
            const R gnV = quick_flux<R>(vq1, bV[1], bV[2], bV[3], bV[4]);
            R SU, SV;
            T_of(bU, bV, bfsU, bfsV, gnU, gnV, uq0, SU, SV);
            const R SUw = lane_prev(SU);
            if (t >= 7 && outB) {  // B's row jb >= j0: no wall face in this wave
                const double unew = bub + rkB * (R(0.5) * (SUw + SU) - c15 * uq0 + bcu);
                const double vnew = bvb + rkB * (R(0.5) * (btV + SV) - c15 * bvn + bcv);
                put(a.uo, jb, unew);
                put(a.vo, jb, vnew);
                if (LAST) {
                    const double du = bub - unew, dv = bvb - vnew;
                    nx_ += fabs(du * du * 1.0);
                    ny_ += fabs(dv * dv * 1.0);
                }
            }
            bfsU = gnU; bfsV = gnV; btV = SV; bvn = vq1;
#pragma unroll
            for (int q = 0; q < 4; ++q) { bU[q] = bU[q + 1]; bV[q] = bV[q + 1]; }
            uq0 = uq1; uq1 = uq2; uq2 = uq3;
            vq1 = vq2; vq2 = vq3;
        }
    }
    if (LAST) {
        for (int o = 32; o > 0; o >>= 1) { nx_ += __shfl_down(nx_, o, 64); ny_ += __shfl_down(ny_, o, 64); }
        if (lane == 0) { a.partials[2 * wid] = nx_; a.partials[2 * wid + 1] = ny_; }
    }
}

template <bool LAST>
__global__ __launch_bounds__(256) void k_stage_pair(StageArgs a)
{
    using R = double;
    const int lane = threadIdx.x & 63;
    int band, chunk, wid;
    if (a.edge_first) {
        if (!pair_wave(a, band, chunk, wid)) return;  // whole wave
    } else {
        wid = wave_id(a);
        if (wid >= a.nbands * a.nchunks) return;  // whole wave
        band = a.order ? wid % a.nbands : wid / a.nchunks;
        chunk = a.order ? wid / a.nbands : wid % a.nchunks;
    }
    const int n = a.n, P = a.P;
    const int i0 = band * WB2, j0 = chunk * a.rows;
    const int i = i0 + lane - 6;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool outA = lane >= 3 && lane <= 60, outB = lane >= 6 && lane <= 57;
    const int jend = min(j0 + a.rows, n);
    const size_t ic = (size_t)(i < 0 ? 0 : i);
    if (a.ctl && a.ctl->mom_done) {  // the same for every wave of the launch
        if (LAST && outB) {
            for (int j = j0; j < jend; ++j) {
                const size_t k = ic + (size_t)P * j;
                if (xface) a.uo[k] = a.ub[k];
                if (xin) a.vo[k] = a.vb[k];
            }
            if (jend == n && xin) a.vo[ic + (size_t)P * n] = a.vb[ic + (size_t)P * n];
        }
        return;
    }
    const bool edge = i0 < 6 || i0 + 58 > n;
    if (!edge && j0 >= 6 && a.fast) {  // the whole stencil inside the domain: the unrolled path
        if (a.rows == 16 && j0 + 16 + 5 <= n) { pair_interior<16, LAST>(a, lane, i, j0, wid); return; }
        if (a.rows == 12 && j0 + 12 + 5 <= n) { pair_interior<12, LAST>(a, lane, i, j0, wid); return; }
    }
    const double *us = a.us, *vs = a.vs;
    const R gx = (xin && a.lid) ? a.lid[i] : a.inflow0;
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren;
    const int ja0 = max(j0 - 3, 0), jaEnd = min(jend + 3, n);

    // ---------------------------------------------------------------- march A
    struct Raw {
        R u0, v0, v1, vn, uw, ub, cu, vb, cv;  // as k_stage64's (u(i + 1) = u0 of lane + 1)
    };
    auto loadA = [&](int jc, Raw &r) {  // everything A's iteration jc consumes
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
        const size_t k = ic + (size_t)P * (jc < 0 ? 0 : jc);
        r.uw = (xface && row) ? us[k] : R(0);
        const bool out = row && jc >= ja0 && outA;
        r.ub = (out && xface) ? a.ub[k] : R(0);
        r.cu = (out && xface) ? a.cu[k] : R(0);
        r.vb = (out && xin) ? a.vb[k] : R(0);
        r.cv = (out && xin) ? a.cv[k] : R(0);
    };
    R aU[5], aV[5];  // A's rolling window: cell (i, ja - 2 + q)
    {
        Raw r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            loadA(ja0 - 5 + q, r);
            finish_cells<R, false>(a, lane, i, ja0 - 3 + q, edge, r.u0, lane_next(r.u0), r.v0, r.v1, gx, aU[q], aV[q]);
        }
    }
    R avn = (xin && ja0 - 1 >= 0) ? vs[i + (size_t)P * (ja0 - 1)] : R(0);
    R afsU = quick_flux<R>(avn, aU[0], aU[1], aU[2], aU[3]), afsV = quick_flux<R>(avn, aV[0], aV[1], aV[2], aV[3]);
    R atV = 0;
    Raw acur, anxt{};
    loadA(ja0 - 1, acur);
    int ja = ja0 - 1;  // A's next iteration
    // the north wall row of y-faces (base and constant, shared by A and B)
    R vbN = 0, cvN = 0;
    if (jend + 2 >= n && xin) {
        vbN = a.vb[ic + (size_t)P * n];
        cvN = a.cv[ic + (size_t)P * n];
    }
    // A's iteration ja (ja < n): its x-face and y-face of row ja (0 off the domain)
    auto stepA = [&](R &uo, R &vo) {
        if (ja + 1 < jaEnd) loadA(ja + 1, anxt);  // prefetch
        finish_cells<R, false>(a, lane, i, ja + 2, edge, acur.u0, lane_next(acur.u0), acur.v0, acur.v1, gx, aU[4],
                               aV[4]);
        const R vnf = acur.vn;
        const R fnU = quick_flux<R>(vnf, aU[1], aU[2], aU[3], aU[4]);
        const R fnV = quick_flux<R>(vnf, aV[1], aV[2], aV[3], aV[4]);
        const R Uc = aU[2], Vc = aV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R uw = acur.uw;
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - afsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - afsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (aU[3] - R(2) * Uc + aU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (aV[3] - R(2) * Vc + aV[1]) * idy2) * iren;
        const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        const double Rx = (i == 0 || i == n) ? 0.0 : R(0.5) * (TUw + TU);
        uo = xface ? acur.ub + a.rk * (Rx - a.c15 * uw + acur.cu) : R(0);
        const double Ry = ja == 0 ? 0.0 : R(0.5) * (atV + TV);
        vo = xin ? acur.vb + a.rk * (Ry - a.c15 * avn + acur.cv) : R(0);
        afsU = fnU; afsV = fnV; atV = TV; avn = vnf;
#pragma unroll
        for (int q = 0; q < 4; ++q) { aU[q] = aU[q + 1]; aV[q] = aV[q + 1]; }
        acur = anxt;
        ++ja;
    };
    // A's next row: a stage row below n, the north wall row at n (R = 0 there)
    auto emitA = [&](R &uo, R &vo) {
        if (ja < n) {
            stepA(uo, vo);
        } else {
            uo = 0;
            vo = (ja == n && xin) ? vbN + a.rk * (0.0 - a.c15 * avn + cvN) : R(0);
            ++ja;
        }
    };

    // ---------------------------------------------------------------- march B
    // cell row jr in [0, n) from A's faces
    auto cellB = [&](int jr, R u, R v0, R v1, R &U, R &V) {
        finish_cells<R, false>(a, lane, i, jr, edge, u, lane_next(u), v0, v1, gx, U, V);
    };
    // the wall ghost of cell row jr (outside [0, n)) from its mirror cell
    auto ghostB = [&](int jr, R Um, R Vm, R &U, R &V) {
        U = Um;
        V = Vm;
        if (xin) ghost_rule<R, false>(jr < 0 ? 2 : 3, jr < 0 ? a.bc_lo1 : a.bc_hi1, Um, Vm, gx, (R)a.inflow1, U, V);
    };
    R bU[5], bV[5];             // B's rolling window: cell (i, jb - 2 + q)
    R uq0, uq1, uq2, uq3 = 0;   // A's x-faces of rows jb .. jb + 3
    R vq1, vq2, vq3 = 0;        // A's y-faces of rows jb + 1 .. jb + 3
    R bvn, bfsU, bfsV, btV = 0; // A's y-face of row jb; B's flux through it; TV of row jb - 1
    int jb;
    {
        R t0, t1;
        stepA(t0, t1);  // A's iteration ja0 - 1 (no row)
    }
    if (j0 == 0) {
        // rows 0..2; B starts at row 0 (its south y-faces are the wall)
        R x0, y0, x1, y1, x2, y2;
        emitA(x0, y0);
        emitA(x1, y1);
        emitA(x2, y2);
        cellB(0, x0, y0, y1, bU[2], bV[2]);
        cellB(1, x1, y1, y2, bU[3], bV[3]);
        ghostB(-2, bU[3], bV[3], bU[0], bV[0]);
        ghostB(-1, bU[2], bV[2], bU[1], bV[1]);
        bvn = y0;
        uq0 = x0; uq1 = x1; uq2 = x2; vq1 = y1; vq2 = y2;
        jb = 0;
    } else {
        // rows j0 - 3 .. j0 + 1 (j0 >= 3); B starts at row j0 - 1
        R x0, y0, x1, y1, x2, y2, x3, y3, x4, y4;
        emitA(x0, y0);
        emitA(x1, y1);
        cellB(j0 - 3, x0, y0, y1, bU[0], bV[0]);
        emitA(x2, y2);
        cellB(j0 - 2, x1, y1, y2, bU[1], bV[1]);
        emitA(x3, y3);
        cellB(j0 - 1, x2, y2, y3, bU[2], bV[2]);
        emitA(x4, y4);
        cellB(j0, x3, y3, y4, bU[3], bV[3]);
        bvn = y2;
        uq0 = x2; uq1 = x3; uq2 = x4; vq1 = y3; vq2 = y4;
        jb = j0 - 1;
    }
    bfsU = quick_flux<R>(bvn, bU[0], bU[1], bU[2], bU[3]);
    bfsV = quick_flux<R>(bvn, bV[0], bV[1], bV[2], bV[3]);
    struct RawB {
        R ub, cu, vb, cv;
    };
    auto loadB = [&](int jc, RawB &r) {  // B's base and constant of row jc
        const bool out = jc >= j0 && jc < jend && outB;
        const size_t k = ic + (size_t)P * (jc < 0 ? 0 : jc);
        r.ub = (out && xface) ? a.ub[k] : R(0);
        r.cu = (out && xface) ? a.cu[k] : R(0);
        r.vb = (out && xin) ? a.vb[k] : R(0);
        r.cv = (out && xin) ? a.cv[k] : R(0);
    };
    RawB bcur, bnxt{};
    loadB(jb, bcur);
    double nx_ = 0.0, ny_ = 0.0;
    for (; jb < jend; ++jb) {
        if (jb + 1 < jend) loadB(jb + 1, bnxt);  // prefetch
        emitA(uq3, vq3);                          // row jb + 3
        const int jr = jb + 2;                    // B's new window cell
        if (jr < n) cellB(jr, uq2, vq2, vq3, bU[4], bV[4]);
        else if (jr == n) ghostB(jr, bU[3], bV[3], bU[4], bV[4]);  // mirror of row n - 1
        else ghostB(jr, bU[1], bV[1], bU[4], bV[4]);               // row n + 1: mirror of row n - 2
        const R vnf = vq1;
        const R fnU = quick_flux<R>(vnf, bU[1], bU[2], bU[3], bU[4]);
        const R fnV = quick_flux<R>(vnf, bV[1], bV[2], bV[3], bV[4]);
        const R Uc = bU[2], Vc = bV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R uw = uq0;
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - bfsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - bfsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (bU[3] - R(2) * Uc + bU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (bV[3] - R(2) * Vc + bV[1]) * idy2) * iren;
        const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        if (jb >= j0 && outB) {
            const size_t k = ic + (size_t)P * jb;
            if (xface) {  // x-face (i, jb); wall faces get R = 0 exactly
                const bool wall = i == 0 || i == n;
                const double Rx = wall ? 0.0 : R(0.5) * (TUw + TU);
                const double unew = bcur.ub + a.rk2 * (Rx - a.c15 * uw + bcur.cu);
                a.uo[k] = unew;
                if (LAST) { const double d = bcur.ub - unew; nx_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
            }
            if (xin) {  // y-face (i, jb)
                const bool wall = jb == 0;
                const double Ry = wall ? 0.0 : R(0.5) * (btV + TV);
                const double vnew = bcur.vb + a.rk2 * (Ry - a.c15 * bvn + bcur.cv);
                a.vo[k] = vnew;
                if (LAST) { const double d = bcur.vb - vnew; ny_ += fabs(d * d * (wall ? 0.5 : 1.0)); }
            }
        }
        bfsU = fnU; bfsV = fnV; btV = TV; bvn = vnf;
#pragma unroll
        for (int q = 0; q < 4; ++q) { bU[q] = bU[q + 1]; bV[q] = bV[q + 1]; }
        uq0 = uq1; uq1 = uq2; uq2 = uq3;
        vq1 = vq2; vq2 = vq3;
        bcur = bnxt;
    }
    if (jend == n && outB && xin) {  // north wall row of y-faces: R = 0
        const double vnew = vbN + a.rk2 * (0.0 - a.c15 * bvn + cvN);
        a.vo[ic + (size_t)P * n] = vnew;
        if (LAST) { const double d = vbN - vnew; ny_ += fabs(d * d * 0.5); }
    }
    if (LAST) {
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

// quick_tl with the base state's upwind interpolant phix given (the FP64
// march's, rounded to FP32): only the increment's interpolant is formed here
__device__ __forceinline__ float quick_tl_given(float vel, float dvel, float phix, float da2, float da1, float db0,
                                                float db1)
{
    const bool pos = vel >= 0.f;
    const float dd = pos ? db0 : da1, du = pos ? da1 : db0, df = pos ? da2 : db1;
    const float phid = 0.375f * dd + 0.75f * du - 0.125f * df;
    return vel * phid + dvel * phix;
}

// quick_flux, also returning the interpolant its upwind side selects
// (vel >= 0: the one from a2, a1, b0), rounded to FP32
__device__ __forceinline__ double quick_flux_sel(double vel, double a2, double a1, double b0, double b1, float &phi)
{
    const double con = 0.5 * vel;
    const double ud = con + fabs(con), up = con - fabs(con);
    const double bn = (1.0 / 8.0) * (-b1 - 2.0 * b0 + 3.0 * a1) + b0;
    const double bp = (1.0 / 8.0) * (-a2 - 2.0 * a1 + 3.0 * b0) + a1;
    phi = (float)(vel >= 0.0 ? bp : bn);
    return up * bn + ud * bp;
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

// One mixed iteration for a wave whose rows lie inside the domain (cells
// j0 - 6 .. j0 + rows + 4: no row of A's march touches the south or north
// wall or its ghost rows).  A runs its rows j0 - 3 .. j0 + rows + 2 (t = 1 ..
// rows + 6 after a first iteration without output), B its rows j0 - 1 .. j0 +
// rows - 1 three rows behind (t >= 6).  Every rolling window, delay line and
// prefetch queue is a ring of K = D + 4 registers indexed by row: the loop
// body is unrolled K times, so each slot index is a compile-time constant and
// no value is copied from one slot to the next.  Loads run D iterations
// ahead, every iteration the same (rows past the needed ones are clamped into
// the arrays), so the memory waits count exactly.  The arithmetic per face is
// k_mixed_pair's (below).  XE false: no lane's stencil reaches a side wall
// (columns i0 - 6 .. i0 + 57 inside the domain), so the wall branches drop
// out.  XE true: a band at the west or east wall; its cells take the side
// walls' ghost rules from their mirror lanes (finish_cells), the wall x-faces
// have R = 0 and half weight in the norm, the loads' columns are clamped into
// the arrays, and only faces inside the domain are stored.
// Iterations t = 0 .. tend in blocks of K = sizeof...(Q): full(t, slot
// constants 0 .. K - 1) while a whole block fits, then part(t, ..) for the
// rest.  The loop holds only whole blocks, so its back edge joins no early
// exit and the ring registers need no copies there.
template <class Full, class Part, int... Q>
__device__ __forceinline__ void ring_loop(Full &full, Part &part, int tend, std::integer_sequence<int, Q...>)
{
    constexpr int K = (int)sizeof...(Q);
    int t = 0;
    for (; t + K - 1 <= tend; t += K) full(t, std::integral_constant<int, Q>{}...);
    if (t <= tend) part(t, std::integral_constant<int, Q>{}...);
}

template <int D, bool XE>
__device__ __forceinline__ void mixed_interior(const StageArgs &a, int lane, int i, int j0, int wid)
{
    using R = double;
    using F = float;
    constexpr int K = D + 4;
    const size_t P = a.P;
    const int n = a.n, rows = a.rows;
    const bool outB = lane >= 6 && lane <= 57;
    const bool xin = !XE || (i >= 0 && i < n), xface = !XE || (i >= 0 && i <= n);
    const bool xwall = XE && (i == 0 || i == n);
    // the lane's column in the x-face and y-face arrays (clamped on the ghost lanes of a wall band)
    const unsigned obx = (unsigned)(XE ? min(max(i, 0), n) : i) * 8u;
    const unsigned oby = (unsigned)(XE ? min(max(i, 0), n - 1) : i) * 8u;
    auto atx = [&](const double *base, int row) -> const double & {
        return *(const double *)((const char *)(base + P * row) + obx);
    };
    auto aty = [&](const double *base, int row) -> const double & {
        return *(const double *)((const char *)(base + P * row) + oby);
    };
    auto put = [&](double *base, int row, double val) {
        *(double *)((char *)(base + P * row) + (unsigned)i * 8u) = val;
    };
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren, c15 = a.c15;
    const F fidx = (F)a.inv_dx, fidy = (F)a.inv_dy, fidx2 = (F)a.inv_dx2, fidy2 = (F)a.inv_dy2, firen = (F)a.inv_ren;
    const double cr = a.cr, cj = a.cj;
    const double *us = a.us, *vs = a.vs;
    // cells of row jr from its x-face u and its y-faces v0 (row jr), v1 (row jr + 1)
    auto cellA = [&](int jr, R u, R v0, R v1, R &U, R &V) {
        if constexpr (XE) {
            finish_cells<R, false>(a, lane, i, jr, true, u, lane_next(u), v0, v1, R(0), U, V);
        } else {
            (void)jr;
            U = R(0.5) * (u + lane_next(u));
            V = R(0.5) * (v0 + v1);
        }
    };
    // Ring slot of row r: (r - (j0 - 4)) mod K, i.e. at iteration t (A's row
    // jc = j0 - 4 + t) row jc + k sits in slot (t + k) mod K.
    const int jA = j0 - 3;  // A's first output row
    R aU[K], aV[K];         // A's cells jc - 2 .. jc + 2
    R ru[K], rv[K];         // u^k faces: x rows jc .. jc + 2 + D, y rows jc .. jc + 3 + D
    R rcu[K], rcv[K];       // the constant, rows jc .. jc + D
    F xb[K], xr[K], yb[K], yr[K];  // A's output faces (FP32 base, FP32 residual), rows jc - 3 .. jc
    // A's upwind QUICK interpolants of the base state (FP32) for its four fluxes
    // of row jc (y-face jc + 1 of U and V, x-face jc of U and V), rows jc - 4 .. jc:
    // B's tangent-linear fluxes take them instead of forming FP32 base cells
    F pnU[K], pnV[K], pwU[K], pwV[K];
    F dU[K], dV[K];                // B's residual cells, rows jc - 5 .. jc - 1
    R obu[K], obv[K];              // the FP64 base faces of B's output rows jb .. jb + D
    {   // cells jA - 3 .. jA (slots K - 2, K - 1, 0, 1)
        const R u0 = atx(us, jA - 3), u1 = atx(us, jA - 2);
        const R v0 = aty(vs, jA - 3), v1 = aty(vs, jA - 2);
#pragma unroll
        for (int q = 0; q <= D + 1; ++q) ru[q] = atx(us, jA - 1 + q);
#pragma unroll
        for (int q = 0; q <= D + 2; ++q) rv[q] = aty(vs, jA - 1 + q);
#pragma unroll
        for (int q = 0; q < D; ++q) { rcu[q] = atx(a.cu, jA - 1 + q); rcv[q] = aty(a.cv, jA - 1 + q); }
        cellA(jA - 3, u0, v0, v1, aU[K - 2], aV[K - 2]);
        cellA(jA - 2, u1, v1, rv[0], aU[K - 1], aV[K - 1]);
        cellA(jA - 1, ru[0], rv[0], rv[1], aU[0], aV[0]);
        cellA(jA, ru[1], rv[1], rv[2], aU[1], aV[1]);
    }
    R afsU = quick_flux<R>(rv[0], aU[K - 2], aU[K - 1], aU[0], aU[1]);
    R afsV = quick_flux<R>(rv[0], aV[K - 2], aV[K - 1], aV[0], aV[1]);
    R atV = 0;
    F fsU = 0, fsV = 0, tVprev = 0;
    double nx_ = 0.0, ny_ = 0.0, sb = 0.0;
    const int tend = rows + 6;  // the last iteration
    auto iter = [&](auto M, const int t) {
        constexpr int m = decltype(M)::value;
        auto sl = [](int k) constexpr { return ((m + k) % K + K) % K; };
        constexpr int a0 = sl(0), a1 = sl(1), a2 = sl(2), a3 = sl(3), am1 = sl(-1), am2 = sl(-2);
        constexpr int b3 = sl(-3), b4 = sl(-4), b5 = sl(-5);
        __builtin_amdgcn_sched_barrier(0);
        const int jc = jA - 1 + t;  // A's row
        const int jb = jc - 3;      // B's row (t >= 6)
        {   // the loads of iteration t + D (rows clamped into the arrays past the last one needed)
            ru[sl(2 + D)] = atx(us, min(jc + 2 + D, n - 1));
            rv[sl(3 + D)] = aty(vs, min(jc + 3 + D, n));
            rcu[sl(D)] = atx(a.cu, min(jc + D, n - 1));
            rcv[sl(D)] = aty(a.cv, min(jc + D, n));
            obu[sl(D - 3)] = atx(us, max(jb + D, 0));
            obv[sl(D - 3)] = aty(vs, max(jb + D, 0));
        }
        // A, row jc: the FP64 residual of its two faces
        cellA(jc + 2, ru[a2], rv[a2], rv[a3], aU[a2], aV[a2]);
        {
            const R fnU = quick_flux_sel(rv[a1], aU[am1], aU[a0], aU[a1], aU[a2], pnU[a0]);
            const R fnV = quick_flux_sel(rv[a1], aV[am1], aV[a0], aV[a1], aV[a2], pnV[a0]);
            const R uw = ru[a0], vw = rv[a0];
            const R Uc = aU[a0], Vc = aV[a0];
            const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
            const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
            const R fwU = quick_flux_sel(uw, Um2, Um1, Uc, Up1, pwU[a0]);
            const R fwV = quick_flux_sel(uw, Vm2, Vm1, Vc, Vp1, pwV[a0]);
            const R feU = lane_next(fwU), feV = lane_next(fwV);
            const R CU = (feU - fwU) * idx + (fnU - afsU) * idy;
            const R CV = (feV - fwV) * idx + (fnV - afsV) * idy;
            const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (aU[a1] - R(2) * Uc + aU[am1]) * idy2) * iren;
            const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (aV[a1] - R(2) * Vc + aV[am1]) * idy2) * iren;
            const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
            const R TUw = lane_prev(TU);
            const double Rx = xwall ? 0.0 : R(0.5) * (TUw + TU);
            const double Ry = R(0.5) * (atV + TV);
            xb[a0] = xface ? (float)uw : 0.f;
            xr[a0] = xface ? (float)(Rx - c15 * uw + rcu[a0]) : 0.f;
            yb[a0] = xin ? (float)vw : 0.f;
            yr[a0] = xin ? (float)(Ry - c15 * vw + rcv[a0]) : 0.f;
            afsU = fnU; afsV = fnV; atV = TV;
        }
        __builtin_amdgcn_sched_barrier(0);  // A's temporaries end here
        // B: residual cell row jc - 1 from A's faces (x row jc - 1, y rows jc - 1, jc)
        if (t >= 2) {
            if constexpr (XE) {
                finish_cells<F, true>(a, lane, i, jc - 1, true, xr[am1], lane_next(xr[am1]), yr[am1], yr[a0], F(0),
                                      dU[am1], dV[am1]);
            } else {
                dU[am1] = F(0.5) * (xr[am1] + lane_next(xr[am1]));
                dV[am1] = F(0.5) * (yr[am1] + yr[a0]);
            }
        }
        if (t >= 6) {  // B's row jb: faces rows jb .. jb + 3 in slots b3, am2, am1, a0; cells jb - 2 .. jb + 2 in b5 .. am1
            // A's interpolants: row jb - 1 (the flux through y-face jb) in b4, row jb in b3
            if (t == 6) {  // the flux through y-face jb
                fsU = quick_tl_given(yb[b3], yr[b3], pnU[b4], dU[b5], dU[b4], dU[b3], dU[am2]);
                fsV = quick_tl_given(yb[b3], yr[b3], pnV[b4], dV[b5], dV[b4], dV[b3], dV[am2]);
            }
            const F fnU = quick_tl_given(yb[am2], yr[am2], pnU[b3], dU[b4], dU[b3], dU[am2], dU[am1]);
            const F fnV = quick_tl_given(yb[am2], yr[am2], pnV[b3], dV[b4], dV[b3], dV[am2], dV[am1]);
            const F dUc = dU[b3], dVc = dV[b3];
            const F dUm1 = lane_prev(dUc), dUm2 = lane_prev(dUm1), dUp1 = lane_next(dUc);
            const F dVm1 = lane_prev(dVc), dVm2 = lane_prev(dVm1), dVp1 = lane_next(dVc);
            const F fwU = quick_tl_given(xb[b3], xr[b3], pwU[b3], dUm2, dUm1, dUc, dUp1);
            const F fwV = quick_tl_given(xb[b3], xr[b3], pwV[b3], dVm2, dVm1, dVc, dVp1);
            const F feU = lane_next(fwU), feV = lane_next(fwV);
            const F CU = (feU - fwU) * fidx + (fnU - fsU) * fidy;
            const F CV = (feV - fwV) * fidx + (fnV - fsV) * fidy;
            const F VU = ((dUp1 - F(2) * dUc + dUm1) * fidx2 + (dU[am2] - F(2) * dUc + dU[b4]) * fidy2) * firen;
            const F VV = ((dVp1 - F(2) * dVc + dVm1) * fidx2 + (dV[am2] - F(2) * dVc + dV[b4]) * fidy2) * firen;
            const F TU = (F(0) - CU) + VU, TV = (F(0) - CV) + VV;
            const F TUw = lane_prev(TU);
            if (t >= 7 && outB) {  // B's row jb >= j0: no wall y-face in this wave
                if (xface) {
                    const float Jx = xwall ? 0.f : 0.5f * (TUw + TU);
                    const double du = cr * (double)xr[b3] + cj * (double)Jx;
                    const double unew = obu[b3] + du;
                    put(a.uo, jb, unew);
                    if constexpr (XE) {
                        nx_ += du * du * (xwall ? 0.5 : 1.0);
                        if (xwall) sb += i == n ? unew * a.inv_dx : -(unew * a.inv_dx);
                    } else {
                        nx_ += fabs(du * du * 1.0);
                    }
                }
                if (xin) {
                    const float Jy = 0.5f * (tVprev + TV);
                    const double dv = cr * (double)yr[b3] + cj * (double)Jy;
                    put(a.vo, jb, obv[b3] + dv);
                    ny_ += fabs(dv * dv * 1.0);
                }
            }
            fsU = fnU; fsV = fnV; tVprev = TV;
        }
    };
    auto full = [&](const int t, auto... Q) { (iter(Q, t + decltype(Q)::value), ...); };
    auto part = [&](const int t, auto... Q) {  // iterations t .. tend (fewer than K)
        bool go = true;
        ((go = go && t + decltype(Q)::value <= tend, go ? (iter(Q, t + decltype(Q)::value), 0) : 0), ...);
    };
    ring_loop(full, part, tend, std::make_integer_sequence<int, K>{});
    for (int o = 32; o > 0; o >>= 1) {
        nx_ += __shfl_down(nx_, o, 64);
        ny_ += __shfl_down(ny_, o, 64);
        sb += __shfl_down(sb, o, 64);
    }
    if (lane == 0) {  // records of three: the norm partials and the wave's boundary flux of u^{k+1}
        a.partials[3 * wid] = nx_;
        a.partials[3 * wid + 1] = ny_;
        a.partials[3 * wid + 2] = sb;
    }
}

// ======================================= one mixed iteration in one pass
// The mixed momentum iteration (k_stage64 mode 3, then k_tangent) in one
// launch.  The layout is k_stage_pair's: a wave owns 52 output columns with a
// 6-lane halo on each side and a chunk of rows.  March A is k_stage64's mode 3,
// the FP64 residual r = R(u^k) - (1.5/dt) u^k + c of the pseudo base, three
// rows ahead of march B, k_tangent's u^{k+1} = u^k + cr r + cj R'(u^k) r.  B
// takes the FP32 base and residual faces (rounded as k_stage64 stores them)
// and the FP64 base faces from A's registers: a delay line of four x-face and
// four y-face rows.  A is exact on lanes 3..60, so B is exact on lanes 6..57;
// B's cells outside the domain are the wall ghosts of its own window cells.
// Per face the arithmetic is the two kernels', in the same precisions, so the
// fields are those of the two launches; the norm partials are summed per wave
// of this layout.  Memory traffic: u^k and c in, u^{k+1} out, 48 B per cell
// (the two launches move 96).  The output must not alias u^k: A reads u^k in
// the halo rows and columns that neighbouring waves update.  After convergence
// (the launch skipped) u^k is copied to the output, so that the caller's
// exchange of the buffers keeps the converged state.  The partials are records
// of three per wave: the two norm sums and the wave's share of the net outflow
// sum (u(n) - u(0))/dx + (v(n) - v(0))/dy of u^{k+1}, whose total fixes the
// mean of the Poisson right-hand side (the cell divergences telescope).
__global__ __launch_bounds__(256) void k_mixed_pair(StageArgs a)
{
    using R = double;
    using F = float;
    const int lane = threadIdx.x & 63;
    int band, chunk, wid;
    if (a.edge_first) {
        if (!pair_wave(a, band, chunk, wid)) return;  // whole wave
    } else {
        wid = wave_id(a);
        if (wid >= a.nbands * a.nchunks) return;  // whole wave
        band = a.order ? wid % a.nbands : wid / a.nchunks;
        chunk = a.order ? wid / a.nbands : wid % a.nchunks;
    }
    const int n = a.n, P = a.P;
    const int i0 = band * WB2, j0 = chunk * a.rows;
    const int i = i0 + lane - 6;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool outA = lane >= 3 && lane <= 60, outB = lane >= 6 && lane <= 57;
    const int jend = min(j0 + a.rows, n);
    const size_t ic = (size_t)(i < 0 ? 0 : i);
    const double *us = a.us, *vs = a.vs;  // u^k: the stencil input and the base
    if (a.ctl && a.ctl->mom_done) {       // the same for every wave of the launch
        if (outB) {
            for (int j = j0; j < jend; ++j) {
                const size_t k = ic + (size_t)P * j;
                if (xface) a.uo[k] = us[k];
                if (xin) a.vo[k] = vs[k];
            }
            if (jend == n && xin) a.vo[ic + (size_t)P * n] = vs[ic + (size_t)P * n];
        }
        return;
    }
    const bool edge = i0 < 6 || i0 + 58 > n;
    if (j0 >= 6 && j0 + a.rows + 5 <= n && a.fast) {  // no row of the stencil at a wall: the ring loop
        if (edge) mixed_interior<2, true>(a, lane, i, j0, wid);
        else mixed_interior<2, false>(a, lane, i, j0, wid);
        return;
    }
    const R gx = (xin && a.lid) ? a.lid[i] : a.inflow0;
    const F gxf = (F)((xin && a.lid) ? a.lid[i] : a.inflow0);
    const R idx = a.inv_dx, idy = a.inv_dy, idx2 = a.inv_dx2, idy2 = a.inv_dy2, iren = a.inv_ren;
    const int ja0 = max(j0 - 3, 0), jaEnd = min(jend + 3, n);

    // ---------------------------------------------------------------- march A
    struct Raw {
        R u0, v0, v1, vn, uw, cu, cv;  // as k_stage64's (u(i + 1) = u0 of lane + 1)
    };
    auto loadA = [&](int jc, Raw &r) {  // everything A's iteration jc consumes
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
        const size_t k = ic + (size_t)P * (jc < 0 ? 0 : jc);
        r.uw = (xface && row) ? us[k] : R(0);
        const bool out = row && jc >= ja0 && outA;
        r.cu = (out && xface) ? a.cu[k] : R(0);
        r.cv = (out && xin) ? a.cv[k] : R(0);
    };
    R aU[5], aV[5];  // A's rolling window: cell (i, ja - 2 + q)
    {
        Raw r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            loadA(ja0 - 5 + q, r);
            finish_cells<R, false>(a, lane, i, ja0 - 3 + q, edge, r.u0, lane_next(r.u0), r.v0, r.v1, gx, aU[q], aV[q]);
        }
    }
    R avn = (xin && ja0 - 1 >= 0) ? vs[i + (size_t)P * (ja0 - 1)] : R(0);
    R afsU = quick_flux<R>(avn, aU[0], aU[1], aU[2], aU[3]), afsV = quick_flux<R>(avn, aV[0], aV[1], aV[2], aV[3]);
    R atV = 0;
    Raw acur, anxt{};
    loadA(ja0 - 1, acur);
    int ja = ja0 - 1;  // A's next iteration
    // the constant of the north wall row of y-faces
    const R cvN = (jaEnd == n && xin) ? a.cv[ic + (size_t)P * n] : R(0);
    // one face of A's output: the FP32 base and residual (k_stage64's stores) and the FP64 base
    struct Face {
        F b, r;
        R b64;
    };
    // A's iteration ja (ja < n): the faces of row ja (0 off the domain)
    auto stepA = [&](Face &xo, Face &yo) {
        if (ja + 1 < jaEnd) loadA(ja + 1, anxt);  // prefetch
        finish_cells<R, false>(a, lane, i, ja + 2, edge, acur.u0, lane_next(acur.u0), acur.v0, acur.v1, gx, aU[4],
                               aV[4]);
        const R vnf = acur.vn;
        const R fnU = quick_flux<R>(vnf, aU[1], aU[2], aU[3], aU[4]);
        const R fnV = quick_flux<R>(vnf, aV[1], aV[2], aV[3], aV[4]);
        const R Uc = aU[2], Vc = aV[2];
        const R Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const R Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const R uw = acur.uw;
        const R fwU = quick_flux<R>(uw, Um2, Um1, Uc, Up1), fwV = quick_flux<R>(uw, Vm2, Vm1, Vc, Vp1);
        const R feU = lane_next(fwU), feV = lane_next(fwV);
        const R CU = (feU - fwU) * idx + (fnU - afsU) * idy;
        const R CV = (feV - fwV) * idx + (fnV - afsV) * idy;
        const R VU = ((Up1 - R(2) * Uc + Um1) * idx2 + (aU[3] - R(2) * Uc + aU[1]) * idy2) * iren;
        const R VV = ((Vp1 - R(2) * Vc + Vm1) * idx2 + (aV[3] - R(2) * Vc + aV[1]) * idy2) * iren;
        const R TU = (R(0) - CU) + VU, TV = (R(0) - CV) + VV;
        const R TUw = lane_prev(TU);
        const double Rx = (i == 0 || i == n) ? 0.0 : R(0.5) * (TUw + TU);
        const double Ry = ja == 0 ? 0.0 : R(0.5) * (atV + TV);
        xo.b64 = xface ? uw : R(0);
        xo.b = xface ? (float)uw : 0.f;
        xo.r = xface ? (float)(Rx - a.c15 * uw + acur.cu) : 0.f;
        yo.b64 = xin ? avn : R(0);
        yo.b = xin ? (float)avn : 0.f;
        yo.r = xin ? (float)(Ry - a.c15 * avn + acur.cv) : 0.f;
        afsU = fnU; afsV = fnV; atV = TV; avn = vnf;
#pragma unroll
        for (int q = 0; q < 4; ++q) { aU[q] = aU[q + 1]; aV[q] = aV[q + 1]; }
        acur = anxt;
        ++ja;
    };
    // A's next row: a residual row below n, the north wall row at n (R = 0 there)
    auto emitA = [&](Face &xo, Face &yo) {
        if (ja < n) {
            stepA(xo, yo);
        } else {
            xo = Face{0.f, 0.f, R(0)};
            if (ja == n && xin) yo = Face{(float)avn, (float)(-a.c15 * avn + cvN), avn};
            else yo = Face{0.f, 0.f, R(0)};
            ++ja;
        }
    };

    // ---------------------------------------------------------------- march B
    // cell row jr in [0, n) from A's faces: the base (wall rules) and the
    // residual (the rules without their constants)
    auto cellB = [&](int jr, const Face &u, const Face &v0, const Face &v1, F &U, F &V, F &dU, F &dV) {
        finish_cells<F, false>(a, lane, i, jr, edge, u.b, lane_next(u.b), v0.b, v1.b, gxf, U, V);
        finish_cells<F, true>(a, lane, i, jr, edge, u.r, lane_next(u.r), v0.r, v1.r, F(0), dU, dV);
    };
    // the wall ghost of cell row jr (outside [0, n)) from its mirror cell
    auto ghostB = [&](int jr, F Um, F Vm, F dUm, F dVm, F &U, F &V, F &dU, F &dV) {
        U = Um; V = Vm; dU = dUm; dV = dVm;
        if (xin) {
            ghost_rule<F, false>(jr < 0 ? 2 : 3, jr < 0 ? a.bc_lo1 : a.bc_hi1, Um, Vm, gxf, (F)a.inflow1, U, V);
            ghost_rule<F, true>(jr < 0 ? 2 : 3, jr < 0 ? a.bc_lo1 : a.bc_hi1, dUm, dVm, F(0), (F)a.inflow1, dU, dV);
        }
    };
    F wU[5], wV[5], dU[5], dV[5];  // B's rolling windows: base and residual cells (i, jb - 2 + q)
    Face uq0, uq1, uq2, uq3{};      // A's x-faces of rows jb .. jb + 3
    Face vq0, vq1, vq2, vq3{};      // A's y-faces of rows jb .. jb + 3
    F fsU, fsV, tVprev = 0;         // B's flux through y-face jb; TV of row jb - 1
    int jb;
    {
        Face t0, t1;
        stepA(t0, t1);  // A's iteration ja0 - 1 (no row)
    }
    if (j0 == 0) {
        // rows 0..2; B starts at row 0 (its south y-faces are the wall)
        Face x0, y0, x1, y1, x2, y2;
        emitA(x0, y0);
        emitA(x1, y1);
        emitA(x2, y2);
        cellB(0, x0, y0, y1, wU[2], wV[2], dU[2], dV[2]);
        cellB(1, x1, y1, y2, wU[3], wV[3], dU[3], dV[3]);
        ghostB(-2, wU[3], wV[3], dU[3], dV[3], wU[0], wV[0], dU[0], dV[0]);
        ghostB(-1, wU[2], wV[2], dU[2], dV[2], wU[1], wV[1], dU[1], dV[1]);
        vq0 = y0;
        uq0 = x0; uq1 = x1; uq2 = x2; vq1 = y1; vq2 = y2;
        jb = 0;
    } else {
        // rows j0 - 3 .. j0 + 1 (j0 >= 3); B starts at row j0 - 1
        Face x0, y0, x1, y1, x2, y2, x3, y3, x4, y4;
        emitA(x0, y0);
        emitA(x1, y1);
        cellB(j0 - 3, x0, y0, y1, wU[0], wV[0], dU[0], dV[0]);
        emitA(x2, y2);
        cellB(j0 - 2, x1, y1, y2, wU[1], wV[1], dU[1], dV[1]);
        emitA(x3, y3);
        cellB(j0 - 1, x2, y2, y3, wU[2], wV[2], dU[2], dV[2]);
        emitA(x4, y4);
        cellB(j0, x3, y3, y4, wU[3], wV[3], dU[3], dV[3]);
        vq0 = y2;
        uq0 = x2; uq1 = x3; uq2 = x4; vq1 = y3; vq2 = y4;
        jb = j0 - 1;
    }
    fsU = quick_tl(vq0.b, vq0.r, wU[0], wU[1], wU[2], wU[3], dU[0], dU[1], dU[2], dU[3]);
    fsV = quick_tl(vq0.b, vq0.r, wV[0], wV[1], wV[2], wV[3], dV[0], dV[1], dV[2], dV[3]);
    const F fidx = (F)a.inv_dx, fidy = (F)a.inv_dy, fidx2 = (F)a.inv_dx2, fidy2 = (F)a.inv_dy2, firen = (F)a.inv_ren;
    const double cr = a.cr, cj = a.cj;
    double nx_ = 0.0, ny_ = 0.0;
    double sb = 0.0;  // the boundary flux of u^{k+1} through this wave's wall faces (net outflow)
    for (; jb < jend; ++jb) {
        emitA(uq3, vq3);        // row jb + 3
        const int jr = jb + 2;  // B's new window cell
        if (jr < n) cellB(jr, uq2, vq2, vq3, wU[4], wV[4], dU[4], dV[4]);
        else if (jr == n) ghostB(jr, wU[3], wV[3], dU[3], dV[3], wU[4], wV[4], dU[4], dV[4]);  // mirror of n - 1
        else ghostB(jr, wU[1], wV[1], dU[1], dV[1], wU[4], wV[4], dU[4], dV[4]);               // mirror of n - 2
        // k_tangent's row: the flux through y-face jb + 1, then the x-faces
        const F fnU = quick_tl(vq1.b, vq1.r, wU[1], wU[2], wU[3], wU[4], dU[1], dU[2], dU[3], dU[4]);
        const F fnV = quick_tl(vq1.b, vq1.r, wV[1], wV[2], wV[3], wV[4], dV[1], dV[2], dV[3], dV[4]);
        const F Uc = wU[2], Vc = wV[2], dUc = dU[2], dVc = dV[2];
        const F Um1 = lane_prev(Uc), Um2 = lane_prev(Um1), Up1 = lane_next(Uc);
        const F Vm1 = lane_prev(Vc), Vm2 = lane_prev(Vm1), Vp1 = lane_next(Vc);
        const F dUm1 = lane_prev(dUc), dUm2 = lane_prev(dUm1), dUp1 = lane_next(dUc);
        const F dVm1 = lane_prev(dVc), dVm2 = lane_prev(dVm1), dVp1 = lane_next(dVc);
        const F fwU = quick_tl(uq0.b, uq0.r, Um2, Um1, Uc, Up1, dUm2, dUm1, dUc, dUp1);
        const F fwV = quick_tl(uq0.b, uq0.r, Vm2, Vm1, Vc, Vp1, dVm2, dVm1, dVc, dVp1);
        const F feU = lane_next(fwU), feV = lane_next(fwV);
        const F CU = (feU - fwU) * fidx + (fnU - fsU) * fidy;
        const F CV = (feV - fwV) * fidx + (fnV - fsV) * fidy;
        const F VU = ((dUp1 - F(2) * dUc + dUm1) * fidx2 + (dU[3] - F(2) * dUc + dU[1]) * fidy2) * firen;
        const F VV = ((dVp1 - F(2) * dVc + dVm1) * fidx2 + (dV[3] - F(2) * dVc + dV[1]) * fidy2) * firen;
        const F TU = (F(0) - CU) + VU, TV = (F(0) - CV) + VV;
        const F TUw = lane_prev(TU);
        if (jb >= j0 && outB) {
            const size_t k = ic + (size_t)P * jb;
            if (xface) {
                const bool wall = i == 0 || i == n;
                const float Jr = wall ? 0.f : 0.5f * (TUw + TU);
                const double du = cr * (double)uq0.r + cj * (double)Jr;
                const double unew = uq0.b64 + du;
                a.uo[k] = unew;
                nx_ += du * du * (wall ? 0.5 : 1.0);
                if (wall) sb += i == n ? unew * a.inv_dx : -(unew * a.inv_dx);
            }
            if (xin) {
                const bool wall = jb == 0;
                const float Jr = wall ? 0.f : 0.5f * (tVprev + TV);
                const double dv = cr * (double)vq0.r + cj * (double)Jr;
                const double vnew = vq0.b64 + dv;
                a.vo[k] = vnew;
                ny_ += dv * dv * (wall ? 0.5 : 1.0);
                if (wall) sb -= vnew * a.inv_dy;
            }
        }
        fsU = fnU; fsV = fnV; tVprev = TV;
#pragma unroll
        for (int q = 0; q < 4; ++q) { wU[q] = wU[q + 1]; wV[q] = wV[q + 1]; dU[q] = dU[q + 1]; dV[q] = dV[q + 1]; }
        uq0 = uq1; uq1 = uq2; uq2 = uq3;
        vq0 = vq1; vq1 = vq2; vq2 = vq3;
    }
    if (jend == n && outB && xin) {  // north wall row: R' r = 0 there
        const double dv = cr * (double)vq0.r;
        const double vnew = vq0.b64 + dv;
        a.vo[ic + (size_t)P * n] = vnew;
        ny_ += dv * dv * 0.5;
        sb += vnew * a.inv_dy;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nx_ += __shfl_down(nx_, o, 64);
        ny_ += __shfl_down(ny_, o, 64);
        sb += __shfl_down(sb, o, 64);
    }
    if (lane == 0) {  // records of three: the norm partials and the wave's boundary flux of u^{k+1}
        a.partials[3 * wid] = nx_;
        a.partials[3 * wid + 1] = ny_;
        a.partials[3 * wid + 2] = sb;
    }
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
    a.edge_first = c.stage_edge_first;
    a.fast = c.stage_fast;
    a.nbands = (g.n + 1 + WB - 1) / WB;
    a.nchunks = (g.n + a.rows - 1) / a.rows;
    return a;
}

static dim3 stage_grid(StageArgs &a)
{
    return dim3((a.nbands * a.nchunks + 3) / 4);
}

// The pair kernels' edge-first wave order (pair_wave): the bands and chunks
// whose waves leave the interior path's stencil box (cells j0 - 6 .. j0 + rows
// + 4 and columns i0 - 6 .. i0 + 57 inside the domain; with sides true the
// interior path also takes the bands at the side walls, so only rows count).
// Without an interior path (interior false) every wave counts as an edge
// wave.  Returns the grid.
static dim3 pair_grid(const Ctx &c, StageArgs &a, bool interior, bool sides)
{
    const int n = c.g.n, nb = a.nbands, nc = a.nchunks;
    auto band_in = [&](int b) { return sides || (b * WB2 >= 6 && b * WB2 + 58 <= n); };
    auto chunk_in = [&](int ch) { return ch * a.rows >= 6 && ch * a.rows + a.rows + 5 <= n; };
    int b0 = 0, b1 = nb, c0 = 0, c1 = nc;
    while (b0 < nb && !band_in(b0)) ++b0;
    while (b1 > b0 && !band_in(b1 - 1)) --b1;
    while (c0 < nc && !chunk_in(c0)) ++c0;
    while (c1 > c0 && !chunk_in(c1 - 1)) --c1;
    if (!interior || b0 >= b1 || c0 >= c1) { b0 = nb; b1 = nb; c0 = nc; c1 = nc; }
    a.eb0 = b0; a.eb1 = nb - b1; a.ec0 = c0; a.ec1 = nc - c1;
    if (!a.edge_first) return stage_grid(a);
    const int I = (b1 - b0) * (c1 - c0), E = nb * nc - I;
    return dim3((E + 3) / 4 + (I + 3) / 4);
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

bool stage_pair_fits(const Ctx &c)
{
    const int nb = (c.g.n + 1 + WB2 - 1) / WB2, nc = (c.g.n + c.pair_rows - 1) / c.pair_rows;
    return c.g.n >= 8 && c.pair_rows >= 4 && (int64_t)nb * nc <= kMaxPartials;
}

hipError_t launch_stage_pair(const Ctx &c, const StageLaunch &s, int *nblocks)
{
    if (s.pair != 1 && s.pair != 2) return hipErrorInvalidValue;
    if (!s.foldp || !stage_pair_fits(c)) return hipErrorInvalidValue;
    if (s.uo == s.us || s.uo == s.ub || s.vo == s.vs || s.vo == s.vb) return hipErrorInvalidValue;
    StageArgs a = base_args(c, s.rk, s.pair == 2 ? 1 : 0, c.pair_rows);
    a.rk2 = s.rk2;
    a.us = s.us; a.vs = s.vs; a.ub = s.ub; a.vb = s.vb; a.cu = s.cu; a.cv = s.cv; a.p = s.p;
    a.uo = s.uo; a.vo = s.vo;
    if (s.unguarded) a.ctl = nullptr;
    a.nbands = (c.g.n + 1 + WB2 - 1) / WB2;
    a.nchunks = (c.g.n + a.rows - 1) / a.rows;
    const dim3 grid = pair_grid(c, a, a.fast && (a.rows == 16 || a.rows == 12), false);
    if (nblocks) *nblocks = a.nbands * a.nchunks;  // one norm partial per wave
    if (s.pair == 2) hipLaunchKernelGGL(k_stage_pair<true>, grid, dim3(256), 0, c.stream, a);
    else hipLaunchKernelGGL(k_stage_pair<false>, grid, dim3(256), 0, c.stream, a);
    return hipGetLastError();
}

bool mixed_pair_fits(const Ctx &c)
{
    const int nb = (c.g.n + 1 + WB2 - 1) / WB2, nc = (c.g.n + c.mpair_rows - 1) / c.mpair_rows;
    return c.g.n >= 8 && c.mpair_rows >= 4 && (int64_t)nb * nc <= kMaxPartials;
}

hipError_t launch_mixed_pair(const Ctx &c, const MixedPairLaunch &s, int *nblocks)
{
    if (!mixed_pair_fits(c)) return hipErrorInvalidValue;
    if (s.uo == s.u64 || s.vo == s.v64) return hipErrorInvalidValue;
    StageArgs a = base_args(c, 0.0, 3, c.mpair_rows);
    a.us = s.u64; a.vs = s.v64; a.cu = s.cu; a.cv = s.cv;
    a.uo = s.uo; a.vo = s.vo;
    a.cr = s.cr; a.cj = s.cj;
    if (s.unguarded) a.ctl = nullptr;
    a.nbands = (c.g.n + 1 + WB2 - 1) / WB2;
    a.nchunks = (c.g.n + a.rows - 1) / a.rows;
    const dim3 grid = pair_grid(c, a, a.fast, false);
    if (nblocks) *nblocks = a.nbands * a.nchunks;  // one norm partial per wave
    hipLaunchKernelGGL(k_mixed_pair, grid, dim3(256), 0, c.stream, a);
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
