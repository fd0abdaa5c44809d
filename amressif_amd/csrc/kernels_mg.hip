// Geometric multigrid for the pressure-Poisson projection, gfx950.
//
// Operator: L phi = (E+W-2P)/dx^2 + (N+S-2P)/dy^2 with homogeneous Neumann
// walls (the reference's MLABecLaplacian with alpha = 0, beta = -1, B = 1 and
// foextrap -> Neumann, Source/src1_rever/poisson.cpp:85-153).  The Neumann
// mirror drops the out-of-domain neighbour from the stencil and from the
// diagonal, so no ghost cells are stored.
//
// Solve (poisson.cpp:181-185): phi_{k+1} = phi_k + V(b - L phi_k) until
// ||b - L phi||inf <= max(tol_abs, tol_rel ||b||inf), with b made solvable by
// removing its mean.  V is one V-cycle on the correction from zero: two
// red-black Gauss-Seidel smooths before and after, 2x2 average restriction,
// piecewise-constant prolongation, dense pseudo-inverse on the coarsest
// level.  V runs in T = double (FP64 step) or T = float (mixed step: FP32
// V-cycle inside FP64 iterative refinement); the outer residual, the
// correction, the norms and the mean stay FP64.
//
// Kernels per V-cycle: on each level larger than 64^2 one fused smoother
// (four dependent half-sweeps in an LDS tile with a 4-cell halo) before and
// one after (with the coarse correction added on load), and one fused
// residual + restriction; the levels from 64^2 down, including the bottom
// solve, run inside one workgroup (k_tail); then one kernel applies the
// correction in FP64 and forms the next residual and its norm partials.
#include "device_common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace amr {

// --------------------------------------------------------- fused smoother
constexpr int SX = 64, SY = 16;  // output tile

// x_out = GSRB^2(x_in + P(ec)) for L x = r on one level: red (i+j even) then
// black, twice, in an LDS region of the tile plus a halo of H cells.  The
// region is split into horizontal cell pairs (one red, one black); thread t
// owns pairs t, t + 256, ... so every half-sweep is a fixed loop of a few
// updates per thread with no index division.  Cells of one colour read only
// the other colour, so the update order inside a half-sweep does not change
// the result.  Each half-sweep invalidates one more ring of the region, so
// after the four the region minus a ring of 4 is exact.
//   MODE 0 (H = 4): x_in null starts from zero, ec null adds no correction.
//   MODE 1 (H = 5, the down leg): from zero; the exact ring around the tile
//     also gives the tile's residual r - L x, averaged over each 2x2 block into
//     the coarse right-hand side rc (the separate residual + restriction pass
//     with the same arithmetic).
//   MODE 2 (H = 5, the up leg on the finest level): the correction e is added
//     to the FP64 iterate, phi_out = phi_in + e, and the next residual
//     r = b - L phi_out is formed in FP64 (stored in T), with block partials of
//     ||r||inf and sum phi_out -- the FP64 outer step fused into the smoother.
template <int H>
struct SmoothShape {
    static constexpr int W = SX + 2 * H, HT = SY + 2 * H;
    static constexpr int PAIRS = W / 2 * HT;
    static constexpr int PER = (PAIRS + 255) / 256;
};

struct OuterArgs {  // MODE 2
    const double *phi_in;
    double *phi_out;
    const double *b;
    double *partials;
    double ax, ay;  // the FP64 operator's coefficients (not the V-cycle's T-rounded ones)
};

// Prolongation of the coarse correction to fine cell (i, j) in coarse cell
// (I, J) = (i / 2, j / 2); ncx x ncy: the coarse level.
//  * FP64 V-cycle: cell-centred bilinear, 9/16 of e(I, J), 3/16 of its x and
//    y neighbours on the fine cell's side and 1/16 of their diagonal
//    neighbour; past a wall the neighbour is the cell itself (the mirror ghost
//    of the Neumann operator).  With the 2 x 2 average restriction its orders
//    sum to 3 (> 2): the FP64 step's cold-started solve takes 7 cycles instead
//    of 8 to 1e-10 at 2048^2.
//  * FP32 V-cycle (the mixed step's preconditioner): e(I, J).  There the
//    contraction of the FP64 refinement is bounded by the FP32 correction's
//    rounding, not by the multigrid rate, and the bilinear form took more
//    cycles (5.0 against 4.35 per step over the first 25 steps at 2048^2).
template <class T>
__device__ __forceinline__ T prolong_at(const T *ec, int pc, int ncx, int ncy, int i, int j)
{
    const int I = i >> 1, J = j >> 1;
    if constexpr (sizeof(T) == 4) {
        (void)ncx;
        (void)ncy;
        return ec[I + (size_t)pc * J];
    }
    const int In = (i & 1) ? min(I + 1, ncx - 1) : max(I - 1, 0);
    const int Jn = (j & 1) ? min(J + 1, ncy - 1) : max(J - 1, 0);
    const size_t r0 = (size_t)pc * J, r1 = (size_t)pc * Jn;
    const T c = ec[I + r0], x = ec[In + r0], y = ec[I + r1], d = ec[In + r1];
    return (T(9) * c + T(3) * (x + y) + d) * T(0.0625);
}

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_smooth(const T *xin, const T *ec, int pc, T *xout, const T *rhs, int nx,
                                                int ny, int pitch, T ax, T ay, T *rc, int ncx, int ncy, int pcc,
                                                OuterArgs o, const Ctl *ctl)
{
    constexpr int H = MODE == 0 ? 4 : 5;
    using S = SmoothShape<H>;
    constexpr int SW = S::W, SHT = S::HT;
    __shared__ T sx[SHT][SW], sr[SHT][SW];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * SX, j0 = blockIdx.y * SY;
    int pl[S::PER], pj[S::PER];  // pair column (cell 2 pl) and row in the region; -1: none
#pragma unroll
    for (int q = 0; q < S::PER; ++q) {
        const int p = tid + 256 * q;
        pj[q] = p < S::PAIRS ? p / (SW / 2) : -1;
        pl[q] = p % (SW / 2);
    }
    // the loads are issued before the control flag is read, so their latency
    // overlaps the flag's (all addresses are in bounds either way)
    T xv[S::PER][2], rv[S::PER][2];
#pragma unroll
    for (int q = 0; q < S::PER; ++q) {
        const int lj = pj[q], j = j0 - H + lj;
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2) {
            const int li = 2 * pl[q] + c2, i = i0 - H + li;
            T x = 0, r = 0;
            if (lj >= 0 && i >= 0 && i < nx && j >= 0 && j < ny) {
                const size_t k = i + (size_t)pitch * j;
                if (xin) x = xin[k];
                if (ec) x += prolong_at(ec, pc, ncx, ncy, i, j);
                r = rhs[k];
            }
            xv[q][c2] = x;
            rv[q][c2] = r;
        }
    }
    if (ctl->mg_done | ctl->gm_stop) return;
#pragma unroll
    for (int q = 0; q < S::PER; ++q) {
        if (pj[q] < 0) continue;
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2) {
            sx[pj[q]][2 * pl[q] + c2] = xv[q][c2];
            sr[pj[q]][2 * pl[q] + c2] = rv[q][c2];
        }
    }
    __syncthreads();
    const T inv_d = T(1) / (T(2) * ax + T(2) * ay);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const int color = h & 1, lo = 1 + h;
#pragma unroll
        for (int q = 0; q < S::PER; ++q) {
            const int lj = pj[q];
            // i0 - H and j0 - H have the parity of H: the colour of (li, lj) is (li + lj + 2H) & 1
            const int li = 2 * pl[q] + ((lj & 1) ^ color);
            if (lj < lo || lj >= SHT - lo || li < lo || li >= SW - lo) continue;
            const int i = i0 - H + li, j = j0 - H + lj;
            if (i < 0 || i >= nx || j < 0 || j >= ny) continue;
            T s = 0;
            if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
                s = ax * sx[lj][li - 1] + ax * sx[lj][li + 1] + ay * sx[lj - 1][li] + ay * sx[lj + 1][li];
                sx[lj][li] = (s - sr[lj][li]) * inv_d;
            } else {
                T d = 0;
                if (i > 0) { s += ax * sx[lj][li - 1]; d += ax; }
                if (i < nx - 1) { s += ax * sx[lj][li + 1]; d += ax; }
                if (j > 0) { s += ay * sx[lj - 1][li]; d += ay; }
                if (j < ny - 1) { s += ay * sx[lj + 1][li]; d += ay; }
                sx[lj][li] = (s - sr[lj][li]) / d;
            }
        }
        __syncthreads();
    }
    if (MODE != 2) {
        for (int e = tid; e < SX * SY; e += 256) {
            const int li = H + (e % SX), lj = H + (e / SX), i = i0 + (e % SX), j = j0 + (e / SX);
            if (i < nx && j < ny) xout[i + (size_t)pitch * j] = sx[lj][li];
        }
    }
    if (MODE == 1) {  // residual of the tile, 2x2 averages: one coarse cell per thread (32 x 8)
        const int I = (i0 >> 1) + (tid & 31), J = (j0 >> 1) + (tid >> 5);
        const int i = 2 * I, j = 2 * J;
        if (I < ncx && J < ncy) {
            auto res = [&](int gi, int gj) {
                const int li = gi - i0 + H, lj = gj - j0 + H;
                const T P = sx[lj][li];
                const T W = gi > 0 ? sx[lj][li - 1] : P;
                const T E = gi < nx - 1 ? sx[lj][li + 1] : P;
                const T Sv = gj > 0 ? sx[lj - 1][li] : P;
                const T N = gj < ny - 1 ? sx[lj + 1][li] : P;
                return sr[lj][li] - ((E + W - T(2) * P) * ax + (N + Sv - T(2) * P) * ay);
            };
            const T t00 = res(i, j), t10 = res(i + 1, j), t01 = res(i, j + 1), t11 = res(i + 1, j + 1);
            rc[I + (size_t)pcc * J] = T(0.25) * (((t00 + t10) + t01) + t11);
        }
    }
    if (MODE == 2) {  // phi_out = phi_in + e, r = b - L phi_out (FP64), partials
        __shared__ double shr[16];
        double m = 0.0, sm = 0.0;
        for (int e = tid; e < SX * SY; e += 256) {
            const int li = H + (e % SX), lj = H + (e / SX), i = i0 + (e % SX), j = j0 + (e / SX);
            if (i >= nx || j >= ny) continue;
            const size_t k = i + (size_t)pitch * j;
            auto x = [&](size_t q, int dl, int dj) -> double { return o.phi_in[q] + (double)sx[lj + dj][li + dl]; };
            const double Pv = x(k, 0, 0);
            const double Wv = i > 0 ? x(k - 1, -1, 0) : Pv;
            const double Ev = i < nx - 1 ? x(k + 1, 1, 0) : Pv;
            const double Sv = j > 0 ? x(k - pitch, 0, -1) : Pv;
            const double Nv = j < ny - 1 ? x(k + pitch, 0, 1) : Pv;
            const double rr = o.b[k] - ((Ev + Wv - 2.0 * Pv) * o.ax + (Nv + Sv - 2.0 * Pv) * o.ay);
            o.phi_out[k] = Pv;
            xout[k] = (T)rr;  // the next residual
            m = (fabs(rr) > m || rr != rr) ? fabs(rr) : m;
            sm += Pv;
        }
        m = block_reduce<true>(m, shr);
        sm = block_reduce<false>(sm, shr + 8);
        if (tid == 0) {
            const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
            o.partials[2 * bid] = m;
            o.partials[2 * bid + 1] = sm;
        }
    }
}

// ------------------------------------------------- wave-column smoother
// The same four red-black half-sweeps (and the same fused restriction / FP64
// outer step per MODE) with no LDS and no barrier: one wavefront owns 64
// adjacent columns, the OC = 64 - 2H output columns plus an H-cell halo on
// each side, and NR = R + 2H rows of them in registers (R output rows plus the
// halo).  x-neighbours come from the adjacent lanes by DPP, y-neighbours are
// the registers of the rows above and below.  Each half-sweep leaves one more
// ring of lanes and rows stale; it skips the rows that are stale already.
// Cells outside the domain hold 0 and are never updated, so the Neumann walls
// only change the diagonal: x = (ax (W + E) + ay (S + N) - r) / d with
// d = cw + ce + cs + cn (ax or 0 per lane, ay or 0 per row).  The sweeps use the
// right-hand side pre-scaled by 1/d and the coefficients ax/d, ay/d.
// Waves are numbered chunk-major (adjacent waves take adjacent bands of one row
// chunk) and workgroups are handed to the 8 XCDs in contiguous ranges, so the
// halo lines two waves share are read through one XCD's L2.
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (w > v || w != w) ? w : v;
    }
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// four red-black half-sweeps of a register window whose first row jb has the
// parity of H (the window starts H rows below an even row): row q holds the
// cells of colour c in the lanes of parity (q + H + c) & 1.  rs = r / d.
// With YINT every updated row is an interior row of the domain.
template <class T, int NR, int H, bool YINT>
__device__ __forceinline__ void gsrb4(T (&x)[NR], const T (&rs)[NR], int jb, int ny, T a_mid, T b_mid, T a_edge,
                                      T b_edge, bool upd_even, bool upd_odd)
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
#pragma unroll
        for (int q = 1 + h; q < NR - 1 - h; ++q) {
            const int j = jb + q;
            T a = a_mid, b = b_mid;
            bool upd = ((q + H + h) & 1) ? upd_odd : upd_even;
            if (!YINT) {
                const bool edge = j == 0 || j == ny - 1;
                a = edge ? a_edge : a_mid;
                b = edge ? b_edge : b_mid;
                upd = upd && j >= 0 && j < ny;
            }
            const T we = lane_prev(x[q]) + lane_next(x[q]);
            const T sn = x[q - 1] + x[q + 1];
            const T nv = a * we + (b * sn - rs[q]);
            x[q] = upd ? nv : x[q];
        }
    }
}

template <class T, int MODE, int R>
__global__ __launch_bounds__(256) void k_gsrb(const T *xin, const T *ec, int pc, T *xout, const T *rhs, int nx,
                                              int ny, int pitch, T ax, T ay, T *rc, int ncx, int ncy, int pcc,
                                              OuterArgs o, const Ctl *ctl, int nbands, int nwaves)
{
    constexpr int H = MODE == 0 ? 4 : 5;
    constexpr int OC = 64 - 2 * H, NR = R + 2 * H;
    const int lane = threadIdx.x & 63;
    const int nblk = gridDim.x;
    const int blk = xcd_block(blockIdx.x, nblk);
    const int w = blk * 4 + (threadIdx.x >> 6);
    // a wave past the last one exits at once; in MODE 2 it stays for the
    // workgroup's reduction (the only barrier) and touches no memory
    const bool live = w < nwaves;
    if (MODE != 2 && !live) return;
    const int band = w % nbands, chunk = w / nbands;
    const int i0 = band * OC, j0 = chunk * R, jb = j0 - H;  // jb: the row of register 0
    const int i = i0 - H + lane;
    const bool in_i = live && i >= 0 && i < nx;
    // loads first (MODE 1 starts from zero); their latency overlaps the flag's.
    // Running offsets (dereferenced only in range) keep one address live.
    const bool yin = jb >= 0 && jb + NR <= ny;  // every window row is a domain row
    T x[NR], r[NR];
    {
        int k = i + pitch * jb;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int j = jb + q;
            T xv = 0, rv = 0;
            if (in_i && (yin || (j >= 0 && j < ny))) {
                rv = rhs[k];
                if (MODE != 1) xv = xin[k] + prolong_at(ec, pc, ncx, ncy, i, j);
            }
            x[q] = xv;
            r[q] = rv;
            k += pitch;
        }
    }
    // MODE 2: the FP64 iterate on the output rows and one row either side, b on the output rows
    constexpr int NP = MODE == 2 ? R + 2 : 1, NB = MODE == 2 ? R : 1;
    double ph[NP], bb[NB];
    if (MODE == 2) {
        int k = i + pitch * (j0 - 1);
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            const int j = j0 - 1 + t;
            ph[t] = (in_i && j >= 0 && j < ny) ? o.phi_in[k] : 0.0;
            if (t >= 1 && t <= NB) bb[t - 1] = (in_i && j < ny) ? o.b[k] : 0.0;
            k += pitch;
        }
    }
    if (ctl->mg_done | ctl->gm_stop) return;
    const T cw = i > 0 ? ax : T(0), ce = i < nx - 1 ? ax : T(0);
    const T d_mid = cw + ce + T(2) * ay, d_edge = cw + ce + ay;  // interior rows, the two wall rows
    const T inv_mid = T(1) / d_mid, inv_edge = T(1) / d_edge;
#pragma unroll
    for (int q = 0; q < NR; ++q) {  // r -> r / d
        const int j = jb + q;
        r[q] *= (j > 0 && j < ny - 1) ? inv_mid : inv_edge;
    }
    const bool odd = (i & 1) != 0, upd_odd = in_i && odd, upd_even = in_i && !odd;
    // most windows hold no wall row and no row outside the domain: their
    // sweeps need no per-row coefficients
    if (yin)
        gsrb4<T, NR, H, true>(x, r, jb, ny, ax * inv_mid, ay * inv_mid, ax * inv_edge, ay * inv_edge, upd_even,
                              upd_odd);
    else
        gsrb4<T, NR, H, false>(x, r, jb, ny, ax * inv_mid, ay * inv_mid, ax * inv_edge, ay * inv_edge, upd_even,
                               upd_odd);
    const bool out_lane = lane >= H && lane < H + OC && i < nx;
    if (MODE != 2) {
#pragma unroll
        for (int q = H; q < H + R; ++q) {
            const int j = jb + q;
            if (out_lane && j < ny) xout[i + (size_t)pitch * j] = x[q];
        }
    }
    if (MODE == 1) {  // residual of the output cells, 2x2 averages into the coarse right-hand side
        auto res = [&](int q) {
            const int j = jb + q;
            const T P = x[q];
            const T Wl = lane_prev(P), El = lane_next(P);
            const T Wv = i > 0 ? Wl : P;
            const T Ev = i < nx - 1 ? El : P;
            const T Sv = j > 0 ? x[q - 1] : P;
            const T Nv = j < ny - 1 ? x[q + 1] : P;
            const T rq = r[q] * ((j > 0 && j < ny - 1) ? d_mid : d_edge);  // r / d * d
            return rq - ((Ev + Wv - T(2) * P) * ax + (Nv + Sv - T(2) * P) * ay);
        };
#pragma unroll
        for (int m = 0; m < R / 2; ++m) {
            const int q = H + 2 * m, j = jb + q;
            const T t00 = res(q), t01 = res(q + 1);
            const T t10 = lane_next(t00), t11 = lane_next(t01);
            const int I = i >> 1, J = j >> 1;
            if (out_lane && !odd && I < ncx && J < ncy)
                rc[I + (size_t)pcc * J] = T(0.25) * (((t00 + t10) + t01) + t11);
        }
    }
    if (MODE == 2) {  // phi_out = phi_in + e, r = b - L phi_out (FP64), wave partials
        double P[NP];
#pragma unroll
        for (int t = 0; t < NP; ++t) P[t] = ph[t] + (double)x[H - 1 + t];
        double m = 0.0, sm = 0.0;
#pragma unroll
        for (int t = 1; t <= R; ++t) {
            const int j = j0 - 1 + t;
            const double Pv = P[t];
            const double Wl = lane_prev(Pv), El = lane_next(Pv);
            const double Wv = i > 0 ? Wl : Pv;
            const double Ev = i < nx - 1 ? El : Pv;
            const double Sv = j > 0 ? P[t - 1] : Pv;
            const double Nv = j < ny - 1 ? P[t + 1] : Pv;
            const double rr = bb[t - 1] - ((Ev + Wv - 2.0 * Pv) * o.ax + (Nv + Sv - 2.0 * Pv) * o.ay);
            if (out_lane && j < ny) {
                const size_t k = i + (size_t)pitch * j;
                o.phi_out[k] = Pv;
                xout[k] = (T)rr;  // the next residual
                m = (fabs(rr) > m || rr != rr) ? fabs(rr) : m;
                sm += Pv;
            }
        }
        m = wave_max(m);
        sm = wave_sum(sm);
        __shared__ double shm[4], shs[4];
        const int wv = threadIdx.x >> 6;
        if (lane == 0) {
            shm[wv] = m;
            shs[wv] = sm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // one partial per workgroup, combined in wave order
            double mm = shm[0], ss = shs[0];
            for (int q = 1; q < 4; ++q) {
                mm = (shm[q] > mm || shm[q] != shm[q]) ? shm[q] : mm;
                ss += shs[q];
            }
            o.partials[2 * (size_t)blockIdx.x] = mm;
            o.partials[2 * (size_t)blockIdx.x + 1] = ss;
        }
    }
}

// ------------------------------------------ paired-column wave smoother
// The wave-column smoother with two adjacent columns per lane: lane l holds
// the even column c = c0 + 2l (x0) and the odd column c + 1 (x1) of 128
// columns (HX halo columns either side, HX even).  In a row, one colour lies in
// the x0 cells and the other in the x1 cells, so each half-sweep updates one
// fixed register per row in every lane (no select, no idle lane): an x0 cell
// reads its west neighbour from lane l - 1's x1 (DPP) and its east one from
// the lane's own x1, an x1 cell the reverse.  Cells outside the domain hold 0
// and get zero coefficients, so they stay 0.  The 2x2 restriction of the down
// leg is lane-local (a coarse cell is one lane's two columns in two rows).
template <class T>
struct Pair;
template <>
struct Pair<float> {
    using V = float2;
};
template <>
struct Pair<double> {
    using V = double2;
};

template <class T, int NR, int H, bool YINT, int HS = 4>
__device__ __forceinline__ void gsrb4_pair(T (&x0)[NR], T (&x1)[NR], const T (&rs0)[NR], const T (&rs1)[NR], int jb,
                                           int ny, const T (&A)[2][2], const T (&B)[2][2], bool in_i)
{
    // A[cell][edge], B[cell][edge]: ax/d and ay/d of the even (0) and odd (1)
    // cell on an interior (0) or wall (1) row; zero outside the domain; HS
    // half-sweeps (colours 0, 1, 0, ..)
#pragma unroll
    for (int h = 0; h < HS; ++h) {
#pragma unroll
        for (int q = 1 + h; q < NR - 1 - h; ++q) {
            const int j = jb + q;
            // the even cell of row q has colour (q + H) & 1: it is updated when that is h's colour
            const bool even = ((q + H + h) & 1) == 0;
            const int cell = even ? 0 : 1;
            T a = A[cell][0], b = B[cell][0];
            if (!YINT) {
                const bool edge = j == 0 || j == ny - 1, rin = j >= 0 && j < ny;
                a = rin ? (edge ? A[cell][1] : A[cell][0]) : T(0);
                b = rin ? (edge ? B[cell][1] : B[cell][0]) : T(0);
            }
            if (even) {
                const T we = lane_prev(x1[q]) + x1[q];
                const T sn = x0[q - 1] + x0[q + 1];
                x0[q] = a * we + (b * sn - rs0[q]);
            } else {
                const T we = x0[q] + lane_next(x0[q]);
                const T sn = x1[q - 1] + x1[q + 1];
                x1[q] = a * we + (b * sn - rs1[q]);
            }
        }
    }
    (void)in_i;
}

// LATE (MODE 2): phi_in and b, needed only by the outer step, are loaded
// after the sweeps, so their registers are not live during them (more waves
// per SIMD, at the cost of one more wait for memory)
template <class T, int MODE, int R, bool LATE = false, int HS = 4>
__global__ __launch_bounds__(256) void k_gsrb2(const T *xin, const T *ec, int pc, T *xout, const T *rhs, int nx,
                                               int ny, int pitch, T ax, T ay, T *rc, int ncx, int ncy, int pcc,
                                               OuterArgs o, const Ctl *ctl, int nbands, int nwaves)
{
    using V = typename Pair<T>::V;
    // halo rows and columns: one per half-sweep, one more for the residual of MODE 1 and 2
    constexpr int H = MODE == 0 ? HS : HS + 1, HX = MODE == 0 ? HS : HS + 2;
    constexpr int OC = 128 - 2 * HX, NR = R + 2 * H;
    const int lane = threadIdx.x & 63;
    const int nblk = gridDim.x;
    const int blk = xcd_block(blockIdx.x, nblk);
    const int w = blk * 4 + (threadIdx.x >> 6);
    // a wave past the last one exits at once; in MODE 2 it stays for the
    // workgroup's reduction (the only barrier) and touches no memory
    const bool live = w < nwaves;
    if (MODE != 2 && !live) return;
    const int band = w % nbands, chunk = w / nbands;
    const int i0 = band * OC, j0 = chunk * R, jb = j0 - H;
    const int ca = i0 - HX + 2 * lane;  // the lane's even column; nx is even, so both or neither are inside
    const bool in_i = live && ca >= 0 && ca < nx;
    const bool yin = jb >= 0 && jb + NR <= ny;
    T x0[NR], x1[NR], r0[NR], r1[NR];
    {
        int k = ca + pitch * jb;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int j = jb + q;
            V xv = {0, 0}, rv = {0, 0};
            if (in_i && (yin || (j >= 0 && j < ny))) {
                rv = *reinterpret_cast<const V *>(rhs + k);
                if (MODE != 1) {
                    xv = *reinterpret_cast<const V *>(xin + k);
                    xv.x += prolong_at(ec, pc, ncx, ncy, ca, j);
                    xv.y += prolong_at(ec, pc, ncx, ncy, ca + 1, j);
                }
            }
            x0[q] = xv.x;
            x1[q] = xv.y;
            r0[q] = rv.x;
            r1[q] = rv.y;
            k += pitch;
        }
    }
    constexpr int NP = MODE == 2 ? R + 2 : 1, NB = MODE == 2 ? R : 1;
    double ph0[NP], ph1[NP], bb0[NB], bb1[NB];
    auto load_outer = [&]() {
        int k = ca + pitch * (j0 - 1);
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            const int j = j0 - 1 + t;
            double2 pv = {0.0, 0.0};
            if (in_i && j >= 0 && j < ny) pv = *reinterpret_cast<const double2 *>(o.phi_in + k);
            ph0[t] = pv.x;
            ph1[t] = pv.y;
            if (t >= 1 && t <= NB) {
                double2 bv = {0.0, 0.0};
                if (in_i && j < ny) bv = *reinterpret_cast<const double2 *>(o.b + k);
                bb0[t - 1] = bv.x;
                bb1[t - 1] = bv.y;
            }
            k += pitch;
        }
    };
    if (MODE == 2 && !LATE) load_outer();
    if (ctl->mg_done | ctl->gm_stop) return;
    // per cell: d = cw + ce + cs + cn; the even cell has a west wall at column 0,
    // the odd one an east wall at column nx - 1
    const T cw0 = ca > 0 ? ax : T(0), ce1 = ca + 1 < nx - 1 ? ax : T(0);
    const T d0m = cw0 + ax + T(2) * ay, d0e = cw0 + ax + ay, d1m = ax + ce1 + T(2) * ay, d1e = ax + ce1 + ay;
    const T i0m = T(1) / d0m, i0e = T(1) / d0e, i1m = T(1) / d1m, i1e = T(1) / d1e;
    T A[2][2], B[2][2];
    A[0][0] = in_i ? ax * i0m : T(0);
    A[0][1] = in_i ? ax * i0e : T(0);
    A[1][0] = in_i ? ax * i1m : T(0);
    A[1][1] = in_i ? ax * i1e : T(0);
    B[0][0] = in_i ? ay * i0m : T(0);
    B[0][1] = in_i ? ay * i0e : T(0);
    B[1][0] = in_i ? ay * i1m : T(0);
    B[1][1] = in_i ? ay * i1e : T(0);
#pragma unroll
    for (int q = 0; q < NR; ++q) {  // r -> r / d
        const int j = jb + q;
        const bool mid = j > 0 && j < ny - 1;
        r0[q] *= mid ? i0m : i0e;
        r1[q] *= mid ? i1m : i1e;
    }
    if (yin)
        gsrb4_pair<T, NR, H, true, HS>(x0, x1, r0, r1, jb, ny, A, B, in_i);
    else
        gsrb4_pair<T, NR, H, false, HS>(x0, x1, r0, r1, jb, ny, A, B, in_i);
    const bool out_lane = in_i && lane >= HX / 2 && lane < HX / 2 + OC / 2;
    if (MODE != 2) {
#pragma unroll
        for (int q = H; q < H + R; ++q) {
            const int j = jb + q;
            if (out_lane && j < ny) *reinterpret_cast<V *>(xout + ca + (size_t)pitch * j) = V{x0[q], x1[q]};
        }
    }
    if (MODE == 1) {  // residual of the output cells, 2x2 averages (lane-local) into the coarse right-hand side
        auto res = [&](int q, T &t0, T &t1) {
            const int j = jb + q;
            const bool mid = j > 0 && j < ny - 1;
            const T P0 = x0[q], P1 = x1[q];
            // the lane shifts run in every lane (a shift inside a branch would
            // read lanes that the branch switched off)
            const T Wl = lane_prev(P1), El = lane_next(P0);
            const T W0 = ca > 0 ? Wl : P0;          // west of the even cell
            const T E1 = ca + 1 < nx - 1 ? El : P1;  // east of the odd cell
            const T S0 = j > 0 ? x0[q - 1] : P0, N0 = j < ny - 1 ? x0[q + 1] : P0;
            const T S1 = j > 0 ? x1[q - 1] : P1, N1 = j < ny - 1 ? x1[q + 1] : P1;
            const T rq0 = r0[q] * (mid ? d0m : d0e), rq1 = r1[q] * (mid ? d1m : d1e);
            t0 = rq0 - ((P1 + W0 - T(2) * P0) * ax + (N0 + S0 - T(2) * P0) * ay);
            t1 = rq1 - ((E1 + P0 - T(2) * P1) * ax + (N1 + S1 - T(2) * P1) * ay);
        };
#pragma unroll
        for (int m = 0; m < R / 2; ++m) {
            const int q = H + 2 * m, j = jb + q;
            T t00, t10, t01, t11;
            res(q, t00, t10);
            res(q + 1, t01, t11);
            const int I = ca >> 1, J = j >> 1;
            if (out_lane && I < ncx && J < ncy) rc[I + (size_t)pcc * J] = T(0.25) * (((t00 + t10) + t01) + t11);
        }
    }
    if (MODE == 2) {  // phi_out = phi_in + e, r = b - L phi_out (FP64), one partial per workgroup
        if (LATE) {
            __builtin_amdgcn_sched_barrier(0);  // keep the loads below the sweeps
            load_outer();
        }
        double P0[NP], P1[NP];
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            P0[t] = ph0[t] + (double)x0[H - 1 + t];
            P1[t] = ph1[t] + (double)x1[H - 1 + t];
        }
        double m = 0.0, sm = 0.0;
#pragma unroll
        for (int t = 1; t <= R; ++t) {
            const int j = j0 - 1 + t;
            const double p0 = P0[t], p1 = P1[t];
            const double Wl = lane_prev(p1), El = lane_next(p0);
            const double W0 = ca > 0 ? Wl : p0;
            const double E1 = ca + 1 < nx - 1 ? El : p1;
            const double S0 = j > 0 ? P0[t - 1] : p0, N0 = j < ny - 1 ? P0[t + 1] : p0;
            const double S1 = j > 0 ? P1[t - 1] : p1, N1 = j < ny - 1 ? P1[t + 1] : p1;
            const double rr0 = bb0[t - 1] - ((p1 + W0 - 2.0 * p0) * o.ax + (N0 + S0 - 2.0 * p0) * o.ay);
            const double rr1 = bb1[t - 1] - ((E1 + p0 - 2.0 * p1) * o.ax + (N1 + S1 - 2.0 * p1) * o.ay);
            if (out_lane && j < ny) {
                const size_t k = ca + (size_t)pitch * j;
                *reinterpret_cast<double2 *>(o.phi_out + k) = double2{p0, p1};
                *reinterpret_cast<V *>(xout + k) = V{(T)rr0, (T)rr1};  // the next residual
                m = (fabs(rr0) > m || rr0 != rr0) ? fabs(rr0) : m;
                m = (fabs(rr1) > m || rr1 != rr1) ? fabs(rr1) : m;
                sm += p0;
                sm += p1;
            }
        }
        m = wave_max(m);
        sm = wave_sum(sm);
        __shared__ double shm[4], shs[4];
        const int wv = threadIdx.x >> 6;
        if (lane == 0) {
            shm[wv] = m;
            shs[wv] = sm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // one partial per workgroup, combined in wave order
            double mm = shm[0], ss = shs[0];
            for (int q = 1; q < 4; ++q) {
                mm = (shm[q] > mm || shm[q] != shm[q]) ? shm[q] : mm;
                ss += shs[q];
            }
            o.partials[2 * (size_t)blockIdx.x] = mm;
            o.partials[2 * (size_t)blockIdx.x + 1] = ss;
        }
    }
}

template <class T>
__device__ __forceinline__ T lap_at(const T *e, int i, int j, int nx, int ny, int pitch, T ax, T ay)
{
    const size_t k = i + (size_t)pitch * j;
    const T P = e[k];
    const T W = i > 0 ? e[k - 1] : P;
    const T E = i < nx - 1 ? e[k + 1] : P;
    const T S = j > 0 ? e[k - pitch] : P;
    const T N = j < ny - 1 ? e[k + pitch] : P;
    return (E + W - T(2) * P) * ax + (N + S - T(2) * P) * ay;
}

// ------------------------------------------------------ single-block tail
struct TailLevels {
    int nlev;          // levels in the tail (the last one is the bottom)
    int nx[12], ny[12];
    int off[12];       // offset of each level in the LDS arrays
    double ax[12], ay[12];
    int total;         // cells over all tail levels
    int mb;            // bottom cells
};

// e / d for 0 <= e < 2^20 and d <= 64 through a float reciprocal: the exact
// quotient's fractional part stays >= 0.5/d away from an integer
__device__ __forceinline__ int qdiv(int e, float inv) { return (int)(((float)e + 0.5f) * inv); }

template <class T>
__device__ void tail_gsrb(T *E, const T *Rr, int nx, int ny, T ax, T ay, int color)
{
    const T inv_d = T(1) / (T(2) * ax + T(2) * ay);
    const int half = (nx + 1) / 2;
    const float ih = 1.0f / (float)half;
    for (int e = threadIdx.x; e < half * ny; e += blockDim.x) {
        const int j = qdiv(e, ih), i = 2 * (e - j * half) + ((j + color) & 1);
        if (i >= nx) continue;
        const int k = i + nx * j;
        T s = 0;
        if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
            s = ax * E[k - 1] + ax * E[k + 1] + ay * E[k - nx] + ay * E[k + nx];
            E[k] = (s - Rr[k]) * inv_d;
        } else {
            T d = 0;
            if (i > 0) { s += ax * E[k - 1]; d += ax; }
            if (i < nx - 1) { s += ax * E[k + 1]; d += ax; }
            if (j > 0) { s += ay * E[k - nx]; d += ay; }
            if (j < ny - 1) { s += ay * E[k + nx]; d += ay; }
            E[k] = (s - Rr[k]) / d;
        }
    }
    __syncthreads();
}

// V-cycle of the coarse levels (all <= 64^2) in one workgroup: r of the top
// tail level in, its correction out.
template <class T>
__global__ __launch_bounds__(1024) void k_tail(const T *rtop, int ptop, T *etop, TailLevels tl, const T *pinv,
                                               const Ctl *ctl)
{
    extern __shared__ unsigned char smem_raw[];
    T *E = reinterpret_cast<T *>(smem_raw);
    T *Rr = E + tl.total;
    const int tid = threadIdx.x, nt = blockDim.x;
    {  // top level r (at most 64^2 = 4 per thread), loaded before the control flag is read
        const int nx = tl.nx[0], ny = tl.ny[0];
        const float inx = 1.0f / (float)nx;
        T v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = tid + q * nt, j = qdiv(k, inx);
            v[q] = k < nx * ny ? rtop[(k - j * nx) + (size_t)ptop * j] : T(0);
        }
        if (ctl->mg_done | ctl->gm_stop) return;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (tid + q * nt < nx * ny) Rr[tid + q * nt] = v[q];
        for (int k = tid + 4 * nt; k < nx * ny; k += nt) Rr[k] = rtop[(k % nx) + (size_t)ptop * (k / nx)];
    }
    __syncthreads();
    for (int l = 0; l < tl.nlev - 1; ++l) {
        const int nx = tl.nx[l], ny = tl.ny[l];
        T *El = E + tl.off[l];
        const T *Rl = Rr + tl.off[l];
        const T ax = (T)tl.ax[l], ay = (T)tl.ay[l];
        for (int k = tid; k < nx * ny; k += nt) El[k] = T(0);
        __syncthreads();
        for (int s = 0; s < 2; ++s) { tail_gsrb(El, Rl, nx, ny, ax, ay, 0); tail_gsrb(El, Rl, nx, ny, ax, ay, 1); }
        const int cx = tl.nx[l + 1], cy = tl.ny[l + 1];
        T *Rc = Rr + tl.off[l + 1];
        const float icx = 1.0f / (float)cx;
        for (int k = tid; k < cx * cy; k += nt) {
            const int J = qdiv(k, icx), I = k - J * cx, i = 2 * I, j = 2 * J;
            const T t00 = Rl[i + nx * j] - lap_at(El, i, j, nx, ny, nx, ax, ay);
            const T t10 = Rl[i + 1 + nx * j] - lap_at(El, i + 1, j, nx, ny, nx, ax, ay);
            const T t01 = Rl[i + nx * (j + 1)] - lap_at(El, i, j + 1, nx, ny, nx, ax, ay);
            const T t11 = Rl[i + 1 + nx * (j + 1)] - lap_at(El, i + 1, j + 1, nx, ny, nx, ax, ay);
            Rc[k] = T(0.25) * (((t00 + t10) + t01) + t11);
        }
        __syncthreads();
    }
    {  // bottom: e = pinv r (pinv symmetric: read it column-wise for coalescing);
       // up to four threads share a row, each over a quarter of the columns
        const int b = tl.nlev - 1, m = tl.mb;
        T *Eb = E + tl.off[b];
        const T *Rb = Rr + tl.off[b];
        T *Pb = Rr + tl.total;  // parts x m partial sums
        const int parts = m > 0 ? max(1, min(4, nt / m)) : 1;
        for (int q = tid; q < m * parts; q += nt) {
            const int row = q % m, part = q / m;
            const int k0 = part * m / parts, k1 = (part + 1) * m / parts;
            T s = 0;
            for (int k = k0; k < k1; ++k) s += pinv[(size_t)k * m + row] * Rb[k];
            Pb[q] = s;
        }
        __syncthreads();
        for (int row = tid; row < m; row += nt) {
            T s = Pb[row];
            for (int part = 1; part < parts; ++part) s += Pb[row + part * m];
            Eb[row] = s;
        }
        __syncthreads();
    }
    for (int l = tl.nlev - 2; l >= 0; --l) {
        const int nx = tl.nx[l], ny = tl.ny[l], cx = tl.nx[l + 1];
        T *El = E + tl.off[l];
        const T *Ec = E + tl.off[l + 1];
        const T *Rl = Rr + tl.off[l];
        const T ax = (T)tl.ax[l], ay = (T)tl.ay[l];
        const float inx = 1.0f / (float)nx;
        const int cy = tl.ny[l + 1];
        for (int k = tid; k < nx * ny; k += nt) {
            const int j = qdiv(k, inx);
            El[k] += prolong_at(Ec, cx, cx, cy, k - j * nx, j);
        }
        __syncthreads();
        for (int s = 0; s < 2; ++s) { tail_gsrb(El, Rl, nx, ny, ax, ay, 0); tail_gsrb(El, Rl, nx, ny, ax, ay, 1); }
    }
    {
        const int nx = tl.nx[0], ny = tl.ny[0];
        for (int k = tid; k < nx * ny; k += nt) etop[(k % nx) + (size_t)ptop * (k / nx)] = E[k];
    }
}

// ------------------------------------------------- exact tail (DCT solve)
// The levels of at most 64 x 64 cells can be replaced by the exact Neumann
// solve on the top one: the 5-point operator with mirror walls is diagonal in
// the orthonormal DCT-II basis, L = Cy D Cx^T with D(m, k) = -ax (2 - 2 cos(pi
// k / nx)) - ay (2 - 2 cos(pi m / ny)), so e = Cy (D^+ (Cy^T r Cx)) Cx^T, the
// (0, 0) mode (the constant) set to zero as the pseudo-inverse does.  Four
// 64 x 64 x 64 matrix products in one workgroup (zero padding to 64), on the
// matrix cores (FP64: or on the VALU); the cosine tables are staged in LDS
// with a padded row stride against bank conflicts.
constexpr int kDct = 64, kDctLd = 65;

// table layout (global): Cx [64 x 64], Cy [64 x 64], D^+ [64 (m) x 64 (k)]
// O = op(A) op(B) (each 64 x 64 in LDS at stride kDctLd), times scale (global,
// 64 x 64) elementwise when it is given.  FP64: VALU, each thread a 4 x 4 block.
template <class T, bool TA, bool TB>
__device__ __forceinline__ void dct_gemm(const T *A, const T *B, T *O, const T *scale)
{
    const int t = threadIdx.x, a0 = 4 * (t >> 4), b0 = 4 * (t & 15);
    T acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = T(0);
#pragma unroll 4
    for (int c = 0; c < kDct; ++c) {
        T av[4], bv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) av[r] = TA ? A[c * kDctLd + a0 + r] : A[(a0 + r) * kDctLd + c];
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = TB ? B[(b0 + q) * kDctLd + c] : B[c * kDctLd + b0 + q];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[r][q] += av[r] * bv[q];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            O[(a0 + r) * kDctLd + b0 + q] = scale ? acc[r][q] * scale[(a0 + r) * kDct + b0 + q] : acc[r][q];
}

// FP32: matrix cores.  v_mfma_f32_16x16x4f32 multiplies a 16 x 4 block of A
// (lane l holds A[l % 16][l / 16]) by a 4 x 16 block of B (lane l holds
// B[l / 16][l % 16]) into a 16 x 16 block of D (lane l holds D[4 (l / 16) + v]
// [l % 16], v = 0..3).  Wave w computes rows 16w..16w+15 of O as four such
// blocks, stepping through k four at a time.
typedef float dct_v4f __attribute__((ext_vector_type(4)));
template <bool TA, bool TB>
__device__ __forceinline__ void dct_gemm_mfma(const float *A, const float *B, float *O, const float *scale)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 15, lk = l >> 4;
    dct_v4f acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = dct_v4f{0.f, 0.f, 0.f, 0.f};
    const int ar = 16 * w + li;
#pragma unroll 4
    for (int k0 = 0; k0 < kDct; k0 += 4) {
        const int kk = k0 + lk;
        const float a = TA ? A[kk * kDctLd + ar] : A[ar * kDctLd + kk];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int bc = 16 * cb + li;
            const float b = TB ? B[bc * kDctLd + kk] : B[kk * kDctLd + bc];
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * w + 4 * lk + v, col = 16 * cb + li;
            O[row * kDctLd + col] = scale ? acc[cb][v] * scale[row * kDct + col] : acc[cb][v];
        }
}

// FP64 on the matrix cores: v_mfma_f64_16x16x4f64.  A and B as for the FP32
// instruction (one double per lane); the 16 x 16 result differs: lane l holds
// D[l / 16 + 4 v][l % 16], v = 0..3.
typedef double dct_v4d __attribute__((ext_vector_type(4)));
template <bool TA, bool TB>
__device__ __forceinline__ void dct_gemm_mfma64(const double *A, const double *B, double *O, const double *scale)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 15, lk = l >> 4;
    dct_v4d acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = dct_v4d{0.0, 0.0, 0.0, 0.0};
    const int ar = 16 * w + li;
#pragma unroll 4
    for (int k0 = 0; k0 < kDct; k0 += 4) {
        const int kk = k0 + lk;
        const double a = TA ? A[kk * kDctLd + ar] : A[ar * kDctLd + kk];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int bc = 16 * cb + li;
            const double b = TB ? B[bc * kDctLd + kk] : B[kk * kDctLd + bc];
            acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * w + lk + 4 * v, col = 16 * cb + li;
            O[row * kDctLd + col] = scale ? acc[cb][v] * scale[row * kDct + col] : acc[cb][v];
        }
}

// FP64 with 16 waves: wave w computes the one 16 x 16 block (w / 4, w % 4) of O
template <bool TA, bool TB>
__device__ __forceinline__ void dct_gemm_mfma64_16(const double *A, const double *B, double *O, const double *scale)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 15, lk = l >> 4;
    const int rb = w >> 2, cb = w & 3;
    dct_v4d acc = dct_v4d{0.0, 0.0, 0.0, 0.0};
    const int ar = 16 * rb + li, bc = 16 * cb + li;
#pragma unroll 4
    for (int k0 = 0; k0 < kDct; k0 += 4) {
        const int kk = k0 + lk;
        const double a = TA ? A[kk * kDctLd + ar] : A[ar * kDctLd + kk];
        const double b = TB ? B[bc * kDctLd + kk] : B[kk * kDctLd + bc];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = 16 * rb + lk + 4 * v, col = 16 * cb + li;
        O[row * kDctLd + col] = scale ? acc[v] * scale[row * kDct + col] : acc[v];
    }
}

// FP32 with 16 waves: wave w computes the one 16 x 16 block (w / 4, w % 4) of O
template <bool TA, bool TB>
__device__ __forceinline__ void dct_gemm_mfma16(const float *A, const float *B, float *O, const float *scale)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, li = l & 15, lk = l >> 4;
    const int rb = w >> 2, cb = w & 3;
    dct_v4f acc = dct_v4f{0.f, 0.f, 0.f, 0.f};
    const int ar = 16 * rb + li, bc = 16 * cb + li;
#pragma unroll 4
    for (int k0 = 0; k0 < kDct; k0 += 4) {
        const int kk = k0 + lk;
        const float a = TA ? A[kk * kDctLd + ar] : A[ar * kDctLd + kk];
        const float b = TB ? B[bc * kDctLd + kk] : B[kk * kDctLd + bc];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = 16 * rb + 4 * lk + v, col = 16 * cb + li;
        O[row * kDctLd + col] = scale ? acc[v] * scale[row * kDct + col] : acc[v];
    }
}

// MF: FP64 products on the matrix cores (FP32 ones always are); NW: waves in
// the workgroup (4, or 16 for the products on the matrix cores)
template <class T, bool TA, bool TB, bool MF, int NW = 4>
__device__ __forceinline__ void dct_product(const T *A, const T *B, T *O, const T *scale)
{
    static_assert(NW == 4 || (NW == 16 && (MF || sizeof(T) == 4)), "16 waves: matrix-core products only");
    if constexpr (sizeof(T) == 8 && MF && NW == 16)
        dct_gemm_mfma64_16<TA, TB>((const double *)A, (const double *)B, (double *)O, (const double *)scale);
    else if constexpr (sizeof(T) == 4 && NW == 16)
        dct_gemm_mfma16<TA, TB>((const float *)A, (const float *)B, (float *)O, (const float *)scale);
    else if constexpr (sizeof(T) == 4)
        dct_gemm_mfma<TA, TB>((const float *)A, (const float *)B, (float *)O, (const float *)scale);
    else if constexpr (MF)
        dct_gemm_mfma64<TA, TB>((const double *)A, (const double *)B, (double *)O, (const double *)scale);
    else
        dct_gemm<T, TA, TB>(A, B, O, scale);
}

// the DCT solve of the workgroup: R in B0 (zero outside ny x nx), the tables
// in Cx, Cy (LDS) and dinv (global); E in B0 on return (B1 overwritten)
template <class T, bool MF = true, int NW = 4>
__device__ __forceinline__ void dct_solve(const T *Cx, const T *Cy, T *B0, T *B1, const T *dinv)
{
    dct_product<T, false, false, MF, NW>(B0, Cx, B1, (const T *)nullptr);  // T1 = R Cx
    __syncthreads();
    dct_product<T, true, false, MF, NW>(Cy, B1, B0, dinv);  // S = (Cy^T T1) D^+
    __syncthreads();
    dct_product<T, false, false, MF, NW>(Cy, B0, B1, (const T *)nullptr);  // T2 = Cy S
    __syncthreads();
    dct_product<T, false, true, MF, NW>(B1, Cx, B0, (const T *)nullptr);  // E = T2 Cx^T
    __syncthreads();
}

// NW: waves in the workgroup (4, or 16 with the products on the matrix cores)
template <class T, bool MF, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_tail_dct(const T *rtop, int ptop, T *etop, int nx, int ny,
                                                      const T *tab, const Ctl *ctl)
{
    __shared__ T Cx[kDct * kDctLd], Cy[kDct * kDctLd], B0[kDct * kDctLd], B1[kDct * kDctLd];
    constexpr int NT = 64 * NW, NQ = kDct * kDct / NT;
    const int t = threadIdx.x;
    // loads first (the tables and r, zero-padded to 64 x 64); their latency overlaps the flag's
    T vx[NQ], vy[NQ], vr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + NT * q, row = e >> 6, col = e & 63;
        vx[q] = tab[e];
        vy[q] = tab[kDct * kDct + e];
        vr[q] = row < ny && col < nx ? rtop[col + (size_t)ptop * row] : T(0);
    }
    if (ctl->mg_done | ctl->gm_stop) return;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + NT * q, row = e >> 6, col = e & 63;
        Cx[row * kDctLd + col] = vx[q];
        Cy[row * kDctLd + col] = vy[q];
        B0[row * kDctLd + col] = vr[q];
    }
    __syncthreads();
    dct_solve<T, MF, NW>(Cx, Cy, B0, B1, tab + 2 * kDct * kDct);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + NT * q, row = e >> 6, col = e & 63;
        if (row < ny && col < nx) etop[col + (size_t)ptop * row] = B0[row * kDctLd + col];
    }
}

// The level above the DCT level (twice its size: ny1 = 4 R1 rows, nx1 <= 128
// columns) in the same workgroup, FP32: wave w owns rows [w R1, (w + 1) R1)
// and every column (two per lane, as k_gsrb2), and trades its first and last
// row with the neighbouring waves through LDS after every half-sweep (double
// buffered: one barrier per exchange).  Down leg: two GSRB sweeps from zero,
// the residual and its 2x2 average as the DCT input; then the DCT solve; up
// leg: the coarse correction added, two sweeps, the level's correction stored.
// Replaces that level's two smoother launches.
template <int R1, int NW>
__device__ __forceinline__ void ext_sweeps(float (&x0)[R1 + 2], float (&x1)[R1 + 2], const float (&rs0)[R1],
                                           const float (&rs1)[R1], const float (&A)[2][2], const float (&B)[2][2],
                                           int j0, int ny1, float (*xch)[NW][2][128], int &par, int w, int lane,
                                           bool last_exchange)
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
#pragma unroll
        for (int q = 1; q <= R1; ++q) {
            const int j = j0 + q - 1;
            const bool even = ((q + 1 + h) & 1) == 0;  // j0 is even: row j has the parity of q - 1
            const int cell = even ? 0 : 1;
            const bool edge = j == 0 || j == ny1 - 1;
            const float a = edge ? A[cell][1] : A[cell][0], b = edge ? B[cell][1] : B[cell][0];
            if (even) {
                const float we = lane_prev(x1[q]) + x1[q];
                const float sn = x0[q - 1] + x0[q + 1];
                x0[q] = a * we + (b * sn - rs0[q - 1]);
            } else {
                const float we = x0[q] + lane_next(x0[q]);
                const float sn = x1[q - 1] + x1[q + 1];
                x1[q] = a * we + (b * sn - rs1[q - 1]);
            }
        }
        if (h < 3 || last_exchange) {
            xch[par][w][0][2 * lane] = x0[1];
            xch[par][w][0][2 * lane + 1] = x1[1];
            xch[par][w][1][2 * lane] = x0[R1];
            xch[par][w][1][2 * lane + 1] = x1[R1];
            __syncthreads();
            if (w > 0) {
                x0[0] = xch[par][w - 1][1][2 * lane];
                x1[0] = xch[par][w - 1][1][2 * lane + 1];
            }
            if (w < NW - 1) {
                x0[R1 + 1] = xch[par][w + 1][0][2 * lane];
                x1[R1 + 1] = xch[par][w + 1][0][2 * lane + 1];
            }
            par ^= 1;
        }
    }
}

// NW: waves in the workgroup (4 or 16; ny1 = NW R1 rows).  With 16 waves
// each SIMD runs four, which hides the latency of the row exchanges and of the
// products better than one.
template <int R1, int NW>
__global__ __launch_bounds__(64 * NW) void k_tail_dct_ext(const float *r1g, int p1, float *e1g, int nx1, float ax1,
                                                          float ay1, int nx, int ny, const float *tab,
                                                          const Ctl *ctl)
{
    __shared__ float Cx[kDct * kDctLd], Cy[kDct * kDctLd], B0[kDct * kDctLd], B1[kDct * kDctLd];
    __shared__ float xch[2][NW][2][128];  // [buffer][wave][first, last own row][column]
    constexpr int ny1 = NW * R1, NT = 64 * NW, NQ = kDct * kDct / NT;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int j0 = w * R1, ca = 2 * lane;
    const bool in_i = ca < nx1;
    // loads first: the tables and the own rows of r; their latency overlaps the flag's
    float vx[NQ], vy[NQ], r0[R1], r1[R1];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        vx[q] = tab[t + NT * q];
        vy[q] = tab[kDct * kDct + t + NT * q];
    }
#pragma unroll
    for (int q = 0; q < R1; ++q) {
        float2 v = {0.f, 0.f};
        if (in_i) v = *reinterpret_cast<const float2 *>(r1g + ca + (size_t)p1 * (j0 + q));
        r0[q] = v.x;
        r1[q] = v.y;
    }
    if (ctl->mg_done | ctl->gm_stop) return;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = t + NT * q, row = e >> 6, col = e & 63;
        Cx[row * kDctLd + col] = vx[q];
        Cy[row * kDctLd + col] = vy[q];
        B0[row * kDctLd + col] = 0.f;  // the DCT input, zero outside ny x nx
    }
    const float cw0 = ca > 0 ? ax1 : 0.f, ce1 = ca + 1 < nx1 - 1 ? ax1 : 0.f;
    const float d0m = cw0 + ax1 + 2.f * ay1, d0e = cw0 + ax1 + ay1, d1m = ax1 + ce1 + 2.f * ay1,
                d1e = ax1 + ce1 + ay1;
    const float i0m = 1.f / d0m, i0e = 1.f / d0e, i1m = 1.f / d1m, i1e = 1.f / d1e;
    float A[2][2], B[2][2];
    A[0][0] = in_i ? ax1 * i0m : 0.f;
    A[0][1] = in_i ? ax1 * i0e : 0.f;
    A[1][0] = in_i ? ax1 * i1m : 0.f;
    A[1][1] = in_i ? ax1 * i1e : 0.f;
    B[0][0] = in_i ? ay1 * i0m : 0.f;
    B[0][1] = in_i ? ay1 * i0e : 0.f;
    B[1][0] = in_i ? ay1 * i1m : 0.f;
    B[1][1] = in_i ? ay1 * i1e : 0.f;
#pragma unroll
    for (int q = 0; q < R1; ++q) {  // r -> r / d
        const int j = j0 + q;
        const bool mid = j > 0 && j < ny1 - 1;
        r0[q] *= mid ? i0m : i0e;
        r1[q] *= mid ? i1m : i1e;
    }
    float x0[R1 + 2], x1[R1 + 2];  // rows j0 - 1 .. j0 + R1 (the outer two from the neighbours)
#pragma unroll
    for (int q = 0; q < R1 + 2; ++q) x0[q] = x1[q] = 0.f;
    int par = 0;
    __syncthreads();  // B0 cleared before the restriction writes it
    ext_sweeps<R1, NW>(x0, x1, r0, r1, A, B, j0, ny1, xch, par, w, lane, true);
    // residual of the own rows, 2x2 averages (lane-local) into the DCT input
#pragma unroll
    for (int m = 0; m < R1 / 2; ++m) {
        float t4[4];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int q = 1 + 2 * m + u, j = j0 + q - 1;
            const bool mid = j > 0 && j < ny1 - 1;
            const float P0 = x0[q], P1 = x1[q];
            const float Wl = lane_prev(P1), El = lane_next(P0);
            const float W0 = ca > 0 ? Wl : P0, E1 = ca + 1 < nx1 - 1 ? El : P1;
            const float S0 = j > 0 ? x0[q - 1] : P0, N0 = j < ny1 - 1 ? x0[q + 1] : P0;
            const float S1 = j > 0 ? x1[q - 1] : P1, N1 = j < ny1 - 1 ? x1[q + 1] : P1;
            const float rq0 = r0[q - 1] * (mid ? d0m : d0e), rq1 = r1[q - 1] * (mid ? d1m : d1e);
            t4[2 * u] = rq0 - ((P1 + W0 - 2.f * P0) * ax1 + (N0 + S0 - 2.f * P0) * ay1);
            t4[2 * u + 1] = rq1 - ((E1 + P0 - 2.f * P1) * ax1 + (N1 + S1 - 2.f * P1) * ay1);
        }
        const int I = lane, J = j0 / 2 + m;
        if (in_i && I < nx && J < ny) B0[J * kDctLd + I] = 0.25f * (((t4[0] + t4[1]) + t4[2]) + t4[3]);
    }
    __syncthreads();
    dct_solve<float, true, NW>(Cx, Cy, B0, B1, tab + 2 * kDct * kDct);  // the coarse correction in B0
    // up leg: add the coarse correction (the halo rows too: they hold the
    // neighbours' down-leg values), two sweeps, store the own rows
#pragma unroll
    for (int q = 0; q < R1 + 2; ++q) {
        const int j = j0 + q - 1;
        if (j >= 0 && j < ny1 && in_i) {
            x0[q] += prolong_at<float>(B0, kDctLd, nx, ny, ca, j);
            x1[q] += prolong_at<float>(B0, kDctLd, nx, ny, ca + 1, j);
        }
    }
    ext_sweeps<R1, NW>(x0, x1, r0, r1, A, B, j0, ny1, xch, par, w, lane, false);
#pragma unroll
    for (int q = 1; q <= R1; ++q)
        if (in_i) *reinterpret_cast<float2 *>(e1g + ca + (size_t)p1 * (j0 + q - 1)) = float2{x0[q], x1[q]};
}

// ------------------------------------------------ FP64 outer iteration
// phi_out = phi_in + e (FP64; e null: no correction and no write),
// r = b - L phi_out in FP64, stored in T for the next V-cycle; block
// partials of ||r||inf and sum phi_out.
template <class T>
__global__ __launch_bounds__(256) void k_correct_residual(const double *phi_in, double *phi_out, const T *e,
                                                          const double *b, T *r, int n, int P, double ax, double ay,
                                                          const Ctl *ctl, double *partials)
{
    __shared__ double sh[16];
    if (ctl->mg_done) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double m = 0.0, s = 0.0;
    if (i < n && j < n) {
        const size_t k = i + (size_t)P * j;
        auto x = [&](size_t q) -> double { return e ? phi_in[q] + (double)e[q] : phi_in[q]; };
        const double Pv = x(k);
        const double Wv = i > 0 ? x(k - 1) : Pv;
        const double Ev = i < n - 1 ? x(k + 1) : Pv;
        const double Sv = j > 0 ? x(k - P) : Pv;
        const double Nv = j < n - 1 ? x(k + P) : Pv;
        const double rr = b[k] - ((Ev + Wv - 2.0 * Pv) * ax + (Nv + Sv - 2.0 * Pv) * ay);
        if (e) phi_out[k] = Pv;
        r[k] = (T)rr;
        m = fabs(rr);
        s = Pv;
    }
    m = block_reduce<true>(m, sh);
    s = block_reduce<false>(s, sh + 8);
    if (threadIdx.x == 0) {
        const size_t bid = blockIdx.x + (size_t)gridDim.x * blockIdx.y;
        partials[2 * bid] = m;
        partials[2 * bid + 1] = s;
    }
}

// require_mom: the step's Poisson phase was enqueued before the host saw the
// pseudo-iteration loop end; if it has not ended, the whole phase is skipped
// (mg_done and mg_skip set) and the host runs it again later.  With the
// predictor faces u, v it also forms the mean of b for k_poisson_start(_rows).
__global__ __launch_bounds__(1024) void k_mg_start(Ctl *ctl, int require_mom, const double *u, const double *v,
                                                   int n, int P, double dx, double dy, double dt)
{
    __shared__ double sh[16];
    __shared__ int skip_s;
    // the flag's load and the boundary faces' loads go out together (the
    // faces are read even when the phase is skipped, and then unused)
    const int mom_done = __hip_atomic_load(&ctl->mom_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // with the predictor faces: the mean of b = (1.5/dt) D(u*) from the
    // boundary fluxes (sum of D over the cells = net outflow), for the start
    double sum = 0.0;
    const int T = blockDim.x, t = threadIdx.x;
    if (u && n <= 2 * T) {
        // two rows of the x-walls' faces and two columns of the y-walls' faces per
        // thread: all eight loads first, then the terms in the loops' order below
        double ue[2], uw[2], vn[2], vs[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int q = t + k * T;
            const bool in = q < n;
            ue[k] = in ? u[(size_t)P * q + n] : 0.0;
            uw[k] = in ? u[(size_t)P * q] : 0.0;
            vn[k] = in ? v[(size_t)P * n + q] : 0.0;
            vs[k] = in ? v[q] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (t + k * T < n) sum += (ue[k] - uw[k]) / dx;
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (t + k * T < n) sum += (vn[k] - vs[k]) / dy;
    } else if (u) {
        for (int j = t; j < n; j += T) {
            const size_t row = (size_t)P * j;
            sum += (u[row + n] - u[row]) / dx;
        }
        for (int i = t; i < n; i += T) sum += (v[(size_t)P * n + i] - v[i]) / dy;
    }
    if (threadIdx.x == 0) {
        const int skip = require_mom && !mom_done;
        ctl->mg_done = skip;
        ctl->mg_skip = skip;
        ctl->gm_stop = 0;
        ctl->mg_iters = 0;
        ctl->mg_started = 0;
        ctl->b_norm_fixed = 0;
        ctl->res_norm = 0.0;
        ctl->b_norm = 0.0;
        ctl->phi_sum = 0.0;
        ctl->phi_mean = 0.0;
        skip_s = skip;
    }
    __syncthreads();
    if (!u || skip_s) return;
    sum = block_reduce<false>(sum, sh);
    if (threadIdx.x == 0) ctl->b_mean = (1.5 / dt) * sum / ((double)n * (double)n);
}

// Start of the Poisson iteration, one pass.  The right-hand side is either b
// (b <- b - mean(b) in place when center is set) or, with FROM_U, formed here
// from the predictor faces u*, v* as b = (1.5/dt) D(u*) - mean (the step's
// poisson_righthand_side_calc, fn_rhs.cpp:71-125, with the mean from
// k_mg_start) and stored.  Then the starting iterate phi0 = 0 (mode 0), phi1
// (mode 1) or the extrapolation 2 phi1 - phi2 (mode 2) of the two previous
// solutions (newest first), r0 = b - L phi0 in FP64 (stored T), and block
// partials of ||r0||inf, sum phi0, ||b||inf and, with FROM_U, ||D(u*)||inf and
// sum (1.5/dt) D(u*) (the stopping test stays relative to ||b||inf whatever
// the start).
constexpr int kStartParts = 5;

template <class T, bool FROM_U>
__global__ __launch_bounds__(256) void k_poisson_start(double *b, const double *u, const double *v, double dx,
                                                       double dy, double dt, const double *phi1, const double *phi2,
                                                       double *phi0, T *r, int n, int P, double ax, double ay,
                                                       int mode, int center, const Ctl *ctl, double *partials,
                                                       int ntx, int nty)
{
    __shared__ double sh[40];
    if (ctl->mg_done) return;  // a skipped phase (k_mg_start)
    const CellTile t = cell_tile(ntx, nty);
    double mr = 0.0, sp = 0.0, mb = 0.0, md = 0.0, sb = 0.0;
    const double bmean = FROM_U || center ? ctl->b_mean : 0.0;
    auto g = [&](size_t q) -> double {
        return mode == 0 ? 0.0 : mode == 1 ? phi1[q] : 2.0 * phi1[q] - phi2[q];
    };
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
        const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + 4 * pass + (threadIdx.x >> 6);
        if (!t.live || i >= n || j >= n) continue;
        const size_t k = i + (size_t)P * j;
        double bc;
        if (FROM_U) {
            const double d = (u[k + 1] - u[k]) / dx + (v[k + P] - v[k]) / dy;
            const double braw = (1.5 / dt) * d;
            bc = braw - bmean;
            md = (fabs(d) > md || d != d) ? fabs(d) : md;
            sb += braw;
        } else {
            bc = b[k] - bmean;
        }
        const double Pv = g(k);
        const double Wv = i > 0 ? g(k - 1) : Pv;
        const double Ev = i < n - 1 ? g(k + 1) : Pv;
        const double Sv = j > 0 ? g(k - P) : Pv;
        const double Nv = j < n - 1 ? g(k + P) : Pv;
        const double rr = bc - ((Ev + Wv - 2.0 * Pv) * ax + (Nv + Sv - 2.0 * Pv) * ay);
        if (FROM_U || center) b[k] = bc;
        phi0[k] = Pv;
        r[k] = (T)rr;
        mr = (fabs(rr) > mr || rr != rr) ? fabs(rr) : mr;
        sp += Pv;
        mb = (fabs(bc) > mb || bc != bc) ? fabs(bc) : mb;
    }
    mr = block_reduce<true>(mr, sh);
    sp = block_reduce<false>(sp, sh + 8);
    mb = block_reduce<true>(mb, sh + 16);
    if (FROM_U) {
        md = block_reduce<true>(md, sh + 24);
        sb = block_reduce<false>(sb, sh + 32);
    }
    if (threadIdx.x == 0) {
        double *o = partials + (size_t)kStartParts * blockIdx.x;
        o[0] = mr;
        o[1] = sp;
        o[2] = mb;
        o[3] = md;
        o[4] = sb;
    }
}

// (1.5/dt) times the mean cell divergence of u, v from the boundary faces alone:
// the cell divergences telescope to the net flux through the walls
// The same start in the row-march layout of the stage kernel: a wave owns a
// band of 62 columns (64 lanes: the band and one halo lane on each side) and
// a chunk of rows, and marches up them.  The guess of the rows j - 1, j, j + 1
// and the south y-face stay in registers, x-neighbours come by DPP, so a cell
// costs one load of each input array (u, v, phi1, phi2 or b) and the three
// stores.  The next row's loads are issued one iteration ahead and a row's
// stores one iteration behind (one wait for memory per row).  The per-cell
// arithmetic is that of k_poisson_start; per-workgroup partials as there.
constexpr int kStartWB = 62;
template <class T, bool FROM_U>
__global__ __launch_bounds__(256) void k_poisson_start_rows(double *b, const double *u, const double *v, double dx,
                                                            double dy, double dt, const double *phi1,
                                                            const double *phi2, double *phi0, T *r, int n, int P,
                                                            double ax, double ay, int mode, int center,
                                                            const Ctl *ctl, double *partials, int nbands, int nwaves,
                                                            int rows)
{
    __shared__ double sh[4 * kStartParts];
    if (ctl->mg_done) return;  // a skipped phase (k_mg_start)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int w = blk * 4 + wv;
    const bool live = w < nwaves;  // a wave past the last one only joins the reduction
    const int band = live ? w % nbands : 0, chunk = live ? w / nbands : 0;
    const int i = band * kStartWB + lane - 1;
    const int j0 = chunk * rows, jend = live ? min(j0 + rows, n) : j0;
    const bool xin = i >= 0 && i < n, xface = i >= 0 && i <= n;
    const bool out = xin && lane >= 1 && lane <= kStartWB;
    const bool keep_b = FROM_U || center;
    const double bmean = keep_b ? ctl->b_mean : 0.0;
    struct Raw {
        double p1, p2;  // guess iterates at (i, j + 1)
        double uu, vn;  // x-face (i, j), y-face (i, j + 1)
        double bb;      // b at (i, j) (b given)
    };
    auto load = [&](int j, Raw &q) {  // what row j consumes
        q.p1 = q.p2 = q.uu = q.vn = q.bb = 0.0;
        const size_t k = (size_t)(i < 0 ? 0 : i) + (size_t)P * j;
        if (xin && j + 1 < n && mode != 0) {
            q.p1 = phi1[k + P];
            if (mode == 2) q.p2 = phi2[k + P];
        }
        if (FROM_U) {
            if (xface) q.uu = u[k];  // also the halo lane: u(i + 1) of the band's last column
            if (out) q.vn = v[k + P];
        } else if (out) {
            q.bb = b[k];
        }
    };
    auto guess = [&](double p1, double p2) -> double { return mode == 0 ? 0.0 : mode == 1 ? p1 : 2.0 * p1 - p2; };
    double gS = 0.0, gC = 0.0, vS = 0.0;
    if (jend > j0 && xin) {
        const size_t k = i + (size_t)P * j0;
        if (mode != 0) {
            gC = guess(phi1[k], mode == 2 ? phi2[k] : 0.0);
            if (j0 > 0) gS = guess(phi1[k - P], mode == 2 ? phi2[k - P] : 0.0);
        }
        if (FROM_U && out) vS = v[k];
    }
    double mr = 0.0, sp = 0.0, mb = 0.0, md = 0.0, sb = 0.0;
    size_t sk = 0;
    bool pend = false;
    double s_b = 0.0, s_phi = 0.0, s_r = 0.0;
    auto flush = [&]() {
        if (pend) {
            if (keep_b) b[sk] = s_b;
            phi0[sk] = s_phi;
            r[sk] = (T)s_r;
        }
    };
    Raw cur, nxt;
    if (jend > j0) load(j0, cur);
    for (int j = j0; j < jend; ++j) {
        if (j + 1 < jend) load(j + 1, nxt);  // prefetch
        flush();
        pend = false;
        const double Pv = gC;
        const double gN = guess(cur.p1, cur.p2);
        const double Wl = lane_prev(Pv), El = lane_next(Pv);  // (every lane shifts: no divergence around DPP)
        const double ue = lane_next(cur.uu);
        const double Wv = i > 0 ? Wl : Pv;
        const double Ev = i < n - 1 ? El : Pv;
        const double Sv = j > 0 ? gS : Pv;
        const double Nv = j < n - 1 ? gN : Pv;
        double bc, d = 0.0, braw = 0.0;
        if (FROM_U) {
            d = (ue - cur.uu) / dx + (cur.vn - vS) / dy;
            braw = (1.5 / dt) * d;
            bc = braw - bmean;
        } else {
            bc = cur.bb - bmean;
        }
        const double rr = bc - ((Ev + Wv - 2.0 * Pv) * ax + (Nv + Sv - 2.0 * Pv) * ay);
        if (out) {
            sk = i + (size_t)P * j;
            pend = true;
            s_b = bc;
            s_phi = Pv;
            s_r = rr;
            mr = (fabs(rr) > mr || rr != rr) ? fabs(rr) : mr;
            sp += Pv;
            mb = (fabs(bc) > mb || bc != bc) ? fabs(bc) : mb;
            if (FROM_U) {
                md = (fabs(d) > md || d != d) ? fabs(d) : md;
                sb += braw;
            }
        }
        gS = gC;
        gC = gN;
        vS = cur.vn;
        cur = nxt;
    }
    flush();
    // wave trees, then the four waves in order
    for (int o = 32; o > 0; o >>= 1) {
        const double a0 = __shfl_down(mr, o, 64), a2 = __shfl_down(mb, o, 64), a3 = __shfl_down(md, o, 64);
        mr = (a0 > mr || a0 != a0) ? a0 : mr;
        mb = (a2 > mb || a2 != a2) ? a2 : mb;
        md = (a3 > md || a3 != a3) ? a3 : md;
        sp += __shfl_down(sp, o, 64);
        sb += __shfl_down(sb, o, 64);
    }
    if (lane == 0) {
        double *o = sh + kStartParts * wv;
        o[0] = mr; o[1] = sp; o[2] = mb; o[3] = md; o[4] = sb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double q[kStartParts];
        for (int e = 0; e < kStartParts; ++e) q[e] = sh[e];
        for (int x = 1; x < 4; ++x) {
            const double *o = sh + kStartParts * x;
            q[0] = (o[0] > q[0] || o[0] != o[0]) ? o[0] : q[0];
            q[1] += o[1];
            q[2] = (o[2] > q[2] || o[2] != o[2]) ? o[2] : q[2];
            q[3] = (o[3] > q[3] || o[3] != o[3]) ? o[3] : q[3];
            q[4] += o[4];
        }
        double *o = partials + (size_t)kStartParts * blockIdx.x;
        for (int e = 0; e < kStartParts; ++e) o[e] = q[e];
    }
}

// the first residual's norm and ||b||inf (and with from_u the divergence
// statistics of u*), and the stopping test on it
__global__ __launch_bounds__(1024) void k_fin_start(const double *partials, int nb, Ctl *ctl, int n, int from_u)
{
    __shared__ double sh[80];
    const int done = __hip_atomic_load(&ctl->mg_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // a skipped phase (k_mg_start)
    double mr = 0.0, sp = 0.0, mb = 0.0, md = 0.0, sb = 0.0;
    for_records<5, kStartParts>(partials, nb, [&](const double *o) {  // (loads issued before the flag is waited for)
        mr = (o[0] > mr || o[0] != o[0]) ? o[0] : mr;
        sp += o[1];
        mb = (o[2] > mb || o[2] != o[2]) ? o[2] : mb;
        md = (o[3] > md || o[3] != o[3]) ? o[3] : md;
        sb += o[4];
    });
    if (done) return;
    double red[5] = {mr, sp, mb, md, sb};
    block_reduce_n<0b01101u>(red, sh);  // max, sum, max, max, sum
    mr = red[0];
    sp = red[1];
    mb = red[2];
    md = red[3];
    sb = red[4];
    if (threadIdx.x == 0) {
        if (from_u) {
            ctl->div_star = md;
            ctl->rhs_sum = sb;
        }
        ctl->b_norm = mb;
        ctl->b_norm_fixed = 1;
        ctl->res_norm = mr;
        ctl->phi_sum = sp;
        ctl->phi_mean = sp / ((double)n * (double)n);
        ctl->mg_started = 1;
        const double thr = fmax(ctl->tol_abs, ctl->tol_rel * mb);
        if (mr <= thr || !(mr == mr)) ctl->mg_done = 1;
    }
}

hipError_t launch_poisson_start(const Ctx &c, int mode, int center, const double *phi1, const double *phi2,
                                double *phi0, int mixed, const double *us, const double *vs)
{
    const int n = c.g.n;
    const Level &L = c.lev[0];
    const double dx = c.g.dx, dy = c.g.dy, dt = c.g.dt;
    int nb = 0;  // (with us, launch_mg_start has formed the mean of b)
    if (c.start_rows > 0) {  // row-march kernel
        const int rows = c.start_rows;
        const int nbands = (n + kStartWB - 1) / kStartWB, nwaves = nbands * ((n + rows - 1) / rows);
        nb = ((nwaves + 3) / 4 + 7) / 8 * 8;  // four waves per workgroup, a multiple of 8 workgroups
        if ((int64_t)nb * kStartParts > (int64_t)kMaxPartials * 4) return hipErrorInvalidValue;
#define AMR_START(T, FU, RR)                                                                                         \
    hipLaunchKernelGGL((k_poisson_start_rows<T, FU>), dim3(nb), dim3(256), 0, c.stream, c.b, us, vs, dx, dy, dt,      \
                       phi1, phi2, phi0, RR, n, c.g.P, L.ax, L.ay, mode, center, (const Ctl *)c.ctl, c.partials,     \
                       nbands, nwaves, rows)
        if (mixed) {
            if (us) AMR_START(float, true, L.r32);
            else AMR_START(float, false, L.r32);
        } else {
            if (us) AMR_START(double, true, L.r);
            else AMR_START(double, false, L.r);
        }
#undef AMR_START
    } else {  // 64 x 16 cell tiles
        const int ntx = (n + 63) / 64, nty = (n + 15) / 16;
        nb = (ntx * nty + 7) / 8 * 8;  // a multiple of 8 workgroups
        if ((int64_t)nb * kStartParts > (int64_t)kMaxPartials * 4) return hipErrorInvalidValue;
#define AMR_START(T, FU, RR)                                                                                         \
    hipLaunchKernelGGL((k_poisson_start<T, FU>), dim3(nb), dim3(256), 0, c.stream, c.b, us, vs, dx, dy, dt, phi1,    \
                       phi2, phi0, RR, n, c.g.P, L.ax, L.ay, mode, center, (const Ctl *)c.ctl, c.partials, ntx, nty)
        if (mixed) {
            if (us) AMR_START(float, true, L.r32);
            else AMR_START(float, false, L.r32);
        } else {
            if (us) AMR_START(double, true, L.r);
            else AMR_START(double, false, L.r);
        }
#undef AMR_START
    }
    hipLaunchKernelGGL(k_fin_start, dim3(1), dim3(1024), 0, c.stream, (const double *)c.partials, nb, c.ctl, n,
                       us ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_mg_start(const Ctx &c, int require_mom, const double *us, const double *vs)
{
    hipLaunchKernelGGL(k_mg_start, dim3(1), dim3(1024), 0, c.stream, c.ctl, require_mom, us, vs, c.g.n, c.g.P, c.g.dx,
                       c.g.dy, c.g.dt);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_fin_resid(const double *partials, int nb, Ctl *ctl, int n)
{
    __shared__ double sh[32];
    const int done = __hip_atomic_load(&ctl->mg_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double m = 0.0, s = 0.0;
    for_records<2, 2>(partials, nb, [&](const double *o) {  // (loads issued before the flag is waited for)
        m = (o[0] > m || o[0] != o[0]) ? o[0] : m;
        s += o[1];
    });
    if (done) return;
    double red[2] = {m, s};
    block_reduce_n<1u>(red, sh);  // max, sum
    m = red[0];
    s = red[1];
    if (threadIdx.x == 0) {
        ctl->res_norm = m;
        ctl->phi_sum = s;
        ctl->phi_mean = s / ((double)n * (double)n);
        if (!ctl->mg_started) {
            if (!ctl->b_norm_fixed) ctl->b_norm = m;  // phi = 0 on the first evaluation: r = b
            ctl->mg_started = 1;
        } else {
            ctl->mg_iters += 1;
        }
        const double thr = fmax(ctl->tol_abs, ctl->tol_rel * ctl->b_norm);
        if (m <= thr || !(m == m)) ctl->mg_done = 1;
        else if (ctl->mg_iters >= ctl->mg_max_iter) { ctl->mg_done = 1; ctl->flags |= AMRESSIF_FLAG_MG_CAP; }
    }
}

hipError_t launch_fin_resid(const Ctx &c, int nblocks)
{
    hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(1024), 0, c.stream, c.partials, nblocks, c.ctl, c.g.n);
    return hipGetLastError();
}

// ------------------------------------------------------------ launchers
template <class T>
static T *lev_e(const Level &L, int which);
template <>
double *lev_e<double>(const Level &L, int which) { return which ? L.e2 : L.e; }
template <>
float *lev_e<float>(const Level &L, int which) { return which ? L.e32b : L.e32; }
template <class T>
static T *lev_r(const Level &L);
template <>
double *lev_r<double>(const Level &L) { return L.r; }
template <>
float *lev_r<float>(const Level &L) { return L.r32; }

static TailLevels tail_levels(const Ctx &c, int lt, size_t *lds_bytes, size_t elem)
{
    TailLevels tl{};
    tl.nlev = c.nlev - lt;
    int off = 0;
    for (int l = 0; l < tl.nlev; ++l) {
        const Level &L = c.lev[lt + l];
        tl.nx[l] = L.nx;
        tl.ny[l] = L.ny;
        tl.ax[l] = L.ax;
        tl.ay[l] = L.ay;
        tl.off[l] = off;
        off += L.nx * L.ny;
    }
    tl.total = off;
    tl.mb = c.mb;
    *lds_bytes = (2 * (size_t)off + 4 * (size_t)c.mb) * elem;  // e, r of every level + bottom partial sums
    return tl;
}

static dim3 tile_grid(const Level &L) { return dim3((L.nx + SX - 1) / SX, (L.ny + SY - 1) / SY); }

// launch shape of one smoother on level l: the LDS tile kernel (one partial per
// workgroup in MODE 2) or the wave-column kernel (one per wave)
struct SmoothShapeRt {
    dim3 grid;
    int nbands = 0, nwaves = 0, rows = 0, wave = 0;
    int parts = 0;  // MODE 2 norm partials: one per workgroup
};

// the finest legs of the FP32 V-cycle with three smoothing sweeps: the up
// leg with mg_post3, the down leg with mg_pre3
static bool post3(const Ctx &c, int l, int mode, size_t tsize)
{
    if (l != 0 || tsize != 4) return false;
    return mode == 2 ? c.mg_post3 && c.mg_late : mode == 1 ? c.mg_pre3 != 0 : false;
}

template <class T>
static SmoothShapeRt smooth_shape(const Ctx &c, int l, int mode)
{
    SmoothShapeRt s;
    const Level &L = c.lev[l];
    s.wave = c.mg_smoother;
    if (!s.wave) {
        s.grid = tile_grid(L);
        s.nwaves = s.parts = (int)(s.grid.x * s.grid.y);
        return s;
    }
    // FP64 finest up leg: the one-column kernel (106 VGPRs, 4 waves per SIMD)
    // beats the paired one (2 waves) there
    if (s.wave == 2 && sizeof(T) == 8 && mode == 2 && l == 0 && c.mg_up64_single) s.wave = 1;
    int OC = s.wave == 2 ? (mode == 0 ? 120 : 116) : (mode == 0 ? 56 : 54);
    if (s.wave == 2 && post3(c, l, mode, sizeof(T))) OC = 128 - 2 * (6 + 2);  // three smoothing sweeps
    s.rows = l == 0 ? c.mg_rows[sizeof(T) == 4][mode] : c.mg_rows_coarse;
    if (s.wave == 2) {  // the paired-column kernel's register limit (pair_rows_max)
        const int cap = sizeof(T) == 4 ? (mode == 2 ? 8 : 16) : (mode == 2 ? 4 : 8);
        s.rows = std::min(s.rows, cap);
    }
    s.nbands = (L.nx + OC - 1) / OC;
    s.nwaves = s.nbands * ((L.ny + s.rows - 1) / s.rows);
    s.grid = dim3((s.nwaves + 3) / 4);
    s.parts = (int)s.grid.x;
    return s;
}

// largest window of the paired-column kernel that fits the registers without
// spilling: the shape falls back to it
template <class T, int MODE>
constexpr int pair_rows_max()
{
    return sizeof(T) == 4 ? (MODE == 2 ? 8 : 16) : (MODE == 2 ? 4 : 8);
}

template <class T, int MODE, int R>
static void launch_gsrb(const Ctx &c, const SmoothShapeRt &s, const T *xin, const T *ec, int pc, T *xout,
                        const T *rhs, const Level &L, T *rc, const Level *Cl, const OuterArgs &o, bool p3)
{
    constexpr int R2 = R < pair_rows_max<T, MODE>() ? R : pair_rows_max<T, MODE>();
    if constexpr (sizeof(T) == 4 && MODE != 0) {
        if (s.wave == 2 && p3) {
            hipLaunchKernelGGL((k_gsrb2<T, MODE, R2, MODE == 2, 6>), s.grid, dim3(256), 0, c.stream, xin, ec, pc, xout,
                               rhs, L.nx, L.ny, L.pitch, (T)L.ax, (T)L.ay, rc, Cl ? Cl->nx : 0, Cl ? Cl->ny : 0,
                               Cl ? Cl->pitch : 0, o, (const Ctl *)c.ctl, s.nbands, s.nwaves);
            return;
        }
    }
    if (s.wave == 2 && MODE == 2 && c.mg_late)
        hipLaunchKernelGGL((k_gsrb2<T, MODE, R2, true>), s.grid, dim3(256), 0, c.stream, xin, ec, pc, xout, rhs, L.nx,
                           L.ny, L.pitch, (T)L.ax, (T)L.ay, rc, Cl ? Cl->nx : 0, Cl ? Cl->ny : 0,
                           Cl ? Cl->pitch : 0, o, (const Ctl *)c.ctl, s.nbands, s.nwaves);
    else if (s.wave == 2)
        hipLaunchKernelGGL((k_gsrb2<T, MODE, R2>), s.grid, dim3(256), 0, c.stream, xin, ec, pc, xout, rhs, L.nx, L.ny,
                           L.pitch, (T)L.ax, (T)L.ay, rc, Cl ? Cl->nx : 0, Cl ? Cl->ny : 0, Cl ? Cl->pitch : 0, o,
                           (const Ctl *)c.ctl, s.nbands, s.nwaves);
    else
        hipLaunchKernelGGL((k_gsrb<T, MODE, R>), s.grid, dim3(256), 0, c.stream, xin, ec, pc, xout, rhs, L.nx, L.ny,
                           L.pitch, (T)L.ax, (T)L.ay, rc, Cl ? Cl->nx : 0, Cl ? Cl->ny : 0, Cl ? Cl->pitch : 0, o,
                           (const Ctl *)c.ctl, s.nbands, s.nwaves);
}

// one smoother on level l (MODE 1: from zero with the fused restriction into
// level l + 1; MODE 0: from xin + P(ec), ec on level l + 1; MODE 2: MODE 0
// plus the FP64 outer step)
template <class T, int MODE>
static void launch_smoother(const Ctx &c, int l, const T *xin, const T *ec, int pc, T *xout, const T *rhs,
                            const OuterArgs &o)
{
    const Level &L = c.lev[l];
    const Level *Cl = &c.lev[l + 1];
    T *rc = MODE == 1 ? lev_r<T>(*Cl) : (T *)nullptr;
    const SmoothShapeRt s = smooth_shape<T>(c, l, MODE);
    const bool p3 = post3(c, l, MODE, sizeof(T));
    if (!s.wave) {
        hipLaunchKernelGGL((k_smooth<T, MODE>), s.grid, dim3(256), 0, c.stream, xin, ec, pc, xout, rhs, L.nx, L.ny,
                           L.pitch, (T)L.ax, (T)L.ay, rc, Cl ? Cl->nx : 0, Cl ? Cl->ny : 0, Cl ? Cl->pitch : 0, o,
                           (const Ctl *)c.ctl);
        return;
    }
    switch (s.rows) {
    case 4: launch_gsrb<T, MODE, 4>(c, s, xin, ec, pc, xout, rhs, L, rc, Cl, o, p3); break;
    case 8: launch_gsrb<T, MODE, 8>(c, s, xin, ec, pc, xout, rhs, L, rc, Cl, o, p3); break;
    case 16: launch_gsrb<T, MODE, 16>(c, s, xin, ec, pc, xout, rhs, L, rc, Cl, o, p3); break;
    default: launch_gsrb<T, MODE, 32>(c, s, xin, ec, pc, xout, rhs, L, rc, Cl, o, p3); break;
    }
}

// ------------------------------------- exact DCT solve of a large bottom
// A coarsest level too large for the dense pseudo-inverse (an odd side above
// 32 cells, e.g. 65 x 65 when n = 130) is solved exactly by the DCT-II, as the
// small levels are: e = C_y (D+ .* (C_y^T r C_x)) C_x^T, here as four dense
// products over the whole grid.  C = op(A) op(B) (.* S): one 16 x 16 output
// tile per workgroup, K in steps of 16 through LDS.
template <class T>
__global__ __launch_bounds__(256) void k_dense_mm(const T *A, int lda, int ta, const T *B, int ldb, int tb, T *C,
                                                  int ldc, int M, int N, int K, const T *S, int lds, const Ctl *ctl)
{
    if (ctl->mg_done | ctl->gm_stop) return;
    __shared__ T sa[16][17], sb[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    T acc = 0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int ak = k0 + tx, bk = k0 + ty;
        sa[ty][tx] = (row < M && ak < K) ? (ta ? A[(size_t)ak * lda + row] : A[(size_t)row * lda + ak]) : T(0);
        sb[ty][tx] = (bk < K && col < N) ? (tb ? B[(size_t)col * ldb + bk] : B[(size_t)bk * ldb + col]) : T(0);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) acc += sa[ty][kk] * sb[kk][tx];
        __syncthreads();
    }
    if (row < M && col < N) C[(size_t)row * ldc + col] = S ? acc * S[(size_t)row * lds + col] : acc;
}

template <class T>
static hipError_t dense_bottom(const Ctx &c, const T *r, int pr, T *e, int pe, const Level &L)
{
    const int nx = L.nx, ny = L.ny;
    const T *Cx = sizeof(T) == 4 ? (const T *)(const void *)c.big_tab32 : (const T *)(const void *)c.big_tab;
    const T *Cy = Cx + (size_t)nx * nx, *Dinv = Cy + (size_t)ny * ny;
    T *W1 = sizeof(T) == 4 ? (T *)(void *)c.big_work32 : (T *)(void *)c.big_work;
    T *W2 = W1 + (size_t)nx * ny;
    const dim3 grid((nx + 15) / 16, (ny + 15) / 16);
    const Ctl *ctl = c.ctl;
    hipLaunchKernelGGL(k_dense_mm<T>, grid, dim3(256), 0, c.stream, r, pr, 0, Cx, nx, 0, W1, nx, ny, nx, nx,
                       (const T *)nullptr, 0, ctl);                                 // r C_x
    hipLaunchKernelGGL(k_dense_mm<T>, grid, dim3(256), 0, c.stream, Cy, ny, 1, (const T *)W1, nx, 0, W2, nx, ny, nx,
                       ny, Dinv, nx, ctl);                                          // D+ .* (C_y^T .)
    hipLaunchKernelGGL(k_dense_mm<T>, grid, dim3(256), 0, c.stream, (const T *)W2, nx, 0, Cx, nx, 1, W1, nx, ny, nx,
                       nx, (const T *)nullptr, 0, ctl);                             // . C_x^T
    hipLaunchKernelGGL(k_dense_mm<T>, grid, dim3(256), 0, c.stream, Cy, ny, 0, (const T *)W1, nx, 0, e, pe, ny, nx, ny,
                       (const T *)nullptr, 0, ctl);                                 // C_y .
    return hipGetLastError();
}

// One V-cycle on the correction from zero.  r0 is the level-0 right-hand side;
// with fuse set (and level 0 smoothed by its own kernels) the last up-leg
// smoother also applies the correction to the FP64 iterate and writes the next
// level-0 right-hand side to r0_next with its norm partials.
template <class T>
static hipError_t vcycle_impl(const Ctx &c, const T *pinv, const T *r0, const OuterArgs *fuse, T *r0_next)
{
    const Ctl *ctl = c.ctl;
    const int ext = sizeof(T) == 4 && c.tail_ext && c.dct_tab32;  // level tail_level - 1 joins the tail kernel
    const int lt = c.tail_level - ext;
    const OuterArgs none{};
    // (the finest-level smoothers of an outer-step cycle are timed when the
    // context's in-step kernel timing is on)
    const int vi = c.kp ? c.kp->mg_idx : 0;
    // algorithmic bytes per finest cell (t = sizeof(T)): down-leg r in, e out, coarse r
    // out (2.25 t); up-leg e, r, coarse e, phi, b in, phi and the next r out (3.25 t + 24)
    const double cells = (double)c.lev[0].nx * c.lev[0].ny, t = (double)sizeof(T);
    for (int l = 0; l < lt; ++l) {  // down: smooth from zero, fused residual + restriction
        if (l == 0 && fuse)
            if (hipError_t e = kprof_begin(c, KP_MG_DOWN, KL_MG, vi, 2.25 * t * cells)) return e;
        launch_smoother<T, 1>(c, l, (const T *)nullptr, (const T *)nullptr, 0, lev_e<T>(c.lev[l], 0),
                              l == 0 ? r0 : (const T *)lev_r<T>(c.lev[l]), none);
        if (l == 0 && fuse)
            if (hipError_t e = kprof_end(c)) return e;
    }
    if (c.big_tab) {
        const Level &Lt = c.lev[lt];
        const T *rt = lt == 0 ? r0 : (const T *)lev_r<T>(Lt);
        if (hipError_t e = dense_bottom<T>(c, rt, Lt.pitch, lev_e<T>(Lt, 1), Lt.pitch, Lt)) return e;
    } else if (ext) {
        const Level &L1 = c.lev[lt], &Lt = c.lev[lt + 1];
        const float *r1 = (const float *)(const void *)lev_r<T>(L1);
        float *e1 = (float *)(void *)lev_e<T>(L1, 1);
#define AMR_EXT(R1, NW)                                                                                              \
    hipLaunchKernelGGL((k_tail_dct_ext<R1, NW>), dim3(1), dim3(64 * NW), 0, c.stream, r1, L1.pitch, e1, L1.nx,        \
                       (float)L1.ax, (float)L1.ay, Lt.nx, Lt.ny, (const float *)c.dct_tab32, ctl)
        if (c.tail_waves == 16) {
            if (L1.ny == 32) AMR_EXT(2, 16);
            else if (L1.ny == 64) AMR_EXT(4, 16);
            else AMR_EXT(8, 16);
        } else {
            if (L1.ny == 32) AMR_EXT(8, 4);
            else if (L1.ny == 64) AMR_EXT(16, 4);
            else AMR_EXT(32, 4);
        }
#undef AMR_EXT
    } else {
        const Level &Lt = c.lev[lt];
        const T *rt = lt == 0 ? r0 : (const T *)lev_r<T>(Lt);
        const T *tab = sizeof(T) == 4 ? (const T *)(const void *)c.dct_tab32 : (const T *)(const void *)c.dct_tab;
        if (tab) {  // the exact solve of the top tail level
            const bool mf = sizeof(T) == 4 || c.dct_mfma64, w16 = mf && c.tail_waves == 16;
            void (*kern)(const T *, int, T *, int, int, const T *, const Ctl *) =
                w16 ? k_tail_dct<T, true, 16> : mf ? k_tail_dct<T, true, 4> : k_tail_dct<T, false, 4>;
            hipLaunchKernelGGL(kern, dim3(1), dim3(w16 ? 1024 : 256), 0, c.stream, rt, Lt.pitch, lev_e<T>(Lt, 1),
                               Lt.nx, Lt.ny, tab, ctl);
        } else {
            size_t lds = 0;
            const TailLevels tl = tail_levels(c, lt, &lds, sizeof(T));
            hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(1024), lds, c.stream, rt, Lt.pitch, lev_e<T>(Lt, 1), tl, pinv,
                               ctl);
        }
    }
    for (int l = lt - 1; l >= 0; --l) {  // up: add the coarse correction, smooth
        const Level &L = c.lev[l], &Cl = c.lev[l + 1];
        const T *rl = l == 0 ? r0 : (const T *)lev_r<T>(L);
        if (l == 0 && fuse) {
            if (hipError_t e = kprof_begin(c, KP_MG_UP, KL_MG, vi, (3.25 * t + 24.0) * cells)) return e;
            launch_smoother<T, 2>(c, l, (const T *)lev_e<T>(L, 0), (const T *)lev_e<T>(Cl, 1), Cl.pitch, r0_next, rl,
                                  *fuse);
            if (hipError_t e = kprof_end(c)) return e;
        } else
            launch_smoother<T, 0>(c, l, (const T *)lev_e<T>(L, 0), (const T *)lev_e<T>(Cl, 1), Cl.pitch,
                                  lev_e<T>(L, 1), rl, none);
    }
    return hipGetLastError();
}

hipError_t launch_vcycle(const Ctx &c, int mixed)
{
    const Level &L0 = c.lev[0];
    return mixed ? vcycle_impl<float>(c, c.pinv32, L0.r32, nullptr, nullptr)
                 : vcycle_impl<double>(c, c.pinv, L0.r, nullptr, nullptr);
}

// One outer step of the Poisson iteration: V-cycle on the level-0 residual
// buffer `par`, phi_out = phi_in + e, the next residual into buffer par ^ 1 and
// its norm partials (then launch_fin_resid).  Falls back to the separate
// correction kernel when level 0 is inside the single-workgroup tail.
hipError_t launch_vcycle_step(const Ctx &c, int mixed, int par, const double *phi_in, double *phi_out, int *nblocks)
{
    const Level &L0 = c.lev[0];
    if (c.tail_level == 0) {
        hipError_t e = launch_vcycle(c, mixed);
        if (e != hipSuccess) return e;
        return launch_correct_residual(c, phi_in, phi_out, 1, mixed, nblocks);
    }
    const int parts = mixed ? smooth_shape<float>(c, 0, 2).parts : smooth_shape<double>(c, 0, 2).parts;
    if (parts > kMaxPartials) return hipErrorInvalidValue;
    *nblocks = parts;
    OuterArgs o{phi_in, phi_out, c.b, c.partials, L0.ax, L0.ay};
    if (mixed) {
        float *rb[2] = {L0.r32, L0.r32b};
        return vcycle_impl<float>(c, c.pinv32, rb[par], &o, rb[par ^ 1]);
    }
    double *rb[2] = {L0.r, L0.rb};
    return vcycle_impl<double>(c, c.pinv, rb[par], &o, rb[par ^ 1]);
}

size_t tail_lds_bytes(const Ctx &c, int mixed)
{
    size_t lds = 0;
    tail_levels(c, c.tail_level, &lds, mixed ? sizeof(float) : sizeof(double));
    return lds;
}

// phi_out = phi_in + e_0 (e_0 = the V-cycle's correction; with e null the
// residual of phi_in alone), then r_0 = b - L phi_out and the norm partials
hipError_t launch_correct_residual(const Ctx &c, const double *phi_in, double *phi_out, int with_correction,
                                   int mixed, int *nblocks)
{
    const int n = c.g.n;
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    if ((int64_t)grid.x * grid.y > kMaxPartials) return hipErrorInvalidValue;
    *nblocks = grid.x * grid.y;
    const Level &L = c.lev[0];
    if (mixed)
        hipLaunchKernelGGL(k_correct_residual<float>, grid, dim3(256), 0, c.stream, phi_in, phi_out,
                           with_correction ? (const float *)L.e32b : (const float *)nullptr, (const double *)c.b,
                           L.r32, n, c.g.P, L.ax, L.ay, c.ctl, c.partials);
    else
        hipLaunchKernelGGL(k_correct_residual<double>, grid, dim3(256), 0, c.stream, phi_in, phi_out,
                           with_correction ? (const double *)L.e2 : (const double *)nullptr, (const double *)c.b,
                           L.r, n, c.g.P, L.ax, L.ay, c.ctl, c.partials);
    return hipGetLastError();
}

// Pick the first level handled by k_tail: the first level of at most 64^2
// cells whose tail fits the workgroup's LDS (raising the dynamic-LDS limit
// where the runtime asks for it); fall back to coarser starts.
hipError_t configure_tail(Ctx &c)
{
    const int mixed = c.in.precision == AMRESSIF_MIXED;
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev)) return e;
    int max_lds = 65536;
    if (hipError_t e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) return e;
    int lt = 0;
    while (lt < c.nlev - 1 && c.lev[lt].nx > 64) ++lt;
    for (; lt < c.nlev; ++lt) {
        c.tail_level = lt;
        const size_t bytes = tail_lds_bytes(c, mixed);
        if (bytes > (size_t)max_lds) continue;
        hipError_t e = mixed ? hipFuncSetAttribute((const void *)k_tail<float>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
                             : hipFuncSetAttribute((const void *)k_tail<double>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e == hipSuccess || bytes <= 65536) {
            (void)hipGetLastError();
            return hipSuccess;
        }
        (void)hipGetLastError();
    }
    return hipErrorInvalidValue;
}

// diagnostic: one fused smoother on the finest level, unguarded (FP64)
__global__ void k_ctl_clear_mg(Ctl *ctl) { ctl->mg_done = 0; }

// diagnostic: one finest-level smoother of the context's V-cycle precision,
// unguarded (the control flag is cleared first), on the buffers of the last
// solve: kind 1 the down-leg smoother with the fused restriction, kind 5 the
// up-leg smoother with the fused FP64 outer step (writing to scratch: phi_out
// into scratch_phi, the next residual into the second level-0 buffer)
template <class T>
static hipError_t smooth_probe(const Ctx &c, int kind, T *r0b, double *scratch_phi)
{
    const Level &L = c.lev[0], &Cl = c.lev[1];
    if (kind == 1)
        launch_smoother<T, 1>(c, 0, (const T *)nullptr, (const T *)nullptr, 0, lev_e<T>(L, 0),
                              (const T *)lev_r<T>(L), OuterArgs{});
    else {
        if (smooth_shape<T>(c, 0, 2).parts > kMaxPartials) return hipErrorInvalidValue;
        OuterArgs o{c.phi, scratch_phi, c.b, c.partials, L.ax, L.ay};
        launch_smoother<T, 2>(c, 0, (const T *)lev_e<T>(L, 0), (const T *)lev_e<T>(Cl, 1), Cl.pitch, r0b,
                              (const T *)lev_r<T>(L), o);
    }
    return hipGetLastError();
}

hipError_t launch_smooth_fine(const Ctx &c, int kind, double *scratch_phi)
{
    if (c.tail_level == 0) return hipErrorInvalidValue;  // level 0 has no smoother kernel
    hipLaunchKernelGGL(k_ctl_clear_mg, dim3(1), dim3(1), 0, c.stream, c.ctl);
    const Level &L = c.lev[0];
    return c.in.precision == AMRESSIF_MIXED ? smooth_probe<float>(c, kind, L.r32b, scratch_phi)
                                            : smooth_probe<double>(c, kind, L.rb, scratch_phi);
}

}  // namespace amr
