// Device helpers shared by the kernel files (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace amr {

// Deterministic block reduction: fixed wave-shuffle tree, then the wave
// partials combined in order by thread 0.  Block size a multiple of 64; every
// thread of the block must call it.  The result is valid in thread 0.
template <bool IsMax>
__device__ inline double block_reduce(double v, double *sh)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_down(v, o, 64);
        if (IsMax) v = (w > v || w != w) ? w : v;
        else v += w;
    }
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sh[0];
        for (int w = 1; w < nw; ++w) {
            if (IsMax) r = (sh[w] > r || sh[w] != sh[w]) ? sh[w] : r;
            else r += sh[w];
        }
    }
    return r;
}

// N block reductions at once: value k is a max (NaN wins) when bit k of
// MAXMASK is set, else a sum.  Each value is reduced in block_reduce's order
// (the same bits), with two barriers in all instead of two per value.  sh
// holds 16 N doubles; the results are valid in thread 0.
template <unsigned MAXMASK, int N>
__device__ inline void block_reduce_n(double (&v)[N], double *sh)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double w = __shfl_down(v[k], o, 64);
            if ((MAXMASK >> k) & 1u) v[k] = (w > v[k] || w != w) ? w : v[k];
            else v[k] += w;
        }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[16 * k + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double r = sh[16 * k];
            for (int w = 1; w < nw; ++w) {
                const double x = sh[16 * k + w];
                if ((MAXMASK >> k) & 1u) r = (x > r || x != x) ? x : r;
                else r += x;
            }
            v[k] = r;
        }
}

// Strided walk of a block over nb records of K doubles (record q at
// p + S q): thread t visits q = t, t + T, t + 2T, ... (T = blockDim.x) in that
// order, as the plain loop does, so a reduction over the records gives the
// same bits; the loads of four records are issued before any is used (one
// memory round trip per four records instead of one per record).
template <int K, int S, class F>
__device__ __forceinline__ void for_records(const double *p, int nb, F &&f)
{
    const int T = blockDim.x;
    for (int q0 = threadIdx.x; q0 < nb; q0 += 4 * T) {
        double a[4][K];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = q0 + u * T;
#pragma unroll
            for (int k = 0; k < K; ++k) a[u][k] = q < nb ? p[(size_t)S * q + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q0 + u * T < nb) f(a[u]);
    }
}

// Workgroup b of a one-dimensional grid of B runs on XCD b mod 8.  The
// logical block number hands each XCD a contiguous range: the first B mod 8
// XCDs take B / 8 + 1 blocks, the others B / 8, in order.  Tiles a stencil
// reads again (the neighbours of a tile) then mostly sit on the same XCD and
// its L2.
__device__ __forceinline__ int xcd_block(int b, int B)
{
    const int q = B >> 3, r = B & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Cell tiles of 64 x 16 for a workgroup of 256 threads (four passes of 64 x 4)
// in row-major order over the XCD-contiguous block numbers.
struct CellTile {
    int i0, j0;
    bool live;  // false: a padding workgroup past the last tile
};
__device__ inline CellTile cell_tile(int ntx, int nty)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    CellTile t;
    t.i0 = (blk % ntx) * 64;
    t.j0 = (blk / ntx) * 16;
    t.live = blk / ntx < nty;
    return t;
}

// QUICK half-node flux of one face (fn_flux_calc.cpp:66-73): vel is the face
// velocity, a2 = U(i-2), a1 = U(i-1), b0 = U(i), b1 = U(i+1) along the face
// normal.  ud = 0.5 vel + |0.5 vel| = max(vel, 0), up = min(vel, 0).
template <class R>
__device__ __forceinline__ R quick_flux(R vel, R a2, R a1, R b0, R b1)
{
    const R con = R(0.5) * vel;
    const R ud = con + fabs(con), up = con - fabs(con);
    return up * ((R(1) / R(8)) * (-b1 - R(2) * b0 + R(3) * a1) + b0) +
           ud * ((R(1) / R(8)) * (-a2 - R(2) * a1 + R(3) * b0) + a1);
}

// Wall ghost rule of cont2cart (utilities.cpp:87-269) for one non-corner ghost
// cell from its mirror cell's values.  side: 0 W, 1 E, 2 S, 3 N.  Every rule
// is affine, ghost = +-mirror + constant; Homog = true drops the constant (the
// rule applied to an increment of the velocity).
template <class R, bool Homog = false>
__device__ __forceinline__ void ghost_rule(int side, int code, R um, R vm, R gx, R gy, R &U, R &V)
{
    const R two = Homog ? R(0) : R(2);
    if (code == 135) { U = -um; V = -vm; return; }
    if (side == 0) {  // W: -135 slip, +-165 unit inflow
        if (code == -135) { U = -um; V = vm; } else { U = two - um; V = two - vm; }
    } else if (side == 1) {  // E: the slip branch tests the code 1; 165 unit inflow
        if (code == 1) { U = -um; V = vm; } else { U = two - um; V = two - vm; }
    } else if (side == 2) {  // S
        if (code == -135) { U = um; V = -vm; } else { U = two - um; V = two - vm; }
    } else {  // N: 165 is the lid, 2 * inflow - mirror
        if (code == -135) { U = um; V = -vm; } else { U = two * gx - um; V = two * gy - vm; }
    }
}

// values held by lane - 1 / lane + 1 (DPP wave shifts; bound_ctrl zero-fills
// the edge lanes, so each 32-bit half is one v_mov_b32_dpp with no prior
// initialisation of the destination)
__device__ __forceinline__ double lane_prev(double x)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_next(double x)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// (float: the compiler can also fold the shift into the consuming VALU
// instruction)
__device__ __forceinline__ float lane_prev(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_next(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xF, 0xF, true));
}

}  // namespace amr
