// Flux fields of one velocity state, for the reference's debugging export
// (Export_Fluxes, utilities.cpp:496-532) and the module-level entry point
// amressif_fluxes.  The step's fused kernels never store these fields; this
// kernel recomputes them from the faces with the reference's operation order
// (no FMA contraction), so they match a plain CPU evaluation bit for bit:
//   fluxHalfN1/N2 on the x- and y-faces: the QUICK half-node fluxes of U and V
//     (fn_flux_calc.cpp:60-105),
//   fluxConvect: their cell divergence (fn_flux_calc.cpp:161-180),
//   fluxViscous: the 5-point Laplacian of U and V over ren (fn_flux_calc.cpp:695-766),
// with the cell velocities of cont2cart and its wall ghost rules
// (utilities.cpp:16-350).  Not on the time step's path.
#include "device_common.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace amr {

namespace {

struct ExportArgs {
    const double *u, *v;  // x-faces, y-faces (pitch P)
    const double *lid;    // north-lid profile or null
    int n, P;
    double dx, dy, ren;
    int bc_lo0, bc_hi0, bc_lo1, bc_hi1;
    double inflow0, inflow1;
    // outputs (pitch P; any may be null)
    double *fxu, *fxv, *fyu, *fyv;  // half-node fluxes of U and V on x- and y-faces
    double *cu, *cv;                // fluxConvect components
    double *vu, *vv;                // fluxViscous components
};

// cell velocity (U, V) of cell (i, j) with at most one index outside the
// domain, by cont2cart's average and ghost rules (ghost layers 1 and 2)
__device__ inline void cell_uv(const ExportArgs &a, int i, int j, double &U, double &V)
{
    const int n = a.n, P = a.P;
    const int mi = i < 0 ? -i - 1 : (i >= n ? 2 * n - 1 - i : i);
    const int mj = j < 0 ? -j - 1 : (j >= n ? 2 * n - 1 - j : j);
    const double um = 0.5 * (a.u[mi + (size_t)P * mj] + a.u[mi + 1 + (size_t)P * mj]);
    const double vm = 0.5 * (a.v[mi + (size_t)P * mj] + a.v[mi + (size_t)P * (mj + 1)]);
    if (mi != i) {
        ghost_rule<double>(i < 0 ? 0 : 1, i < 0 ? a.bc_lo0 : a.bc_hi0, um, vm, 0.0, 0.0, U, V);
    } else if (mj != j) {
        const double gx = (a.lid && j >= n) ? a.lid[i] : a.inflow0;
        ghost_rule<double>(j < 0 ? 2 : 3, j < 0 ? a.bc_lo1 : a.bc_hi1, um, vm, gx, a.inflow1, U, V);
    } else {
        U = um;
        V = vm;
    }
}

__device__ inline double quick_ref(double vel, double a2, double a1, double b0, double b1)
{
    const double con = 0.5 * vel;
    const double ud = con + fabs(con), up = con - fabs(con);
    return up * ((1.0 / 8.0) * (-b1 - 2 * b0 + 3 * a1) + b0) + ud * ((1.0 / 8.0) * (-a2 - 2 * a1 + 3 * b0) + a1);
}

// half-node fluxes of U and V on x-face (i, j), i in [0, n], j in [0, n)
__device__ inline void xface_flux(const ExportArgs &a, int i, int j, double &fu, double &fv)
{
    double U[4], V[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cell_uv(a, i - 2 + q, j, U[q], V[q]);
    const double w = a.u[i + (size_t)a.P * j];
    fu = quick_ref(w, U[0], U[1], U[2], U[3]);
    fv = quick_ref(w, V[0], V[1], V[2], V[3]);
}

// half-node fluxes on y-face (i, j), i in [0, n), j in [0, n]
__device__ inline void yface_flux(const ExportArgs &a, int i, int j, double &fu, double &fv)
{
    double U[4], V[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cell_uv(a, i, j - 2 + q, U[q], V[q]);
    const double w = a.v[i + (size_t)a.P * j];
    fu = quick_ref(w, U[0], U[1], U[2], U[3]);
    fv = quick_ref(w, V[0], V[1], V[2], V[3]);
}

__global__ __launch_bounds__(256) void k_flux_fields(ExportArgs a)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n = a.n;
    if (i > n || j > n) return;
    const size_t k = i + (size_t)a.P * j;
    if (j < n && a.fxu) {
        double fu, fv;
        xface_flux(a, i, j, fu, fv);
        a.fxu[k] = fu;
        a.fxv[k] = fv;
    }
    if (i < n && a.fyu) {
        double fu, fv;
        yface_flux(a, i, j, fu, fv);
        a.fyu[k] = fu;
        a.fyv[k] = fv;
    }
    if (i >= n || j >= n) return;
    if (a.cu) {
        double wu, wv, eu, ev, su, sv, nu, nv;
        xface_flux(a, i, j, wu, wv);
        xface_flux(a, i + 1, j, eu, ev);
        yface_flux(a, i, j, su, sv);
        yface_flux(a, i, j + 1, nu, nv);
        a.cu[k] = (eu - wu) / a.dx + (nu - su) / a.dy;
        a.cv[k] = (ev - wv) / a.dx + (nv - sv) / a.dy;
    }
    if (a.vu) {
        double Uc, Vc, Uw, Vw, Ue, Ve, Us, Vs, Un, Vn;
        cell_uv(a, i, j, Uc, Vc);
        cell_uv(a, i - 1, j, Uw, Vw);
        cell_uv(a, i + 1, j, Ue, Ve);
        cell_uv(a, i, j - 1, Us, Vs);
        cell_uv(a, i, j + 1, Un, Vn);
        a.vu[k] = ((Ue - 2 * Uc + Uw) / (a.dx * a.dx) + (Un - 2 * Uc + Us) / (a.dy * a.dy)) / a.ren;
        a.vv[k] = ((Ve - 2 * Vc + Vw) / (a.dx * a.dx) + (Vn - 2 * Vc + Vs) / (a.dy * a.dy)) / a.ren;
    }
}

}  // namespace

hipError_t launch_flux_fields(const Ctx &c, const double *u, const double *v, double *fxu, double *fxv, double *fyu,
                              double *fyv, double *cu, double *cv, double *vu, double *vv)
{
    ExportArgs a{};
    a.u = u; a.v = v; a.lid = c.lid;
    a.n = c.g.n; a.P = c.g.P;
    a.dx = c.g.dx; a.dy = c.g.dy; a.ren = c.g.ren;
    a.bc_lo0 = c.g.bc_lo0; a.bc_hi0 = c.g.bc_hi0; a.bc_lo1 = c.g.bc_lo1; a.bc_hi1 = c.g.bc_hi1;
    a.inflow0 = c.g.inflow0; a.inflow1 = c.g.inflow1;
    a.fxu = fxu; a.fxv = fxv; a.fyu = fyu; a.fyv = fyv;
    a.cu = cu; a.cv = cv; a.vu = vu; a.vv = vv;
    const int m = c.g.n + 1;
    hipLaunchKernelGGL(k_flux_fields, dim3((m + 63) / 64, (m + 3) / 4), dim3(256), 0, c.stream, a);
    return hipGetLastError();
}

}  // namespace amr
