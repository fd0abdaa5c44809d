// The general path: the fractional step on n^dim cells (dim 2 or 3) with
// periodic directions (phy_bc 111) beside the wall codes, FP64.  Used by a
// context whose inputs are 3D or name a periodic direction; the 2D wall-only
// configurations keep the fused single-level kernels.
//
// Layout in HBM (one box, i fastest, then j, then k; no padding):
//   face array of direction d: n + 1 along d, n along the others
//   cell arrays (p, phi, b, multigrid levels, fluxTotal comps): n^dim
//   cell velocities X: dim comps with two ghost layers, (n + 4)^dim each
//   (in 2D the k extent is 1 and has no ghosts)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "internal.hpp"

namespace amr {
namespace gen {

constexpr int kBlock = 256;
constexpr int kMaxGenLevels = 16;

// Ghost rule of a wall side (cont2cart, utilities.cpp:87-346): 0 every comp
// -mirror (135), 1 slip: the normal comp -mirror, the others +mirror (-135;
// the east side's code 1), 2 unit inflow 2 - mirror (165 on W, E, S, B, F,
// -165 on W), 3 the north side's 2 inflow - mirror (165).
struct Geom {
    int dim, n, nk;        // nk: cells along k (n in 3D, 1 in 2D)
    int per[3];            // periodic directions
    int kind_lo[3], kind_hi[3];
    double inflow[3];
    double idx[3], idx2[3];  // 1/dx, 1/dx^2
    double iren, c15, dt;
    int pseudo;            // PSEUDO_TIMESTEPPING
    int64_t nf[3];         // faces per direction
    int64_t ncell;         // n^dim
    int64_t xs;            // cells per component of X ((n + 4)^dim)
};

struct Level {
    int n;            // cells per side
    double a[3];      // 1/h_d^2 of this level
    double *e, *r;    // correction and right-hand side (n^dim)
    int64_t ncell;
};

struct Ctx {
    Geom g{};
    double rk[4]{};
    double *U[3]{}, *Up[3]{}, *C[3]{}, *B[3]{}, *S[3]{};  // u^n, u^{n-1}, BDF2 constant, pseudo base, u*
    double *X = nullptr;   // cell velocities with ghosts
    double *T = nullptr;   // fluxTotal (dim comps, n^dim)
    double *p = nullptr, *phi = nullptr, *b = nullptr;
    Level lev[kMaxGenLevels]{};
    int nlev = 0;
    double *pinv = nullptr;  // bottom pseudo-inverse (m x m, symmetric)
    int mb = 0;
    double *partials = nullptr;
    int64_t npartials = 0;
    int nb_face[3]{};        // blocks per face direction in k_gen_update
    int last_P = 8, last_V = 8;
    int flux_z = 1;          // 3D: fluxTotal by the z-march kernel (0: one thread per cell)
    int flux_zc = 16;        // z-march kernel: planes per workgroup
};

// ---- launchers (kernels_gen.hip); every guarded kernel returns at once once
// its loop's flag in the control block is set
hipError_t cells(const Ctx &c, hipStream_t s, double *const *faces, double *X, const Ctl *guard);
hipError_t flux(const Ctx &c, hipStream_t s, double *const *faces, const double *X, double *T, const Ctl *guard);
// stage update of every face: u* <- u^k + rk (R - c15 u* + C); last: the
// norm partials of u^k - u* and u^k <- u* as well
hipError_t update(const Ctx &c, hipStream_t s, double rk, int last, const Ctl *guard);
hipError_t fin_pseudo(const Ctx &c, hipStream_t s, Ctl *ctl);
hipError_t bdf2(const Ctx &c, hipStream_t s);
// b = (1.5/dt) D(faces) (b may be null: the norm only); partials of max|D|, sum b
hipError_t divergence(const Ctx &c, hipStream_t s, double *const *faces, double *b, int *nblocks);
// which 0: div_star, rhs_sum and b's mean (the step's Poisson RHS), 1: div_new
hipError_t fin_div(const Ctx &c, hipStream_t s, int nblocks, Ctl *ctl, int which);
hipError_t mg_start(const Ctx &c, hipStream_t s, Ctl *ctl);
hipError_t resid0(const Ctx &c, hipStream_t s, const Ctl *ctl, int *nblocks);
hipError_t fin_resid(const Ctx &c, hipStream_t s, int nblocks, Ctl *ctl);
hipError_t vcycle(const Ctx &c, hipStream_t s, const Ctl *ctl);
// phi's mean into the control block, then u = u* - (dt/1.5) G phi and p += phi - mean
hipError_t project(const Ctx &c, hipStream_t s, Ctl *ctl);
hipError_t fin_step(const Ctx &c, hipStream_t s, Ctl *ctl, amressif_step_stats *log);
hipError_t fill(hipStream_t s, double *dst, int64_t count, double value);
// Array4 views <-> internal arrays: a box region [0, ext) with the view's lo/hi
hipError_t pack(hipStream_t s, const amr_array4 &a, int comp, double *dst, int ex, int ey, int ez);
hipError_t unpack(hipStream_t s, const double *src, const amr_array4 &a, int comp, int ex, int ey, int ez);
// the window [lo, hi] of a view <-> the same cells of an internal array of
// extents ex x ey x .. (unpack 0: view -> array)
hipError_t pack_window(hipStream_t s, const amr_array4 &a, int comp, double *dst, int ex, int ey, const int *lo,
                       const int *hi, int unpack);
// userCtx / velCart views with ghosts: the cells the view holds, ghosts by the
// rules (p, phi: Neumann mirror / periodic image; phi shifted by -mean)
hipError_t unpack_cells_ghosted(const Ctx &c, hipStream_t s, const double *src, const amr_array4 &a, int comp,
                                const double *shift);
hipError_t unpack_velcart(const Ctx &c, hipStream_t s, const double *X, const amr_array4 &a);
hipError_t phi_mean(const Ctx &c, hipStream_t s, Ctl *ctl);

}  // namespace gen
}  // namespace amr
