// Allocation-free variant of amressif_poisson_rhs (poisson_righthand_side_calc,
// Source/src1_rever/fn_rhs.cpp:71-125) for callers that cannot allocate
// device memory: code under HIP graph capture and tight serving loops.
//
// It reads the caller's Array4 views directly (no pitched working copy) and
// keeps its block partials in the file-scope device array `scratch`, whose
// address on each device is taken once, at the entry point's first call on
// that device (with it current), with hipGetSymbolAddress; the host never
// passes the array's name itself to a kernel.  The results go to a
// caller-owned device buffer, so the call neither allocates nor synchronises.
//
// One block per two rows of cells: at most kMaxRows / 2 blocks, so scratch
// holds 8192 x 2 doubles (128 KiB per device, held until the process ends).
// Every block writes its partial on every call, so nothing needs clearing
// between calls; calls that overlap on one device (two streams) share the
// array and race.
#include <mutex>

#include "device_common.hpp"
#include "internal.hpp"

namespace {

constexpr int kRowsPerBlock = 2;
constexpr int kMaxRows = 16384;  // the largest n_cell amressif_create accepts
constexpr int kMaxBlocks = kMaxRows / kRowsPerBlock;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[2 * kMaxBlocks];

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // scratch's address on each device

// Array4 element (i, j) of component 0: ptr[(i - lo0) + js (j - lo1)]
struct View {
    const double *p;
    int lo0, lo1;
    int64_t js;
};
struct OutView {
    double *p;
    int lo0, lo1;
    int64_t js;
};

__global__ __launch_bounds__(256) void k_prhs_views(View u, View v, OutView b, int n, double dx, double dy,
                                                    double c15, double *part)
{
    __shared__ double sh[16];
    const int j0 = blockIdx.x * kRowsPerBlock;
    double m = 0.0, s = 0.0;
    for (int jj = 0; jj < kRowsPerBlock; ++jj) {
        const int j = j0 + jj;
        if (j >= n) break;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double uw = u.p[(i - u.lo0) + u.js * (j - u.lo1)], ue = u.p[(i + 1 - u.lo0) + u.js * (j - u.lo1)];
            const double vs = v.p[(i - v.lo0) + v.js * (j - v.lo1)], vn = v.p[(i - v.lo0) + v.js * (j + 1 - v.lo1)];
            const double d = (ue - uw) / dx + (vn - vs) / dy;
            const double bb = c15 * d;
            b.p[(i - b.lo0) + b.js * (j - b.lo1)] = bb;
            m = (fabs(d) > m || d != d) ? fabs(d) : m;
            s += bb;
        }
    }
    m = amr::block_reduce<true>(m, sh);
    s = amr::block_reduce<false>(s, sh + 8);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = m;
        part[2 * blockIdx.x + 1] = s;
    }
}

__global__ __launch_bounds__(1024) void k_prhs_views_fin(const double *part, int nb, double *out)
{
    __shared__ double sh[32];
    double m = 0.0, s = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const double pm = part[2 * k];
        m = (pm > m || pm != pm) ? pm : m;
        s += part[2 * k + 1];
    }
    m = amr::block_reduce<true>(m, sh);
    s = amr::block_reduce<false>(s, sh + 16);
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = s;
    }
}

// scratch's address on the current device (taken once per device)
hipError_t scratch_ptr(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

bool covers(const amr_array4 &a, int lo0, int lo1, int hi0, int hi1)
{
    return a.ptr && a.dtype == 0 && a.ncomp >= 1 && a.hi[2] >= a.lo[2] && a.lo[0] <= lo0 && a.lo[1] <= lo1 &&
           a.hi[0] >= hi0 && a.hi[1] >= hi1;
}

}  // namespace

extern "C" int amressif_poisson_rhs_noalloc(amressif_ctx *ctx, const amr_array4 vel[2], amr_array4 *rhs,
                                            double *dev_out)
{
    if (ctx && ctx->c.gen)
        return amr_fail(const_cast<amressif_ctx *>(ctx), AMRESSIF_ERR_UNSUPPORTED, "amressif_poisson_rhs_noalloc: not available for 3D or periodic inputs");
    if (!ctx || !vel || !rhs || !dev_out) return amr_fail(ctx, AMRESSIF_ERR_ARG, "poisson_rhs_noalloc: null argument");
    const amr::Ctx &c = ctx->c;
    const int n = c.g.n;
    if (n > kMaxRows)
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "poisson_rhs_noalloc: n_cell " + std::to_string(n) +
                                                   " exceeds the scratch limit of " + std::to_string(kMaxRows) + " rows");
    if (!covers(vel[0], 0, 0, n, n - 1) || !covers(vel[1], 0, 0, n - 1, n) || !covers(*rhs, 0, 0, n - 1, n - 1))
        return amr_fail(ctx, AMRESSIF_ERR_ARG, "poisson_rhs_noalloc: a view does not cover the box (double views only)");
    double *part = nullptr;
    if (hipError_t e = scratch_ptr(&part)) return amr_fail(ctx, AMRESSIF_ERR_HIP, hipGetErrorString(e));
    auto view = [](const amr_array4 &a) {
        return View{(const double *)a.ptr, a.lo[0], a.lo[1], (int64_t)(a.hi[0] - a.lo[0] + 1)};
    };
    const OutView ob{(double *)rhs->ptr, rhs->lo[0], rhs->lo[1], (int64_t)(rhs->hi[0] - rhs->lo[0] + 1)};
    const int nb = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(k_prhs_views, dim3(nb), dim3(256), 0, c.stream, view(vel[0]), view(vel[1]), ob, n, c.g.dx,
                       c.g.dy, 1.5 / c.g.dt, part);
    hipLaunchKernelGGL(k_prhs_views_fin, dim3(1), dim3(1024), 0, c.stream, (const double *)part, nb, dev_out);
    if (hipError_t e = hipGetLastError()) return amr_fail(ctx, AMRESSIF_ERR_HIP, hipGetErrorString(e));
    return AMRESSIF_OK;
}
