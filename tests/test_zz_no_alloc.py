"""amressif_poisson_rhs_noalloc, the allocation-free variant of
amressif_poisson_rhs (fn_rhs.cpp:71-125), through the C-ABI: the same
right-hand side (bit for bit), ||D||inf (bit for bit) and sum (to rounding)
as the existing entry point, on consecutive calls that share its static
scratch, the second with a smaller grid; device memory in use does not change
across a call; views that do not cover the box are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(n, seed):
    import torch
    from amressif_amd import api
    rng = np.random.default_rng(seed)
    ctx = api.Context(api.cavity_inputs(n, 100.0))
    u = torch.tensor(rng.uniform(-1, 1, (n, n + 1)), device="cuda")
    v = torch.tensor(rng.uniform(-1, 1, (n + 1, n)), device="cuda")
    return api, ctx, u, v


def _noalloc(api, ctx, u, v, rhs, out):
    n = ctx.n
    cv = api.cell_view(rhs, n, 1)
    return api._lib.amressif_poisson_rhs_noalloc(ctx._h, api.faces_view(u, v, n), ctypes.byref(cv),
                                                  ctypes.c_void_p(out.data_ptr()))


def test_noalloc_poisson_rhs_matches_the_allocating_entry_point():
    import torch
    for call, (n, seed) in enumerate(((384, 1), (200, 2), (64, 3))):  # each call smaller than the one before
        api, ctx, u, v = _setup(n, seed)
        ref, linf, s = ctx.poisson_righthand_side_calc(u, v)
        rhs = torch.zeros((n + 2, n + 2), dtype=torch.float64, device="cuda")
        out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = _noalloc(api, ctx, u, v, rhs, out)
        torch.cuda.synchronize()
        assert rc == 0, api._lib.amressif_last_error(ctx._h)
        if call > 0:  # the first call on a device may load the kernels' code object
            assert torch.cuda.mem_get_info()[0] == free0
        assert torch.equal(rhs, ref)
        o = out.cpu().numpy()
        assert o[0] == linf
        assert abs(o[1] - s) <= 1e-13 * float(ref.abs().sum())
        ctx.close()


def test_noalloc_poisson_rhs_refuses_short_views():
    import torch
    api, ctx, u, v = _setup(32, 4)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    small = torch.zeros((31, 31), dtype=torch.float64, device="cuda")
    cv = api._a4(small, (0, 0), (30, 30), 1)
    rc = api._lib.amressif_poisson_rhs_noalloc(ctx._h, api.faces_view(u, v, 32), ctypes.byref(cv),
                                               ctypes.c_void_p(out.data_ptr()))
    assert rc == -1 and "does not cover" in api._lib.amressif_last_error(ctx._h).decode()
