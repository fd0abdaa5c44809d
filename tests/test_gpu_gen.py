"""GPU parity of the general path (3D, and 2D with periodic directions)
against the general-path CPU FP64 oracle (oracle/refgen.c), through the C-ABI.

The general path keeps the reference's step (main.cpp:875-1202) on n^dim
cells with every boundary code; its periodic semantics are the oracle's
(refgen.h): periodic images in every ghost layer, face n = face 0.
Tolerances: 1e-10 relative for the velocity after N steps and 1e-7 for the
mean-removed pressure, bit equality for the cell velocities of cont2cart and
the pseudo-iteration counts equal.  The GPU's Poisson solve stops at
1e-12 ||b||inf, the oracle's at 1e-13; p sums the steps' phi, and the oracle's
own p moves by up to 2.7e-8 (relative) between Poisson tolerances 1e-12 and
1e-14 on the 2D periodic cases below (walls in x, periodic in y: the smallest
Neumann eigenvalue), while the velocities move by 2e-15.
"""
import numpy as np
import pytest

from oracle.refgen import GenOracle

pytestmark = pytest.mark.gpu


def _api():
    from amressif_amd import api
    return api


def _rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def _cuda(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _inputs(api, dim, n, ren, bc_lo, bc_hi, inflow=(1.0, 0.0, 0.0), length=1.0, **kw):
    args = dict(n_cell=n, max_grid_size=(n, n, n), IterNum=100, cfl=0.5, ren=ren, vis=1.0 / ren,
                phy_bc_lo=bc_lo, phy_bc_hi=bc_hi, inflow_waveform=inflow, momentum_tolerance=1e-10,
                spacedim=dim, hi_phy_dim=(length, length, length))
    args.update(kw)
    return api.default_inputs(**args)


def _random_faces(o, rng, amp, bc_lo, bc_hi):
    """random faces; wall-normal faces 0 on wall sides, face n = face 0 on periodic ones"""
    dim, n = o.dim, o.n
    f = []
    for d in range(dim):
        a = amp * rng.uniform(-1, 1, o.face_shape(d))
        ax = dim - 1 - d
        v = np.moveaxis(a, ax, 0)
        if bc_lo[d] == 111:
            v[-1] = v[0]
        else:
            v[0] = 0.0
            v[-1] = 0.0
        f.append(np.ascontiguousarray(a))
    return f


def _run_pair(dim, n, ren, bc_lo, bc_hi, steps, init="random", inflow=(1.0, 0.0, 0.0), length=1.0, pseudo=1,
              seed=0, amp=0.2, state=None):
    api = _api()
    import torch
    # the Poisson solves of both sides to 1e-12 / 1e-13 ||b||inf: the random starts are far from
    # divergence-free, and phi's error (which p accumulates) is its residual times the operator's
    # condition number
    inp = _inputs(api, dim, n, ren, bc_lo, bc_hi, inflow=inflow, length=length, PSEUDO_TIMESTEPPING=pseudo,
                  poisson_tol_rel=1e-12)
    dt = api.dt_of(inp)
    o = GenOracle(dim, n, ren, dt, length=(length,) * 3, bc_lo=bc_lo, bc_hi=bc_hi, inflow=inflow,
                  iter_num=inp.IterNum, mom_tol=inp.momentum_tolerance, pseudo=pseudo)
    rng = np.random.default_rng(seed)
    if state is not None:
        f, fp, ctx = state
    elif init == "random":
        f = _random_faces(o, rng, amp, bc_lo, bc_hi)
        fp, ctx = [a.copy() for a in f], o.ctx()
    else:
        f = [np.zeros(o.face_shape(d)) for d in range(dim)]
        fp, ctx = [a.copy() for a in f], o.ctx()
    ref = {"u": [a.copy() for a in f], "uprev": [a.copy() for a in fp], "ctx": ctx.copy()}
    ref_stats = o.advance(ref, steps)
    ctxg = api.Context(inp)
    ctxg.load_faces([_cuda(a) for a in f], [_cuda(a) for a in fp], _cuda(ctx))
    stats = ctxg.advance(steps)
    got_f = [torch.zeros(api.face_shape(d, n, dim), dtype=torch.float64, device="cuda") for d in range(dim)]
    got_c = torch.zeros(o.cell_shape(1, 2), dtype=torch.float64, device="cuda")
    ctxg.store_faces(got_f, None, got_c)
    got = {"u": [a.cpu().numpy() for a in got_f], "ctx": got_c.cpu().numpy()}
    return ref, ref_stats, got, stats, ctxg, o


def _check(ref, ref_stats, got, stats, dim, tol_u=1e-10, tol_p=1e-7):
    print("pseudo-iterations GPU", [s["pseudo_iters"] for s in stats], "oracle",
          [s["pseudo_iters"] for s in ref_stats], "V-cycles", [s["vcycles"] for s in stats])
    assert [s["pseudo_iters"] for s in stats] == [s["pseudo_iters"] for s in ref_stats]
    for d in range(dim):
        e = _rel(got["u"][d], ref["u"][d])
        print(f"velocity comp {d}: rel err {e:.2e}")
        assert e < tol_u, d
    sl = (0,) + (slice(1, -1),) * dim
    pg, pr = got["ctx"][sl], ref["ctx"][sl]
    e = _rel(pg - pg.mean(), pr - pr.mean())
    print(f"pressure (mean removed): rel err {e:.2e}")
    assert e < tol_p
    for s in stats:
        assert s["flags"] == 0
        assert s["div_new"] <= 1e-10 * s["div_star"] + 1e-12 * max(1.0, s["div_star"]), s


@pytest.mark.parametrize("bc_lo,bc_hi", [((111, 135, 0), (111, 165, 0)), ((111, 111, 0), (111, 111, 0)),
                                         ((135, 111, 0), (135, 111, 0))])
def test_2d_periodic_steps_match_oracle(bc_lo, bc_hi):
    ref, ref_stats, got, stats, _, _ = _run_pair(2, 64, 400.0, bc_lo, bc_hi, 4, inflow=(0.5, 0.0, 0.0))
    _check(ref, ref_stats, got, stats, 2)


def test_2d_taylor_green_matches_oracle():
    n = 64
    L = 2 * np.pi
    h = L / n
    xf, xc = np.arange(n + 1) * h, (np.arange(n) + 0.5) * h
    u = np.sin(xf)[None, :] * np.cos(xc)[:, None]
    v = -np.cos(xc)[None, :] * np.sin(xf)[:, None]
    ctx = np.zeros((2, n + 2, n + 2))
    ctx[0, 1:-1, 1:-1] = 0.25 * (np.cos(2 * xc)[None, :] + np.cos(2 * xc)[:, None])
    o = GenOracle(2, n, 1.0, 1.0, length=(L, L, L), bc_lo=(111,) * 3, bc_hi=(111,) * 3)
    o.userctx_ghosts(ctx)
    ref, ref_stats, got, stats, _, _ = _run_pair(2, n, 1.0, (111, 111, 0), (111, 111, 0), 5, length=L,
                                                 state=([u, v], [u.copy(), v.copy()], ctx))
    _check(ref, ref_stats, got, stats, 2)


@pytest.mark.parametrize("bc_lo,bc_hi,init", [
    ((135, 135, 135), (135, 165, 135), "rest"),       # the 3D lid-driven cavity from rest
    ((135, 135, 135), (135, 165, 135), "random"),
    ((-135, 135, 135), (135, -135, 165), "random"),   # slip and unit-inflow sides
    ((111, 111, 111), (111, 111, 111), "random"),
    ((135, 111, 135), (135, 111, 165), "random"),
])
def test_3d_steps_match_oracle(bc_lo, bc_hi, init):
    ref, ref_stats, got, stats, _, _ = _run_pair(3, 24, 200.0, bc_lo, bc_hi, 3, init=init, inflow=(1.0, 0.0, 0.5))
    _check(ref, ref_stats, got, stats, 3)


@pytest.mark.parametrize("dim", [2, 3])
def test_explicit_time_marching_matches_oracle(dim):
    bc_lo = (111, 135, 135)
    bc_hi = (111, 165, 135)
    ref, ref_stats, got, stats, _, _ = _run_pair(dim, 24 if dim == 3 else 64, 100.0, bc_lo, bc_hi, 3, pseudo=0)
    assert all(s["pseudo_iters"] == 0 for s in stats)
    for d in range(dim):
        assert _rel(got["u"][d], ref["u"][d]) < 1e-10


def test_3d_abc_flow_matches_oracle():
    n, ren = 32, 1.0
    L = 2 * np.pi
    h = L / n
    c = (np.arange(n) + 0.5) * h
    ones = np.ones(n + 1)
    u = (np.sin(c)[:, None, None] + np.cos(c)[None, :, None]) * ones[None, None, :]
    v = (np.sin(c)[None, None, :] + np.cos(c)[:, None, None]) * ones[None, :, None]
    w = (np.sin(c)[None, :, None] + np.cos(c)[None, None, :]) * ones[:, None, None]
    f = [np.ascontiguousarray(a) for a in (u, v, w)]
    ctx = np.zeros((2, n + 2, n + 2, n + 2))
    ref, ref_stats, got, stats, _, _ = _run_pair(3, n, ren, (111,) * 3, (111,) * 3, 3, length=L,
                                                 state=(f, [a.copy() for a in f], ctx))
    _check(ref, ref_stats, got, stats, 3)


@pytest.mark.parametrize("bc_lo,bc_hi", [((135, 135, 135), (135, 165, 135)), ((111, -135, 135), (111, 135, 165))])
def test_3d_store_state_ghosts_equal_the_oracle(bc_lo, bc_hi):
    """velCart with two ghost layers (cont2cart, corners in the reference's W,
    E, S, N, B, F order) and userCtx with one (Neumann mirror / periodic
    image), bit for bit."""
    api = _api()
    import torch
    n, dim = 12, 3
    inp = _inputs(api, dim, n, 100.0, bc_lo, bc_hi, inflow=(0.3, 0.7, -0.4))
    o = GenOracle(dim, n, 100.0, 1e-3, bc_lo=bc_lo, bc_hi=bc_hi, inflow=(0.3, 0.7, -0.4))
    rng = np.random.default_rng(4)
    f = [rng.uniform(-1, 1, o.face_shape(d)) for d in range(dim)]
    for d in range(dim):
        if bc_lo[d] == 111:
            v = np.moveaxis(f[d], dim - 1 - d, 0)
            v[-1] = v[0]
    ctx = o.ctx()
    ctx[(slice(None),) + (slice(1, -1),) * 3] = rng.uniform(-1, 1, (2, n, n, n))
    ctx[1] -= ctx[1][1:-1, 1:-1, 1:-1].mean()
    c = api.Context(inp)
    c.load_faces([_cuda(a) for a in f], [_cuda(a) for a in f], _cuda(ctx))
    vc = torch.zeros(o.cell_shape(2, dim), dtype=torch.float64, device="cuda")
    uc = torch.zeros(o.cell_shape(1, 2), dtype=torch.float64, device="cuda")
    c.store_faces(None, None, uc, vc)
    ref_vc = o.cont2cart(f)
    ref_ctx = ctx.copy()
    o.userctx_ghosts(ref_ctx)
    assert np.array_equal(vc.cpu().numpy(), ref_vc)
    got = uc.cpu().numpy()
    assert np.array_equal(got[0], ref_ctx[0])
    assert np.abs(got[1] - ref_ctx[1]).max() < 1e-15


def test_general_path_refusals():
    api = _api()
    inp = _inputs(api, 3, 16, 100.0, (135, 135, 135), (135, 165, 135))
    c = api.Context(inp)
    import torch
    with pytest.raises(api.AmressifError) as e:
        c.poisson_righthand_side_calc(torch.zeros((16, 17), dtype=torch.float64, device="cuda"),
                                      torch.zeros((17, 16), dtype=torch.float64, device="cuda"))
    assert e.value.code == -3
    with pytest.raises(api.AmressifError) as e:
        api.Tape(c, 4)
    assert e.value.code == -3


def test_general_path_kernel_timing():
    api = _api()
    inp = _inputs(api, 3, 64, 100.0, (135, 135, 135), (135, 165, 135))
    c = api.Context(inp)
    c.init_state()
    c.advance(1)
    for which in (7, 8, 9):
        ms, nbytes = c.time_kernel(which, reps=5)
        print(f"kernel {which}: {ms * 1e3:.1f} us, {nbytes / ms / 1e6:.0f} GB/s")
        assert ms > 0 and nbytes > 0


@pytest.mark.parametrize("dim,n,mgs,bc_lo,bc_hi", [
    (3, 20, (8, 12, 20), (135, -135, 135), (135, 165, 135)),   # ragged x boxes, walls, slip, lid
    (3, 16, (8, 8, 8), (111, 135, 111), (111, 165, 111)),      # periodic x and z
    (2, 40, (16, 24), (111, 135, 0), (111, 165, 0))])          # 2D periodic x
def test_multibox_general_path_matches_one_box(dim, n, mgs, bc_lo, bc_hi):
    """amressif_load_state_mf / _store_state_mf on the general path: a run loaded
    from and stored to the boxes of max_grid_size (main.cpp:317-324) is bit
    for bit the one-box run, ghosts of userCtx and velCart included (inside the
    domain the neighbouring box's values, outside it the wall rules or the
    periodic image)."""
    api = _api()
    import torch
    inp = _inputs(api, dim, n, 200.0, bc_lo, bc_hi, inflow=(1.0, 0.0, 0.5))
    o = GenOracle(dim, n, 200.0, api.dt_of(inp), bc_lo=bc_lo, bc_hi=bc_hi, inflow=(1.0, 0.0, 0.5))
    rng = np.random.default_rng(11)
    f = _random_faces(o, rng, 0.2, bc_lo, bc_hi)
    fp = [a.copy() for a in f]
    ctx = o.ctx()
    ctx[(slice(None),) + (slice(1, -1),) * dim] = rng.uniform(-1, 1, (2,) + (n,) * dim)
    steps = 3
    one = api.Context(inp)
    one.load_faces([_cuda(a) for a in f], [_cuda(a) for a in fp], _cuda(ctx))
    one.advance(steps)
    gf = [torch.zeros(api.face_shape(d, n, dim), dtype=torch.float64, device="cuda") for d in range(dim)]
    gp = [torch.zeros(api.face_shape(d, n, dim), dtype=torch.float64, device="cuda") for d in range(dim)]
    gc = torch.zeros(o.cell_shape(1, 2), dtype=torch.float64, device="cuda")
    gv = torch.zeros(o.cell_shape(2, dim), dtype=torch.float64, device="cuda")
    one.store_faces(gf, gp, gc, gv)
    ref = {"f": [a.cpu().numpy() for a in gf], "fp": [a.cpu().numpy() for a in gp], "ctx": gc.cpu().numpy(),
           "vc": gv.cpu().numpy()}

    boxes = api.box_array_nd(n, dim, mgs)
    assert len(boxes) > 1
    MF = api.MultiFabND
    mf_f = [MF(boxes, dim, d).fill_from(f[d]) for d in range(dim)]
    mf_p = [MF(boxes, dim, d).fill_from(fp[d]) for d in range(dim)]
    mf_c = MF(boxes, dim, "cell", 2, 1).fill_from(ctx, src_ghosts=1)
    many = api.Context(inp)
    many.load_faces_mf(mf_f, mf_p, mf_c)
    stats = many.advance(steps)
    assert all(s["flags"] == 0 for s in stats)
    out_f = [MF(boxes, dim, d) for d in range(dim)]
    out_p = [MF(boxes, dim, d) for d in range(dim)]
    out_c = MF(boxes, dim, "cell", 2, 1)
    out_v = MF(boxes, dim, "cell", dim, 2)
    many.store_faces_mf(out_f, out_p, out_c, out_v)

    def window(glob, mf, b, ghosts):
        vlo, vhi = mf.window(*mf.boxes[b])
        sl = tuple(slice(vlo[d] + ghosts, vhi[d] + 1 + ghosts) for d in range(dim))[::-1]
        return glob[(slice(None),) + sl] if mf.kind == "cell" else glob[sl]

    for b in range(len(boxes)):
        for d in range(dim):
            assert np.array_equal(out_f[d].fabs[b].cpu().numpy(), window(ref["f"][d], out_f[d], b, 0)), (b, d)
            assert np.array_equal(out_p[d].fabs[b].cpu().numpy(), window(ref["fp"][d], out_p[d], b, 0)), (b, d)
        assert np.array_equal(out_c.fabs[b].cpu().numpy(), window(ref["ctx"], out_c, b, 1)), b
        assert np.array_equal(out_v.fabs[b].cpu().numpy(), window(ref["vc"], out_v, b, 2)), b
    # a box array that does not tile the domain is refused, the state unchanged
    with pytest.raises(api.AmressifError) as e:
        many.load_faces_mf([MF(boxes[:-1], dim, d) for d in range(dim)], mf_p, mf_c)
    assert e.value.code == -1
