"""GPU parity of the HIP path (through the C-ABI) against the CPU FP64 oracle.

Module-level entry points are compared one by one with the oracle's
restatement of the reference functions; whole steps are compared after N
steps.  Tolerances are written per test: relative to the reference field's
max-norm, 1e-12 for single kernels (FMA contraction and operation order are
the only differences), 1e-10 for velocity and mean-removed pressure after N
steps (the parity bar).
"""
import numpy as np
import pytest

from amressif_amd import workloads as W
from oracle.refcpu import Oracle

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _api():
    from amressif_amd import api
    return api


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def make_pair(n, ren, bc_lo=(135, 135), bc_hi=(135, 165), inflow=(1.0, 0.0), precision=0, poisson="mg",
              poisson_tol=1e-10, lid=None, **kw):
    api = _api()
    inp = api.cavity_inputs(n, ren, precision=precision, phy_bc_lo=(bc_lo[0], bc_lo[1], 135),
                            phy_bc_hi=(bc_hi[0], bc_hi[1], 135), inflow_waveform=(inflow[0], inflow[1], 0.0),
                            poisson_tol_rel=poisson_tol, **kw)
    dt = api.dt_of(inp)
    o = Oracle(n, ren, dt, bc_lo=bc_lo, bc_hi=bc_hi, inflow=inflow, poisson=poisson,
               poisson_tol_rel=1e-13 if poisson == "mg" else 1e-10, lid=lid,
               iter_num=inp.IterNum, mom_tol=inp.momentum_tolerance)
    return inp, o


def cuda(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def random_faces(n, rng, walls_zero=False):
    u = rng.uniform(-1, 1, (n, n + 1))
    v = rng.uniform(-1, 1, (n + 1, n))
    if walls_zero:
        u[:, 0] = u[:, -1] = 0.0
        v[0, :] = v[-1, :] = 0.0
    return u, v


BCS = [((135, 135), (135, 165)), ((-135, -135), (135, -135)), ((165, 135), (165, 165)), ((135, -135), (135, 135))]


@pytest.mark.parametrize("n", [64, 100])
@pytest.mark.parametrize("bc", BCS)
def test_cont2cart_parity(n, bc):
    rng = np.random.default_rng(1)
    inp, o = make_pair(n, 100.0, bc_lo=bc[0], bc_hi=bc[1], inflow=(0.7, -0.2))
    api = _api()
    ctx = api.Context(inp)
    u, v = random_faces(n, rng)
    vc_ref = o.cont2cart(u, v)
    vc = ctx.cont2cart(cuda(u), cuda(v)).cpu().numpy()
    assert rel(vc, vc_ref) == 0.0  # same arithmetic, corners included


@pytest.mark.parametrize("n", [64, 100, 120])
@pytest.mark.parametrize("bc", BCS[:3])
def test_momentum_rhs_parity(n, bc):
    rng = np.random.default_rng(2)
    inp, o = make_pair(n, 100.0, bc_lo=bc[0], bc_hi=bc[1])
    ctx = _api().Context(inp)
    u, v = random_faces(n, rng)
    uc = np.zeros((2, n + 2, n + 2))
    uc[:, 1:-1, 1:-1] = rng.uniform(-1, 1, (2, n, n))
    o.userctx_ghosts(uc)
    rx_ref, ry_ref = o.stage_rhs(u, v, uc)
    rx, ry = ctx.momentum_rhs(cuda(u), cuda(v), cuda(uc))
    assert rel(rx.cpu().numpy(), rx_ref) < 1e-12
    assert rel(ry.cpu().numpy(), ry_ref) < 1e-12
    # wall-normal faces: R = 0 exactly
    assert np.all(rx.cpu().numpy()[:, [0, -1]] == 0.0)
    assert np.all(ry.cpu().numpy()[[0, -1], :] == 0.0)


@pytest.mark.parametrize("sub", [0, 1, 2, 3])
def test_rk_stage_parity(sub):
    n = 96
    rng = np.random.default_rng(3 + sub)
    inp, o = make_pair(n, 400.0)
    ctx = _api().Context(inp)
    us, vs = random_faces(n, rng, walls_zero=True)
    uk, vk = random_faces(n, rng, walls_zero=True)
    un, vn = random_faces(n, rng, walls_zero=True)
    um, vm = random_faces(n, rng, walls_zero=True)
    uc = np.zeros((2, n + 2, n + 2))
    uc[0, 1:-1, 1:-1] = rng.uniform(-1, 1, (n, n))
    o.userctx_ghosts(uc)
    # oracle: stage RHS then the dual-time update (momentum.cpp:68-87)
    rx, ry = o.stage_rhs(us, vs, uc)
    ust, vst = us.copy(), vs.copy()
    rk = W.rk_coefficients(o.prm.dt)
    o.rk4_stage(rk[sub], rx, ry, ust, vst, uk, vk, un - um, vn - vm, un, vn)
    tu, tv = cuda(us), cuda(vs)
    ctx.runge_kutta4_pseudo_time_stepping(sub, tu, tv, cuda(uk), cuda(vk), cuda(un), cuda(vn), cuda(um), cuda(vm),
                                          cuda(uc))
    assert rel(tu.cpu().numpy(), ust) < 1e-12
    assert rel(tv.cpu().numpy(), vst) < 1e-12


def test_pseudo_norm_parity():
    n = 128
    rng = np.random.default_rng(7)
    inp, o = make_pair(n, 100.0)
    ctx = _api().Context(inp)
    a, b = random_faces(n, rng)
    c, d = random_faces(n, rng)
    ref = o.error_computation(a, b, c, d)
    got = ctx.Error_Computation(cuda(a), cuda(b), cuda(c), cuda(d))
    assert abs(got - ref) <= 1e-14 * ref


def test_poisson_rhs_parity():
    n = 100
    rng = np.random.default_rng(8)
    inp, o = make_pair(n, 100.0)
    ctx = _api().Context(inp)
    u, v = random_faces(n, rng)
    rhs_ref, linf_ref, sum_ref = o.poisson_rhs(u, v)
    rhs, linf, s = ctx.poisson_righthand_side_calc(cuda(u), cuda(v))
    assert rel(rhs.cpu().numpy()[1:-1, 1:-1], rhs_ref[1:-1, 1:-1]) == 0.0
    assert linf == linf_ref
    assert abs(s - sum_ref) <= 1e-12 * np.abs(rhs_ref).sum()


@pytest.mark.parametrize("n", [8, 64, 100, 256])
def test_poisson_solve_vs_exact(n):
    rng = np.random.default_rng(9)
    inp, o = make_pair(n, 100.0, poisson="dct")
    ctx = _api().Context(inp)
    b = np.zeros((n + 2, n + 2))
    b[1:-1, 1:-1] = rng.standard_normal((n, n))
    sol_ref, _, _ = o.poisson_solve(b)  # exact (DCT), mean-zero
    sol, iters, res = ctx.poisson_advance(cuda(b), tol_rel=1e-12)
    assert res <= 1e-12
    assert iters <= 30
    assert rel(sol.cpu().numpy()[1:-1, 1:-1], sol_ref[1:-1, 1:-1]) < 1e-9


def test_project_parity():
    n = 96
    rng = np.random.default_rng(10)
    inp, o = make_pair(n, 100.0)
    ctx = _api().Context(inp)
    us, vs = random_faces(n, rng, walls_zero=True)
    uc = np.zeros((2, n + 2, n + 2))
    uc[:, 1:-1, 1:-1] = rng.uniform(-1, 1, (2, n, n))
    o.userctx_ghosts(uc)
    uc_ref = uc.copy()
    u_ref, v_ref, _ = o.update_solution(uc_ref, us, vs)
    tuc = cuda(uc)
    u, v, d = ctx.update_solution(cuda(us), cuda(vs), tuc)
    assert rel(u.cpu().numpy(), u_ref) < 1e-14
    assert rel(v.cpu().numpy(), v_ref) < 1e-14
    assert rel(tuc.cpu().numpy()[0], uc_ref[0]) < 1e-15


def run_both(n, ren, steps, init="zero", **kw):
    api = _api()
    inp, o = make_pair(n, ren, **kw)
    st = W.initial_state(n, init)
    ref = {k: v.copy() for k, v in st.items()}
    ref_stats = o.advance(ref, steps)
    run = api.Run.from_numpy(inp, st)
    if kw.get("lid") is not None:
        run.ctx.set_lid_profile(cuda(kw["lid"]))
    stats = run.advance(steps)
    got = run.fetch()
    return ref, ref_stats, got, stats


def div_floor(state):
    # the round-off floor of the discrete divergence: a few ulps of |u| / h
    n = state["u"].shape[0]
    return max(1e-12, 16 * np.finfo(np.float64).eps * max(np.abs(state["u"]).max(), np.abs(state["v"]).max()) * n)


def check_step_parity(ref, ref_stats, got, stats, tol=1e-10):
    assert rel(got["u"], ref["u"]) < tol
    assert rel(got["v"], ref["v"]) < tol
    p, pr = got["ctx"][0, 1:-1, 1:-1], ref["ctx"][0, 1:-1, 1:-1]
    assert rel(p - p.mean(), pr - pr.mean()) < tol
    assert rel(got["uprev"], ref["uprev"]) < tol
    floor = div_floor(got)
    for s, r in zip(stats, ref_stats):
        assert s["pseudo_iters"] == r["pseudo_iters"]
        assert s["flags"] == 0
        assert s["div_new"] <= 1e-10 * s["div_star"] + floor
    # wall-normal faces stay exactly zero
    assert np.all(got["u"][:, [0, -1]] == 0.0) and np.all(got["v"][[0, -1], :] == 0.0)


def test_advance_cavity128_re100_zero_start():
    ref, rs, got, st = run_both(128, 100.0, 20)
    check_step_parity(ref, rs, got, st)
    assert rel(got["vc"][:, 2:-2, 2:-2], ref["vc"][:, 2:-2, 2:-2]) < 1e-10


def test_c1_cavity128_re100_full_200_steps():
    # the plain CPU-reference workload in full: 128^2, Re = 100, from rest, 200 steps
    ref, rs, got, st = run_both(128, 100.0, 200)
    check_step_parity(ref, rs, got, st)


def test_c2_cavity1024_re1000_first_steps():
    # the FP64 GPU workload at its full size (1024^2, Re = 1000, seeded), first steps
    ref, rs, got, st = run_both(1024, 1000.0, 2, init="seeded")
    check_step_parity(ref, rs, got, st)


def test_advance_warm_started_poisson_matches_oracle():
    # the warm-started solve stops on the same ||r||inf <= 1e-10 ||b||inf test
    ref, rs, got, st = run_both(128, 400.0, 12, init="seeded", poisson_warm_start=1)
    check_step_parity(ref, rs, got, st)
    _, _, _, cold = run_both(128, 400.0, 12, init="seeded", poisson_warm_start=0)
    assert sum(s["vcycles"] for s in st[2:]) < sum(s["vcycles"] for s in cold[2:])


def test_advance_seeded_256_re1000():
    ref, rs, got, st = run_both(256, 1000.0, 6, init="seeded")
    check_step_parity(ref, rs, got, st)


@pytest.mark.parametrize("bc", BCS[1:3])
def test_advance_other_walls(bc):
    ref, rs, got, st = run_both(64, 200.0, 5, init="seeded", bc_lo=bc[0], bc_hi=bc[1])
    assert rel(got["u"], ref["u"]) < 1e-10
    assert rel(got["v"], ref["v"]) < 1e-10
    for s, r in zip(st, rs):
        assert s["pseudo_iters"] == r["pseudo_iters"]


def test_advance_lid_profile():
    n = 64
    lid = W.lid_profile(n)
    ref, rs, got, st = run_both(n, 500.0, 5, init="seeded", lid=lid)
    check_step_parity(ref, rs, got, st)


@pytest.mark.parametrize("n", [8, 100])
def test_advance_odd_sizes(n):
    ref, rs, got, st = run_both(n, 100.0, 4, init="seeded")
    check_step_parity(ref, rs, got, st)


def _with_env(env, fn):
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("n,rows,init,bc,lid", [
    (8, 4, "seeded", 0, False), (64, 4, "seeded", 0, False), (64, 5, "seeded", 1, False),
    (100, 16, "seeded", 2, True), (130, 7, "seeded", 0, False), (256, 16, "seeded", 3, False),
    (256, 64, "zero", 0, True), (300, 12, "seeded", 0, False)])
def test_two_stage_pass_matches_four_stage_launches(n, rows, init, bc, lid):
    """The FP64 pseudo-iteration as two launches of two stages each (the first
    stage of a pair kept in registers), and as four single-stage launches with
    the unrolled interior march, against four single-stage launches on the
    general march: the same pseudo-iteration counts and the same fields to
    rounding, for row chunks that do and do not divide n, every wall set, a
    lid profile, and batches of pseudo-iterations that overshoot convergence
    (the skipped last pair copies the base)."""
    api = _api()
    prof = W.lid_profile(n) if lid else None

    def run():
        inp = api.cavity_inputs(n, 400.0, phy_bc_lo=(BCS[bc][0][0], BCS[bc][0][1], 135),
                                phy_bc_hi=(BCS[bc][1][0], BCS[bc][1][1], 135), inflow_waveform=(0.8, -0.1, 0.0))
        r = api.Run.from_numpy(inp, W.initial_state(n, init))
        if prof is not None:
            r.ctx.set_lid_profile(cuda(prof))
        st = r.advance(5)
        return r.fetch(), st

    srows = "16" if rows >= 16 else "8"  # the unrolled march's chunk heights
    f4, s4 = _with_env({"AMRESSIF_STAGE_PAIR": "0", "AMRESSIF_STAGE_FAST": "0", "AMRESSIF_STAGE_ROWS": srows}, run)
    f1, s1 = _with_env({"AMRESSIF_STAGE_PAIR": "0", "AMRESSIF_STAGE_FAST": "1", "AMRESSIF_STAGE_ROWS": srows}, run)
    f2, s2 = _with_env({"AMRESSIF_STAGE_PAIR": "1", "AMRESSIF_PAIR_ROWS": str(rows)}, run)
    for f, st, name in ((f1, s1, "unrolled"), (f2, s2, "pair")):
        print(f"n={n} rows={rows} {name}: pseudo-iterations {[s['pseudo_iters'] for s in st]}, "
              f"u {rel(f['u'], f4['u']):.1e} v {rel(f['v'], f4['v']):.1e}")
        assert [s["pseudo_iters"] for s in st] == [s["pseudo_iters"] for s in s4], name
        assert all(s["flags"] == 0 for s in st)
        for k in ("u", "v", "uprev"):
            assert rel(f[k], f4[k]) < 1e-12, (name, k)
        p2, p4 = f["ctx"][0, 1:-1, 1:-1], f4["ctx"][0, 1:-1, 1:-1]
        assert rel(p2 - p2.mean(), p4 - p4.mean()) < 1e-12, name


@pytest.mark.parametrize("n,rows,init,bc,lid", [
    (8, 4, "seeded", 0, False), (64, 4, "seeded", 0, False), (64, 5, "seeded", 1, False),
    (100, 16, "seeded", 2, True), (130, 7, "seeded", 0, False), (256, 16, "seeded", 3, False),
    (256, 64, "zero", 0, True), (300, 12, "seeded", 0, False)])
def test_one_pass_mixed_iteration_matches_two_launches(n, rows, init, bc, lid):
    """The mixed iteration as one k_mixed_pair launch (the residual kept in
    registers, u^{k+1} chained through two arrays, the projection copying
    nothing) against k_stage64 mode 3 + k_tangent: the same pseudo-iteration
    and V-cycle counts and the same fields to rounding, after two advance calls
    (the state at a call boundary, velContPrev included), for row chunks that do
    and do not divide n, every wall set, a lid profile, and batches that
    overshoot convergence (the skipped launches copy u^k)."""
    api = _api()
    prof = W.lid_profile(n) if lid else None

    def run():
        inp = api.cavity_inputs(n, 400.0, precision=1, phy_bc_lo=(BCS[bc][0][0], BCS[bc][0][1], 135),
                                phy_bc_hi=(BCS[bc][1][0], BCS[bc][1][1], 135), inflow_waveform=(0.8, -0.1, 0.0))
        r = api.Run.from_numpy(inp, W.initial_state(n, init))
        if prof is not None:
            r.ctx.set_lid_profile(cuda(prof))
        st = r.advance(5)
        f1 = r.fetch()
        st += r.advance(3)
        return f1, r.fetch(), st

    a1, a2, sa = _with_env({"AMRESSIF_MIXED_PAIR": "0"}, run)
    b1, b2, sb = _with_env({"AMRESSIF_MIXED_PAIR": "1", "AMRESSIF_MPAIR_ROWS": str(rows)}, run)
    print(f"n={n} rows={rows}: pseudo-iterations {[s['pseudo_iters'] for s in sb]}, V-cycles "
          f"{[s['vcycles'] for s in sb]}, u {rel(b2['u'], a2['u']):.1e} v {rel(b2['v'], a2['v']):.1e}")
    assert [s["pseudo_iters"] for s in sb] == [s["pseudo_iters"] for s in sa]
    assert [s["vcycles"] for s in sb] == [s["vcycles"] for s in sa]
    assert all(s["flags"] == 0 for s in sb)
    for fa, fb in ((a1, b1), (a2, b2)):
        for k in ("u", "v", "uprev", "vprev"):
            assert rel(fb[k], fa[k]) < 1e-12, k
        pa, pb = fa["ctx"][0, 1:-1, 1:-1], fb["ctx"][0, 1:-1, 1:-1]
        assert rel(pb - pb.mean(), pa - pa.mean()) < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n,rows", [(300, 16), (257, 12), (130, 17)])
def test_edge_first_wave_order_gives_the_same_steps(precision, n, rows):
    """The pair kernels dispatch the waves on the general march (at a wall or an
    array edge) first and the interior waves after them, each XCD a contiguous
    range; the other order is the plain chunk-major one.  Only the order of the
    norm partials' sum differs: the same pseudo-iteration and V-cycle counts,
    the same fields to rounding, for sizes and chunk heights that leave
    partial bands and chunks (and, for 17-row chunks in the FP64 step, no
    unrolled interior path: every wave an edge wave)."""
    api = _api()

    def run():
        inp = api.cavity_inputs(n, 400.0, precision=precision)
        r = api.Run.from_numpy(inp, W.initial_state(n, "seeded"))
        st = r.advance(4)
        return r.fetch(), st

    rk = "AMRESSIF_MPAIR_ROWS" if precision == 1 else "AMRESSIF_PAIR_ROWS"
    fa, sa = _with_env({"AMRESSIF_EDGE_FIRST": "0", rk: str(rows)}, run)
    fb, sb = _with_env({"AMRESSIF_EDGE_FIRST": "1", rk: str(rows)}, run)
    assert [s["pseudo_iters"] for s in sb] == [s["pseudo_iters"] for s in sa]
    assert [s["vcycles"] for s in sb] == [s["vcycles"] for s in sa]
    assert all(s["flags"] == 0 for s in sb)
    for k in ("u", "v", "uprev", "vprev"):
        assert rel(fb[k], fa[k]) < 1e-12, k
    pa, pb = fa["ctx"][0, 1:-1, 1:-1], fb["ctx"][0, 1:-1, 1:-1]
    assert rel(pb - pb.mean(), pa - pa.mean()) < 1e-12


@pytest.mark.parametrize("n", [256, 300, 130])
def test_three_sweep_finest_legs_of_the_fp32_cycle(n):
    """The mixed step's FP32 V-cycle with three red-black sweeps on each finest
    leg (the default) against two (AMRESSIF_MG_PRE3=AMRESSIF_MG_POST3=0): every
    step converges to the same Poisson tolerance in no more V-cycles, and the
    fields agree to the 1e-6 bar of the mixed-against-FP64 comparison (the two
    solves stop at different iterates inside the tolerance)."""
    api = _api()

    def run():
        inp = api.cavity_inputs(n, 1000.0, precision=1)
        r = api.Run.from_numpy(inp, W.initial_state(n, "seeded"))
        st = r.advance(6)
        return r.fetch(), st

    f2, s2 = _with_env({"AMRESSIF_MG_PRE3": "0", "AMRESSIF_MG_POST3": "0"}, run)
    f3, s3 = _with_env({"AMRESSIF_MG_PRE3": "1", "AMRESSIF_MG_POST3": "1"}, run)
    print(f"n={n}: V-cycles two sweeps {[s['vcycles'] for s in s2]}, three {[s['vcycles'] for s in s3]}")
    assert all(s["flags"] == 0 and s["poisson_res"] <= 1e-10 for s in s3)
    assert sum(s["vcycles"] for s in s3) <= sum(s["vcycles"] for s in s2)
    for k in ("u", "v"):
        assert rel(f3[k], f2[k]) < 1e-6, k
    p2, p3 = f2["ctx"][0, 1:-1, 1:-1], f3["ctx"][0, 1:-1, 1:-1]
    assert rel(p3 - p3.mean(), p2 - p2.mean()) < 1e-6


@pytest.mark.parametrize("n", [130, 98])
def test_odd_bottom_poisson_solve_matches_the_exact_solve(n):
    """n = 130 coarsens to an odd 65 x 65 bottom (98 to 49 x 49), above the
    dense pseudo-inverse's 1024 cells: that level is solved exactly by the DCT
    as four dense products.  The solve reaches the exact mean-zero solution."""
    api = _api()
    rng = np.random.default_rng(n)
    rhs = np.zeros((n + 2, n + 2))
    rhs[1:-1, 1:-1] = rng.standard_normal((n, n))
    xe, _, _ = Oracle(n, 100.0, 1e-4, poisson="dct").poisson_solve(rhs)
    for prec, tol, bar in ((0, 1e-12, 1e-9), (1, 1e-10, 1e-7)):
        ctx = api.Context(api.cavity_inputs(n, 100.0, precision=prec))
        sol, it, res = ctx.poisson_advance(cuda(rhs), tol_rel=tol)
        assert res <= tol and it <= 40, (prec, it, res)
        assert rel(sol.cpu().numpy()[1:-1, 1:-1], xe[1:-1, 1:-1]) < bar, prec


def test_odd_bottom_steps_match_oracle():
    ref, rs, got, st = run_both(130, 400.0, 4, init="seeded")
    check_step_parity(ref, rs, got, st)


def test_unsupported_multigrid_size_reported():
    api = _api()
    with pytest.raises(api.AmressifError) as e:
        api.Context(api.cavity_inputs(4098, 100.0))  # coarsens to a 2049 x 2049 bottom level
    assert e.value.code == -3


def test_iternum_cap_from_host():
    api = _api()
    n = 32
    inp = api.cavity_inputs(n, 100.0, IterNum=0)
    run = api.Run.from_numpy(inp, W.initial_state(n, "seeded"))
    st = run.advance(2)
    assert all(s["pseudo_iters"] == 1 and (s["flags"] & api.FLAG_CAP) for s in st)


def test_nonfinite_reported():
    api = _api()
    n = 32
    inp = api.cavity_inputs(n, 100.0)
    state = W.initial_state(n, "seeded")
    state["u"][5, 7] = np.nan
    run = api.Run.from_numpy(inp, state)
    with pytest.raises(api.AmressifError) as e:
        run.advance(3)
    assert e.value.code == -5


# ------------------------------------------------------------- mixed precision
def _run_gpu(n, ren, steps, precision, init="seeded", **kw):
    api = _api()
    inp = api.cavity_inputs(n, ren, precision=precision, **kw)
    run = api.Run.from_numpy(inp, W.initial_state(n, init))
    stats = run.advance(steps)
    return run.fetch(), stats


@pytest.mark.parametrize("mm", [1, 0])
@pytest.mark.parametrize("n,ren,steps", [(128, 100.0, 30), (256, 1000.0, 20)])
def test_mixed_matches_fp64_within_1e_6(n, ren, steps, mm):
    api = _api()
    a, sa = _run_gpu(n, ren, steps, api.AMRESSIF_MIXED, mixed_momentum=mm)
    b, sb = _run_gpu(n, ren, steps, api.AMRESSIF_FP64)
    # per-field bar of the mixed form: 1e-6 relative to the FP64 field's max-norm
    assert rel(a["u"], b["u"]) < 1e-6
    assert rel(a["v"], b["v"]) < 1e-6
    pa, pb = a["ctx"][0, 1:-1, 1:-1], b["ctx"][0, 1:-1, 1:-1]
    assert rel(pa - pa.mean(), pb - pb.mean()) < 1e-6
    for s, r in zip(sa, sb):
        if mm == 0:
            # the RK4 recursion to first order in the Jacobian contracts by
            # |1 + z + z^2/2| instead of the RK4 polynomial (z ~ -0.6 here): up
            # to ~15% more iterations than the FP64 dual-time loop
            assert r["pseudo_iters"] - 1 <= s["pseudo_iters"] <= 1.25 * r["pseudo_iters"] + 2
        else:
            # Newton corrections with |R'| dt / 1.5 << 1: a handful of iterations
            assert 1 <= s["pseudo_iters"] <= 6 and s["pseudo_iters"] < r["pseudo_iters"]
        assert s["flags"] == 0
        # the outer Poisson residual and the divergence check stay FP64
        assert s["div_new"] <= 1e-10 * s["div_star"] + div_floor(a)
    assert np.all(a["u"][:, [0, -1]] == 0.0) and np.all(a["v"][[0, -1], :] == 0.0)


def test_mixed_matches_oracle_within_1e_6():
    api = _api()
    n, ren, steps = 128, 400.0, 10
    got, _ = _run_gpu(n, ren, steps, api.AMRESSIF_MIXED)
    inp, o = make_pair(n, ren)
    ref = W.initial_state(n, "seeded")
    o.advance(ref, steps)
    assert rel(got["u"], ref["u"]) < 1e-6 and rel(got["v"], ref["v"]) < 1e-6


def test_kernel_timing_diagnostic():
    api = _api()
    for prec in (api.AMRESSIF_FP64, api.AMRESSIF_MIXED):
        run = api.Run.from_numpy(api.cavity_inputs(256, 1000.0, precision=prec), W.initial_state(256, "seeded"))
        run.advance(1)
        kinds = (0, 1, 3, 4) if prec == api.AMRESSIF_MIXED else (0, 1, 6)
        for which in kinds:
            ms, nbytes = run.ctx.time_kernel(which, reps=5)
            assert ms > 0 and nbytes > 0
        if prec == api.AMRESSIF_FP64:
            with pytest.raises(api.AmressifError):
                run.ctx.time_kernel(3, reps=1)
    ms, nbytes = run.ctx.time_kernel(2, reps=5, copy_bytes=1 << 24)
    assert ms > 0 and nbytes == 2 * (1 << 24)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("batch", ["0", "1"])
def test_in_step_kernel_timing_counts_and_leaves_results_unchanged(prec, batch):
    """The in-step timing hook counts exactly the launches that did work (from
    the step statistics) and does not change the step's results; with the first
    V-cycle batch capped at one cycle, the projections enqueued behind unfinished
    solves (which do nothing) are not counted either."""
    import os
    api = _api()
    n, steps = 256, 4
    state = W.initial_state(n, "seeded")
    plain = api.Run.from_numpy(api.cavity_inputs(n, 1000.0, precision=prec), state)
    plain.advance(steps)
    ref = plain.fetch()
    old = os.environ.get("AMRESSIF_MG_BATCH")
    os.environ["AMRESSIF_MG_BATCH"] = batch
    try:
        run = api.Run.from_numpy(api.cavity_inputs(n, 1000.0, precision=prec), state)
    finally:
        if old is None:
            os.environ.pop("AMRESSIF_MG_BATCH")
        else:
            os.environ["AMRESSIF_MG_BATCH"] = old
    run.ctx.kernel_timing(True)
    stats = run.advance(steps)
    t = run.ctx.kernel_times()
    run.ctx.kernel_timing(False)
    got = run.fetch()
    for k in ("u", "v", "ctx"):
        assert np.array_equal(got[k], ref[k]), k
    pi = sum(s["pseudo_iters"] for s in stats)
    vc = sum(s["vcycles"] for s in stats)
    if prec == 0:  # two stages per pass
        assert t["pair"][1] == pi and t["pair_last"][1] == pi
        assert t["stage"][1] == 0 and t["stage_last"][1] == 0
        assert t["resid"][1] == 0 and t["tangent"][1] == 0
    else:  # the mixed iteration in one pass
        assert t["mpair"][1] == pi
        assert t["resid"][1] == 0 and t["tangent"][1] == 0
        assert t["stage"][1] == 0
    assert t["mg_down"][1] == vc and t["mg_up"][1] == vc
    assert t["project"][1] == steps and t["poisson_start"][1] == steps
    cells = float(n * n)
    per_cell = {"stage": 64, "stage_last": 64, "resid": 48, "tangent": 48, "pair": 64, "pair_last": 64,
                "mpair": 48}
    for name, (ms, cnt, nbytes) in t.items():
        assert (ms > 0) == (cnt > 0) and (nbytes > 0) == (cnt > 0), name
        assert ms < 50.0, name  # a 256^2 launch takes microseconds
        if cnt and name in per_cell:
            assert nbytes == per_cell[name] * cells, name
    # disabled: the totals stay, nothing more is counted
    run.advance(1)
    assert run.ctx.kernel_times() == t


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("var", ["AMRESSIF_START_ROWS", "AMRESSIF_PROJECT_ROWS"])
@pytest.mark.parametrize("n", [62, 256])
def test_row_march_kernels_in_the_step(prec, var, n):
    """Whole steps with the row-march Poisson start (b from u*, warm start in
    the mixed step) or the row-march projection (with the next step's BDF2
    constant) match those with the cell-tile kernels to rounding; n = 62 puts
    the east wall face on the first lane past a band."""
    import os
    api = _api()
    steps = 4
    state = W.initial_state(n, "seeded")
    out = {}
    for rows in ("0", "8", "16"):
        old = os.environ.get(var)
        os.environ[var] = rows
        try:
            run = api.Run.from_numpy(api.cavity_inputs(n, 1000.0, precision=prec), state)
        finally:
            if old is None:
                os.environ.pop(var)
            else:
                os.environ[var] = old
        stats = run.advance(steps)
        assert all(s["flags"] == 0 for s in stats)
        out[rows] = run.fetch()
    b = out["0"]
    for rows in ("8", "16"):
        a = out[rows]
        assert rel(a["u"], b["u"]) < 1e-12 and rel(a["v"], b["v"]) < 1e-12, rows
        pa, pb = a["ctx"][0, 1:-1, 1:-1], b["ctx"][0, 1:-1, 1:-1]
        assert rel(pa - pa.mean(), pb - pb.mean()) < 1e-10, rows


@pytest.mark.parametrize("prec", [0, 1])
def test_short_first_vcycle_batch_gives_the_same_steps(prec):
    """A first batch of one V-cycle per solve (AMRESSIF_MG_BATCH=1): every step's
    projection, enqueued behind the batch, finds the solve unfinished and does
    nothing; the host undoes its buffer swaps and runs more cycles. The steps
    are identical to those of the default batches."""
    import os
    api = _api()
    n, steps = 128, 6
    state = W.initial_state(n, "seeded")
    out = {}
    for batch in ("0", "1"):
        old = os.environ.get("AMRESSIF_MG_BATCH")
        os.environ["AMRESSIF_MG_BATCH"] = batch
        try:
            run = api.Run.from_numpy(api.cavity_inputs(n, 1000.0, precision=prec), state)
        finally:
            if old is None:
                os.environ.pop("AMRESSIF_MG_BATCH")
            else:
                os.environ["AMRESSIF_MG_BATCH"] = old
        stats = run.advance(steps)
        assert all(s["flags"] == 0 for s in stats) and all(s["vcycles"] > 1 for s in stats)
        out[batch] = (run.fetch(), [(s["pseudo_iters"], s["vcycles"]) for s in stats])
    (a, sa), (b, sb) = out["1"], out["0"]
    assert sa == sb
    for k in ("u", "v", "ctx"):
        assert np.array_equal(a[k], b[k]), k


# ------------------------------------------- PSEUDO_TIMESTEPPING = 0 (explicit)
@pytest.mark.parametrize("n,init,steps", [(64, "zero", 10), (128, "seeded", 6), (100, "seeded", 4)])
def test_explicit_time_marching_matches_oracle(n, init, steps):
    """main.cpp:900-926 + momentum.cpp:96-191: one explicit update per step,
    countIter 0, and the reference's wall faces: with no wall-ghost pass on
    fluxTotal the lid row's wall-normal faces move (the oracle's closed form
    pins that), so the projected field keeps a uniform divergence equal to the
    net wall flux; both sides agree on all of it."""
    api = _api()
    ren = 400.0
    inp = api.cavity_inputs(n, ren, PSEUDO_TIMESTEPPING=0)
    o = Oracle(n, ren, api.dt_of(inp), poisson="mg", poisson_tol_rel=1e-13, pseudo=0)
    st = W.initial_state(n, init)
    ref = {k: v.copy() for k, v in st.items()}
    rs = o.advance(ref, steps)
    run = api.Run.from_numpy(inp, st)
    gs = run.advance(steps)
    got = run.fetch()
    assert rel(got["u"], ref["u"]) < 1e-10 and rel(got["v"], ref["v"]) < 1e-10
    assert rel(got["uprev"], ref["uprev"]) < 1e-10
    p, pr = got["ctx"][0, 1:-1, 1:-1], ref["ctx"][0, 1:-1, 1:-1]
    assert rel(p - p.mean(), pr - pr.mean()) < 1e-10
    for s, r in zip(gs, rs):
        assert s["pseudo_iters"] == 0 == r["pseudo_iters"] and s["flags"] == 0
        assert s["norm_error"] == pytest.approx(r["norm_error"], rel=1e-9)
        assert s["div_new"] == pytest.approx(r["div_new"], rel=1e-6, abs=1e-9 * r["div_star"])
    assert np.abs(got["u"][-1, [0, -1]]).min() > 0  # the lid row's wall faces moved, as in the reference


def test_explicit_time_marching_refuses_tape_and_two_level():
    api = _api()
    run = api.Run.from_numpy(api.cavity_inputs(32, 100.0, PSEUDO_TIMESTEPPING=0), W.initial_state(32, "zero"))
    with pytest.raises(api.AmressifError) as e:
        run.ctx.tape(2)
    assert e.value.code == -3
    with pytest.raises(api.AmressifError) as e:
        api.TwoLevel(api.cavity_inputs(32, 100.0, PSEUDO_TIMESTEPPING=0), 24, W.seeded_two_level(32, 24))
    assert e.value.code == -3
