"""GPU checks of the tangent-linear and adjoint step (C-ABI pair
amressif_step_fwd / amressif_step_bwd / amressif_step_tlm).

The reference has no adjoint, so these are pinned by mathematics, not by a
restatement:
  * dot-product test  <TLM(dx), l> = <dx, ADJ(l)>  over several steps;
  * the tangent against central finite differences of the forward step, both
    the HIP forward and the CPU oracle's forward;
  * the adjoint gradient of the cavity cost J (lower-left kinetic energy)
    through the checkpointed torch.autograd harness against central finite
    differences along seeded directions of (lid profile, initial velocity).
"""
import numpy as np
import pytest
import torch

from amressif_amd import workloads as W

pytestmark = pytest.mark.gpu


def _api():
    from amressif_amd import api
    return api


def _inputs(n, ren, **kw):
    api = _api()
    args = dict(momentum_tolerance=1e-14, IterNum=400, poisson_tol_rel=1e-13)
    args.update(kw)
    return api.cavity_inputs(n, ren, **args)


def _faces(rng, n, dev="cuda"):
    u = rng.standard_normal((n, n + 1))
    v = rng.standard_normal((n + 1, n))
    u[:, [0, n]] = 0.0  # wall-normal faces are not free
    v[[0, n], :] = 0.0
    return (torch.tensor(u, device=dev), torch.tensor(v, device=dev))


def _cells(rng, n, dev="cuda"):
    p = np.zeros((n + 2, n + 2))
    p[1:-1, 1:-1] = rng.standard_normal((n, n))
    return torch.tensor(p, device=dev)


def _dual(n, dev="cuda"):
    z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)  # noqa: E731
    return [z(n, n + 1), z(n + 1, n), z(n, n + 1), z(n + 1, n), z(n + 2, n + 2), z(n)]


def _inner(a, b):
    return sum(float((x * y).sum()) for x, y in zip(a, b))


def _seeded_run(n, ren, lid=None, **kw):
    api = _api()
    run = api.Run.from_numpy(_inputs(n, ren, **kw), W.initial_state(n, "seeded"))
    if lid is not None:
        run.ctx.set_lid_profile(torch.tensor(lid, device="cuda"))
    return run


@pytest.mark.parametrize("n,steps", [(32, 1), (32, 3), (64, 2)])
def test_dot_product_tangent_vs_adjoint(n, steps):
    rng = np.random.default_rng(7 + n + steps)
    run = _seeded_run(n, 1000.0, lid=W.lid_profile(n))
    ctx = run.ctx
    run.advance(2)  # a state with history
    tape = ctx.tape(steps)
    ctx.step_fwd(steps, tape)
    # tangent of a random perturbation (velocity, history, pressure, lid)
    du, dv = _faces(rng, n)
    dup, dvp = _faces(rng, n)
    dp = _cells(rng, n)
    dlid = torch.tensor(rng.standard_normal(n), device="cuda")
    ctx.load_dual(du, dv, dup, dvp, dp, dlid)
    tl_stats = ctx.step_tlm(tape, 0, steps)
    out = _dual(n)
    ctx.store_dual(*out)
    # adjoint of a random output weight
    lu, lv = _faces(rng, n)
    lup, lvp = _faces(rng, n)
    lp = _cells(rng, n)
    ctx.load_dual(lu, lv, lup, lvp, lp, None)
    ad_stats = ctx.step_bwd(tape, steps)
    adj = _dual(n)
    ctx.store_dual(*adj)
    assert tape.steps == 0
    lhs = _inner(out[:5], [lu, lv, lup, lvp, lp])
    rhs = _inner([du, dv, dup, dvp, dp, dlid], adj)
    scale = max(abs(lhs), abs(rhs))
    assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs)
    for s in tl_stats + ad_stats:
        assert s["flags"] == 0 and 1 <= s["lin_iters"] <= 30, s


def _forward_state(n, ren, lid, state, steps):
    api = _api()
    run = api.Run.from_numpy(_inputs(n, ren), state)
    run.ctx.set_lid_profile(torch.tensor(lid, device="cuda"))
    run.advance(steps)
    return run.fetch()


def test_tangent_matches_finite_differences_of_the_forward():
    n, ren, steps, eps = 32, 400.0, 2, 1e-6
    rng = np.random.default_rng(11)
    lid = W.lid_profile(n)
    base = W.initial_state(n, "seeded")
    base["uprev"] = base["uprev"] * 0.98  # nonzero BDF2 history
    d = {"u": _faces(rng, n, "cpu")[0].numpy() * 1e-2, "v": _faces(rng, n, "cpu")[1].numpy() * 1e-2}
    d["uprev"], d["vprev"] = d["u"] * 0.5, d["v"] * 0.5
    dp = _cells(rng, n, "cpu").numpy()
    dlid = rng.standard_normal(n) * 1e-2

    def shifted(sign):
        s = {k: v.copy() for k, v in base.items()}
        for k in ("u", "v", "uprev", "vprev"):
            s[k] = s[k] + sign * eps * d[k]
        s["ctx"][0] = s["ctx"][0] + sign * eps * dp
        return s
    fp = _forward_state(n, ren, lid + eps * dlid, shifted(1), steps)
    fm = _forward_state(n, ren, lid - eps * dlid, shifted(-1), steps)
    # tangent along the same direction at the base trajectory
    api = _api()
    run = api.Run.from_numpy(_inputs(n, ren), base)
    run.ctx.set_lid_profile(torch.tensor(lid, device="cuda"))
    tape = run.ctx.tape(steps)
    run.ctx.step_fwd(steps, tape)
    t = lambda a: torch.tensor(a, device="cuda")  # noqa: E731
    run.ctx.load_dual(t(d["u"]), t(d["v"]), t(d["uprev"]), t(d["vprev"]), t(dp), t(dlid))
    run.ctx.step_tlm(tape, 0, steps)
    out = _dual(n)
    run.ctx.store_dual(*out)
    for key, got in (("u", out[0]), ("v", out[1]), ("uprev", out[2]), ("vprev", out[3])):
        fd = (fp[key] - fm[key]) / (2 * eps)
        err = np.abs(got.cpu().numpy() - fd).max() / np.abs(fd).max()
        assert err < 1e-6, (key, err)
    pfd = (fp["ctx"][0, 1:-1, 1:-1] - fm["ctx"][0, 1:-1, 1:-1]) / (2 * eps)
    pgot = out[4].cpu().numpy()[1:-1, 1:-1]
    pfd, pgot = pfd - pfd.mean(), pgot - pgot.mean()
    assert np.abs(pgot - pfd).max() / np.abs(pfd).max() < 1e-6


def test_tangent_matches_finite_differences_of_the_oracle():
    from oracle.refcpu import Oracle
    api = _api()
    n, ren, eps = 16, 100.0, 1e-6
    rng = np.random.default_rng(5)
    base = W.initial_state(n, "seeded")
    du, dv = (x.numpy() * 1e-2 for x in _faces(rng, n, "cpu"))
    inp = _inputs(n, ren)
    o = Oracle(n, ren, api.dt_of(inp), poisson="dct", mom_tol=1e-14, iter_num=400)
    fds = []
    for sign in (1, -1):
        s = {k: v.copy() for k, v in base.items()}
        s["u"] = s["u"] + sign * eps * du
        s["v"] = s["v"] + sign * eps * dv
        s["uprev"], s["vprev"] = s["u"].copy(), s["v"].copy()
        o.advance(s, 1)
        fds.append(s)
    run = api.Run.from_numpy(inp, base)
    tape = run.ctx.tape(1)
    run.ctx.step_fwd(1, tape)
    t = lambda a: torch.tensor(a, device="cuda")  # noqa: E731
    run.ctx.load_dual(t(du), t(dv), t(du), t(dv), None, None)
    run.ctx.step_tlm(tape, 0, 1)
    out = _dual(n)
    run.ctx.store_dual(*out)
    for key, got in (("u", out[0]), ("v", out[1])):
        fd = (fds[0][key] - fds[1][key]) / (2 * eps)
        err = np.abs(got.cpu().numpy() - fd).max() / np.abs(fd).max()
        assert err < 1e-6, (key, err)


@pytest.mark.parametrize("every", [4, 16])
def test_gradient_of_cavity_cost_matches_finite_differences(every):
    from amressif_amd.adjoint import CavityRollout, CheckpointedRollout, kinetic_energy_lower_left
    n, ren, steps, eps = 32, 1000.0, 12, 1e-6
    inp = _inputs(n, ren)
    lid0 = torch.tensor(W.lid_profile(n), device="cuda", requires_grad=True)
    st = W.initial_state(n, "seeded")
    u0 = torch.tensor(st["u"], device="cuda", requires_grad=True)
    v0 = torch.tensor(st["v"], device="cuda", requires_grad=True)
    ro = CheckpointedRollout(inp, steps, every=every)
    u, v = CavityRollout.apply(lid0, u0, v0, ro)
    J = kinetic_energy_lower_left(u, v)
    J.backward()
    assert len(ro.checkpoints) == -(-steps // every)

    def cost(lid, uu, vv):
        r = CheckpointedRollout(inp, steps, every=steps)
        a, b = r.forward(lid, uu, vv)
        return float(kinetic_energy_lower_left(a, b))
    rng = np.random.default_rng(2024)
    for _ in range(3):
        dl = torch.tensor(rng.standard_normal(n), device="cuda")
        # divergence-free initial-velocity direction from a seeded streamfunction
        du, dv = (torch.tensor(x, device="cuda") for x in W.seeded_velocity(n, int(rng.integers(1 << 30))))
        with torch.no_grad():
            jp = cost(lid0 + eps * dl, u0 + eps * du, v0 + eps * dv)
            jm = cost(lid0 - eps * dl, u0 - eps * du, v0 - eps * dv)
        fd = (jp - jm) / (2 * eps)
        ad = float((lid0.grad * dl).sum() + (u0.grad * du).sum() + (v0.grad * dv).sum())
        assert abs(ad - fd) <= 1e-6 * abs(fd) + 1e-14, (ad, fd)


def test_tape_bounds_are_checked():
    api = _api()
    run = _seeded_run(16, 100.0)
    tape = run.ctx.tape(2)
    with pytest.raises(api.AmressifError):
        run.ctx.step_fwd(3, tape)
    run.ctx.step_fwd(2, tape)
    with pytest.raises(api.AmressifError):
        run.ctx.step_bwd(tape, 3)
    with pytest.raises(api.AmressifError):
        run.ctx.step_tlm(tape, 1, 2)


@pytest.mark.parametrize("prec", [0, 1])
def test_linear_solves_leave_the_forward_phi_and_history_untouched(prec):
    """amressif_step_tlm / amressif_step_bwd run their projections' Poisson
    solves on their own iterates: store_state afterwards returns the forward
    phi, and later steps (whose warm start, in the mixed step, extrapolates
    from the phi history) are bit-identical to those of a context that ran no
    tangent or adjoint."""
    api = _api()
    n = 32
    rng = np.random.default_rng(31)
    runs = []
    for with_linear in (True, False):
        run = api.Run.from_numpy(api.cavity_inputs(n, 1000.0, precision=prec), W.initial_state(n, "seeded"))
        run.advance(3)
        tape = run.ctx.tape(2)
        run.ctx.step_fwd(2, tape)
        before = run.fetch()
        if with_linear:
            du, dv = _faces(rng, n)
            run.ctx.load_dual(du, dv, None, None, None, None)
            run.ctx.step_tlm(tape, 0, 2)
            run.ctx.load_dual(du, dv, None, None, None, None)
            run.ctx.step_bwd(tape, 2)
            after = run.fetch()
            for k in ("u", "v", "uprev", "vprev", "ctx"):
                assert np.array_equal(after[k], before[k]), k
        run.advance(3)
        runs.append(run.fetch())
    for k in ("u", "v", "ctx"):
        assert np.array_equal(runs[0][k], runs[1][k]), k
