"""GPU parity of the FGMRES Poisson solver (poisson_solver = 1): the same
Neumann operator and stopping test as the multigrid solve, checked against the
exact (DCT) oracle solve and, inside whole steps, against the CPU oracle."""
import numpy as np
import pytest

from amressif_amd import workloads as W

from test_gpu_parity import _api, check_step_parity, cuda, make_pair, rel, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,restart", [(8, 30), (64, 30), (100, 8), (256, 30), (256, 3)])
def test_gmres_poisson_solve_vs_exact(n, restart):
    rng = np.random.default_rng(19 + n)
    inp, o = make_pair(n, 100.0, poisson="dct", poisson_solver=1, gmres_restart=restart)
    ctx = _api().Context(inp)
    b = np.zeros((n + 2, n + 2))
    b[1:-1, 1:-1] = rng.standard_normal((n, n))
    sol_ref, _, _ = o.poisson_solve(b)
    sol, iters, res = ctx.poisson_advance(cuda(b), tol_rel=1e-12)
    assert res <= 1e-12
    assert 1 <= iters <= 60
    assert rel(sol.cpu().numpy()[1:-1, 1:-1], sol_ref[1:-1, 1:-1]) < 1e-9


def test_gmres_needs_fewer_preconditioner_calls_than_the_v_cycle_iteration():
    n = 256
    rng = np.random.default_rng(3)
    b = np.zeros((n + 2, n + 2))
    b[1:-1, 1:-1] = rng.standard_normal((n, n))
    its = {}
    for solver in (0, 1):
        inp, _ = make_pair(n, 100.0, poisson_solver=solver)
        _, it, _ = _api().Context(inp).poisson_advance(cuda(b), tol_rel=1e-10)
        its[solver] = it
    assert its[1] <= its[0]


def test_advance_with_gmres_matches_oracle():
    ref, rs, got, st = run_both(128, 400.0, 8, init="seeded", poisson_solver=1)
    check_step_parity(ref, rs, got, st)


def test_mixed_with_gmres_within_1e_6_of_fp64():
    api = _api()
    n, steps = 128, 10
    out = {}
    for prec in (api.AMRESSIF_FP64, api.AMRESSIF_MIXED):
        inp = api.cavity_inputs(n, 1000.0, precision=prec, poisson_solver=1)
        run = api.Run.from_numpy(inp, W.initial_state(n, "seeded"))
        st = run.advance(steps)
        assert all(s["flags"] == 0 and s["div_new"] <= 1e-10 * s["div_star"] + 1e-12 for s in st)
        out[prec] = run.fetch()
    a, b = out[api.AMRESSIF_MIXED], out[api.AMRESSIF_FP64]
    assert rel(a["u"], b["u"]) < 1e-6 and rel(a["v"], b["v"]) < 1e-6


def test_gmres_restart_bounds_are_checked():
    api = _api()
    with pytest.raises(api.AmressifError) as e:
        api.Context(api.cavity_inputs(32, 100.0, poisson_solver=1, gmres_restart=65))
    assert e.value.code == -1
