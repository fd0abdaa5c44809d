"""GPU checks of the multigrid smoothers.

The wave-column smoothers (the column window in registers, x-neighbours by DPP
lane shifts; one or two columns per lane) and the LDS tile smoother run the
same red-black Gauss-Seidel half-sweeps, the same fused residual + restriction
and the same fused FP64 outer step.  So after a fixed number of V-cycles a Poisson solve with either
has the same residual and the same iterate up to rounding, for every window
height (output rows per wave).  Sizes cover partial last bands and row chunks
(100, 200 and 400 are not multiples of the band widths 54 and 56, and 100 and
200 not of every row chunk) as well as the 2048^2 workload's hierarchy.
"""
import os

import numpy as np
import pytest

from test_gpu_parity import _api, cuda, rel

pytestmark = pytest.mark.gpu

ENV_KEYS = ("AMRESSIF_MG_SMOOTHER", "AMRESSIF_MG_ROWS32", "AMRESSIF_MG_ROWS64", "AMRESSIF_MG_ROWS_COARSE",
            "AMRESSIF_TAIL_DCT", "AMRESSIF_TAIL_EXT", "AMRESSIF_DCT_MFMA64",
            "AMRESSIF_START_ROWS", "AMRESSIF_MG_UP64", "AMRESSIF_TAIL_WAVES", "AMRESSIF_MG_PRE3", "AMRESSIF_MG_POST3")


def _solve(n, precision, b, cycles, tol=0.0, **env):
    """cycles V-cycles of the step's Poisson iteration from zero (tolerance tol,
    cycle cap), with the smoother settings in env; returns (phi, ||r||/||b||)"""
    api = _api()
    saved = {k: os.environ.get(k) for k in ENV_KEYS}
    for k in ENV_KEYS:
        os.environ.pop(k, None)
    os.environ.update(env)
    try:
        ctx = api.Context(api.cavity_inputs(n, 100.0, precision=precision, mg_max_iter=cycles))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    sol, it, res = ctx.poisson_advance(cuda(b), tol_rel=tol)
    assert it == cycles if tol == 0.0 else it <= cycles
    return sol.cpu().numpy()[1:-1, 1:-1], res


def _rhs(n, seed):
    rng = np.random.default_rng(seed)
    b = np.zeros((n + 2, n + 2))
    b[1:-1, 1:-1] = rng.standard_normal((n, n))
    return b


@pytest.mark.parametrize("smoother", ["1", "2"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [100, 200, 400, 1024])
def test_wave_column_smoother_matches_tile_smoother(n, precision, smoother):
    # the same cycle in every kernel: two sweeps per leg (the paired kernel's
    # three-sweep finest FP32 legs are off)
    b = _rhs(n, 11 + n)
    two = dict(AMRESSIF_MG_PRE3="0", AMRESSIF_MG_POST3="0")
    ref, res_ref = _solve(n, precision, b, 3, AMRESSIF_MG_SMOOTHER="0", **two)
    got, res = _solve(n, precision, b, 3, AMRESSIF_MG_SMOOTHER=smoother, **two)
    # three V-cycles reduce the residual by orders of magnitude; the two
    # kernels differ only in rounding: ~1e-16 (FP64) or ~1e-7 (FP32 V-cycle)
    # of the correction per cycle, and ||r|| / ||b|| by as much
    assert res_ref < 1e-2
    tol = 1e-10 if precision == 0 else 1e-5
    assert abs(res - res_ref) <= (1e-12 if precision == 0 else 1e-5), (res, res_ref)
    assert rel(got, ref) < tol


@pytest.mark.parametrize("smoother", ["1", "2"])
@pytest.mark.parametrize("rows", ["4,4,4", "8,8,8", "16,16,16", "32,32,8", "8,32,16"])
def test_wave_column_window_heights_agree(rows, smoother):
    n = 200
    b = _rhs(n, 5)
    ref, res_ref = _solve(n, 0, b, 3, AMRESSIF_MG_SMOOTHER="0")
    for coarse in ("4", "8", "32"):
        got, res = _solve(n, 0, b, 3, AMRESSIF_MG_SMOOTHER=smoother, AMRESSIF_MG_ROWS64=rows,
                          AMRESSIF_MG_ROWS_COARSE=coarse)
        assert abs(res - res_ref) <= 1e-12, (rows, coarse)
        assert rel(got, ref) < 1e-10, (rows, coarse)


@pytest.mark.parametrize("smoother", ["1", "2"])
def test_wave_column_smoother_converges_at_2048(smoother):
    n = 2048
    b = _rhs(n, 3)
    for precision in (0, 1):
        _, res = _solve(n, precision, b, 6, AMRESSIF_MG_SMOOTHER=smoother)
        assert res < 1e-6, (precision, res)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [64, 100, 256, 1024])
def test_exact_tail_solve_matches_tail_v_cycle(n, precision):
    """The levels of at most 64^2 cells solved exactly (DCT) instead of by the
    V-cycle below them: the same converged solution, in no more V-cycles, and
    each cycle reduces the residual at least as much."""
    b = _rhs(n, 23 + n)
    ref, res_ref = _solve(n, precision, b, 60, tol=1e-10, AMRESSIF_TAIL_DCT="0")
    got, res = _solve(n, precision, b, 60, tol=1e-10, AMRESSIF_TAIL_DCT="1")
    assert res <= 1e-10 and res_ref <= 1e-10
    assert rel(got, ref) < 1e-8
    _, r3_ref = _solve(n, precision, b, 3, AMRESSIF_TAIL_DCT="0")
    _, r3 = _solve(n, precision, b, 3, AMRESSIF_TAIL_DCT="1")
    assert r3 <= 1.5 * r3_ref, (r3, r3_ref)


@pytest.mark.parametrize("n", [256, 512, 1024, 2048])
@pytest.mark.parametrize("waves", ["4", "16"])
def test_tail_kernel_with_the_level_above_matches_separate_launches(n, waves):
    """FP32 V-cycle: the 128^2 level smoothed inside the tail kernel (rows
    traded between its 4 or 16 waves through LDS) gives the same cycles as its
    two smoother launches and the DCT-only tail kernel of the same wave count,
    up to FP32 rounding."""
    b = _rhs(n, 31 + n)
    ref, res_ref = _solve(n, 1, b, 3, AMRESSIF_TAIL_EXT="0", AMRESSIF_TAIL_WAVES=waves)
    got, res = _solve(n, 1, b, 3, AMRESSIF_TAIL_EXT="1", AMRESSIF_TAIL_WAVES=waves)
    assert res_ref < 1e-2
    assert abs(res - res_ref) <= 1e-5, (res, res_ref)
    assert rel(got, ref) < 1e-5


@pytest.mark.parametrize("n", [64, 100, 256, 1024])
@pytest.mark.parametrize("waves", ["4", "16"])
def test_fp64_dct_on_matrix_cores_matches_valu(n, waves):
    """FP64 exact solve of the small levels: the products on the matrix cores
    (v_mfma_f64_16x16x4f64; one 16 x 16 output block per wave with 16 waves,
    four with 4) and on the VALU give the same cycles up to FP64 rounding (a
    wrong operand or result layout would not)."""
    b = _rhs(n, 41 + n)
    ref, res_ref = _solve(n, 0, b, 3, AMRESSIF_DCT_MFMA64="0")
    got, res = _solve(n, 0, b, 3, AMRESSIF_DCT_MFMA64="1", AMRESSIF_TAIL_WAVES=waves)
    assert abs(res - res_ref) <= 1e-9 * max(res_ref, 1e-300) + 1e-14, (res, res_ref)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("n", [8, 100, 256, 1024])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("rows", ["4", "16"])
def test_row_march_poisson_start_matches_cell_tile_kernel(n, precision, rows):
    """The Poisson start (centred b, first residual, partials) in the row-march
    layout gives the cell-tile kernel's cycles: the same per-cell arithmetic;
    only the order of the two partial sums differs."""
    b = _rhs(n, 53 + n)
    ref, res_ref = _solve(n, precision, b, 3, AMRESSIF_START_ROWS="0")
    got, res = _solve(n, precision, b, 3, AMRESSIF_START_ROWS=rows)
    assert abs(res - res_ref) <= 1e-12 * max(res_ref, 1e-300), (res, res_ref)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("n", [100, 256, 1024])
def test_fp64_up_leg_kernels_agree(n):
    """FP64 finest up leg on the one-column kernel (the default) or the paired
    one: the same cycles up to FP64 rounding."""
    b = _rhs(n, 61 + n)
    ref, res_ref = _solve(n, 0, b, 3, AMRESSIF_MG_UP64="0")
    got, res = _solve(n, 0, b, 3, AMRESSIF_MG_UP64="1")
    assert abs(res - res_ref) <= 1e-9 * max(res_ref, 1e-300) + 1e-14, (res, res_ref)
    assert rel(got, ref) < 1e-12
