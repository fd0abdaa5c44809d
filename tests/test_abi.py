"""The C-ABI library loads and exports every entry point include/amressif.h
declares, and its host-only logic (inputs parsing, dt, argument validation)
behaves -- CPU only, no device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "amressif.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(amressif_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def api():
    from amressif_amd import api
    return api


def test_every_declared_symbol_is_exported(api):
    names = declared_functions()
    assert len(names) >= 18
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_parse_inputs_file(api):
    inp = api.parse_inputs(os.path.join(GOLDEN, "inputs_cavity128"))
    assert inp.n_cell == 128 and inp.IterNum == 100 and inp.nsteps == 200
    assert tuple(inp.phy_bc_lo)[:2] == (135, 135) and tuple(inp.phy_bc_hi)[:2] == (135, 165)
    assert inp.ren == 100.0 and inp.cfl == 0.5 and inp.fixed_dt == -1.0
    assert inp.inflow_waveform[0] == 1.0 and inp.momentum_tolerance == 1e-10
    assert inp.PSEUDO_TIMESTEPPING == 1 and inp.precision == 0
    assert api.dt_of(inp) == 0.5 / (4 * 128 ** 2)


def test_parse_overrides_and_defaults(api):
    inp = api.parse_inputs(os.path.join(GOLDEN, "inputs_cavity128"), ["ren=1000", "amressif.precision = 1"])
    assert inp.ren == 1000.0 and inp.precision == 1
    inp = api.parse_inputs(None, ["n_cell=64", "max_grid_size_x=64", "max_grid_size_y=64", "max_grid_size_z=64",
                                  "IterNum=5", "ren=10", "vis=0.1"])
    assert inp.cfl == 0.9 and inp.nsteps == 1 and inp.fixed_dt == -1.0  # main.cpp:89-157 defaults
    assert inp.poisson_tol_rel == 1e-10 and inp.mg_max_iter == 2000    # poisson.cpp:159-182


def test_parse_errors(api, tmp_path):
    bad = tmp_path / "inputs"
    bad.write_text("n_cell = 32\nchk_int = = 10\n")  # the malformed line of the reference's Exec/inputs
    with pytest.raises(api.AmressifError) as e:
        api.parse_inputs(str(bad))
    assert e.value.code == -6
    with pytest.raises(api.AmressifError, match="missing required key n_cell"):
        api.parse_inputs(None, ["ren=1"])


def _create_status(api, **kw):
    inp = api.cavity_inputs(32, 100.0, **kw)
    h = ctypes.c_void_p()
    rc = api._lib.amressif_create(ctypes.byref(inp), None, ctypes.byref(h))
    msg = api._lib.amressif_last_error(h).decode() if h.value else ""
    api._lib.amressif_destroy(h)
    return rc, msg


def test_create_rejects_codes_the_reference_aborts_on(api):
    # validation runs before any device allocation
    assert _create_status(api, phy_bc_lo=(0, 135, 135))[0] == -2      # main.cpp:822 Invalid bc_lo
    assert _create_status(api, phy_bc_lo=(195, 135, 135))[0] == -2    # utilities.cpp:128 Not yet implemented
    assert _create_status(api, phy_bc_hi=(135, -165, 135))[0] == -2
    # periodic needs 111 on both sides of a direction (is_periodic, main.cpp:159-166)
    rc, msg = _create_status(api, phy_bc_lo=(111, 135, 135), phy_bc_hi=(135, 165, 135))
    assert rc == -3 and "periodic" in msg
    assert _create_status(api, spacedim=4)[0] == -1
    # the general path (3D, periodic) is FP64 with the multigrid solve
    assert _create_status(api, spacedim=3, precision=1)[0] == -3
    assert _create_status(api, phy_bc_lo=(111, 135, 135), phy_bc_hi=(111, 165, 135), poisson_solver=1)[0] == -3
    assert _create_status(api, spacedim=3, phy_bc_lo=(135, 135, -165))[0] == -2  # utilities.cpp:296 Not yet implemented
    assert _create_status(api, PSEUDO_TIMESTEPPING=2)[0] == -1     # only 0 (explicit) and 1 (dual time)
    assert _create_status(api, precision=7)[0] == -1


def test_box_array_mirrors_max_size_chopping(api):
    # the reference's Exec/inputs: n_cell 32, max_grid_size_x 32, max_grid_size_y 16
    assert api.box_array(32, (32, 16)) == [(0, 0, 31, 15), (0, 16, 31, 31)]
    # uneven cuts: the first pieces are one cell longer
    assert api.box_array(10, (4, 10)) == [(0, 0, 3, 9), (4, 0, 6, 9), (7, 0, 9, 9)]
    for n, m in ((48, (20, 32)), (100, 32), (64, 64)):
        boxes = api.box_array(n, m)
        cover = np.zeros((n, n), dtype=int)
        for x0, y0, x1, y1 in boxes:
            cover[y0:y1 + 1, x0:x1 + 1] += 1
            mx, my = (m, m) if isinstance(m, int) else m
            assert x1 - x0 + 1 <= mx and y1 - y0 + 1 <= my
        assert np.all(cover == 1)
