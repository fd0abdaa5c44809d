"""Initial state, checkpoint/restart and plotfile entry points through the
C-ABI (amressif_init_state, amressif_write_checkpoint /
amressif_read_checkpoint, amressif_write_plotfile).

* hybrid_grid_init (fn_init.cpp:30-631): the cavity at rest, and the seeded
  streamfunction start of the larger workloads against the numpy recipe;
* SaveCheckpoint / LoadCheckpoint (utilities.cpp:792-1022) with the restart of
  main.cpp:386-800: advancing N steps equals advancing k, saving, loading into
  a fresh context and advancing N - k, bit for bit (FP64); the files follow
  AMReX's checkpoint layout, read back here by an independent numpy reader.
  Byte-level identity with an AMReX build's files is not checked (AMReX is not
  available): parity unpinned for the bytes, pinned for the values.
* Export_Flow_Field (utilities.cpp:534-570): the plotfile holds p and the cell
  averages of the faces.
"""
import os

import numpy as np
import pytest

from amressif_amd import workloads as W

import amrex_files as AF

pytestmark = pytest.mark.gpu


def _api():
    from amressif_amd import api
    return api


def _fetch(ctx, n):
    import torch
    t = lambda *s: torch.zeros(s, dtype=torch.float64, device="cuda")  # noqa: E731
    u, v, up, vp, uc = t(n, n + 1), t(n + 1, n), t(n, n + 1), t(n + 1, n), t(2, n + 2, n + 2)
    ctx.store_state(u, v, up, vp, uc)
    return {k: a.cpu().numpy() for k, a in (("u", u), ("v", v), ("uprev", up), ("vprev", vp), ("ctx", uc))}


def test_init_state_rest_and_seeded():
    api = _api()
    n = 96
    ctx = api.Context(api.cavity_inputs(n, 100.0))
    ctx.init_state()
    st = _fetch(ctx, n)
    assert all(np.all(st[k] == 0.0) for k in st)
    ctx.init_state(W.streamfunction_coeffs())
    st = _fetch(ctx, n)
    u, v = W.seeded_velocity(n)
    scale = max(np.abs(u).max(), np.abs(v).max())
    assert np.abs(st["u"] - u).max() <= 1e-13 * scale and np.abs(st["v"] - v).max() <= 1e-13 * scale
    assert np.array_equal(st["uprev"], st["u"]) and np.array_equal(st["vprev"], st["v"])
    assert np.all(st["u"][:, [0, -1]] == 0.0) and np.all(st["v"][[0, -1], :] == 0.0)
    div = np.diff(st["u"], axis=1) * n + np.diff(st["v"], axis=0) * n
    assert np.abs(div).max() <= 1e-12 * scale * n
    assert np.all(st["ctx"] == 0.0)


@pytest.mark.parametrize("prec", [0, 1])
def test_checkpoint_restart(prec, tmp_path):
    api = _api()
    n, k, rest = 128, 5, 6
    inp = api.cavity_inputs(n, 1000.0, precision=prec, max_grid_size=(64, 32, 32))
    dt = api.dt_of(inp)
    a = api.Context(inp)
    a.init_state(W.streamfunction_coeffs())
    a.advance(k)
    path = a.write_checkpoint(str(tmp_path), k, k * dt)
    saved = _fetch(a, n)
    a.advance(rest)
    ref = _fetch(a, n)

    # the files: AMReX checkpoint layout, values equal to the state
    h = AF.read_checkpoint_header(path)
    assert h["title"] == "Checkpoint file for AMRESSIF" and h["time"] == pytest.approx(k * dt, rel=1e-14)
    assert sorted((b[0], b[1], b[2], b[3]) for b in h["boxes"]) == sorted(api.box_array(n, (64, 32)))
    for name, key, shape, nodal in (("pressure", None, (n, n), None), ("vel_xCont", "u", (n, n + 1), 0),
                                    ("vel_yCont", "v", (n + 1, n), 1), ("vel_xContPrev", "uprev", (n, n + 1), 0),
                                    ("vel_yContPrev", "vprev", (n + 1, n), 1)):
        hdr, parts = AF.read_multifab(os.path.join(path, "Level_0", name))
        assert hdr["ncomp"] == 1 and hdr["ngrow"] == 0 and len(parts) == 8
        if nodal is not None:
            assert all(b[4 + nodal] == 1 for b in hdr["boxes"])
        g = AF.gather(parts, shape)[0]
        want = saved["ctx"][0, 1:-1, 1:-1] if key is None else saved[key]
        assert np.array_equal(g, want), name
        assert np.allclose(hdr["minmax"][0][:, 0], [p[1].min() for p in parts], rtol=0, atol=0)

    # restart into a fresh context
    b = api.Context(inp)
    t = b.read_checkpoint(path)
    assert t == pytest.approx(k * dt, rel=1e-14)
    back = _fetch(b, n)
    for key in ("u", "v", "uprev", "vprev"):
        assert np.array_equal(back[key], saved[key]), key
    assert np.array_equal(back["ctx"][0], saved["ctx"][0])
    b.advance(rest)
    got = _fetch(b, n)
    if prec == 0:
        # phi is not saved and the FP64 step does not read it: bit-identical
        for key in ("u", "v", "uprev", "vprev"):
            assert np.array_equal(got[key], ref[key]), key
        assert np.array_equal(got["ctx"][0], ref["ctx"][0])
    else:
        # the mixed step's warm start loses its phi history at a load: the
        # first solve starts from 0 and stops at a different point within
        # its tolerance
        for key in ("u", "v"):
            assert np.abs(got[key] - ref[key]).max() <= 1e-8 * np.abs(ref[key]).max(), key


def test_read_checkpoint_rejects_missing_and_bad_files(tmp_path):
    api = _api()
    n = 32
    ctx = api.Context(api.cavity_inputs(n, 100.0))
    with pytest.raises(api.AmressifError):
        ctx.read_checkpoint(str(tmp_path / "checkpoint00001"))
    path = ctx.write_checkpoint(str(tmp_path), 3, 0.5)
    other = api.Context(api.cavity_inputs(2 * n, 100.0))
    with pytest.raises(api.AmressifError):
        other.read_checkpoint(path)  # boxes of a 32^2 domain do not tile 64^2
    with open(os.path.join(path, "Level_0", "vel_xCont_H"), "w") as f:
        f.write("garbage\n")
    with pytest.raises(api.AmressifError):
        ctx.read_checkpoint(path)


def test_plotfile_holds_pressure_and_cell_velocities(tmp_path):
    api = _api()
    n = 64
    inp = api.cavity_inputs(n, 400.0, max_grid_size=(32, 64, 32))
    ctx = api.Context(inp)
    ctx.init_state(W.streamfunction_coeffs())
    ctx.advance(3)
    st = _fetch(ctx, n)
    path = ctx.write_plotfile(str(tmp_path), "pltResults", 3, 0.25)
    assert os.path.basename(path) == "pltResults00003"
    h = AF.read_plotfile_header(path)
    assert h["version"] == "HyperCLaw-V1.1" and h["names"] == ["pressure", "U", "V"] and h["dim"] == 2
    assert h["time"] == 0.25 and h["step"] == 3 and h["finest_level"] == 0 and h["coord"] == 0
    assert h["prob_lo"] == [0.0, 0.0] and h["prob_hi"] == [1.0, 1.0] and h["dx"] == [1.0 / n, 1.0 / n]
    assert h["domain"] == f"((0,0) ({n - 1},{n - 1}) (0,0))" and h["nboxes"] == 2 and h["mf_path"] == "Level_0/Cell"
    hdr, parts = AF.read_multifab(os.path.join(path, "Level_0", "Cell"))
    g = AF.gather(parts, (n, n), 3)
    assert np.array_equal(g[0], st["ctx"][0, 1:-1, 1:-1])
    assert np.array_equal(g[1], 0.5 * (st["u"][:, :-1] + st["u"][:, 1:]))
    assert np.array_equal(g[2], 0.5 * (st["v"][:-1, :] + st["v"][1:, :]))


INPUTS = """nsteps = 6
n_cell = 32
max_grid_size_x = 32
max_grid_size_y = 16
max_grid_size_z = 16
plot_int = 3
chk_int = 3
PSEUDO_TIMESTEPPING = 1
momentum_tolerance = 1.e-10
cfl = 0.5
IterNum = 100
ren = 100
vis = 1
phy_bc_lo = 135 135 135
phy_bc_hi = 135 165 135
inflow_waveform = 1.0 0.0 0.0
"""


def test_driver_program_checkpoint_cadence_and_restart(tmp_path):
    """amressif_run (main_main on the C-ABI): from rest, checkpoints and
    plotfiles every 3 steps; a restart from checkpoint00003 (chk_out = 3)
    reproduces checkpoint00006 bit for bit; the run equals the same chunks of
    amressif_advance through the Python binding."""
    import shutil
    import subprocess
    import sys
    api = _api()
    exe = os.path.join(os.path.dirname(api.LIB_PATH), "amressif_run")
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    (a / "inputs").write_text(INPUTS)
    r = subprocess.run([exe, "inputs"], cwd=a, capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    for d in ("checkpoint00003", "checkpoint00006", "pltInit00000", "pltResults00003", "pltResults00006",
              "pltFlux00003", "pltFlux00006"):
        assert (a / d / "Header").exists(), d
    assert r.stdout.count("STEP| ") == 6
    shutil.copytree(a / "checkpoint00003", b / "checkpoint00003")
    (b / "inputs").write_text(INPUTS)
    r = subprocess.run([exe, "inputs", "chk_out=3"], cwd=b, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("STEP| ") == 3 and "restart from checkpoint00003" in r.stdout
    for name in ("pressure", "vel_xCont", "vel_yCont", "vel_xContPrev", "vel_yContPrev"):
        da = (a / "checkpoint00006" / "Level_0" / f"{name}_D_00000").read_bytes()
        db = (b / "checkpoint00006" / "Level_0" / f"{name}_D_00000").read_bytes()
        assert da == db, name
    ha, hb = AF.read_checkpoint_header(str(a / "checkpoint00006")), AF.read_checkpoint_header(str(b / "checkpoint00006"))
    assert ha["time"] == pytest.approx(hb["time"], rel=1e-12) and ha["boxes"] == hb["boxes"]
    # the same chunks through the binding
    n = 32
    ctx = api.Context(api.parse_inputs(str(a / "inputs")))
    ctx.init_state()
    ctx.advance(3)
    ctx.advance(3)
    st = _fetch(ctx, n)
    _, parts = AF.read_multifab(str(a / "checkpoint00006" / "Level_0" / "vel_xCont"))
    assert np.array_equal(AF.gather(parts, (n, n + 1))[0], st["u"])
