"""GPU checks of the two-level composite step (amressif_amr_*).

The reference is single-level, so there is no oracle for this path; it is
pinned by what the composite discretisation fixes:
  * the composite projection leaves the composite divergence at the solver
    tolerance, and every coarse face on or above the interface is the mean of
    the fine faces on it (flux matching);
  * a cavity at rest with a resting lid stays exactly at rest;
  * inside the refined region (away from the interface) the two-level result
    is an order of magnitude closer to a uniform run at the fine resolution
    than a uniform run at the coarse resolution is, and over the whole refined
    region it is no further;
  * the mixed form stays within 1e-6 of the FP64 form.
"""
import numpy as np
import pytest

from amressif_amd import workloads as W

pytestmark = pytest.mark.gpu


def _api():
    from amressif_amd import api
    return api


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _two_level(n0, jc0, ren, precision=0, **kw):
    api = _api()
    inp = api.cavity_inputs(n0, ren, precision=precision, **kw)
    return api.TwoLevel(inp, jc0, W.seeded_two_level(n0, jc0))


@pytest.mark.parametrize("n0,jc0", [(32, 24), (64, 48), (64, 40)])
def test_composite_projection_divergence_and_flux_matching(n0, jc0):
    tl = _two_level(n0, jc0, 2000.0)
    st = tl.advance(4)
    for s in st:
        assert s["flags"] == 0, s
        assert s["poisson_res"] <= 1e-10
        assert s["div_new"] <= 1e-9 * s["div_star"] + 1e-10, s
        assert 1 <= s["poisson_iters"] <= 60
    f = tl.fetch()
    u0, v0, u1, v1 = f["u0"], f["v0"], f["u1"], f["v1"]
    # coarse faces on and above the interface are the means of the fine faces
    assert np.array_equal(v0[jc0:, :], 0.5 * (v1[0::2, 0::2] + v1[0::2, 1::2]))
    assert np.array_equal(u0[jc0:, :], 0.5 * (u1[0::2, 0::2] + u1[1::2, 0::2]))
    # walls stay closed
    assert np.all(u1[:, [0, -1]] == 0.0) and np.all(v1[-1, :] == 0.0)
    assert np.all(u0[:, [0, -1]] == 0.0) and np.all(v0[[0, -1], :] == 0.0)


def test_rest_state_stays_at_rest():
    api = _api()
    n0, jc0 = 32, 24
    inp = api.cavity_inputs(n0, 100.0, inflow_waveform=(0.0, 0.0, 0.0))
    z = {"u0": np.zeros((n0, n0 + 1)), "v0": np.zeros((n0 + 1, n0)),
         "u1": np.zeros((2 * (n0 - jc0), 2 * n0 + 1)), "v1": np.zeros((2 * (n0 - jc0) + 1, 2 * n0))}
    tl = api.TwoLevel(inp, jc0, z)
    tl.advance(3)
    f = tl.fetch()
    for k in ("u0", "v0", "u1", "v1"):
        assert np.all(f[k] == 0.0), k


def _coarsen_faces(u, v):
    """uniform fine faces -> coarse faces (means of the fine faces on them)"""
    return 0.5 * (u[0::2, 0::2] + u[1::2, 0::2]), 0.5 * (v[0::2, 0::2] + v[0::2, 1::2])


def test_refined_region_is_closer_to_the_fine_uniform_run():
    api = _api()
    n0, jc0, ren, steps = 64, 48, 100.0, 10
    dt = 0.5 / (4.0 * (2 * n0) ** 2)  # the fine level's dt, used by all three runs
    kw = dict(fixed_dt=dt)
    tl = _two_level(n0, jc0, ren, **kw)
    tl.advance(steps)
    amr = tl.fetch()
    fine = api.Run.from_numpy(api.cavity_inputs(2 * n0, ren, **kw), W.initial_state(2 * n0, "seeded"))
    fine.advance(steps)
    f = fine.fetch()
    coarse = api.Run.from_numpy(api.cavity_inputs(n0, ren, **kw), W.initial_state(n0, "seeded"))
    coarse.advance(steps)
    c = coarse.fetch()
    # compare on the coarse faces of the refined region: away from the
    # interface (whose faces come from the coarse predictor) the patch carries
    # the fine run's accuracy; over the whole region it is no worse
    uf, vf = _coarsen_faces(f["u"], f["v"])

    def err(u, v, rows):
        return max(np.abs(u[rows] - uf[rows]).max(), np.abs(v[rows] - vf[rows]).max())
    inner, region = slice(jc0 + 4, None), slice(jc0, None)
    e_amr, e_crs = err(amr["u0"], amr["v0"], inner), err(c["u"], c["v"], inner)
    assert e_amr < 0.1 * e_crs, (e_amr, e_crs)
    e_amr, e_crs = err(amr["u0"], amr["v0"], region), err(c["u"], c["v"], region)
    assert e_amr <= e_crs, (e_amr, e_crs)


def test_mixed_two_level_within_1e_6_of_fp64():
    api = _api()
    n0, jc0, steps = 64, 48, 5
    a = _two_level(n0, jc0, 2000.0, precision=api.AMRESSIF_MIXED)
    b = _two_level(n0, jc0, 2000.0, precision=api.AMRESSIF_FP64)
    sa, sb = a.advance(steps), b.advance(steps)
    fa, fb = a.fetch(), b.fetch()
    for k in ("u0", "v0", "u1", "v1"):
        assert rel(fa[k], fb[k]) < 1e-6, k
    assert all(s["div_new"] <= 1e-9 * s["div_star"] + 1e-10 for s in sa)


def test_patch_bounds_are_checked():
    api = _api()
    with pytest.raises(api.AmressifError):
        _two_level(32, 31, 100.0)
    with pytest.raises(api.AmressifError):
        _two_level(32, 1, 100.0)


@pytest.mark.parametrize("prec", [0, 1])
def test_device_composite_fgmres_matches_the_host_driven_one(prec):
    """The composite FGMRES with its Arnoldi process on the device (one host
    wait per batch of steps) takes the same steps as the host-driven form
    (AMRESSIF_AMR_HOST_GMRES=1, one wait per dot-product pass) and gives the
    same fields to rounding."""
    import os
    api = _api()
    n0, jc0, steps = 64, 48, 4
    out = {}
    for host in ("1", "0"):
        old = os.environ.get("AMRESSIF_AMR_HOST_GMRES")
        os.environ["AMRESSIF_AMR_HOST_GMRES"] = host
        try:
            tl = _two_level(n0, jc0, 2000.0, precision=prec)
        finally:
            if old is None:
                os.environ.pop("AMRESSIF_AMR_HOST_GMRES")
            else:
                os.environ["AMRESSIF_AMR_HOST_GMRES"] = old
        st = tl.advance(steps)
        out[host] = (tl.fetch(), st)
    (fh, sh), (fd, sd) = out["1"], out["0"]
    assert [s["poisson_iters"] for s in sd] == [s["poisson_iters"] for s in sh]
    for s in sd:
        assert s["flags"] == 0 and s["poisson_res"] <= 1e-10 and s["div_new"] <= 1e-9 * s["div_star"] + 1e-10
    for k in ("u0", "v0", "u1", "v1"):
        assert rel(fd[k], fh[k]) < 1e-9, k
