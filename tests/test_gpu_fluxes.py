"""Flux fields through the C-ABI (amressif_fluxes) and the flux plotfile
(amressif_write_flux_plotfile) against the CPU FP64 oracle.

* convective_flux_calc_new_quick (fn_flux_calc.cpp:11-202) and
  viscous_flux_calc (:695-766): fluxConvect, fluxViscous and the half-node
  fluxes fluxHalfN1 / fluxHalfN2, from cont2cart's cell velocities with the
  wall rules.  The kernel keeps the reference's operation order with no FMA
  contraction, so the bar is bit equality with the oracle's restatement.
* Export_Fluxes (utilities.cpp:496-532): pltFluxNNNNN holds conv_flux_x/y,
  visc_flux_x/y and press_grad_x/y (0: the reference never fills
  fluxPrsGrad) of the step's predictor u*, rebuilt here from the stored state
  as u* = u^{n+1} + (dt/1.5) G phi; that rebuild rounds, so the bar there is
  1e-10 relative.
"""
import os

import numpy as np
import pytest

from amressif_amd import workloads as W
from oracle.refcpu import Oracle

import amrex_files as AF

pytestmark = pytest.mark.gpu

BCS = [((135, 135), (135, 165)), ((-135, -135), (135, -135)), ((165, 135), (165, 165)), ((135, -135), (135, 135))]


def _api():
    from amressif_amd import api
    return api


def _cuda(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _oracle_fluxes(o, u, v):
    vc = o.cont2cart(u, v)
    fxu, fxv, fyu, fyv, fconv, _ = o.convective_flux(vc, u, v)
    fvisc, _ = o.viscous_flux(vc)
    return fconv, fvisc, (fxu, fyu, fxv, fyv)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("n", [64, 100])
def test_fluxes_bit_exact_with_oracle(n, bc):
    api = _api()
    rng = np.random.default_rng(n + 7 * len(str(bc)))
    u = rng.uniform(-1, 1, (n, n + 1))
    v = rng.uniform(-1, 1, (n + 1, n))
    inp = api.cavity_inputs(n, 400.0, phy_bc_lo=(bc[0][0], bc[0][1], 135), phy_bc_hi=(bc[1][0], bc[1][1], 135),
                            inflow_waveform=(0.75, -0.25, 0.0))
    ctx = api.Context(inp)
    o = Oracle(n, 400.0, api.dt_of(inp), bc_lo=bc[0], bc_hi=bc[1], inflow=(0.75, -0.25))
    fc, fv, half = ctx.fluxes(_cuda(u), _cuda(v))
    rc, rv, rhalf = _oracle_fluxes(o, u, v)
    print(f"n={n} bc={bc}: |fluxConvect| {np.abs(rc).max():.3e}, |fluxViscous| {np.abs(rv).max():.3e}")
    assert np.array_equal(fc.cpu().numpy(), rc)
    assert np.array_equal(fv.cpu().numpy(), rv)
    for g, r in zip(half, rhalf):
        assert np.array_equal(g.cpu().numpy(), r)


def test_fluxes_with_a_lid_profile():
    api = _api()
    import torch
    n = 64
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, (n, n + 1))
    v = rng.uniform(-1, 1, (n + 1, n))
    lid = 1.0 + 0.1 * np.sin(np.linspace(0, 3, n))
    inp = api.cavity_inputs(n, 1000.0)
    ctx = api.Context(inp)
    lid_t = _cuda(lid)
    ctx.set_lid_profile(lid_t)
    o = Oracle(n, 1000.0, api.dt_of(inp), lid=lid)
    fc, fv, half = ctx.fluxes(_cuda(u), _cuda(v))
    rc, rv, rhalf = _oracle_fluxes(o, u, v)
    torch.cuda.synchronize()
    assert np.array_equal(fc.cpu().numpy(), rc) and np.array_equal(fv.cpu().numpy(), rv)
    for g, r in zip(half, rhalf):
        assert np.array_equal(g.cpu().numpy(), r)


def test_flux_plotfile_holds_the_predictor_fluxes(tmp_path):
    api = _api()
    n = 64
    ren = 400.0
    inp = api.cavity_inputs(n, ren, max_grid_size=(32, 64, 32))
    ctx = api.Context(inp)
    ctx.init_state(W.streamfunction_coeffs())
    with pytest.raises(api.AmressifError):
        ctx.write_flux_plotfile(str(tmp_path), 0, 0.0)  # no step since the state was set
    ctx.advance(2)
    path = ctx.write_flux_plotfile(str(tmp_path), 2, 0.125)
    assert os.path.basename(path) == "pltFlux00002"
    h = AF.read_plotfile_header(path)
    assert h["names"] == ["conv_flux_x", "conv_flux_y", "visc_flux_x", "visc_flux_y", "press_grad_x",
                          "press_grad_y"]
    assert h["step"] == 2 and h["time"] == 0.125 and h["nboxes"] == 2
    _, parts = AF.read_multifab(os.path.join(path, "Level_0", "Cell"))
    g = AF.gather(parts, (n, n), 6)
    # the predictor from the stored state: u* = u^{n+1} + (dt/1.5) G phi
    import torch
    u = torch.zeros((n, n + 1), dtype=torch.float64, device="cuda")
    v = torch.zeros((n + 1, n), dtype=torch.float64, device="cuda")
    c = torch.zeros((2, n + 2, n + 2), dtype=torch.float64, device="cuda")
    ctx.store_state(u=u, v=v, ctx=c)
    u, v, phi = u.cpu().numpy(), v.cpu().numpy(), c[1].cpu().numpy()
    dt, h_ = api.dt_of(inp), 1.0 / n
    us = u + (dt / 1.5) * (phi[1:-1, 1:] - phi[1:-1, :-1]) / h_
    vs = v + (dt / 1.5) * (phi[1:, 1:-1] - phi[:-1, 1:-1]) / h_
    o = Oracle(n, ren, dt)
    rc, rv, _ = _oracle_fluxes(o, us, vs)
    for q, ref in enumerate((rc[0], rc[1], rv[0], rv[1])):
        err = np.abs(g[q] - ref).max() / np.abs(ref).max()
        print(f"pltFlux component {h['names'][q]}: rel err {err:.2e} against the rebuilt predictor")
        assert err <= 1e-10, h["names"][q]
    assert not np.any(g[4]) and not np.any(g[5])
