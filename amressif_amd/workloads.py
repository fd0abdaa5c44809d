"""Synthetic inputs of the lid-driven-cavity workloads (host-side numpy).

These build the *inputs* of a run -- the t = 0 state and the optional lid
profile -- in the reference's single-box Array4 layout; they are not part of
the time step.  The reference itself only starts from rest
(Source/src1_rever/fn_init.cpp:111-112, 150-151, 202-213); the seeded
streamfunction start is this project's addition for the larger workloads.

Layout (numpy, C order; i fastest as in AMReX Array4):
  x-face arrays have shape (n, n+1), y-face arrays (n+1, n),
  userCtx (p, phi) has shape (2, n+2, n+2) (one ghost).
"""
from __future__ import annotations

import math

import numpy as np

SEED = 20261202

# The five workloads (unit square, phy_bc_lo = 135 135, phy_bc_hi = 135 165,
# lid inflow_waveform = 1 0, momentum_tolerance 1e-10, IterNum 100, cfl 0.5).
CONFIGS = {
    "cavity128_re100_zero": dict(n=128, ren=100.0, steps=200, init="zero", precision="fp64"),
    "cavity1024_re1000_seeded": dict(n=1024, ren=1000.0, steps=500, init="seeded", precision="fp64"),
    "cavity2048_re5000_mixed": dict(n=2048, ren=5000.0, steps=1000, init="seeded", precision="mixed"),
    "cavity512_re1000_adjoint": dict(n=512, ren=1000.0, steps=1000, init="seeded", precision="fp64",
                                     checkpoint_every=32, fd_directions=16),
    "cavity512_amr2_re2000_mixed": dict(n=512, ren=2000.0, steps=500, init="seeded", precision="mixed"),
}


def dt_from_cfl(n: int, cfl: float = 0.5, dx: float | None = None, dy: float | None = None,
                fixed_dt: float = -1.0) -> float:
    """dt = cfl / (2 * sum_d 1/dx_d^2) unless fixed_dt != -1 (main.cpp:284-294)."""
    dx = 1.0 / n if dx is None else dx
    dy = 1.0 / n if dy is None else dy
    coeff = 1.0 / (dx * dx) + 1.0 / (dy * dy)
    dt = cfl / (2.0 * coeff)
    if fixed_dt != -1.0:
        dt = fixed_dt
    return dt


def rk_coefficients(dt: float):
    """rk = dtau * {1/4, 1/3, 1/2, 1}, dtau = 0.4 dt (main.cpp:860-870)."""
    d_tau = 0.4 * dt
    return [d_tau * 0.25, d_tau * (1.0 / 3.0), d_tau * 0.5, d_tau * 1.0]


def streamfunction_coeffs(seed: int = SEED, kmax: int = 8, sigma: float = 0.1) -> np.ndarray:
    """a[kx-1, ky-1] ~ N(0, sigma^2), numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.normal(0.0, sigma, size=(kmax, kmax))


def seeded_velocity(n: int, seed: int = SEED, kmax: int = 8, sigma: float = 0.1):
    """Discretely divergence-free face velocities from
    psi = sum_{kx,ky<=kmax} a sin(kx pi x) sin(ky pi y) sin^2(pi x) sin^2(pi y),
    evaluated at nodes (psi = 0 on the boundary nodes) and differenced to faces:
    u = (psi(i,j+1) - psi(i,j))/dy, v = -(psi(i+1,j) - psi(i,j))/dx.
    Wall-normal faces come out exactly 0."""
    a = streamfunction_coeffs(seed, kmax, sigma)
    x = np.arange(n + 1) / n
    k = np.arange(1, kmax + 1)
    sx = np.sin(np.pi * np.outer(k, x))  # (kmax, n+1)
    env = np.sin(np.pi * x) ** 2
    # psi[j, i] = sum_{kx,ky} a[kx,ky] sx[kx,i] sy[ky,j] env[i] env[j]
    psi = (sx.T @ a @ sx).T * np.outer(env, env)  # (n+1 rows j, n+1 cols i)
    psi[0, :] = psi[-1, :] = psi[:, 0] = psi[:, -1] = 0.0
    h = 1.0 / n
    u = np.ascontiguousarray((psi[1:, :] - psi[:-1, :]) / h)  # (n, n+1): rows j, faces i
    v = np.ascontiguousarray(-(psi[:, 1:] - psi[:, :-1]) / h)  # (n+1, n)
    return u, v


def lid_profile(n: int, seed: int = SEED + 1, amp: float = 0.05, passes: int = 4) -> np.ndarray:
    """g = smooth(1 + amp * xi), xi ~ N(0,1), smoothed by `passes` passes of
    [1/4, 1/2, 1/4] with replicated ends."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = 1.0 + amp * rng.standard_normal(n)
    for _ in range(passes):
        gp = np.concatenate([g[:1], g, g[-1:]])
        g = 0.25 * gp[:-2] + 0.5 * gp[1:-1] + 0.25 * gp[2:]
    return np.ascontiguousarray(g)


def _psi_nodes(xs: np.ndarray, ys: np.ndarray, seed: int, kmax: int, sigma: float) -> np.ndarray:
    """psi[j, i] at the nodes (xs[i], ys[j]) of the seeded streamfunction."""
    a = streamfunction_coeffs(seed, kmax, sigma)
    k = np.arange(1, kmax + 1)
    sx = np.sin(np.pi * np.outer(k, xs))
    sy = np.sin(np.pi * np.outer(k, ys))
    psi = (sx.T @ a @ sy).T * np.outer(np.sin(np.pi * ys) ** 2, np.sin(np.pi * xs) ** 2)
    return psi


def seeded_two_level(n0: int, jc0: int, seed: int = SEED, kmax: int = 8, sigma: float = 0.1):
    """The seeded field on a coarse n0^2 grid and on a ratio-2 patch over the
    coarse rows [jc0, n0): both levels difference the same nodal psi, so the
    patch is discretely divergence-free, its wall-normal faces are 0 and every
    coarse face under the patch is the mean of the fine faces on it."""
    u0, v0 = seeded_velocity(n0, seed, kmax, sigma)
    nx, ny = 2 * n0, 2 * (n0 - jc0)
    h = 1.0 / nx
    xs = np.arange(nx + 1) * h
    ys = jc0 / n0 + np.arange(ny + 1) * h
    psi = _psi_nodes(xs, ys, seed, kmax, sigma)
    psi[:, 0] = psi[:, -1] = 0.0  # west / east walls
    psi[-1, :] = 0.0              # north wall
    u1 = np.ascontiguousarray((psi[1:, :] - psi[:-1, :]) / h)   # (ny, nx+1)
    v1 = np.ascontiguousarray(-(psi[:, 1:] - psi[:, :-1]) / h)  # (ny+1, nx)
    return {"u0": u0, "v0": v0, "u1": u1, "v1": v1}


def initial_state(n: int, init: str = "zero", seed: int = SEED):
    """t = 0 state: faces u, v, their previous-step copies (velContPrev =
    velCont, so the BDF2 history term is 0 at step 1) and userCtx = 0."""
    if init == "zero":
        u = np.zeros((n, n + 1))
        v = np.zeros((n + 1, n))
    elif init == "seeded":
        u, v = seeded_velocity(n, seed)
    else:
        raise ValueError(f"unknown init {init!r}")
    return {"u": u, "v": v, "uprev": u.copy(), "vprev": v.copy(), "ctx": np.zeros((2, n + 2, n + 2))}


def mcell_steps_per_s(n: int, steps: int, seconds: float) -> float:
    return n * n * steps / seconds / 1e6


def kinetic_energy_lower_left(U: np.ndarray, V: np.ndarray, h: float) -> float:
    """J = sum over cells in [0,1/2]^2 of (U^2+V^2)/2 dx dy (cell-centred U, V)."""
    n = U.shape[0]
    q = n // 2
    return float(0.5 * np.sum(U[:q, :q] ** 2 + V[:q, :q] ** 2) * h * h)


