"""Pins of the general-path oracle (oracle/refgen.c: 2D and 3D, wall and
periodic directions), on the CPU.

What fixes each part without the reference (which needs AMReX):
* 2D walls: equal to the 2D oracle (oracle/refcpu.c, pinned in
  test_oracle.py) function by function and over whole steps;
* the third direction: a z-independent 3D state with z periodic steps as the
  2D state does, and a cyclic permutation of the axes commutes with the 3D
  step (a slip in any z term breaks one of the two);
* the 3D ghost rules: (ghost + mirror)/2 is the wall value, per side and code
  (utilities.cpp:87-346);
* periodic directions: translation along them commutes with the step, and the
  Poisson solve equals the exact FFT/DCT solution;
* the whole periodic step: Taylor-Green (2D, utilities.cpp:572-643) and the
  Arnold-Beltrami-Childress flow (3D, an exact Navier-Stokes solution whose
  nonlinear term is a gradient) decay at second order in the cell size.
"""
import numpy as np
import pytest
import scipy.fft as sfft

from amressif_amd import workloads as W
from oracle.refcpu import Oracle
from oracle.refgen import GenOracle


def _seeded2d(n):
    s = W.initial_state(n, "seeded")
    return s


def _rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


# ------------------------------------------------------------ 2D = refcpu
WALLS = [((135, 135), (135, 165)), ((-135, -135), (135, -135)), ((165, 135), (165, 165)), ((135, -135), (1, 135))]


@pytest.mark.parametrize("bc", WALLS)
def test_modules_equal_the_2d_oracle(bc):
    n = 24
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (n, n + 1))
    v = rng.uniform(-1, 1, (n + 1, n))
    ctx = np.zeros((2, n + 2, n + 2))
    ctx[0, 1:-1, 1:-1] = rng.uniform(-1, 1, (n, n))
    o2 = Oracle(n, 300.0, 1e-4, bc_lo=bc[0], bc_hi=bc[1], inflow=(0.7, -0.3))
    og = GenOracle(2, n, 300.0, 1e-4, bc_lo=bc[0] + (135,), bc_hi=bc[1] + (135,), inflow=(0.7, -0.3, 0.0))
    import ctypes
    from oracle.refcpu import _p as p2
    o2.L.rc_userctx_ghosts(ctypes.byref(o2.prm), p2(ctx))
    ctxg = ctx.copy()
    og.userctx_ghosts(ctxg)
    assert np.array_equal(ctx, ctxg)
    vc2 = o2.cont2cart(u, v)
    vcg = og.cont2cart([u, v])
    # corners: the 2D oracle writes them in the reference's W, E, S, N order too
    assert np.array_equal(vc2, vcg)
    fxu, fxv, fyu, fyv, fconv, ftot = o2.convective_flux(vc2, u, v)
    fh, fconvg, ftotg = og.convective_flux(vcg, [u, v])
    for a, b in zip((fxu, fxv, fyu, fyv), fh):
        assert np.array_equal(a, b)
    assert np.array_equal(fconv, fconvg) and np.array_equal(ftot, ftotg)
    fv2, _ = o2.viscous_flux(vc2)
    fvg, _ = og.viscous_flux(vcg)
    assert np.array_equal(fv2, fvg)
    rx, ry = o2.stage_rhs(u, v, ctx)
    rg = og.stage_rhs([u, v], ctx)
    assert np.array_equal(rx, rg[0]) and np.array_equal(ry, rg[1])
    uk = [u + 1e-3 * rng.standard_normal(u.shape), v + 1e-3 * rng.standard_normal(v.shape)]
    assert og.error_computation(uk, [u, v]) == pytest.approx(o2.error_computation(uk[0], uk[1], u, v), rel=1e-14)
    b2, m2, s2 = o2.poisson_rhs(u, v)
    bg, mg, sg = og.poisson_rhs([u, v])
    assert np.array_equal(b2, bg) and m2 == mg and s2 == pytest.approx(sg, rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("pseudo", [1, 0])
def test_steps_equal_the_2d_oracle(pseudo):
    n, ren, steps = 32, 400.0, 3
    s = _seeded2d(n)
    dt = 0.5 / (4 * n * n)
    o2 = Oracle(n, ren, dt, poisson_tol_rel=1e-13, pseudo=pseudo)
    og = GenOracle(2, n, ren, dt, pseudo=pseudo)
    a = {k: v.copy() for k, v in s.items()}
    st2 = o2.advance(a, steps)
    g = {"u": [s["u"].copy(), s["v"].copy()], "uprev": [s["uprev"].copy(), s["vprev"].copy()], "ctx": s["ctx"].copy()}
    stg = og.advance(g, steps)
    assert [x["pseudo_iters"] for x in st2] == [x["pseudo_iters"] for x in stg]
    assert _rel(g["u"][0], a["u"]) < 1e-12 and _rel(g["u"][1], a["v"]) < 1e-12
    p2 = a["ctx"][0, 1:-1, 1:-1]
    pg = g["ctx"][0, 1:-1, 1:-1]
    assert _rel(pg - pg.mean(), p2 - p2.mean()) < 1e-10


# ------------------------------------------------------------ 3D ghost rules
@pytest.mark.parametrize("side", range(6))
@pytest.mark.parametrize("code", [135, -135, 165])
def test_3d_ghost_rules_closed_form(side, code):
    n = 6
    d, hi = side // 2, side % 2
    if d == 0 and hi and code == -135:
        code = 1  # the east slip branch tests the code 1 (utilities.cpp:150)
    bc_lo, bc_hi = [135, 135, 135], [135, 135, 135]
    (bc_hi if hi else bc_lo)[d] = code
    infl = (0.3, 0.6, -0.2)
    o = GenOracle(3, n, 100.0, 1e-3, bc_lo=tuple(bc_lo), bc_hi=tuple(bc_hi), inflow=infl)
    rng = np.random.default_rng(side * 7 + abs(code))
    f = [rng.uniform(-1, 1, o.face_shape(k)) for k in range(3)]
    vc = o.cont2cart(f)
    # axis order in the array: (comp, k, j, i) with 2 ghosts
    ax = 3 - d
    for g in (1, 2):
        gi = (n + 1 + g) if hi else (2 - g)
        mi = (n + 2 - g) if hi else (1 + g)
        ghost = np.take(vc, gi, axis=ax)[:, 2:-2, 2:-2]
        mirr = np.take(vc, mi, axis=ax)[:, 2:-2, 2:-2]
        for c in range(3):
            avg = 0.5 * (ghost[c] + mirr[c])
            if code == 135:
                assert np.allclose(avg, 0.0, atol=1e-15)
            elif code in (-135, 1):
                if c == d:
                    assert np.allclose(avg, 0.0, atol=1e-15)
                else:
                    assert np.array_equal(ghost[c], mirr[c])
            else:
                wall = infl[c] if (d == 1 and hi) else 1.0
                assert np.allclose(avg, wall, atol=1e-15)


# ------------------------------------------------------------ z-independent 3D = 2D
@pytest.mark.parametrize("bc", [((135, 135), (135, 165)), ((111, 135), (111, 165))])
def test_z_independent_3d_steps_as_2d(bc):
    n, ren, steps = 16, 200.0, 2
    dt = 0.5 / (4 * n * n)
    s = _seeded2d(n)
    if bc[0][0] == 111:  # x periodic: make face n equal face 0
        for key in ("u", "uprev"):
            s[key][:, -1] = s[key][:, 0]
    o2 = GenOracle(2, n, ren, dt, bc_lo=bc[0] + (135,), bc_hi=bc[1] + (135,))
    o3 = GenOracle(3, n, ren, dt, bc_lo=bc[0] + (111,), bc_hi=bc[1] + (111,))
    g2 = {"u": [s["u"].copy(), s["v"].copy()], "uprev": [s["uprev"].copy(), s["vprev"].copy()], "ctx": s["ctx"].copy()}

    def lift(a, extra):
        return np.ascontiguousarray(np.broadcast_to(a, (n + extra,) + a.shape))

    g3 = {"u": [lift(s["u"], 0), lift(s["v"], 0), np.zeros(o3.face_shape(2))],
          "uprev": [lift(s["uprev"], 0), lift(s["vprev"], 0), np.zeros(o3.face_shape(2))],
          "ctx": np.zeros(o3.cell_shape(1, 2))}
    st2 = o2.advance(g2, steps)
    st3 = o3.advance(g3, steps)
    assert [x["pseudo_iters"] for x in st2] == [x["pseudo_iters"] for x in st3]
    for k in range(n):
        assert _rel(g3["u"][0][k], g2["u"][0]) < 1e-11
        assert _rel(g3["u"][1][k], g2["u"][1]) < 1e-11
    assert np.abs(g3["u"][2]).max() < 1e-12 * np.abs(g2["u"][0]).max()
    p2 = g2["ctx"][0, 1:-1, 1:-1]
    for k in range(n):
        p3 = g3["ctx"][0, k + 1, 1:-1, 1:-1]
        assert _rel(p3 - p3.mean(), p2 - p2.mean()) < 1e-9


# ------------------------------------------------------------ axis permutation
def _random_state_3d(o, rng, walls_zero):
    f = [rng.uniform(-1, 1, o.face_shape(d)) for d in range(3)]
    if walls_zero:
        for d in range(3):
            ax = 2 - d
            idx = [slice(None)] * 3
            idx[ax] = 0
            f[d][tuple(idx)] = 0
            idx[ax] = -1
            f[d][tuple(idx)] = 0
    return f


def _perm_face(f, d_src):
    """faces of direction d_src under the cyclic relabelling x->y->z->x:
    new (k', j', i') = old (j, i, k), i.e. new axis order (z', y', x') = (y, x, z)"""
    return np.ascontiguousarray(np.transpose(f, (1, 2, 0)))


@pytest.mark.parametrize("bc", ["walls", "periodic"])
def test_axis_permutation_commutes_with_the_3d_step(bc):
    n, ren = 12, 150.0
    dt = 1.5e-4
    code = 135 if bc == "walls" else 111
    kw = dict(bc_lo=(code,) * 3, bc_hi=(code,) * 3, inflow=(0.0, 0.0, 0.0))
    o = GenOracle(3, n, ren, dt, **kw)
    rng = np.random.default_rng(11)
    f = _random_state_3d(o, rng, walls_zero=bc == "walls")
    if bc == "periodic":
        for d in range(3):
            ax = 2 - d
            a = np.moveaxis(f[d], ax, 0)
            a[-1] = a[0]
    st = {"u": [x.copy() for x in f], "uprev": [x.copy() for x in f], "ctx": np.zeros(o.cell_shape(1, 2))}
    # relabelled: new x = old z, new y = old x, new z = old y; new face dir d' = (d + 1) % 3
    # array (z, y, x) -> (y', x', ...) : new[z', y', x'] = old[x = y', y = z', z = x'] ... built by transposes
    def relabel(a):
        # old axes (z, y, x); new coords x' = z, y' = x, z' = y  ->  new axes (z', y', x') = (y, x, z)
        return np.ascontiguousarray(np.transpose(a, (1, 2, 0)))
    fp = [None, None, None]
    for d in range(3):
        fp[(d + 1) % 3] = relabel(f[d])
    sp = {"u": [x.copy() for x in fp], "uprev": [x.copy() for x in fp], "ctx": np.zeros(o.cell_shape(1, 2))}
    s1 = o.advance(st, 1)
    s2 = o.advance(sp, 1)
    assert s1[0]["pseudo_iters"] == s2[0]["pseudo_iters"]
    for d in range(3):
        assert _rel(sp["u"][(d + 1) % 3], relabel(st["u"][d])) < 1e-11
    p1 = st["ctx"][0, 1:-1, 1:-1, 1:-1]
    p2 = sp["ctx"][0, 1:-1, 1:-1, 1:-1]
    assert _rel(p2 - p2.mean(), relabel(p1 - p1.mean())) < 1e-9


# ------------------------------------------------------------ periodic translation
@pytest.mark.parametrize("dim", [2, 3])
def test_translation_along_a_periodic_direction_commutes(dim):
    n, ren, shift = 16, 250.0, 5
    dt = 2e-4
    bc_lo = (111, 135, 111) if dim == 3 else (111, 135, 135)
    bc_hi = (111, 165, 111) if dim == 3 else (111, 165, 135)
    o = GenOracle(dim, n, ren, dt, bc_lo=bc_lo, bc_hi=bc_hi, inflow=(0.5, 0.0, 0.0))
    rng = np.random.default_rng(2)
    f = [rng.uniform(-1, 1, o.face_shape(d)) for d in range(dim)]
    # y walls: wall-normal faces 0; x periodic: face n = face 0 (z likewise in 3D)
    f[1][..., 0, :] = f[1][..., -1, :] = 0.0
    f[0][..., -1] = f[0][..., 0]
    if dim == 3:
        f[2][-1] = f[2][0]
    st = {"u": [x.copy() for x in f], "uprev": [x.copy() for x in f], "ctx": np.zeros(o.cell_shape(1, 2))}

    def roll(a, d):  # shift the cells by `shift` along x
        if d == 0:
            b = np.roll(a[..., :-1], shift, axis=-1)
            return np.ascontiguousarray(np.concatenate([b, b[..., :1]], axis=-1))
        return np.ascontiguousarray(np.roll(a, shift, axis=-1))

    fr = [roll(f[d], d) for d in range(dim)]
    sr = {"u": [x.copy() for x in fr], "uprev": [x.copy() for x in fr], "ctx": np.zeros(o.cell_shape(1, 2))}
    a = o.advance(st, 2)
    b = o.advance(sr, 2)
    assert [x["pseudo_iters"] for x in a] == [x["pseudo_iters"] for x in b]
    for d in range(dim):
        assert _rel(sr["u"][d], roll(st["u"][d], d)) < 1e-11
        if d == 0:
            assert np.array_equal(sr["u"][0][..., -1], sr["u"][0][..., 0])


# ------------------------------------------------------------ Poisson vs FFT/DCT
def _exact_poisson(b, h, periodic_axes):
    """mean-zero solution of the 5/7-point Laplacian: FFT along periodic axes,
    DCT-II along Neumann axes (array axes)"""
    bh = b.copy()
    lam = np.zeros(b.shape)
    for ax in range(b.ndim):
        m = b.shape[ax]
        k = np.arange(m)
        if ax in periodic_axes:
            bh = sfft.fft(bh, axis=ax)
            l1 = -4.0 / h ** 2 * np.sin(np.pi * k / m) ** 2
        else:
            bh = sfft.dct(bh, type=2, norm="ortho", axis=ax)
            l1 = -4.0 / h ** 2 * np.sin(np.pi * k / (2 * m)) ** 2
        shp = [1] * b.ndim
        shp[ax] = m
        lam = lam + l1.reshape(shp)
    lam.flat[0] = 1.0
    xh = bh / lam
    xh.flat[0] = 0.0
    for ax in reversed(range(b.ndim)):
        if ax in periodic_axes:
            xh = sfft.ifft(xh, axis=ax)
        else:
            xh = sfft.idct(xh, type=2, norm="ortho", axis=ax)
    return np.real(xh)


@pytest.mark.parametrize("dim,bc_lo,bc_hi", [
    (2, (111, 135, 135), (111, 165, 135)),
    (2, (111, 111, 135), (111, 111, 135)),
    (3, (111, 111, 111), (111, 111, 111)),
    (3, (135, 111, 135), (135, 111, 165)),
    (3, (135, 135, 135), (135, 165, 135)),
])
def test_poisson_solve_equals_the_exact_solution(dim, bc_lo, bc_hi):
    n = 16 if dim == 3 else 32
    o = GenOracle(dim, n, 100.0, 1e-3, bc_lo=bc_lo, bc_hi=bc_hi)
    rng = np.random.default_rng(dim)
    shape = (n,) * dim
    b = rng.standard_normal(shape)
    rhs = np.zeros((n + 2,) * dim)
    rhs[(slice(1, -1),) * dim] = b
    sol, it, res = o.poisson_solve(rhs)
    x = sol[(slice(1, -1),) * dim]
    periodic_axes = {dim - 1 - d for d in range(dim) if bc_lo[d] == 111}
    ref = _exact_poisson(b - b.mean(), 1.0 / n, periodic_axes)
    assert res <= 1e-13 and it > 0
    assert _rel(x, ref) < 1e-11
    assert abs(x.mean()) < 1e-14 * np.abs(x).max()


# ------------------------------------------------------------ known answers
def _tg_faces(n, t):
    L = 2 * np.pi
    h = L / n
    xf = np.arange(n + 1) * h
    xc = (np.arange(n) + 0.5) * h
    e = np.exp(-2 * t)
    u = np.sin(xf)[None, :] * np.cos(xc)[:, None] * e        # (j, i) x-faces
    v = -np.cos(xc)[None, :] * np.sin(xf)[:, None] * e       # (j, i) y-faces
    return u, v


def _tg_error(n, T):
    L = 2 * np.pi
    h = L / n
    dt = 0.5 / (2 * 2 / h ** 2)
    steps = int(round(T / dt))
    dt = T / steps
    o = GenOracle(2, n, 1.0, dt, length=(L, L, L), bc_lo=(111, 111, 111), bc_hi=(111, 111, 111),
                  mom_tol=1e-12)
    u, v = _tg_faces(n, 0.0)
    xc = (np.arange(n) + 0.5) * h
    ctx = np.zeros((2, n + 2, n + 2))
    ctx[0, 1:-1, 1:-1] = 0.25 * (np.cos(2 * xc)[None, :] + np.cos(2 * xc)[:, None])
    o.userctx_ghosts(ctx)  # the reference fills them at initialisation (fn_init.cpp:550-629)
    st = {"u": [u.copy(), v.copy()], "uprev": [u.copy(), v.copy()], "ctx": ctx}
    # the BDF2 history of the first step: u^{n-1} = u(-dt)
    um, vm = _tg_faces(n, -dt)
    st["uprev"] = [um, vm]
    o.advance(st, steps)
    ue, ve = _tg_faces(n, T)
    return max(np.abs(st["u"][0] - ue).max(), np.abs(st["u"][1] - ve).max()) / np.exp(-2 * T)


def test_taylor_green_decays_at_second_order():
    """utilities.cpp:572-643: u = sin x cos y e^{-2t}, v = -cos x sin y e^{-2t},
    p = (cos 2x + cos 2y) e^{-4t}/4 on [0, 2 pi]^2 periodic with ren = 1."""
    T = 0.05
    e1, e2 = _tg_error(32, T), _tg_error(64, T)
    print(f"Taylor-Green errors {e1:.3e} (32^2), {e2:.3e} (64^2), ratio {e1 / e2:.2f}")
    assert e2 < 1e-3 and 3.5 < e1 / e2 < 4.5


def _abc_faces(n, t, ren):
    L = 2 * np.pi
    h = L / n
    f = np.arange(n + 1) * h
    c = (np.arange(n) + 0.5) * h
    e = np.exp(-t / ren)
    # u(x, y, z) = sin z + cos y at x-faces: arrays (k, j, i)
    u = (np.sin(c)[:, None, None] + np.cos(c)[None, :, None]) * np.ones(n + 1)[None, None, :] * e
    v = (np.sin(c)[None, None, :] + np.cos(c)[:, None, None]) * np.ones(n + 1)[None, :, None] * e
    w = (np.sin(c)[None, :, None] + np.cos(c)[None, None, :]) * np.ones(n + 1)[:, None, None] * e
    u = u * 1.0
    # x-faces sit at x = f (the field does not depend on x), y-faces at y = f, z-faces at z = f
    return [np.ascontiguousarray(u), np.ascontiguousarray(v), np.ascontiguousarray(w)]


def _abc_error(n, T, ren):
    L = 2 * np.pi
    h = L / n
    dt = 0.5 / (2 * 3 / h ** 2)
    steps = int(round(T / dt))
    dt = T / steps
    o = GenOracle(3, n, ren, dt, length=(L, L, L), bc_lo=(111,) * 3, bc_hi=(111,) * 3, mom_tol=1e-12)
    f0 = _abc_faces(n, 0.0, ren)
    # p = -|u|^2/2 at the cell centres (the gradient that balances u . grad u)
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    ctx = np.zeros(o.cell_shape(1, 2))
    ctx[0, 1:-1, 1:-1, 1:-1] = -0.5 * ((np.sin(z) + np.cos(y)) ** 2 + (np.sin(x) + np.cos(z)) ** 2 +
                                       (np.sin(y) + np.cos(x)) ** 2)
    o.userctx_ghosts(ctx)
    st = {"u": [a.copy() for a in f0], "uprev": _abc_faces(n, -dt, ren), "ctx": ctx}
    o.advance(st, steps)
    fe = _abc_faces(n, T, ren)
    return max(np.abs(st["u"][d] - fe[d]).max() for d in range(3)) / np.exp(-T / ren)


def test_abc_flow_decays_at_second_order():
    """u = (sin z + cos y, sin x + cos z, sin y + cos x) e^{-t/Re} on [0, 2 pi]^3:
    curl u = u, so u . grad u = grad(|u|^2/2) and the flow decays viscously."""
    ren, T = 1.0, 0.1
    e1, e2 = _abc_error(12, T, ren), _abc_error(24, T, ren)
    print(f"ABC errors {e1:.3e} (12^3), {e2:.3e} (24^3), ratio {e1 / e2:.2f}")
    assert e2 < 5e-2 and 3.2 < e1 / e2 < 5.0
