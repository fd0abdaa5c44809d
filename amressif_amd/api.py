"""Python host side of the MI355X fractional step: ctypes bindings of the
C-ABI in include/amressif.h (libamressif_hip.so), with torch used only for
device memory and the HIP stream.

The functions mirror the reference's module functions (Source/src1_rever)
one for one -- same names, argument meaning and in-place semantics -- on
torch device tensors laid out like the reference's single-box MultiFabs:

  x-face  (velCont[0], ...)  : float64 (n, n+1)
  y-face  (velCont[1], ...)  : float64 (n+1, n)
  userCtx (p, phi), 1 ghost   : float64 (2, n+2, n+2)
  velCart, 2 ghosts           : float64 (2, n+4, n+4)
  Poisson rhs/sol, 1 ghost    : float64 (n+2, n+2)

There is no CPU fallback: importing this module raises if the HIP library is
missing, and every entry point raises on a non-zero status.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamressif_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python -m amressif_amd.build`")

_lib = ctypes.CDLL(LIB_PATH)

AMRESSIF_FP64 = 0
AMRESSIF_MIXED = 1
FLAG_CAP, FLAG_DIVERGED, FLAG_NONFINITE, FLAG_MG_CAP, FLAG_LIN_CAP = 1, 2, 4, 8, 16
ERRORS = {-1: "ERR_ARG", -2: "ERR_BC", -3: "ERR_UNSUPPORTED", -4: "ERR_HIP", -5: "ERR_NONFINITE", -6: "ERR_PARSE"}


class Inputs(ctypes.Structure):
    _fields_ = [
        ("n_cell", ctypes.c_int),
        ("max_grid_size", ctypes.c_int * 3),
        ("lo_phy_dim", ctypes.c_double * 3),
        ("hi_phy_dim", ctypes.c_double * 3),
        ("IterNum", ctypes.c_int),
        ("nsteps", ctypes.c_int),
        ("cfl", ctypes.c_double),
        ("fixed_dt", ctypes.c_double),
        ("ren", ctypes.c_double),
        ("vis", ctypes.c_double),
        ("phy_bc_lo", ctypes.c_int * 3),
        ("phy_bc_hi", ctypes.c_int * 3),
        ("inflow_waveform", ctypes.c_double * 3),
        ("target_resolution", ctypes.c_int),
        ("momentum_tolerance", ctypes.c_double),
        ("PSEUDO_TIMESTEPPING", ctypes.c_int),
        ("plot_int", ctypes.c_int),
        ("chk_int", ctypes.c_int),
        ("txt_int", ctypes.c_int),
        ("chk_out", ctypes.c_int),
        ("poisson_tol_rel", ctypes.c_double),
        ("poisson_tol_abs", ctypes.c_double),
        ("mg_max_iter", ctypes.c_int),
        ("mg_bottom_max", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("poisson_solver", ctypes.c_int),
        ("gmres_restart", ctypes.c_int),
        ("mixed_momentum", ctypes.c_int),
        ("poisson_warm_start", ctypes.c_int),
        ("spacedim", ctypes.c_int),
    ]


class Array4(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("lo", ctypes.c_int * 3),
        ("hi", ctypes.c_int * 3),
        ("ncomp", ctypes.c_int),
        ("dtype", ctypes.c_int),
    ]


class MultiFabC(ctypes.Structure):
    _fields_ = [
        ("nbox", ctypes.c_int),
        ("fab", ctypes.POINTER(Array4)),
        ("ngrow", ctypes.c_int),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("pseudo_iters", ctypes.c_int),
        ("vcycles", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("norm_error", ctypes.c_double),
        ("div_star", ctypes.c_double),
        ("rhs_sum", ctypes.c_double),
        ("poisson_res", ctypes.c_double),
        ("div_new", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class LinStats(ctypes.Structure):
    _fields_ = [
        ("lin_iters", ctypes.c_int),
        ("vcycles", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("lin_rel", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class AmrStats(ctypes.Structure):
    _fields_ = [
        ("coarse_iters", ctypes.c_int),
        ("fine_iters", ctypes.c_int),
        ("poisson_iters", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("div_star", ctypes.c_double),
        ("div_new", ctypes.c_double),
        ("poisson_res", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_P = ctypes.POINTER
_A4 = _P(Array4)
_c = ctypes
_sig = {
    "amressif_inputs_default": (None, [_P(Inputs)]),
    "amressif_parse_inputs": (_c.c_int, [_c.c_char_p, _c.c_int, _P(_c.c_char_p), _P(Inputs), _c.c_char_p, _c.c_int]),
    "amressif_dt": (_c.c_double, [_P(Inputs)]),
    "amressif_create": (_c.c_int, [_P(Inputs), _c.c_void_p, _P(_c.c_void_p)]),
    "amressif_destroy": (_c.c_int, [_c.c_void_p]),
    "amressif_last_error": (_c.c_char_p, [_c.c_void_p]),
    "amressif_workspace_bytes": (_c.c_int64, [_c.c_void_p]),
    "amressif_set_lid_profile": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "amressif_load_state": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4]),
    "amressif_store_state": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _A4]),
    "amressif_load_state_mf": (_c.c_int, [_c.c_void_p, _P(MultiFabC), _P(MultiFabC), _P(MultiFabC)]),
    "amressif_store_state_mf": (_c.c_int, [_c.c_void_p, _P(MultiFabC), _P(MultiFabC), _P(MultiFabC), _P(MultiFabC)]),
    "amressif_advance": (_c.c_int, [_c.c_void_p, _c.c_int, _P(StepStats)]),
    "amressif_init_state": (_c.c_int, [_c.c_void_p, _P(_c.c_double), _c.c_int]),
    "amressif_write_checkpoint": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_int, _c.c_double]),
    "amressif_read_checkpoint": (_c.c_int, [_c.c_void_p, _c.c_char_p, _P(_c.c_double)]),
    "amressif_write_plotfile": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_char_p, _c.c_int, _c.c_double]),
    "amressif_write_flux_plotfile": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_int, _c.c_double]),
    "amressif_cont2cart": (_c.c_int, [_c.c_void_p, _A4, _A4]),
    "amressif_fluxes": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _A4, _A4]),
    "amressif_momentum_rhs": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4]),
    "amressif_rk_stage": (_c.c_int, [_c.c_void_p, _c.c_int, _A4, _A4, _A4, _A4, _A4]),
    "amressif_pseudo_norm": (_c.c_int, [_c.c_void_p, _A4, _A4, _P(_c.c_double)]),
    "amressif_poisson_rhs": (_c.c_int, [_c.c_void_p, _A4, _A4, _P(_c.c_double), _P(_c.c_double)]),
    "amressif_poisson_solve": (_c.c_int, [_c.c_void_p, _A4, _A4, _c.c_double, _c.c_double, _P(_c.c_int),
                                          _P(_c.c_double)]),
    "amressif_project": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _P(_c.c_double)]),
    "amressif_poisson_rhs_noalloc": (_c.c_int, [_c.c_void_p, _A4, _A4, _c.c_void_p]),
    "amressif_time_kernel": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int64, _P(_c.c_double),
                                        _P(_c.c_double)]),
    "amressif_kernel_timing": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "amressif_kernel_time": (_c.c_int, [_c.c_void_p, _c.c_int, _P(_c.c_double), _P(_c.c_int64),
                                        _P(_c.c_double)]),
    "amressif_tape_create": (_c.c_int, [_c.c_void_p, _c.c_int, _P(_c.c_void_p)]),
    "amressif_tape_destroy": (_c.c_int, [_c.c_void_p]),
    "amressif_tape_reset": (_c.c_int, [_c.c_void_p]),
    "amressif_tape_steps": (_c.c_int, [_c.c_void_p]),
    "amressif_step_fwd": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_void_p, _P(StepStats)]),
    "amressif_step_bwd": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _P(LinStats)]),
    "amressif_step_tlm": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _P(LinStats)]),
    "amressif_load_dual": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _c.c_void_p]),
    "amressif_store_dual": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _c.c_void_p]),
    "amressif_set_adjoint_solver": (_c.c_int, [_c.c_void_p, _c.c_double, _c.c_int]),
    "amressif_amr_create": (_c.c_int, [_P(Inputs), _c.c_int, _c.c_void_p, _P(_c.c_void_p)]),
    "amressif_amr_destroy": (_c.c_int, [_c.c_void_p]),
    "amressif_amr_last_error": (_c.c_char_p, [_c.c_void_p]),
    "amressif_amr_dt": (_c.c_int, [_c.c_void_p, _P(_c.c_double)]),
    "amressif_amr_load_state": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _A4]),
    "amressif_amr_store_state": (_c.c_int, [_c.c_void_p, _A4, _A4, _A4, _A4]),
    "amressif_amr_advance": (_c.c_int, [_c.c_void_p, _c.c_int, _P(AmrStats)]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class AmressifError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def default_inputs(**kw) -> Inputs:
    inp = Inputs()
    _lib.amressif_inputs_default(ctypes.byref(inp))
    for k, v in kw.items():
        cur = getattr(inp, k)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(inp, k, v)
    return inp


def cavity_inputs(n: int, ren: float, precision: int = AMRESSIF_FP64, cfl: float = 0.5, **kw) -> Inputs:
    """The workloads' inputs: unit square, phy_bc_lo = 135 135, phy_bc_hi =
    135 165, lid inflow 1 0, IterNum 100, momentum_tolerance 1e-10, one box."""
    args = dict(n_cell=n, max_grid_size=(n, n, n), IterNum=100, cfl=cfl, ren=ren, vis=1.0 / ren,
                phy_bc_lo=(135, 135, 135), phy_bc_hi=(135, 165, 135), inflow_waveform=(1.0, 0.0, 0.0),
                momentum_tolerance=1e-10, precision=precision)
    args.update(kw)
    return default_inputs(**args)


def parse_inputs(path: str | None, overrides=()) -> Inputs:
    inp = Inputs()
    arr = (ctypes.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
    err = ctypes.create_string_buffer(512)
    rc = _lib.amressif_parse_inputs(path.encode() if path else None, len(overrides), arr, ctypes.byref(inp), err, 512)
    if rc:
        raise AmressifError(rc, err.value.decode())
    return inp


def dt_of(inp: Inputs) -> float:
    return _lib.amressif_dt(ctypes.byref(inp))


# ------------------------------------------------------------ Array4 views
def _a4(t: torch.Tensor, lo, hi, ncomp) -> Array4:
    if t.dtype != torch.float64 or not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError("expected a contiguous float64 cuda tensor")
    a = Array4()
    a.ptr = t.data_ptr()
    a.lo[0], a.lo[1], a.lo[2] = lo[0], lo[1], 0
    a.hi[0], a.hi[1], a.hi[2] = hi[0], hi[1], 0
    a.ncomp = ncomp
    a.dtype = 0
    return a


def xface_view(t, n):
    assert tuple(t.shape) == (n, n + 1), t.shape
    return _a4(t, (0, 0), (n, n - 1), 1)


def yface_view(t, n):
    assert tuple(t.shape) == (n + 1, n), t.shape
    return _a4(t, (0, 0), (n - 1, n), 1)


def faces_view(u, v, n):
    arr = (Array4 * 2)()
    arr[0] = xface_view(u, n)
    arr[1] = yface_view(v, n)
    return arr


def face_shape(d: int, n: int, dim: int) -> tuple:
    """torch shape (slowest axis first) of the face array of direction d on n^dim cells"""
    ext = [n + (d == 0), n + (d == 1), n + (d == 2)]
    return (ext[2], ext[1], ext[0]) if dim == 3 else (ext[1], ext[0])


def _a4_nd(t: torch.Tensor, lo, hi, ncomp) -> Array4:
    if t.dtype != torch.float64 or not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError("expected a contiguous float64 cuda tensor")
    a = Array4()
    a.ptr = t.data_ptr()
    for k in range(3):
        a.lo[k], a.hi[k] = lo[k], hi[k]
    a.ncomp = ncomp
    a.dtype = 0
    return a


def faces_view_nd(faces, n: int, dim: int):
    """Array4 views of dim face arrays (velCont[0..dim-1]) on n^dim cells"""
    arr = (Array4 * dim)()
    for d in range(dim):
        assert tuple(faces[d].shape) == face_shape(d, n, dim), (d, faces[d].shape)
        hi = [n - 1 + (d == k) if k < dim else 0 for k in range(3)]
        arr[d] = _a4_nd(faces[d], (0, 0, 0), hi, 1)
    return arr


def cell_view_nd(t, n: int, dim: int, ghosts: int):
    """Array4 view of a cell array (comps, [n+2g,] n+2g, n+2g) with g ghost layers"""
    ncomp = t.shape[0] if t.dim() == dim + 1 else 1
    lo = [-ghosts if k < dim else 0 for k in range(3)]
    hi = [n - 1 + ghosts if k < dim else 0 for k in range(3)]
    return _a4_nd(t, lo, hi, ncomp)


def cell_view(t, n, ghosts):
    ncomp = 1 if t.dim() == 2 else t.shape[0]
    assert tuple(t.shape[-2:]) == (n + 2 * ghosts, n + 2 * ghosts), t.shape
    return _a4(t, (-ghosts, -ghosts), (n - 1 + ghosts, n - 1 + ghosts), ncomp)


# ------------------------------------------------------ multi-box MultiFabs
def box_array(n: int, max_grid_size) -> list:
    """Cell boxes (lo0, lo1, hi0, hi1) of the n x n domain chopped so that no
    side exceeds max_grid_size (x, y): BoxArray::maxSize (main.cpp:317-320).
    A side of length L > m is cut into k = ceil(L / m) pieces as even as
    possible, the first L mod k of them one cell longer; boxes are listed with
    x varying fastest."""
    def cuts(length, m):
        k = -(-length // m)
        size, extra = divmod(length, k)
        out, lo = [], 0
        for b in range(k):
            w = size + (1 if b < extra else 0)
            out.append((lo, lo + w - 1))
            lo += w
        return out
    mx, my = (max_grid_size, max_grid_size) if isinstance(max_grid_size, int) else tuple(max_grid_size)[:2]
    return [(x0, y0, x1, y1) for (y0, y1) in cuts(n, my) for (x0, x1) in cuts(n, mx)]


def box_array_nd(n: int, dim: int, max_grid_size) -> list:
    """Cell boxes ((lo x, lo y[, lo z]), (hi x, hi y[, hi z])) of the n^dim domain
    chopped as box_array does per direction (x fastest in the list)."""
    def cuts(length, m):
        k = -(-length // m)
        size, extra = divmod(length, k)
        out, lo = [], 0
        for b in range(k):
            w = size + (1 if b < extra else 0)
            out.append((lo, lo + w - 1))
            lo += w
        return out
    m = (max_grid_size,) * 3 if isinstance(max_grid_size, int) else tuple(max_grid_size)
    cx, cy = cuts(n, m[0]), cuts(n, m[1])
    cz = cuts(n, m[2]) if dim == 3 else [(0, 0)]
    out = []
    for (z0, z1) in cz:
        for (y0, y1) in cy:
            for (x0, x1) in cx:
                lo, hi = (x0, y0, z0), (x1, y1, z1)
                out.append((lo[:dim], hi[:dim]))
    return out


class MultiFabND:
    """MultiFab over box_array_nd boxes in 2D or 3D: kind = face direction d
    (nodal in d: both bounding faces per box) or "cell" (ncomp components,
    ngrow ghosts).  Tensors are C-ordered with the slowest axis first:
    (comps, [z,] y, x) for cells, ([z,] y, x) for faces."""

    def __init__(self, boxes, dim: int, kind, ncomp: int = 1, ngrow: int = 0, device="cuda"):
        self.boxes, self.dim, self.kind, self.ncomp, self.ngrow = list(boxes), dim, kind, ncomp, ngrow
        self.fabs = []
        arr = (Array4 * len(self.boxes))()
        for k, (lo, hi) in enumerate(self.boxes):
            vlo, vhi = self.window(lo, hi)
            ext = [vhi[d] - vlo[d] + 1 for d in range(dim)][::-1]
            shape = tuple(([ncomp] if kind == "cell" else []) + ext)
            t = torch.zeros(shape, dtype=torch.float64, device=device)
            self.fabs.append(t)
            a = Array4()
            a.ptr = t.data_ptr()
            for d in range(3):
                a.lo[d], a.hi[d] = (vlo[d], vhi[d]) if d < dim else (0, 0)
            a.ncomp = ncomp
            a.dtype = 0
            arr[k] = a
        self._arr = arr
        self.c = MultiFabC(len(self.boxes), ctypes.cast(arr, ctypes.POINTER(Array4)), ngrow)

    def window(self, lo, hi):
        """the view's bounds (ghosts included) of a valid cell box"""
        g = self.ngrow
        vlo = [lo[d] - g for d in range(self.dim)]
        vhi = [hi[d] + g + (1 if self.kind == d else 0) for d in range(self.dim)]
        return vlo, vhi

    def fill_from(self, glob, src_ghosts: int = 0):
        """each box's view from a global array (faces: face_shape; cells: (comps, [n + 2 g,] ...))"""
        for (lo, hi), t in zip(self.boxes, self.fabs):
            vlo, vhi = self.window(lo, hi)
            sl = tuple(slice(vlo[d] + src_ghosts, vhi[d] + 1 + src_ghosts) for d in range(self.dim))[::-1]
            part = glob[(slice(None),) + sl] if self.kind == "cell" else glob[sl]
            t.copy_(torch.as_tensor(np.ascontiguousarray(part), dtype=torch.float64))
        return self


class MultiFab:
    """Per-box torch tensors of one field over a box_array, with the ctypes
    amr_multifab that views them.  kind: "x" / "y" faces (nodal in that
    direction, both bounding faces per box) or "cell" (ncomp components, ngrow
    ghosts).  Tensor shapes: x faces (ny_b, nx_b + 1), y faces (ny_b + 1, nx_b),
    cells (ncomp, ny_b + 2 g, nx_b + 2 g)."""

    def __init__(self, boxes, kind: str, ncomp: int = 1, ngrow: int = 0, device="cuda", tensors=None):
        self.boxes, self.kind, self.ncomp, self.ngrow = list(boxes), kind, ncomp, ngrow
        g = ngrow
        if tensors is None:
            tensors = [torch.zeros(self.shape(b), dtype=torch.float64, device=device) for b in self.boxes]
        self.fabs = list(tensors)
        arr = (Array4 * len(self.boxes))()
        for k, ((x0, y0, x1, y1), t) in enumerate(zip(self.boxes, self.fabs)):
            assert tuple(t.shape) == self.shape(self.boxes[k]), (t.shape, self.shape(self.boxes[k]))
            ex, ey = (1, 0) if kind == "x" else ((0, 1) if kind == "y" else (0, 0))
            arr[k] = _a4(t, (x0 - g, y0 - g), (x1 + ex + g, y1 + ey + g), ncomp)
        self._arr = arr
        self.c = MultiFabC(len(self.boxes), ctypes.cast(arr, ctypes.POINTER(Array4)), ngrow)

    def shape(self, box):
        x0, y0, x1, y1 = box
        nx, ny, g = x1 - x0 + 1, y1 - y0 + 1, self.ngrow
        if self.kind == "x":
            return (ny + 2 * g, nx + 1 + 2 * g)
        if self.kind == "y":
            return (ny + 1 + 2 * g, nx + 2 * g)
        return (self.ncomp, ny + 2 * g, nx + 2 * g)

    @classmethod
    def from_global(cls, boxes, kind, a, ncomp=1, ngrow=0, src_ghosts=0, device="cuda"):
        """split a global numpy array (the single-box layout, src_ghosts ghost
        layers for cells) into per-box device tensors"""
        g, sg = ngrow, src_ghosts
        parts = []
        for (x0, y0, x1, y1) in boxes:
            if kind == "x":
                blk = a[y0:y1 + 1, x0:x1 + 2]
            elif kind == "y":
                blk = a[y0:y1 + 2, x0:x1 + 1]
            else:
                blk = a[:, y0 - g + sg:y1 + 1 + g + sg, x0 - g + sg:x1 + 1 + g + sg]
            parts.append(torch.as_tensor(np.ascontiguousarray(blk), dtype=torch.float64, device=device))
        return cls(boxes, kind, ncomp, ngrow, device, parts)


# ----------------------------------------------------------------- context
class Context:
    """One amressif_ctx on the current torch HIP stream."""

    def __init__(self, inp: Inputs, stream: torch.cuda.Stream | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("amressif_amd needs a GPU: no HIP device is visible")
        self.inp = inp
        self.n = inp.n_cell
        self.stream = stream or torch.cuda.current_stream()
        h = ctypes.c_void_p()
        rc = _lib.amressif_create(ctypes.byref(inp), ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self._h = h
        if rc:
            msg = _lib.amressif_last_error(h).decode() if h.value else ""
            _lib.amressif_destroy(h)
            self._h = ctypes.c_void_p()
            raise AmressifError(rc, msg)
        self._lid = None

    def close(self):
        if self._h and self._h.value:
            _lib.amressif_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise AmressifError(rc, _lib.amressif_last_error(self._h).decode())

    @property
    def dt(self):
        return dt_of(self.inp)

    def workspace_bytes(self):
        return _lib.amressif_workspace_bytes(self._h)

    def set_lid_profile(self, lid: torch.Tensor | None):
        if lid is not None:
            assert lid.dtype == torch.float64 and lid.numel() == self.n and lid.is_cuda
            self._lid = lid.contiguous()
            self._check(_lib.amressif_set_lid_profile(self._h, ctypes.c_void_p(self._lid.data_ptr())))
        else:
            self._lid = None
            self._check(_lib.amressif_set_lid_profile(self._h, None))

    # ---- state on n^dim cells (2D or 3D, any boundary codes)
    @property
    def dim(self) -> int:
        return 3 if self.inp.spacedim == 3 else 2

    def load_faces(self, faces, faces_prev, ctx=None):
        """amressif_load_state with dim face arrays per state (face_shape) and userCtx
        (2, [n+2g,] n+2g, n+2g) with g >= 0 ghost layers, or None (p = phi = 0)"""
        n, dim = self.n, self.dim
        uc = None
        if ctx is not None:
            g = (ctx.shape[-1] - n) // 2
            uc = cell_view_nd(ctx, n, dim, g)
        self._check(_lib.amressif_load_state(self._h, faces_view_nd(faces, n, dim), faces_view_nd(faces_prev, n, dim),
                                             ctypes.byref(uc) if uc is not None else None))

    def store_faces(self, faces=None, faces_prev=None, ctx=None, velCart=None):
        """amressif_store_state into dim face arrays, userCtx (<= 1 ghost) and velCart (<= 2 ghosts)"""
        n, dim = self.n, self.dim
        vc = faces_view_nd(faces, n, dim) if faces is not None else None
        vp = faces_view_nd(faces_prev, n, dim) if faces_prev is not None else None
        uc = cell_view_nd(ctx, n, dim, (ctx.shape[-1] - n) // 2) if ctx is not None else None
        cc = cell_view_nd(velCart, n, dim, (velCart.shape[-1] - n) // 2) if velCart is not None else None
        self._check(_lib.amressif_store_state(self._h, vc, vp, ctypes.byref(uc) if uc is not None else None,
                                              ctypes.byref(cc) if cc is not None else None))

    # ---- state (2D)
    def load_state(self, u, v, uprev, vprev, ctx=None):
        n = self.n
        vc = faces_view(u, v, n)
        vp = faces_view(uprev, vprev, n)
        uc = cell_view(ctx, n, (ctx.shape[-1] - n) // 2) if ctx is not None else None
        self._check(_lib.amressif_load_state(self._h, vc, vp, ctypes.byref(uc) if uc is not None else None))

    def store_state(self, u=None, v=None, uprev=None, vprev=None, ctx=None, velCart=None):
        n = self.n
        vc = faces_view(u, v, n) if u is not None else None
        vp = faces_view(uprev, vprev, n) if uprev is not None else None
        uc = cell_view(ctx, n, (ctx.shape[-1] - n) // 2) if ctx is not None else None
        cc = cell_view(velCart, n, (velCart.shape[-1] - n) // 2) if velCart is not None else None
        self._check(_lib.amressif_store_state(self._h, vc, vp, ctypes.byref(uc) if uc is not None else None,
                                              ctypes.byref(cc) if cc is not None else None))

    def load_state_mf(self, u: MultiFab, v: MultiFab, uprev: MultiFab, vprev: MultiFab, ctx: MultiFab | None = None):
        """amressif_load_state_mf: the state from multi-box MultiFabs"""
        vc = (MultiFabC * 2)(u.c, v.c)
        vp = (MultiFabC * 2)(uprev.c, vprev.c)
        self._check(_lib.amressif_load_state_mf(self._h, vc, vp, ctypes.byref(ctx.c) if ctx is not None else None))

    def store_state_mf(self, u=None, v=None, uprev=None, vprev=None, ctx=None, velCart=None):
        """amressif_store_state_mf: the state into multi-box MultiFabs (any may be None)"""
        vc = (MultiFabC * 2)(u.c, v.c) if u is not None else None
        vp = (MultiFabC * 2)(uprev.c, vprev.c) if uprev is not None else None
        self._check(_lib.amressif_store_state_mf(self._h, vc, vp, ctypes.byref(ctx.c) if ctx is not None else None,
                                                 ctypes.byref(velCart.c) if velCart is not None else None))

    def load_faces_mf(self, faces, faces_prev, ctx=None):
        """amressif_load_state_mf with dim MultiFabND per face state (3D and periodic contexts too)"""
        d = self.dim
        vc = (MultiFabC * d)(*[m.c for m in faces])
        vp = (MultiFabC * d)(*[m.c for m in faces_prev])
        self._check(_lib.amressif_load_state_mf(self._h, vc, vp, ctypes.byref(ctx.c) if ctx is not None else None))

    def store_faces_mf(self, faces=None, faces_prev=None, ctx=None, velCart=None):
        """amressif_store_state_mf into dim MultiFabND per face state, userCtx and velCart (any may be None)"""
        d = self.dim
        vc = (MultiFabC * d)(*[m.c for m in faces]) if faces is not None else None
        vp = (MultiFabC * d)(*[m.c for m in faces_prev]) if faces_prev is not None else None
        self._check(_lib.amressif_store_state_mf(self._h, vc, vp, ctypes.byref(ctx.c) if ctx is not None else None,
                                                 ctypes.byref(velCart.c) if velCart is not None else None))

    def init_state(self, psi_coeffs=None):
        """amressif_init_state: the cavity at rest (hybrid_grid_init), or the
        seeded streamfunction field of the kmax x kmax coefficients a[kx-1, ky-1]"""
        if psi_coeffs is None:
            self._check(_lib.amressif_init_state(self._h, None, 0))
        else:
            a = np.ascontiguousarray(psi_coeffs, dtype=np.float64)
            assert a.ndim == 2 and a.shape[0] == a.shape[1]
            self._check(_lib.amressif_init_state(self._h, a.ctypes.data_as(_P(_c.c_double)), a.shape[0]))

    def write_checkpoint(self, directory: str, step: int, time: float):
        """SaveCheckpoint: <directory>/checkpoint%05d (AMReX Header + VisMF Level_0 MultiFabs)"""
        self._check(_lib.amressif_write_checkpoint(self._h, directory.encode(), step, time))
        return os.path.join(directory, f"checkpoint{step:05d}")

    def read_checkpoint(self, path: str) -> float:
        """LoadCheckpoint: the state from a checkpoint directory; returns its time"""
        t = ctypes.c_double()
        self._check(_lib.amressif_read_checkpoint(self._h, path.encode(), ctypes.byref(t)))
        return t.value

    def write_plotfile(self, directory: str, name: str, step: int, time: float):
        """Export_Flow_Field: <directory>/<name>%05d with pressure, U, V"""
        self._check(_lib.amressif_write_plotfile(self._h, directory.encode(), name.encode(), step, time))
        return os.path.join(directory, f"{name}{step:05d}")

    def write_flux_plotfile(self, directory: str, step: int, time: float):
        """Export_Fluxes (utilities.cpp:496-532): pltFluxNNNNN of the last step's predictor."""
        self._check(_lib.amressif_write_flux_plotfile(self._h, directory.encode(), step, time))
        return os.path.join(directory, f"pltFlux{step:05d}")

    def advance(self, nsteps: int, with_stats: bool = True):
        st = (StepStats * max(1, nsteps))() if with_stats else None
        self._check(_lib.amressif_advance(self._h, nsteps, st))
        return [st[k].as_dict() for k in range(nsteps)] if with_stats else None

    def time_kernel(self, which: int, reps: int = 20, copy_bytes: int = 0):
        """(average ms per launch, algorithmic bytes per launch) of one kernel, HIP events on this stream"""
        ms, b = ctypes.c_double(), ctypes.c_double()
        self._check(_lib.amressif_time_kernel(self._h, which, reps, copy_bytes, ctypes.byref(ms), ctypes.byref(b)))
        return ms.value, b.value

    # in-step kernel timing (HIP events around the step's main launches)
    KERNELS = ("stage", "stage_last", "resid", "tangent", "mg_down", "mg_up", "project", "poisson_start", "pair",
               "pair_last", "mpair")

    def kernel_timing(self, enable: bool) -> None:
        """start (clearing the totals) or stop per-launch timing inside advance()"""
        self._check(_lib.amressif_kernel_timing(self._h, 1 if enable else 0))

    def kernel_times(self) -> dict:
        """{kernel: (average ms per counted launch, counted launches, algorithmic bytes per launch)}"""
        out = {}
        for k, name in enumerate(self.KERNELS):
            ms, cnt, nb = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
            self._check(_lib.amressif_kernel_time(self._h, k, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(nb)))
            out[name] = (ms.value, cnt.value, nb.value)
        return out

    # ---- discrete adjoint (forward/backward pair)
    def tape(self, max_steps: int) -> "Tape":
        return Tape(self, max_steps)

    def step_fwd(self, nsteps: int, tape: "Tape"):
        """advance nsteps, recording each step's converged predictor u* on the tape"""
        st = (StepStats * max(1, nsteps))()
        self._check(_lib.amressif_step_fwd(self._h, nsteps, tape._h, st))
        return [st[k].as_dict() for k in range(nsteps)]

    def step_bwd(self, tape: "Tape", nsteps: int):
        """adjoint of the tape's last nsteps steps (newest first) on the dual state"""
        st = (LinStats * max(1, nsteps))()
        self._check(_lib.amressif_step_bwd(self._h, tape._h, nsteps, st))
        return [st[k].as_dict() for k in range(nsteps)]

    def step_tlm(self, tape: "Tape", first: int, nsteps: int):
        """tangent-linear of tape steps [first, first + nsteps) on the dual state"""
        st = (LinStats * max(1, nsteps))()
        self._check(_lib.amressif_step_tlm(self._h, tape._h, first, nsteps, st))
        return [st[k].as_dict() for k in range(nsteps)]

    def load_dual(self, du=None, dv=None, dup=None, dvp=None, dp=None, dlid=None):
        """dual state in: faces (du, dv), (dup, dvp); cells dp (n+2g square); lid dlid (n); None = zeros"""
        n = self.n
        a = faces_view(du, dv, n) if du is not None else None
        b = faces_view(dup, dvp, n) if dup is not None else None
        c = cell_view(dp, n, (dp.shape[-1] - n) // 2) if dp is not None else None
        lid = ctypes.c_void_p(dlid.data_ptr()) if dlid is not None else None
        self._check(_lib.amressif_load_dual(self._h, a, b, ctypes.byref(c) if c is not None else None, lid))

    def store_dual(self, du=None, dv=None, dup=None, dvp=None, dp=None, dlid=None):
        n = self.n
        a = faces_view(du, dv, n) if du is not None else None
        b = faces_view(dup, dvp, n) if dup is not None else None
        c = cell_view(dp, n, (dp.shape[-1] - n) // 2) if dp is not None else None
        lid = ctypes.c_void_p(dlid.data_ptr()) if dlid is not None else None
        self._check(_lib.amressif_store_dual(self._h, a, b, ctypes.byref(c) if c is not None else None, lid))

    def set_adjoint_solver(self, tol: float = 1e-14, max_sweeps: int = 200):
        self._check(_lib.amressif_set_adjoint_solver(self._h, tol, max_sweeps))

    # ---- module functions (reference names)
    def cont2cart(self, u, v, velCart=None):
        """utilities.cpp:16-350"""
        n = self.n
        velCart = torch.zeros((2, n + 4, n + 4), dtype=torch.float64, device=u.device) if velCart is None else velCart
        cc = cell_view(velCart, n, 2)
        self._check(_lib.amressif_cont2cart(self._h, faces_view(u, v, n), ctypes.byref(cc)))
        return velCart

    def momentum_rhs(self, ustar, vstar, userCtx):
        """One stage's momentum RHS: gradient_calc_approach2 + convective_flux_calc_new_quick +
        viscous_flux_calc + flux wall ghosts + momentum_righthand_side_calc."""
        n = self.n
        rx = torch.empty((n, n + 1), dtype=torch.float64, device=ustar.device)
        ry = torch.empty((n + 1, n), dtype=torch.float64, device=ustar.device)
        uc = cell_view(userCtx, n, (userCtx.shape[-1] - n) // 2)
        self._check(_lib.amressif_momentum_rhs(self._h, faces_view(ustar, vstar, n), ctypes.byref(uc),
                                               faces_view(rx, ry, n)))
        return rx, ry

    def fluxes(self, ustar, vstar):
        """convective_flux_calc_new_quick / viscous_flux_calc outputs (fn_flux_calc.cpp:11-202, 695-766):
        returns fluxConvect (2, n, n), fluxViscous (2, n, n) and the half-node fluxes
        (N1 x-face, N1 y-face, N2 x-face, N2 y-face)."""
        n = self.n
        dev = ustar.device
        fc = torch.zeros((2, n, n), dtype=torch.float64, device=dev)
        fv = torch.zeros((2, n, n), dtype=torch.float64, device=dev)
        h1 = (torch.zeros((n, n + 1), dtype=torch.float64, device=dev),
              torch.zeros((n + 1, n), dtype=torch.float64, device=dev))
        h2 = (torch.zeros((n, n + 1), dtype=torch.float64, device=dev),
              torch.zeros((n + 1, n), dtype=torch.float64, device=dev))
        cv, vv = cell_view(fc, n, 0), cell_view(fv, n, 0)
        self._check(_lib.amressif_fluxes(self._h, faces_view(ustar, vstar, n), ctypes.byref(cv), ctypes.byref(vv),
                                         faces_view(*h1, n), faces_view(*h2, n)))
        return fc, fv, (h1[0], h1[1], h2[0], h2[1])

    def runge_kutta4_pseudo_time_stepping(self, sub, ustar, vstar, ucont, vcont, uprev, vprev, uprevprev,
                                          vprevprev, userCtx):
        """momentum.cpp:9-94 with the stage RHS: velStar updated in place."""
        n = self.n
        uc = cell_view(userCtx, n, (userCtx.shape[-1] - n) // 2)
        self._check(_lib.amressif_rk_stage(self._h, sub, faces_view(ustar, vstar, n), faces_view(ucont, vcont, n),
                                           faces_view(uprev, vprev, n), faces_view(uprevprev, vprevprev, n),
                                           ctypes.byref(uc)))

    def Error_Computation(self, ucont, vcont, ustar, vstar) -> float:
        """utilities.cpp:421-493"""
        out = ctypes.c_double()
        self._check(_lib.amressif_pseudo_norm(self._h, faces_view(ucont, vcont, self.n),
                                              faces_view(ustar, vstar, self.n), ctypes.byref(out)))
        return out.value

    def poisson_righthand_side_calc(self, u, v, rhs=None):
        """fn_rhs.cpp:71-125: returns (rhs, ||D||inf, sum rhs)"""
        n = self.n
        rhs = torch.zeros((n + 2, n + 2), dtype=torch.float64, device=u.device) if rhs is None else rhs
        rv = cell_view(rhs, n, 1)
        linf, s = ctypes.c_double(), ctypes.c_double()
        self._check(_lib.amressif_poisson_rhs(self._h, faces_view(u, v, n), ctypes.byref(rv), ctypes.byref(linf),
                                              ctypes.byref(s)))
        return rhs, linf.value, s.value

    def poisson_advance(self, rhs, tol_rel=1e-10, tol_abs=0.0, sol=None):
        """poisson.cpp:41-186: returns (sol, V-cycles, ||r||inf/||b||inf); sol mean-zero."""
        n = self.n
        sol = torch.zeros((n + 2, n + 2), dtype=torch.float64, device=rhs.device) if sol is None else sol
        sv, rv = cell_view(sol, n, 1), cell_view(rhs, n, 1)
        it, res = ctypes.c_int(), ctypes.c_double()
        self._check(_lib.amressif_poisson_solve(self._h, ctypes.byref(sv), ctypes.byref(rv), tol_rel, tol_abs,
                                                ctypes.byref(it), ctypes.byref(res)))
        return sol, it.value, res.value

    def update_solution(self, ustar, vstar, userCtx):
        """poisson.cpp:192-392: returns (u, v, ||D(u)||inf); userCtx p updated in place."""
        n = self.n
        u = torch.empty((n, n + 1), dtype=torch.float64, device=ustar.device)
        v = torch.empty((n + 1, n), dtype=torch.float64, device=ustar.device)
        uc = cell_view(userCtx, n, (userCtx.shape[-1] - n) // 2)
        d = ctypes.c_double()
        self._check(_lib.amressif_project(self._h, faces_view(ustar, vstar, n), ctypes.byref(uc),
                                          faces_view(u, v, n), ctypes.byref(d)))
        return u, v, d.value


class Tape:
    """Device record of the converged predictors u* of consecutive steps."""

    def __init__(self, ctx: Context, max_steps: int):
        h = ctypes.c_void_p()
        ctx._check(_lib.amressif_tape_create(ctx._h, max_steps, ctypes.byref(h)))
        self._h, self.ctx, self.max_steps = h, ctx, max_steps

    @property
    def steps(self) -> int:
        return _lib.amressif_tape_steps(self._h)

    def reset(self):
        _lib.amressif_tape_reset(self._h)

    def close(self):
        if self._h and self._h.value:
            _lib.amressif_tape_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TwoLevel:
    """Coarse n0^2 level + a ratio-2 patch over the coarse rows [jc0, n0)
    (amressif_amr_*), state held in torch device tensors:
    u0 (n0, n0+1), v0 (n0+1, n0), ctx0 (2, n0+2, n0+2); u1 (ny, 2n0+1),
    v1 (ny+1, 2n0), p1 (ny, 2n0) with ny = 2 (n0 - jc0)."""

    def __init__(self, inp: Inputs, jc0: int, state: dict, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("amressif_amd needs a GPU: no HIP device is visible")
        self.n0, self.jc0 = inp.n_cell, jc0
        self.nx, self.ny = 2 * self.n0, 2 * (self.n0 - jc0)
        h = ctypes.c_void_p()
        rc = _lib.amressif_amr_create(ctypes.byref(inp), jc0, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                      ctypes.byref(h))
        self._h = h
        if rc:
            msg = _lib.amressif_amr_last_error(h).decode() if h.value else ""
            _lib.amressif_amr_destroy(h)
            self._h = ctypes.c_void_p()
            raise AmressifError(rc, msg)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)  # noqa: E731
        n0, nx, ny = self.n0, self.nx, self.ny
        self.u0, self.v0 = t(state["u0"]), t(state["v0"])
        self.ctx0 = t(state.get("ctx0", np.zeros((2, n0 + 2, n0 + 2))))
        self.u1, self.v1 = t(state["u1"]), t(state["v1"])
        self.p1 = t(state.get("p1", np.zeros((ny, nx))))
        self._check(_lib.amressif_amr_load_state(self._h, faces_view(self.u0, self.v0, n0),
                                                 ctypes.byref(cell_view(self.ctx0, n0, 1)), self._fine_faces(),
                                                 ctypes.byref(self._fine_p())))

    def _fine_faces(self):
        arr = (Array4 * 2)()
        arr[0] = _a4(self.u1, (0, 0), (self.nx, self.ny - 1), 1)
        arr[1] = _a4(self.v1, (0, 0), (self.nx - 1, self.ny), 1)
        return arr

    def _fine_p(self):
        return _a4(self.p1, (0, 0), (self.nx - 1, self.ny - 1), 1)

    def _check(self, rc):
        if rc:
            raise AmressifError(rc, _lib.amressif_amr_last_error(self._h).decode())

    @property
    def dt(self) -> float:
        d = ctypes.c_double()
        self._check(_lib.amressif_amr_dt(self._h, ctypes.byref(d)))
        return d.value

    def advance(self, nsteps: int):
        st = (AmrStats * max(1, nsteps))()
        self._check(_lib.amressif_amr_advance(self._h, nsteps, st))
        return [st[k].as_dict() for k in range(nsteps)]

    def fetch(self):
        n0 = self.n0
        self._check(_lib.amressif_amr_store_state(self._h, faces_view(self.u0, self.v0, n0),
                                                  ctypes.byref(cell_view(self.ctx0, n0, 1)), self._fine_faces(),
                                                  ctypes.byref(self._fine_p())))
        return {"u0": self.u0.cpu().numpy(), "v0": self.v0.cpu().numpy(), "ctx0": self.ctx0.cpu().numpy(),
                "u1": self.u1.cpu().numpy(), "v1": self.v1.cpu().numpy(), "p1": self.p1.cpu().numpy()}

    def close(self):
        if self._h and self._h.value:
            _lib.amressif_amr_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Run:
    """A whole single-box run: state tensors on the GPU and a context."""
    ctx: Context
    u: torch.Tensor
    v: torch.Tensor
    uprev: torch.Tensor
    vprev: torch.Tensor
    userCtx: torch.Tensor

    @classmethod
    def from_numpy(cls, inp: Inputs, state: dict, device="cuda"):
        ctx = Context(inp)
        t = {k: torch.as_tensor(state[k], dtype=torch.float64, device=device).contiguous()
             for k in ("u", "v", "uprev", "vprev", "ctx")}
        r = cls(ctx, t["u"], t["v"], t["uprev"], t["vprev"], t["ctx"])
        ctx.load_state(r.u, r.v, r.uprev, r.vprev, r.userCtx)
        return r

    def advance(self, nsteps):
        return self.ctx.advance(nsteps)

    def fetch(self):
        """state back to host numpy (same keys as the oracle's state dict)"""
        n = self.ctx.n
        velCart = torch.zeros((2, n + 4, n + 4), dtype=torch.float64, device=self.u.device)
        self.ctx.store_state(self.u, self.v, self.uprev, self.vprev, self.userCtx, velCart)
        return {"u": self.u.cpu().numpy(), "v": self.v.cpu().numpy(), "uprev": self.uprev.cpu().numpy(),
                "vprev": self.vprev.cpu().numpy(), "ctx": self.userCtx.cpu().numpy(), "vc": velCart.cpu().numpy()}
