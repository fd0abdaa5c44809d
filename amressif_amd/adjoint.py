"""Discrete adjoint of the cavity rollout: a thin torch.autograd.Function over
the C-ABI forward/backward pair (amressif_step_fwd / amressif_step_bwd) with
checkpointing every k steps.

Forward: N steps from (u0, v0) (velContPrev = velCont at t = 0, p = 0) under
the per-column lid profile g, storing the state (u, v, uprev, vprev, p) at the
start of every k-step segment.  Backward: for each segment, newest first,
restore its checkpoint, re-run it recording u* on a tape, and sweep the adjoint
back through it.  The recomputation is bitwise the forward (the step's
reductions are deterministic), so the adjoint is that of the forward run.

Controls: the lid profile g (n values) and the initial velocity (u0, v0); the
gradient with respect to the initial velocity is the sum of the adjoints of
velCont and velContPrev at t = 0, which both start from it.
"""
from __future__ import annotations

import torch

from . import api


class CheckpointedRollout:
    """N forward steps of one context with a checkpoint every `every` steps."""

    def __init__(self, inp: api.Inputs, nsteps: int, every: int = 32, adjoint_tol: float = 1e-14):
        self.inp = inp
        self.n = inp.n_cell
        self.nsteps = nsteps
        self.every = max(1, min(every, nsteps))
        self.ctx = api.Context(inp)
        self.ctx.set_adjoint_solver(adjoint_tol, 200)
        self.tape = self.ctx.tape(self.every)
        self.checkpoints: list[dict] = []
        self.fwd_stats: list[dict] = []
        self.bwd_stats: list[dict] = []

    def _zeros_state(self, dev):
        n = self.n
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)  # noqa: E731
        return {"u": z(n, n + 1), "v": z(n + 1, n), "uprev": z(n, n + 1), "vprev": z(n + 1, n),
                "ctx": z(2, n + 2, n + 2)}

    def _store(self):
        s = self._zeros_state("cuda")
        self.ctx.store_state(s["u"], s["v"], s["uprev"], s["vprev"], s["ctx"])
        return s

    def _load(self, s):
        self.ctx.load_state(s["u"], s["v"], s["uprev"], s["vprev"], s["ctx"])

    def forward(self, lid: torch.Tensor, u0: torch.Tensor, v0: torch.Tensor):
        n = self.n
        self.lid = lid.detach().to(torch.float64).contiguous()
        self.ctx.set_lid_profile(self.lid)
        s = self._zeros_state("cuda")
        s["u"].copy_(u0.detach())
        s["v"].copy_(v0.detach())
        s["uprev"].copy_(u0.detach())
        s["vprev"].copy_(v0.detach())
        self._load(s)
        self.checkpoints, self.fwd_stats = [], []
        done = 0
        while done < self.nsteps:
            k = min(self.every, self.nsteps - done)
            self.checkpoints.append((done, k, self._store()))
            self.fwd_stats += self.ctx.advance(k)
            done += k
        out = self._store()
        assert out["u"].shape == (n, n + 1)
        return out["u"], out["v"]

    def backward(self, gu: torch.Tensor, gv: torch.Tensor):
        """adjoint of (u_N, v_N) -> (g_bar, u0_bar, v0_bar)"""
        n = self.n
        dev = gu.device
        glid = torch.zeros(n, dtype=torch.float64, device=dev)
        self.ctx.load_dual(gu.detach().contiguous(), gv.detach().contiguous(), None, None, None, glid)
        self.bwd_stats = []
        for start, k, state in reversed(self.checkpoints):
            self._load(state)
            self.tape.reset()
            self.ctx.step_fwd(k, self.tape)
            self.bwd_stats += self.ctx.step_bwd(self.tape, k)
        z = self._zeros_state(dev)
        self.ctx.store_dual(z["u"], z["v"], z["uprev"], z["vprev"], None, glid)
        return glid, z["u"] + z["uprev"], z["v"] + z["vprev"]


class CavityRollout(torch.autograd.Function):
    """(lid, u0, v0) -> (u_N, v_N) through the HIP step; backward = discrete adjoint."""

    @staticmethod
    def forward(fctx, lid, u0, v0, rollout: CheckpointedRollout):
        fctx.rollout = rollout
        return rollout.forward(lid, u0, v0)

    @staticmethod
    def backward(fctx, gu, gv):
        if gu is None:
            gu = torch.zeros((fctx.rollout.n, fctx.rollout.n + 1), dtype=torch.float64, device="cuda")
        if gv is None:
            gv = torch.zeros((fctx.rollout.n + 1, fctx.rollout.n), dtype=torch.float64, device="cuda")
        glid, gu0, gv0 = fctx.rollout.backward(gu.contiguous(), gv.contiguous())
        return glid, gu0, gv0, None


def kinetic_energy_lower_left(u: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """J = sum over the cells of [0,1/2]^2 of (U^2 + V^2)/2 dx dy with the cell
    velocities U = (u_i + u_{i+1})/2, V = (v_j + v_{j+1})/2 (unit square)."""
    n = u.shape[0]
    q = n // 2
    U = 0.5 * (u[:q, :q] + u[:q, 1:q + 1])
    V = 0.5 * (v[:q, :q] + v[1:q + 1, :q])
    return 0.5 * (U * U + V * V).sum() / (n * n)
