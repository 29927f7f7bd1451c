"""Differentiable fixed-dt rollouts of the 2D Euler step for torch.

``Rollout`` runs K fixed-dt RK2 steps of a patch mesh (or of a batch of
independent problems) on the GPU through the C ABI and is differentiable
with respect to the initial conserved state: the backward pass is the
library's hand-written discrete adjoint (nest_hip_sim2d_rollout_adjoint),
which reverses the stored rollout. By default every state of the rollout
is kept in device memory when 2K states fit in half of the free memory (the
reverse then recomputes nothing); otherwise it goes interval by interval,
recomputing each interval's states from its checkpoint (intervals of
ceil(sqrt(K)) steps; the forward keeps the states of the last four).
Derivatives are exact for the discrete map, branch by branch (limiter cases,
HLLE upwind / subsonic branch, Davis speeds, floors, the periodic /
zero-gradient halo fill).

State tensors are float64 CUDA tensors in the problem-major interior layout:
[4, nx, ny] for a mesh, [B, 4, nx, ny] for a batch of B equal-size problems
(variables rho, rho*u, rho*v, E; y contiguous). The library's work runs on a
stream of the Rollout's own, ordered after torch's current stream on entry
and before it on exit (no host synchronisation).
"""
from __future__ import annotations

import torch

from .euler import Sim2D


def _check_bc(bc: str) -> None:
    """The reverse sweep differentiates periodic and outflow boundaries; the
    library rejects a reflecting wall too (NEST_ERR_UNSUPPORTED), this says so
    before any device work."""
    if bc not in ("periodic", "outflow"):
        raise ValueError(f"adjoint rollouts support bc 'periodic' or 'outflow', not {bc!r}")


class Rollout:
    """K fixed-dt RK2 steps as a differentiable map U0 -> U_K."""

    def __init__(self, sim: Sim2D, shape, nsteps: int, dt: float, interval: int = 0):
        self.sim = sim
        self.shape = tuple(shape)
        self.nsteps, self.dt, self.interval = int(nsteps), float(dt), int(interval)
        self._auto = None
        self._token = 0
        self.stream = torch.cuda.Stream()  # default priority
        sim.set_stream(self.stream.cuda_stream)
        n = 1
        for s in self.shape:
            n *= s
        if n != sim.global_size():
            raise ValueError(f"shape {self.shape} does not hold the handle's {sim.global_size()} interior values")

    @classmethod
    def mesh(cls, nx, ny, npx=1, npy=1, nsteps=1, dt=1e-3, bc="periodic", limiter="minmod", mode="fast",
             interval=0, gamma=1.4, tune=False):
        _check_bc(bc)
        sim = Sim2D.mesh(nx, ny, npx, npy, gamma=gamma, bcx=bc, bcy=bc, limiter=limiter, mode=mode)
        ro = cls(sim, (4, nx, ny), nsteps, dt, interval)
        if tune:
            ro.tune([(nx, ny)])
        return ro

    @classmethod
    def batch(cls, nprob, nx, ny, nsteps=1, dt=1e-3, bc="periodic", limiter="minmod", mode="fast", interval=0,
              gamma=1.4, tune=False):
        _check_bc(bc)
        sim = Sim2D.batch([(nx, ny)] * nprob, gamma=gamma, bc=bc, limiter=limiter, mode=mode)
        ro = cls(sim, (nprob, 4, nx, ny), nsteps, dt, interval)
        if tune:
            ro.tune([(nx, ny)] * nprob)
        return ro

    def rollout_interval(self) -> int:
        """The interval handed to the library: `interval` as given (> 0:
        checkpoints every `interval` steps; < 0: every state kept), except 0,
        the automatic choice: every state kept (-1) when the 2K states it
        needs fit in half of the free device memory, else checkpoints every
        ceil(sqrt(K)) steps (0). The choice is made at the first call and kept,
        so later calls reuse the buffers it allocated."""
        if self.interval != 0:
            return self.interval
        if self._auto is None:
            need = 2 * self.nsteps * self.sim.total * 8
            free, _ = torch.cuda.mem_get_info()
            self._auto = -1 if need <= free // 2 else 0
        return self._auto

    def tune(self, sizes, log=None):
        """Autotune the forward kernels the rollouts and their recomputes run:
        the one-launch step kernel (which also writes the stage-1 states the
        reverse sweep reads; rollouts whose intervals are all kept, or that
        keep a single checkpoint) and the two stage kernels (rollouts whose
        intervals are recomputed alongside the reverse sweep); and the adjoint
        kernels' tiling on this handle (results cached per device and
        problem-size signature). The
        handle's state and recorded rollout are overwritten (every forward
        call reloads the state); tune before a forward, not between a forward
        and its backward."""
        from . import autotune
        self.sim.start(cfl=0.4)
        fwd = autotune.tune(self.sim, sizes, log=log, kernels="stages")
        fwd.update(autotune.tune(self.sim, sizes, log=log, kernels="step"))
        adj = autotune.tune_adjoint(self.sim, sizes, self.dt, log=log, keep_all=self.rollout_interval() < 0)
        return fwd, adj

    def __call__(self, U0: torch.Tensor) -> torch.Tensor:
        if U0.dtype != torch.float64 or not U0.is_cuda or tuple(U0.shape) != self.shape:
            raise ValueError(f"expected a float64 CUDA tensor of shape {self.shape}")
        return _RolloutFn.apply(U0, self)

    def close(self) -> None:
        self.sim.close()


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, ro):
        sim = ro.sim
        U0c = U0.detach().contiguous()
        out = torch.empty_like(U0c)
        cur = torch.cuda.current_stream(U0.device)
        ro.stream.wait_stream(cur)
        Uptr, _ = sim.device_ptrs()
        sim.unpack(U0c.data_ptr(), Uptr)
        sim.rollout(ro.nsteps, ro.dt, ro.rollout_interval())
        sim.pack(Uptr, out.data_ptr())
        for t in (U0c, out):
            t.record_stream(ro.stream)
        cur.wait_stream(ro.stream)
        ro._token += 1
        ctx.ro = ro
        ctx.token = ro._token
        return out

    @staticmethod
    def backward(ctx, g):
        ro = ctx.ro
        if ro._token != ctx.token:
            raise RuntimeError("this rollout's checkpoints were replaced by a later forward call on the same "
                               "Rollout; run backward before the next forward, or use one Rollout per graph")
        sim = ro.sim
        gc = g.detach().contiguous()
        Ub = torch.zeros(sim.total, dtype=torch.float64, device=g.device)
        g0 = torch.empty_like(gc)
        cur = torch.cuda.current_stream(g.device)
        ro.stream.wait_stream(cur)
        sim.unpack(gc.data_ptr(), Ub.data_ptr())
        sim.rollout_adjoint(Ub.data_ptr())
        sim.pack(Ub.data_ptr(), g0.data_ptr())
        for t in (gc, Ub, g0):
            t.record_stream(ro.stream)
        cur.wait_stream(ro.stream)
        return g0, None


def density_misfit(U: torch.Tensor, target_rho: torch.Tensor, cell_area: float) -> torch.Tensor:
    """0.5 * sum((rho - rho_target)^2) * cell area: the L2 misfit of the
    final density (dim -3 of U is the variable)."""
    d = U.select(-3, 0) - target_rho
    return 0.5 * cell_area * torch.sum(d * d)
