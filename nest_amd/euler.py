"""Python front end of the gfx950 Euler backend (over the C ABI).

Mirrors the reference's host objects for this path:
  * ``Solver1D``  — the CUDA 1D solver object behind
    src/backend_cuda/euler1d_solver.hpp:27-87 (create / upload / download /
    compute_dt / rk2_step / run);
  * ``Sim2D``     — the 2D patch-mesh driver loop of
    examples/sod2d/sod2d.cpp:318-356 (exchange, x_sweep_t, exchange,
    y_sweep_t; euler2d.hpp:67-239) with compute_dt_2d (euler2d.hpp:242-271),
    stepped as SSP-RK2 over the X-then-Y split stage; ``Sim2D.batch`` holds
    many independent problems stepped together.

Arrays are float64 numpy arrays in field_t layout (variable-major, last
index contiguous) over each patch's extended space (halo 2).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC, LIMITER, MODE, check

_pd = C.POINTER(C.c_double)


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pd)


class Solver1D:
    """Single 1D patch, n interior cells, halo 2, 3 conserved variables."""

    def __init__(self, n: int, dx: float, gamma: float = 1.4, bc: str = "outflow", limiter: str = "minmod",
                 mode: str = "exact"):
        self.L = _lib.load()
        self.n, self.dx = n, dx
        h = C.c_void_p()
        check(self.L.nest_hip_solver1d_create(n, dx, gamma, BC[bc], LIMITER[limiter], MODE[mode], C.byref(h)))
        self.h = h

    @property
    def size(self) -> int:
        return 3 * (self.n + 4)

    def upload(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        assert U.size == self.size
        check(self.L.nest_hip_solver1d_upload(self.h, _dptr(U)))

    def download(self) -> np.ndarray:
        U = np.empty(self.size)
        check(self.L.nest_hip_solver1d_download(self.h, _dptr(U)))
        return U

    def compute_dt(self, cfl: float) -> float:
        dt = C.c_double()
        check(self.L.nest_hip_solver1d_compute_dt(self.h, cfl, C.byref(dt)))
        return dt.value

    def rk2_step(self, dt: float) -> None:
        check(self.L.nest_hip_solver1d_rk2_step(self.h, dt))

    def write_dat(self, directory: str, time: float, step: int, x_min: float = 0.0, x_max: float = 1.0,
                  gamma: float = 1.4) -> str:
        """sod_NNNN.dat of the current state, byte-compatible with the
        reference's sod1d writer (nest_amd.dat.write_sod1d)."""
        from . import dat
        return dat.write_sod1d(directory, self.download().reshape(3, -1)[:, 2:-2], time, step, x_min, x_max, gamma)

    def set_step_kernel(self, kind: str = "auto") -> None:
        """Kernel of the fused device-dt steps: "auto", "tile" (LDS tile per
        block), "lanes" (each wave on its own 56-cell segments, one cell per
        lane) or "pairs" (120-cell segments, two cells per lane)."""
        kinds = {"auto": 0, "tile": 1, "lanes": 2, "pairs": 3}
        check(self.L.nest_hip_solver1d_set_step_kernel(self.h, kinds[kind]))

    def stage(self, stage: int, dt: float) -> None:
        """One SSP-RK2 stage with the given dt (1: U1 = S(U); 2: U = (U + S(U1)) / 2)."""
        check(self.L.nest_hip_solver1d_stage(self.h, stage, dt))

    def exchange(self, buffer: int = 0) -> None:
        """Ghost fill of U (0) or U1 (1) through the boundary condition."""
        check(self.L.nest_hip_solver1d_exchange(self.h, buffer))

    def start(self, cfl: float = 0.4, t_end: float = float("inf")) -> None:
        check(self.L.nest_hip_solver1d_start(self.h, cfl, t_end))

    def advance(self, nsteps: int) -> None:
        """nsteps device-dt steps (no host round trip); steps past t_end are no-ops."""
        check(self.L.nest_hip_solver1d_advance(self.h, nsteps))

    def status(self):
        t, dt = C.c_double(), C.c_double()
        steps, done = C.c_int(), C.c_int()
        check(self.L.nest_hip_solver1d_status(self.h, C.byref(t), C.byref(dt), C.byref(steps), C.byref(done)))
        return {"t": t.value, "dt": dt.value, "steps": steps.value, "done": done.value}

    def run(self, t_final: float, cfl: float) -> int:
        k = C.c_int()
        check(self.L.nest_hip_solver1d_run(self.h, t_final, cfl, C.byref(k)))
        return k.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nest_hip_solver1d_destroy(self.h)
            self.h = None

    __del__ = close


class Sim2D:
    """A 2D patch mesh (one problem) or a batch of independent problems."""

    def __init__(self, handle: C.c_void_p):
        self.L = _lib.load()
        self.h = handle
        total, npatch, nprob, nitems = C.c_size_t(), C.c_int(), C.c_int(), C.c_int()
        check(self.L.nest_hip_sim2d_info(self.h, C.byref(total), C.byref(npatch), C.byref(nprob), C.byref(nitems)))
        self.total, self.npatch, self.nprob, self.nitems = total.value, npatch.value, nprob.value, nitems.value
        self.patches = []
        for p in range(self.npatch):
            off, nx, ny, prob = C.c_longlong(), C.c_int(), C.c_int(), C.c_int()
            check(self.L.nest_hip_sim2d_patch_info(self.h, p, C.byref(off), C.byref(nx), C.byref(ny),
                                                   C.byref(prob)))
            self.patches.append((off.value, nx.value, ny.value, prob.value))

    @classmethod
    def mesh(cls, nx, ny, npx=1, npy=1, dx=None, dy=None, gamma=1.4, bcx="outflow", bcy="outflow",
             limiter="minmod", mode="exact"):
        L = _lib.load()
        dx = 1.0 / nx if dx is None else dx
        dy = 1.0 / ny if dy is None else dy
        h = C.c_void_p()
        check(L.nest_hip_mesh2d_create(nx, ny, npx, npy, dx, dy, gamma, BC[bcx], BC[bcy], LIMITER[limiter],
                                       MODE[mode], C.byref(h)))
        sim = cls(h)
        sim.grid = (nx, ny, npx, npy)
        return sim

    @classmethod
    def batch(cls, sizes, dx=None, dy=None, gamma=1.4, bc="outflow", limiter="minmod", mode="exact"):
        L = _lib.load()
        n = len(sizes)
        nx = (C.c_int * n)(*[s[0] for s in sizes])
        ny = (C.c_int * n)(*[s[1] for s in sizes])
        dxs = (C.c_double * n)(*(dx if dx is not None else [1.0 / s[0] for s in sizes]))
        dys = (C.c_double * n)(*(dy if dy is not None else [1.0 / s[1] for s in sizes]))
        h = C.c_void_p()
        check(L.nest_hip_batch2d_create(n, nx, ny, dxs, dys, gamma, BC[bc], LIMITER[limiter], MODE[mode],
                                        C.byref(h)))
        return cls(h)

    # ---- data movement ----
    def upload(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        assert U.size == self.total, (U.size, self.total)
        check(self.L.nest_hip_sim2d_upload(self.h, _dptr(U)))

    def download(self) -> np.ndarray:
        U = np.empty(self.total)
        check(self.L.nest_hip_sim2d_download(self.h, _dptr(U)))
        return U

    def patch_view(self, U: np.ndarray, p: int) -> np.ndarray:
        off, nx, ny, _ = self.patches[p]
        return U[off: off + 4 * (nx + 4) * (ny + 4)].reshape(4, nx + 4, ny + 4)

    def _origins(self):
        """(x0, y0) of every patch of a mesh (subspace() partition)."""
        nx, ny, npx, npy = self.grid

        def starts(total, n):
            base, extra = divmod(total, n)
            return [p * base + min(p, extra) for p in range(n)]

        xs, ys = starts(nx, npx), starts(ny, npy)
        return [(xs[p % npx], ys[p // npx]) for p in range(self.npatch)]

    def scatter(self, G: np.ndarray) -> np.ndarray:
        """Global interior array G[4, nx, ny] -> concatenated patch buffer (halos zero)."""
        U = np.zeros(self.total)
        for p, (x0, y0) in enumerate(self._origins()):
            _, lx, ly, _ = self.patches[p]
            self.patch_view(U, p)[:, 2:2 + lx, 2:2 + ly] = G[:, x0:x0 + lx, y0:y0 + ly]
        return U

    def gather(self, U: np.ndarray) -> np.ndarray:
        """Concatenated patch buffer -> global interior array G[4, nx, ny]."""
        nx, ny, _, _ = self.grid
        G = np.empty((4, nx, ny))
        for p, (x0, y0) in enumerate(self._origins()):
            _, lx, ly, _ = self.patches[p]
            G[:, x0:x0 + lx, y0:y0 + ly] = self.patch_view(U, p)[:, 2:2 + lx, 2:2 + ly]
        return G

    def device_ptrs(self):
        a, b = C.c_void_p(), C.c_void_p()
        check(self.L.nest_hip_sim2d_device_ptrs(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def write_dat(self, directory: str, time: float, step: int, x_min: float = 0.0, x_max: float = 1.0,
                  y_min: float = 0.0, y_max: float = 0.1, gamma: float = 1.4, xslice: bool = True):
        """sod2d_NNNN.dat (and sod2d_xslice_NNNN.dat) of a mesh's current
        state, byte-compatible with the reference's sod2d writers
        (nest_amd.dat); returns the paths written."""
        from . import dat
        nx, ny, npx, npy = self.grid
        G = self.gather(self.download())
        paths = [dat.write_sod2d(directory, G, time, step, npx, npy, x_min, x_max, y_min, y_max, gamma)]
        if xslice:
            paths.append(dat.write_sod2d_xslice(directory, G, time, step, x_min, x_max, gamma))
        return paths

    def bind(self, dev_U: int | None, dev_U1: int | None = None) -> None:
        check(self.L.nest_hip_sim2d_bind(self.h, C.c_void_p(dev_U or 0), C.c_void_p(dev_U1 or 0)))

    def set_stream(self, stream: int | None) -> None:
        check(self.L.nest_hip_sim2d_set_stream(self.h, C.c_void_p(stream or 0)))

    def set_tiling(self, rows=0, waves_per_block=0, persistent=0) -> None:
        check(self.L.nest_hip_sim2d_set_tiling(self.h, rows, waves_per_block, persistent))
        total, npatch, nprob, nitems = C.c_size_t(), C.c_int(), C.c_int(), C.c_int()
        check(self.L.nest_hip_sim2d_info(self.h, C.byref(total), C.byref(npatch), C.byref(nprob), C.byref(nitems)))
        self.nitems = nitems.value

    def set_stage_kernel(self, kind: str = "auto") -> None:
        """Which 2D stage kernel runs: "auto", "register" or "lds_ring"."""
        check(self.L.nest_hip_sim2d_set_stage_kernel(self.h, {"auto": 0, "register": 1, "lds_ring": 2}[kind]))

    def set_step_kernel(self, kind: str = "auto") -> None:
        """How a whole RK2 step runs: "auto" (one launch per step, U1 kept
        on chip, where it applies) or "two_stage" (stage 1 writes U1, stage
        2 reads it)."""
        check(self.L.nest_hip_sim2d_set_step_kernel(self.h, {"auto": 0, "two_stage": 1}[kind]))

    @property
    def one_launch(self) -> bool:
        """True when rk2_step / advance run the one-launch step kernel."""
        v = C.c_int()
        check(self.L.nest_hip_sim2d_get_step_kernel(self.h, C.byref(v)))
        return bool(v.value)

    def set_stage_tiling(self, stage: int, rows: int = 0, waves_per_block: int = 0) -> None:
        """Rows per work item and waves per block of stage 1 or 2 of the
        two-launch step, or of the one-launch step kernel (stage 3)."""
        check(self.L.nest_hip_sim2d_set_stage_tiling(self.h, stage, rows, waves_per_block))

    def get_step_tiling(self):
        """(rows, waves per block) of the one-launch step kernel."""
        r, w = C.c_int(), C.c_int()
        check(self.L.nest_hip_sim2d_get_step_tiling(self.h, C.byref(r), C.byref(w)))
        return r.value, w.value

    def set_problem_rows(self, stage: int, rows) -> None:
        """Rows per work item for each problem of a batch (None clears), of
        stage 1, 2 or the one-launch step (3)."""
        if rows is None:
            check(self.L.nest_hip_sim2d_set_problem_rows(self.h, stage, None))
            return
        arr = (C.c_int * self.nprob)(*[int(r) for r in rows])
        check(self.L.nest_hip_sim2d_set_problem_rows(self.h, stage, arr))

    def get_tiling(self):
        """((rows stage 1, rows stage 2), (waves/block stage 1, stage 2)) of the fused stages."""
        r, w = (C.c_int * 2)(), (C.c_int * 2)()
        check(self.L.nest_hip_sim2d_get_tiling(self.h, r, w))
        return (r[0], r[1]), (w[0], w[1])

    # ---- numerics ----
    def fill_halos(self) -> None:
        check(self.L.nest_hip_sim2d_fill_halos(self.h))

    def compute_dt(self, cfl: float) -> np.ndarray:
        dt = np.zeros(self.nprob)
        check(self.L.nest_hip_sim2d_compute_dt(self.h, cfl, _dptr(dt)))
        return dt

    def stage(self, stage: int, dt: float, patch: int = -1) -> None:
        check(self.L.nest_hip_sim2d_stage(self.h, stage, dt, patch))

    def rk2_step(self, dt: float) -> None:
        check(self.L.nest_hip_sim2d_rk2_step(self.h, dt))

    def start(self, cfl: float = 0.4, t_end: float = float("inf"), fixed_dt: float = 0.0) -> None:
        check(self.L.nest_hip_sim2d_start(self.h, cfl, t_end, fixed_dt))

    def advance(self, nsteps: int, graph: bool = False) -> None:
        check(self.L.nest_hip_sim2d_advance(self.h, nsteps, int(graph)))

    def advance_per_patch(self, nsteps: int, cfl: float) -> None:
        check(self.L.nest_hip_sim2d_advance_per_patch(self.h, nsteps, cfl))

    def status(self):
        t, dt = np.zeros(self.nprob), np.zeros(self.nprob)
        steps, done = np.zeros(self.nprob, dtype=np.int32), np.zeros(self.nprob, dtype=np.int32)
        check(self.L.nest_hip_sim2d_status(self.h, _dptr(t), _dptr(dt),
                                           steps.ctypes.data_as(C.POINTER(C.c_int)),
                                           done.ctypes.data_as(C.POINTER(C.c_int))))
        return {"t": t, "dt": dt, "steps": steps, "done": done}

    def advance_timed(self, nsteps: int):
        """nsteps device-dt steps with HIP events around each launch; returns
        (per-component ms [finalize, stage 1, stage 2], total ms); with the
        one-launch step [0, step kernel, 0]."""
        ms = np.zeros(3)
        tot = C.c_double()
        check(self.L.nest_hip_sim2d_advance_timed(self.h, nsteps, _dptr(ms), C.byref(tot)))
        return ms, tot.value

    def synchronize(self) -> None:
        check(self.L.nest_hip_sim2d_synchronize(self.h))

    # ---- discrete adjoint (device pointers in the patch layout) ----
    def global_size(self) -> int:
        """Doubles in the problem-major interior layout [problem][var][x][y]."""
        n = C.c_longlong()
        check(self.L.nest_hip_sim2d_global_size(self.h, C.byref(n)))
        return n.value

    def pack(self, src: int, dst: int) -> None:
        """dst (interior layout) <- src (patch layout), device pointers."""
        check(self.L.nest_hip_sim2d_pack(self.h, C.c_void_p(src), C.c_void_p(dst), 0))

    def unpack(self, src: int, dst: int) -> None:
        """dst (patch layout, interior cells) <- src (interior layout)."""
        check(self.L.nest_hip_sim2d_pack(self.h, C.c_void_p(src), C.c_void_p(dst), 1))

    def set_adjoint_tiling(self, rows_y=0, waves_y=0, rows_x=0, waves_x=0) -> None:
        check(self.L.nest_hip_sim2d_set_adjoint_tiling(self.h, rows_y, waves_y, rows_x, waves_x))

    def adjoint_stage(self, U: int, G: int, g: float, H: int | None, h: float, out: int, dt: float) -> None:
        """out = S'(U)^T (g G) + h H for one split stage."""
        check(self.L.nest_hip_sim2d_adjoint_stage(self.h, C.c_void_p(U), C.c_void_p(G), g, C.c_void_p(H or 0), h,
                                                  C.c_void_p(out), dt))

    def rk2_vjp(self, U0: int, Ubar: int, dt: float) -> None:
        """Ubar <- (dU'/dU)^T Ubar for one fixed-dt RK2 step from state U0."""
        check(self.L.nest_hip_sim2d_rk2_vjp(self.h, C.c_void_p(U0), C.c_void_p(Ubar), dt))

    def rollout(self, nsteps: int, dt: float, interval: int = 0) -> None:
        """nsteps fixed-dt steps of the state with checkpoints every
        `interval` steps (0: ceil(sqrt(nsteps)); < 0: every state kept, so
        the reverse recomputes nothing; >= nsteps: a plain forward)."""
        check(self.L.nest_hip_sim2d_rollout(self.h, nsteps, dt, interval))

    def rollout_adjoint(self, Ubar: int) -> None:
        """Cotangent of the last rollout's final state -> of its initial state, in place."""
        check(self.L.nest_hip_sim2d_rollout_adjoint(self.h, C.c_void_p(Ubar)))

    def rollout_info(self):
        k, m, dt = C.c_int(), C.c_int(), C.c_double()
        check(self.L.nest_hip_sim2d_rollout_info(self.h, C.byref(k), C.byref(m), C.byref(dt)))
        return k.value, m.value, dt.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nest_hip_sim2d_destroy(self.h)
            self.h = None

    __del__ = close
