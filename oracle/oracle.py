"""TEST INFRASTRUCTURE ONLY — ctypes bindings of the parity oracle.

Loads ``oracle/liboracle.so`` (the CPU restatement of nest's Euler numerics,
built by ``oracle/Makefile``) and, when present, ``oracle/_ref/libnest_ref.so``
(the reference's own headers driven by ``oracle/ref_harness.cpp``). Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module, and only as the comparison reference / baseline — never as the thing
measured or shipped.

All arrays are float64 numpy arrays in field_t SoA layout (variable-major,
last index fastest), extended space including the halo.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnest_ref.so")

LIM = {"minmod": 0, "mc": 1, "none": 2}
BC = {"periodic": 0, "outflow": 1, "reflecting": 2}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_d, _i, _sz = C.c_double, C.c_int, C.c_size_t


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and the reference harness if asked and present)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Oracle:
    """Thin numpy front end over liboracle.so."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.lib = L
        s = lambda n, r, a: _sig(L, n, r, a)  # noqa: E731
        self.set_threads = s("or_set_threads", None, [_i])
        self.get_threads = s("or_get_threads", _i, [])
        self.eos_pressure = s("or_eos_pressure", _d, [_d, _d, _d, _d])
        self.eos_total_energy = s("or_eos_total_energy", _d, [_d, _d, _d, _d])
        self.eos_sound_speed = s("or_eos_sound_speed", _d, [_d, _d, _d])
        self._c2p = s("or_cons_to_prim", None, [_dp, _dp, _d])
        self._p2c = s("or_prim_to_cons", None, [_dp, _dp, _d])
        self.minmod = s("or_minmod", _d, [_d, _d])
        self.mc = s("or_mc", _d, [_d, _d])
        self.limiter = s("or_limiter", _d, [_d, _d, _i])
        self.interface_velocity = s("or_interface_velocity", _d, [_d, _d, _d, _d])
        self.passive_flux = s("or_passive_flux", _d, [_d, _d, _d])
        self._hlle_b = s("or_hlle_batch", None, [_i, _dp, _dp, _d, _dp])
        self._plm_b = s("or_plm_batch", None, [_i, _dp, _dp, _dp, _i, _dp, _dp])
        self._sweep_cell = s("or_sweep_cell", None, [_dp, _d, _i, _i, _d, _dp])
        self._fill_1d = s("or_fill_1d", None, [_dp, _i, _i, _i, _i])
        self._stage_1d = s("or_stage_1d", None, [_dp, _dp, _i, _i, _d, _d, _d, _i])
        self._max_speed_1d = s("or_max_speed_1d", _d, [_dp, _i, _i, _d])
        self._rk2_1d = s("or_rk2_step_1d", None, [_dp, _i, _i, _i, _d, _d, _d, _i])
        self._run_1d = s("or_run_1d", _i, [_dp, _i, _i, _i, _d, _d, _d, _d, _i, _i, _dp])
        self._sweep_2d = s("or_sweep_2d", None, [_dp, _dp, _i, _i, _i, _d, _d, _d, _i, _i])
        self._dt2d = s("or_dt_2d_partial", None, [_dp, _i, _i, _i, _d, _d, _d, _dp])
        self._msize = s("or_mesh_size", _sz, [_i, _i, _i, _i, _i])
        self._mfill = s("or_mesh_fill", None, [_dp, _i, _i, _i, _i, _i, _i, _i])
        self._msweep = s("or_mesh_sweep", None, [_dp, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _i, _i])
        self._msplit = s("or_mesh_split_stage", None, [_dp, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _i])
        self._mrk2 = s("or_mesh_rk2_step", None, [_dp, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _i])
        self._mdt = s("or_mesh_dt", _d, [_dp, _i, _i, _i, _i, _i, _d, _d, _d, _d])
        self._mrun = s("or_mesh_run", _i, [_dp, _i, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _dp])
        self._mjvp = s("or_mesh_rk2_jvp", None,
                       [_dp, _dp, _i, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _i])
        self._msjvp = s("or_mesh_stage_jvp", None,
                        [_dp, _dp, _i, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _i])

    # ---- scalar / batched numerics ----
    def cons_to_prim(self, U, gamma=1.4):
        W = np.zeros(3)
        self._c2p(np.ascontiguousarray(U, dtype=np.float64), W, gamma)
        return W

    def prim_to_cons(self, W, gamma=1.4):
        U = np.zeros(3)
        self._p2c(np.ascontiguousarray(W, dtype=np.float64), U, gamma)
        return U

    def hlle(self, WL, WR, gamma=1.4):
        WL = np.ascontiguousarray(WL, dtype=np.float64).reshape(-1, 3)
        WR = np.ascontiguousarray(WR, dtype=np.float64).reshape(-1, 3)
        F = np.zeros_like(WL)
        self._hlle_b(WL.shape[0], WL, WR, gamma, F)
        return F

    def plm(self, Wm, W0, Wp, lim="minmod"):
        Wm, W0, Wp = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3) for a in (Wm, W0, Wp))
        WL, WR = np.zeros_like(W0), np.zeros_like(W0)
        self._plm_b(W0.shape[0], Wm, W0, Wp, LIM[lim], WL, WR)
        return WL, WR

    def sweep_cell(self, S, dtdx, xsweep=True, gamma=1.4, lim="minmod"):
        out = np.zeros(4)
        self._sweep_cell(np.ascontiguousarray(S, dtype=np.float64).reshape(5, 4), gamma, LIM[lim],
                         int(xsweep), dtdx, out)
        return out

    # ---- 1D ----
    def fill_1d(self, U, n, h=2, bc="outflow", nvars=3):
        self._fill_1d(U, n, h, nvars, BC[bc])

    def stage_1d(self, Uin, n, dt, dx, h=2, gamma=1.4, lim="minmod"):
        Uout = Uin.copy()
        self._stage_1d(Uin, Uout, n, h, dt, dx, gamma, LIM[lim])
        return Uout

    def max_speed_1d(self, U, n, h=2, gamma=1.4):
        return self._max_speed_1d(U, n, h, gamma)

    def rk2_step_1d(self, U, n, dt, dx, h=2, bc="outflow", gamma=1.4, lim="minmod"):
        self._rk2_1d(U, n, h, BC[bc], dt, dx, gamma, LIM[lim])

    def run_1d(self, U, n, t_end, cfl, dx, h=2, bc="outflow", gamma=1.4, lim="minmod", max_steps=1 << 20):
        dts = np.zeros(max_steps)
        k = self._run_1d(U, n, h, BC[bc], t_end, cfl, dx, gamma, LIM[lim], max_steps, dts)
        return k, dts[: max(k, 0)]

    # ---- 2D single patch ----
    def sweep_2d(self, Uin, nx, ny, dt, dxy, xsweep=True, h=2, gamma=1.4, lim="minmod"):
        Uout = Uin.copy()
        self._sweep_2d(Uin, Uout, nx, ny, h, dt, dxy, gamma, LIM[lim], int(xsweep))
        return Uout

    def dt_2d_partial(self, U, nx, ny, dx, dy, h=2, gamma=1.4):
        out = np.zeros(2)
        self._dt2d(U, nx, ny, h, dx, dy, gamma, out)
        return out

    # ---- 2D patch mesh ----
    def mesh_size(self, nx, ny, npx, npy, h=2):
        return int(self._msize(nx, ny, npx, npy, h))

    def mesh_fill(self, U, m):
        self._mfill(U, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy])

    def mesh_sweep(self, U, m, dt, dxy, xsweep, gamma=1.4, lim="minmod"):
        self._msweep(U, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], dt, dxy, gamma, LIM[lim],
                     int(xsweep))

    def mesh_split_stage(self, U, m, dt, dx, dy, gamma=1.4, lim="minmod"):
        self._msplit(U, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], dt, dx, dy, gamma, LIM[lim])

    def mesh_rk2_step(self, U, m, dt, dx, dy, gamma=1.4, lim="minmod"):
        self._mrk2(U, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], dt, dx, dy, gamma, LIM[lim])

    def mesh_dt(self, U, m, dx, dy, cfl, gamma=1.4):
        return self._mdt(U, m.nx, m.ny, m.npx, m.npy, m.h, dx, dy, cfl, gamma)

    def mesh_run(self, U, m, nsteps, dx, dy, cfl=0.4, t_end=float("inf"), fixed_dt=0.0, gamma=1.4,
                 lim="minmod"):
        dts = np.zeros(max(nsteps, 1))
        k = self._mrun(U, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], nsteps, t_end, cfl, fixed_dt,
                       dx, dy, gamma, LIM[lim], dts)
        return k, dts[:k]

    def mesh_rk2_jvp(self, U, dU, m, K, dt, dx, dy, gamma=1.4, lim="minmod"):
        self._mjvp(U, dU, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], K, dt, dx, dy, gamma, LIM[lim])

    def mesh_stage_jvp(self, U, dU, m, dt, dx, dy, gamma=1.4, lim="minmod"):
        """One split stage S = Y o fill o X o fill on dual numbers: U <- S(U), dU <- S'(U) dU."""
        self._msjvp(U, dU, m.nx, m.ny, m.npx, m.npy, m.h, BC[m.bcx], BC[m.bcy], dt, dx, dy, gamma, LIM[lim])


class RefHarness:
    """The reference's own headers (built into oracle/_ref, container only)."""

    def __init__(self, path: str = REF_PATH):
        L = C.CDLL(path)
        s = lambda n, r, a: _sig(L, n, r, a)  # noqa: E731
        self.eos_pressure = s("ref_eos_pressure", _d, [_d, _d, _d, _d])
        self.eos_total_energy = s("ref_eos_total_energy", _d, [_d, _d, _d, _d])
        self.eos_sound_speed = s("ref_eos_sound_speed", _d, [_d, _d, _d])
        self.minmod = s("ref_minmod", _d, [_d, _d])
        self.mc = s("ref_mc", _d, [_d, _d])
        self.limiter = s("ref_limiter", _d, [_d, _d, _i])
        self.interface_velocity = s("ref_interface_velocity", _d, [_d, _d, _d, _d])
        self.passive_flux = s("ref_passive_flux", _d, [_d, _d, _d])
        self._hlle_b = s("ref_hlle_batch", None, [_i, _dp, _dp, _d, _dp])
        self._plm_b = s("ref_plm_batch", None, [_i, _dp, _dp, _dp, _i, _dp, _dp])
        self._rhs_1d = s("ref_rhs_1d", None, [_dp, _dp, _i, _i, _d, _d, _d, _i])
        self._sweep_2d = s("ref_sweep_2d", None, [_dp, _dp, _i, _i, _i, _d, _d, _d, _i, _i])
        self._pmesh = s("ref_periodic_mesh", _d, [_dp, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _d, _i])

    def hlle(self, WL, WR, gamma=1.4):
        WL = np.ascontiguousarray(WL, dtype=np.float64).reshape(-1, 3)
        WR = np.ascontiguousarray(WR, dtype=np.float64).reshape(-1, 3)
        F = np.zeros_like(WL)
        self._hlle_b(WL.shape[0], WL, WR, gamma, F)
        return F

    def plm(self, Wm, W0, Wp, lim="minmod"):
        Wm, W0, Wp = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3) for a in (Wm, W0, Wp))
        WL, WR = np.zeros_like(W0), np.zeros_like(W0)
        self._plm_b(W0.shape[0], Wm, W0, Wp, LIM[lim], WL, WR)
        return WL, WR

    def rhs_1d(self, Uin, n, dt, dx, h=2, gamma=1.4, lim="minmod"):
        Uout = Uin.copy()
        self._rhs_1d(Uin, Uout, n, h, dt, dx, gamma, LIM[lim])
        return Uout

    def sweep_2d(self, Uin, nx, ny, dt, dxy, xsweep=True, h=2, gamma=1.4, lim="minmod"):
        Uout = Uin.copy()
        self._sweep_2d(Uin, Uout, nx, ny, h, dt, dxy, gamma, LIM[lim], int(xsweep))
        return Uout

    def periodic_mesh(self, U, nx, ny, npx, npy, mode, dt=0.0, dx=1.0, dy=1.0, cfl=0.4, h=2, gamma=1.4,
                      lim="minmod"):
        return self._pmesh(U, nx, ny, npx, npy, h, mode, dt, dx, dy, cfl, gamma, LIM[lim])


class MeshSpec:
    """Plain description of a 2D patch mesh (tensor-product partition)."""

    def __init__(self, nx, ny, npx=1, npy=1, h=2, bcx="outflow", bcy="outflow"):
        self.nx, self.ny, self.npx, self.npy, self.h = nx, ny, npx, npy, h
        self.bcx, self.bcy = bcx, bcy

    def part(self, total, n):
        base, extra = divmod(total, n)
        starts, lens = [], []
        for p in range(n):
            starts.append(p * base + min(p, extra))
            lens.append(base + (1 if p < extra else 0))
        return starts, lens

    def patches(self):
        """[(index, x0, lx, y0, ly, offset)] in patch-index order (py*npx + px)."""
        xs, lxs = self.part(self.nx, self.npx)
        ys, lys = self.part(self.ny, self.npy)
        out, off = [], 0
        for py in range(self.npy):
            for px in range(self.npx):
                out.append((py * self.npx + px, xs[px], lxs[px], ys[py], lys[py], off))
                off += 4 * (lxs[px] + 2 * self.h) * (lys[py] + 2 * self.h)
        return out

    @property
    def total(self):
        p = self.patches()[-1]
        return p[5] + 4 * (p[2] + 2 * self.h) * (p[4] + 2 * self.h)

    def to_patches(self, G):
        """Scatter a global interior array G[4, nx, ny] into the concatenated mesh buffer."""
        h = self.h
        U = np.zeros(self.total)
        for (_, x0, lx, y0, ly, off) in self.patches():
            a = U[off: off + 4 * (lx + 2 * h) * (ly + 2 * h)].reshape(4, lx + 2 * h, ly + 2 * h)
            a[:, h:h + lx, h:h + ly] = G[:, x0:x0 + lx, y0:y0 + ly]
        return U

    def to_global(self, U):
        h = self.h
        G = np.zeros((4, self.nx, self.ny))
        for (_, x0, lx, y0, ly, off) in self.patches():
            a = U[off: off + 4 * (lx + 2 * h) * (ly + 2 * h)].reshape(4, lx + 2 * h, ly + 2 * h)
            G[:, x0:x0 + lx, y0:y0 + ly] = a[:, h:h + lx, h:h + ly]
        return G
