"""Text outputs byte-compatible with the reference's example programs.

The reference's Sod drivers write their state as whitespace-separated text
(`.dat`), which its plotting scripts read:

* ``write_sod1d``  — ``sod_NNNN.dat`` of examples/sod1d/sod1d.cpp:50-91:
  columns x, rho, u, p, e per interior cell;
* ``write_sod2d``  — ``sod2d_NNNN.dat`` of examples/sod2d/sod2d.cpp:58-100:
  columns x, y, rho, u, v, p, cells in patch order (py outer, px inner),
  each patch's interior with the last index fastest;
* ``write_sod2d_xslice`` — ``sod2d_xslice_NNNN.dat`` of sod2d.cpp:102-152:
  the row j = ny // 2, columns x, rho, u, p, sorted by x.

Derived quantities follow the reference's arithmetic operation by operation
(cons_to_prim, euler1d.hpp:69-84, with the p >= 0 clamp in 1D; the raw
formulas of the 2D writers, no floors), and numbers are printed as
``std::scientific`` with precision 10 prints them (``%.10e``), so the files
are byte-identical to the reference's for the same state. Inputs are the
interior conserved states as float64 arrays: [3, n] (rho, rho*u, E) in 1D and
[4, nx, ny] (rho, rho*u, rho*v, E) in 2D (Sim2D.gather / Solver1D downloads
give them).
"""
from __future__ import annotations

import os

import numpy as np

_FMT = "%.10e"


def _partition(total: int, n: int):
    """subspace() partition of `total` cells into `n` parts: (starts, lengths)."""
    base, extra = divmod(total, n)
    starts = [p * base + min(p, extra) for p in range(n)]
    lens = [base + (1 if p < extra else 0) for p in range(n)]
    return starts, lens


def _rows(columns) -> str:
    """Lines of space-separated %.10e numbers from equal-length 1D arrays."""
    if len(columns[0]) == 0:
        return ""
    stacked = np.column_stack(columns)
    line = " ".join([_FMT] * len(columns)) + "\n"
    return "".join(line % tuple(r) for r in stacked.tolist())


def write_sod1d(directory: str, U: np.ndarray, time: float, step: int, x_min: float = 0.0, x_max: float = 1.0,
                gamma: float = 1.4) -> str:
    """sod_NNNN.dat of the 1D state U[3, n]; returns the path written."""
    U = np.asarray(U, dtype=np.float64)
    n = U.shape[1]
    dx = (x_max - x_min) / n
    rho, mom, erg = U[0], U[1], U[2]
    u = mom / rho
    # eos_t::pressure (euler1d.hpp:31-44): ke = 0.5 * rho * u * u, p >= 0
    ke = 0.5 * rho * u * u
    p = (gamma - 1.0) * (erg - ke)
    p = np.where(p < 0.0, 0.0, p)
    e = p / (rho * (gamma - 1.0))
    x = x_min + (np.arange(n, dtype=np.float64) + 0.5) * dx
    path = os.path.join(directory, "sod_%04d.dat" % step)
    with open(path, "w") as f:
        f.write("# Sod shock tube at t = " + _FMT % time + "\n")
        f.write("# x rho u p e\n")
        f.write(_rows([x, rho, u, p, e]))
    return path


def _prims2d(U, gamma):
    rho = U[0]
    u = U[1] / rho
    v = U[2] / rho
    ke = 0.5 * rho * (u * u + v * v)
    p = (gamma - 1.0) * (U[3] - ke)
    return rho, u, v, p


def write_sod2d(directory: str, G: np.ndarray, time: float, step: int, npx: int = 1, npy: int = 1,
                x_min: float = 0.0, x_max: float = 1.0, y_min: float = 0.0, y_max: float = 0.1,
                gamma: float = 1.4) -> str:
    """sod2d_NNNN.dat of the 2D state G[4, nx, ny] on an npx x npy patch
    mesh (the patch order of the output); returns the path written."""
    G = np.asarray(G, dtype=np.float64)
    nx, ny = G.shape[1], G.shape[2]
    dx = (x_max - x_min) / nx
    dy = (y_max - y_min) / ny
    rho, u, v, p = _prims2d(G, gamma)
    xs, lxs = _partition(nx, npx)
    ys, lys = _partition(ny, npy)
    path = os.path.join(directory, "sod2d_%04d.dat" % step)
    with open(path, "w") as f:
        f.write("# 2D Sod shock tube at t = " + _FMT % time + "\n")
        f.write("# x y rho u v p\n")
        for py in range(npy):
            for px in range(npx):
                i = np.arange(xs[px], xs[px] + lxs[px])
                j = np.arange(ys[py], ys[py] + lys[py])
                I, J = np.meshgrid(i, j, indexing="ij")
                I, J = I.ravel(), J.ravel()
                x = x_min + (I.astype(np.float64) + 0.5) * dx
                y = y_min + (J.astype(np.float64) + 0.5) * dy
                f.write(_rows([x, y, rho[I, J], u[I, J], v[I, J], p[I, J]]))
    return path


def write_sod2d_xslice(directory: str, G: np.ndarray, time: float, step: int, x_min: float = 0.0,
                       x_max: float = 1.0, gamma: float = 1.4) -> str:
    """sod2d_xslice_NNNN.dat: the row j = ny // 2 of G[4, nx, ny], sorted by
    x; returns the path written."""
    G = np.asarray(G, dtype=np.float64)
    nx, ny = G.shape[1], G.shape[2]
    dx = (x_max - x_min) / nx
    j = ny // 2
    rho, u, _, p = _prims2d(G[:, :, j:j + 1], gamma)
    x = x_min + (np.arange(nx, dtype=np.float64) + 0.5) * dx
    path = os.path.join(directory, "sod2d_xslice_%04d.dat" % step)
    with open(path, "w") as f:
        f.write("# x-slice at y_mid, t = " + _FMT % time + "\n")
        f.write("# x rho u p\n")
        f.write(_rows([x, rho[:, 0], u[:, 0], p[:, 0]]))
    return path
