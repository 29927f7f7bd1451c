"""Reflecting walls in the oracle, pinned by what the mathematics fixes.

The reference has periodic and outflow boundaries only; the reflecting wall is
an extension (ghost cells mirror the interior across the wall, the
wall-normal momentum negated). Nothing in the reference's tests covers it, so
the oracle's wall is pinned by:

- the method of images: a box with reflecting walls evolves exactly like the
  periodic problem on the mirror-doubled box whose second half is the mirror
  image of the first (up to the round-off by which the scheme is not bitwise
  mirror-symmetric);
- a closed box conserves mass and energy: the wall flux of each vanishes for
  mirror-image face states;
- ghost values against a brute-force image construction.

A ghost taken from the wrong interior cell, a missed or extra negation, or a
wall that copies instead of mirrors fails one of these.
"""
import math

import numpy as np
import pytest

from nest_amd import problems as P
from oracle.oracle import MeshSpec

GAM = 1.4


def _prim_state(nx, ny, seed):
    """A smooth state with nonzero velocity at every wall."""
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.0, 6)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * (a[0] * X + a[1] * Y))
    u = 0.4 + 0.3 * np.cos(2 * np.pi * (a[2] * X - 0.3 * Y))
    v = -0.3 + 0.2 * np.sin(2 * np.pi * (0.4 * X + a[3] * Y))
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (a[4] * X + a[5] * Y))
    return np.stack([rho, rho * u, rho * v, p / (GAM - 1) + 0.5 * rho * (u * u + v * v)])


def _mirror_x(Gl):
    """Image across the high-x wall: cells reversed in x, x momentum negated."""
    M = Gl[:, ::-1, :].copy()
    M[1] = -M[1]
    return M


def _mirror_y(Gl):
    M = Gl[:, :, ::-1].copy()
    M[2] = -M[2]
    return M


def _image_ghost(Gl, gi, gj, nx, ny, bcx, bcy):
    """Ghost value by explicit image construction: unfold the box into the
    periodic tiling of its mirror images (period 2n in a reflecting
    direction), clamp for outflow, wrap for periodic."""
    v = np.array(Gl[:, 0, 0])
    ix, flipx = gi, False
    if bcx == "reflecting":
        ix = gi % (2 * nx)
        if ix >= nx:
            ix, flipx = 2 * nx - 1 - ix, True
    elif bcx == "periodic":
        ix = gi % nx
    else:
        ix = min(max(gi, 0), nx - 1)
    iy, flipy = gj, False
    if bcy == "reflecting":
        iy = gj % (2 * ny)
        if iy >= ny:
            iy, flipy = 2 * ny - 1 - iy, True
    elif bcy == "periodic":
        iy = gj % ny
    else:
        iy = min(max(gj, 0), ny - 1)
    v[:] = Gl[:, ix, iy]
    if flipx:
        v[1] = -v[1]
    if flipy:
        v[2] = -v[2]
    return v


@pytest.mark.parametrize("bcx,bcy", [("reflecting", "outflow"), ("periodic", "reflecting"),
                                     ("reflecting", "reflecting")])
def test_reflecting_ghosts_are_images(oracle, bcx, bcy):
    nx, ny = 9, 7
    m = MeshSpec(nx, ny, 2, 3, bcx=bcx, bcy=bcy)
    Gl = np.random.default_rng(5).uniform(-1, 1, size=(4, nx, ny))
    U = m.to_patches(Gl)
    oracle.mesh_fill(U, m)
    for (_, x0, lx, y0, ly, off) in m.patches():
        a = U[off: off + 4 * (lx + 4) * (ly + 4)].reshape(4, lx + 4, ly + 4)
        for ei in range(lx + 4):
            for ej in range(ly + 4):
                want = _image_ghost(Gl, x0 + ei - 2, y0 + ej - 2, nx, ny, bcx, bcy)
                assert np.array_equal(a[:, ei, ej], want), (ei, ej)


@pytest.mark.parametrize("lim", ["minmod", "mc", "none"])
def test_reflecting_box_equals_mirror_doubled_periodic(oracle, lim):
    """Method of images in both directions: walls on all four sides against
    the periodic 2nx x 2ny box tiled with the state and its three images."""
    nx, ny = 24, 18
    Gl = _prim_state(nx, ny, 7)
    top = np.concatenate([Gl, _mirror_x(Gl)], axis=1)
    D = np.concatenate([top, _mirror_y(top)], axis=2)
    mr = MeshSpec(nx, ny, 2, 2, bcx="reflecting", bcy="reflecting")
    mp = MeshSpec(2 * nx, 2 * ny, 2, 2, bcx="periodic", bcy="periodic")
    Ur, Up = mr.to_patches(Gl), mp.to_patches(D)
    oracle.mesh_fill(Ur, mr)
    oracle.mesh_fill(Up, mp)
    dt = 0.3 / (nx * 2.5)
    for _ in range(4):
        oracle.mesh_rk2_step(Ur, mr, dt, 1 / nx, 1 / ny, lim=lim)
        oracle.mesh_rk2_step(Up, mp, dt, 1 / nx, 1 / ny, lim=lim)
    R, Q = mr.to_global(Ur), mp.to_global(Up)[:, :nx, :ny]
    assert np.allclose(R, Q, rtol=1e-13, atol=1e-13)
    # the doubled problem stays a mirror image of itself: the wall is a plane of symmetry
    Qf = mp.to_global(Up)
    assert np.allclose(Qf[:, nx:, :ny], _mirror_x(Qf[:, :nx, :ny]), rtol=1e-13, atol=1e-13)


def test_reflecting_x_outflow_y_equals_doubled_in_x(oracle):
    nx, ny = 20, 12
    Gl = _prim_state(nx, ny, 8)
    D = np.concatenate([Gl, _mirror_x(Gl)], axis=1)
    mr = MeshSpec(nx, ny, 3, 2, bcx="reflecting", bcy="outflow")
    mp = MeshSpec(2 * nx, ny, 3, 2, bcx="periodic", bcy="outflow")
    Ur, Up = mr.to_patches(Gl), mp.to_patches(D)
    oracle.mesh_fill(Ur, mr)
    oracle.mesh_fill(Up, mp)
    dt = 0.3 / (nx * 2.5)
    for _ in range(3):
        oracle.mesh_rk2_step(Ur, mr, dt, 1 / nx, 1 / ny)
        oracle.mesh_rk2_step(Up, mp, dt, 1 / nx, 1 / ny)
    assert np.allclose(mr.to_global(Ur), mp.to_global(Up)[:, :nx, :], rtol=1e-13, atol=1e-13)


def test_closed_box_conserves_mass_energy(oracle):
    """All four walls reflecting: mass and energy totals stay constant (to the
    round-off of summing), the flux of each through a wall being exactly zero
    for mirror-image face states. The same state with outflow walls loses
    mass, so the check has teeth."""
    n = 32
    Gl = _prim_state(n, n, 9)
    totals = {}
    for bc in ("reflecting", "outflow"):
        m = MeshSpec(n, n, 2, 2, bcx=bc, bcy=bc)
        U = m.to_patches(Gl)
        oracle.mesh_fill(U, m)
        t0 = [math.fsum(x) for x in m.to_global(U).reshape(4, -1)]
        for _ in range(6):
            dt = oracle.mesh_dt(U, m, 1 / n, 1 / n, 0.4)
            oracle.mesh_rk2_step(U, m, dt, 1 / n, 1 / n)
        t1 = [math.fsum(x) for x in m.to_global(U).reshape(4, -1)]
        totals[bc] = (t0, t1)
    t0, t1 = totals["reflecting"]
    for k in (0, 3):
        assert abs(t1[k] - t0[k]) <= 1e-12 * abs(t0[k]), k
    # momentum is not conserved in a box: the walls push back
    assert abs(t1[1] - t0[1]) > 1e-6 * n * n
    o0, o1 = totals["outflow"]
    assert abs(o1[0] - o0[0]) > 1e-6 * abs(o0[0])


def test_reflecting_1d_equals_mirror_doubled_periodic(oracle):
    n, h = 40, 2
    x = (np.arange(n) + 0.5) / n
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x)
    u = 0.5 + 0.2 * np.cos(2 * np.pi * 1.3 * x)
    p = 1.0 + 0.1 * np.sin(2 * np.pi * 0.7 * x)
    W = np.stack([rho, rho * u, p / (GAM - 1) + 0.5 * rho * u * u])
    Wm = W[:, ::-1].copy()
    Wm[1] = -Wm[1]
    D = np.concatenate([W, Wm], axis=1)

    def ext(A, m):
        E = np.zeros((3, m + 2 * h))
        E[:, h:h + m] = A
        return E.reshape(-1).copy()

    Ur, Up = ext(W, n), ext(D, 2 * n)
    oracle.fill_1d(Ur, n, bc="reflecting")
    oracle.fill_1d(Up, 2 * n, bc="periodic")
    # ghosts of the walled problem are the doubled problem's cells across the wall
    R = Ur.reshape(3, -1)
    assert np.array_equal(R[:, n + h:n + 2 * h], D[:, n:n + h])
    assert np.array_equal(R[:, :h], D[:, 2 * n - h:])
    dt = 0.4 / (n * 2.0)
    for _ in range(5):
        oracle.rk2_step_1d(Ur, n, dt, 1 / n, bc="reflecting")
        oracle.rk2_step_1d(Up, 2 * n, dt, 1 / n, bc="periodic")
    assert np.allclose(Ur.reshape(3, -1)[:, h:h + n], Up.reshape(3, -1)[:, h:h + n],
                       rtol=1e-13, atol=1e-13)


def test_reflecting_sod_wall_reflects_shock(oracle):
    """1D Sod against a wall: mass is conserved exactly to summation
    round-off and the wall cell's velocity stays near zero while the shock
    reflects (run past the shock's arrival at x = 1)."""
    n, h = 200, 2
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc="reflecting")
    m0 = math.fsum(U.reshape(3, -1)[0, h:h + n])
    k, _ = oracle.run_1d(U, n, 0.45, 0.4, 1 / n, bc="reflecting")
    assert k > 0
    V = U.reshape(3, -1)
    assert abs(math.fsum(V[0, h:h + n]) - m0) <= 1e-13 * m0
    # behind the reflected shock the gas is at rest against the wall
    assert abs(V[1, h + n - 1] / V[0, h + n - 1]) < 0.05
    assert V[0, h + n - 1] > 0.3  # compressed twice: well above the right state's 0.125
