"""CPU: the oracle's forward-mode dual-number instantiation (the reference
that pins the adjoint kernels) checked against its own primal.

* dual JVP of K RK2 steps vs central differences of the oracle's forward
  map, elementwise, at a generic smooth state along a smooth direction (a
  rough per-cell direction crosses some limiter / Davis-speed / HLLE-branch
  kink within the finite step, where central differences average branches),
  step 1e-6: 1e-8 relative for the unlimited scheme and for minmod;
* the value part of the dual run equals the plain oracle run bitwise;
* linearity in the tangent, and the stage JVP composes into the step JVP
  (U1 = S(U), U' = (U + S(U1)) / 2).
"""
import numpy as np
import pytest

from nest_amd import problems as P
from oracle.oracle import MeshSpec


def smooth_state(nx, ny, seed):
    """Generic smooth state: rho, u, v, p each a constant plus seeded smooth
    periodic fields (no exact ties in any difference)."""
    def field(k):
        r = np.random.default_rng(seed + k)
        x = (np.arange(nx) + 0.5) / nx
        y = (np.arange(ny) + 0.5) / ny
        X, Y = np.meshgrid(x, y, indexing="ij")
        f = np.zeros((nx, ny))
        for kx in range(1, 3):
            for ky in range(1, 3):
                a, ph = r.uniform(-1, 1), r.uniform(0, 2 * np.pi)
                f += a * np.sin(2 * np.pi * (kx * X + ky * Y) + ph)
        return f / np.max(np.abs(f))
    rho = 1.0 + 0.2 * field(0)
    u = 0.3 * field(1)
    v = -0.2 + 0.3 * field(2)
    p = 1.0 + 0.2 * field(3)
    return np.stack([rho, rho * u, rho * v, P.energy2d(rho, u, v, p, 1.4)])


CASES = [(32, 24, 2, 2, "periodic", "periodic"), (30, 26, 2, 1, "outflow", "outflow"),
         (28, 28, 1, 2, "periodic", "outflow")]


def _rollout(oracle, U, m, K, dt, dx, dy, lim):
    X, dX = U.copy(), np.zeros_like(U)
    oracle.mesh_rk2_jvp(X, dX, m, K, dt, dx, dy, lim=lim)
    return X


@pytest.mark.parametrize("lim, tol", [("none", 1e-8), ("minmod", 1e-8)])
@pytest.mark.parametrize("case", CASES)
def test_dual_jvp_matches_central_differences(oracle, case, lim, tol):
    nx, ny, npx, npy, bcx, bcy = case
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    dx, dy = 1.0 / nx, 1.0 / ny
    dt = 0.3 * min(dx, dy) / 2.0
    K = 4
    U = m.to_patches(smooth_state(nx, ny, 11))
    v = m.to_patches(smooth_state(nx, ny, 77) - smooth_state(nx, ny, 78))
    X, dX = U.copy(), v.copy()
    oracle.mesh_rk2_jvp(X, dX, m, K, dt, dx, dy, lim=lim)
    # the value part is the plain oracle run, bitwise
    Y = U.copy()
    for _ in range(K):
        oracle.mesh_rk2_step(Y, m, dt, dx, dy, lim=lim)
    assert np.array_equal(m.to_global(X), m.to_global(Y))
    eps = 1e-6
    fp = m.to_global(_rollout(oracle, U + eps * v, m, K, dt, dx, dy, lim))
    fm = m.to_global(_rollout(oracle, U - eps * v, m, K, dt, dx, dy, lim))
    fd = (fp - fm) / (2 * eps)
    jv = m.to_global(dX)
    err = np.max(np.abs(fd - jv)) / np.max(np.abs(jv))
    assert err <= tol, err


def test_dual_jvp_is_linear_and_stage_composes(oracle):
    nx, ny = 24, 20
    m = MeshSpec(nx, ny, 2, 2, bcx="periodic", bcy="outflow")
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.004
    U = m.to_patches(smooth_state(nx, ny, 3))
    rng = np.random.default_rng(9)
    v, w = (m.to_patches(rng.standard_normal((4, nx, ny))) for _ in range(2))

    def jvp_step(t):
        X, dX = U.copy(), t.copy()
        oracle.mesh_rk2_jvp(X, dX, m, 1, dt, dx, dy)
        return m.to_global(dX)

    a, b = 0.7, -1.3
    lhs = jvp_step(a * v + b * w)
    rhs = a * jvp_step(v) + b * jvp_step(w)
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    # step JVP = (v + S'(U1) S'(U) v) / 2 with U1 = S(U)
    U1, t1 = U.copy(), v.copy()
    oracle.mesh_stage_jvp(U1, t1, m, dt, dx, dy)
    U2, t2 = U1.copy(), t1.copy()
    oracle.mesh_stage_jvp(U2, t2, m, dt, dx, dy)
    comp = 0.5 * (m.to_global(v) + m.to_global(t2))
    step = jvp_step(v)
    assert np.max(np.abs(comp - step)) <= 1e-13 * np.max(np.abs(step))
