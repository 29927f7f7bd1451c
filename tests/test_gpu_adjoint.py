"""GPU adjoint: the hand-written reverse kernels (through the C ABI) against
forward-mode dual numbers run through the oracle's restatement of the
reference numerics.

* tiny meshes: the whole vector S'(U)^T w (stage) / step'(U)^T w (RK2 step)
  is compared entry by entry with J^T w, J assembled column by column from
  dual-number JVPs (every entry of the transpose is pinned);
* larger meshes and K-step rollouts with checkpointing: the dot-product
  identity <J^T w, v> == <w, J v> for random v, w.
Rough random states make every branch occur (limiter cases, subsonic and
both supersonic HLLE branches, both passive-flux signs). Tolerance 1e-11
relative: the two sides differ only in summation order.
"""
import numpy as np
import pytest

from nest_amd import Sim2D, problems as P
from oracle.oracle import MeshSpec
from tests.gpu_util import print_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-11


def rough_state(nx, ny, seed, vmax=1.6):
    """Random per-cell state: rho, p ~ U[0.5, 1.5], u, v ~ U[-vmax, vmax]
    (sound speed ~1.2, so some faces are supersonic each way)."""
    r = np.random.default_rng(seed)
    rho = r.uniform(0.5, 1.5, (nx, ny))
    p = r.uniform(0.5, 1.5, (nx, ny))
    u = r.uniform(-vmax, vmax, (nx, ny))
    v = r.uniform(-vmax, vmax, (nx, ny))
    return np.stack([rho, rho * u, rho * v, P.energy2d(rho, u, v, p, 1.4)])


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def gpu_stage_vjp(sim, U, w, dt):
    """S'(U)^T w (patch layout in, patch layout out) on the GPU."""
    Ud, Wd = dev(U), dev(w)
    out = torch.zeros_like(Ud)
    torch.cuda.synchronize()
    sim.adjoint_stage(Ud.data_ptr(), Wd.data_ptr(), 1.0, None, 0.0, out.data_ptr(), dt)
    sim.synchronize()
    return host(out)


def gpu_step_vjp(sim, U, w, dt):
    Ud, Wd = dev(U), dev(w)
    torch.cuda.synchronize()
    sim.rk2_vjp(Ud.data_ptr(), Wd.data_ptr(), dt)
    sim.synchronize()
    return host(Wd)


CASES = [  # nx, ny, npx, npy, bcx, bcy
    (12, 10, 2, 2, "periodic", "periodic"),
    (12, 10, 2, 2, "outflow", "outflow"),
    (9, 14, 1, 3, "outflow", "periodic"),
    (14, 9, 3, 1, "periodic", "outflow"),
]


def _jacobian_transpose_w(oracle, U, w_glob, m, sim, jvp):
    """J^T w by columns: one dual-number JVP per interior entry."""
    nx, ny = m.nx, m.ny
    out = np.zeros((4, nx, ny))
    for k in range(4):
        for i in range(nx):
            for j in range(ny):
                e = np.zeros((4, nx, ny))
                e[k, i, j] = 1.0
                Uc = U.copy()
                dU = sim.scatter(e)
                jvp(Uc, dU)
                out[k, i, j] = np.sum(w_glob * sim.gather(dU))
    return out


@pytest.mark.parametrize("lim", ["minmod", "mc", "none"])
@pytest.mark.parametrize("case", CASES)
def test_stage_adjoint_full_transpose(oracle, case, lim):
    nx, ny, npx, npy, bcx, bcy = case
    print_case("adjoint stage full transpose", case, lim)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy, limiter=lim, mode="exact")
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.2 / max(nx, ny) / 3.0
    U = sim.scatter(rough_state(nx, ny, seed=nx * 100 + ny))
    w = np.random.default_rng(7).standard_normal((4, nx, ny))
    got = sim.gather(gpu_stage_vjp(sim, U, sim.scatter(w), dt))

    def jvp(Uc, dU):
        oracle.mesh_stage_jvp(Uc, dU, m, dt, dx, dy, lim=lim)

    want = _jacobian_transpose_w(oracle, U, w, m, sim, jvp)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err <= TOL, err
    sim.close()


@pytest.mark.parametrize("bcs", [("outflow", "outflow"), ("periodic", "outflow"), ("outflow", "periodic")])
def test_stage_adjoint_short_strips(oracle, bcs):
    """Adjoint work items of 1..5 rows (the X kernel's march with and
    without an interior part, items that start or end at an outflow end, a
    one-row item touching both its patch's ends), against the full dual
    J^T w: every tiling gives the same vector."""
    bcx, bcy = bcs
    nx, ny, npx, npy = 14, 12, 2, 2
    print_case("adjoint short strips", bcs)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy, mode="exact")
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.2 / max(nx, ny) / 3.0
    U = sim.scatter(rough_state(nx, ny, seed=41))
    w = np.random.default_rng(42).standard_normal((4, nx, ny))

    def jvp(Uc, dU):
        oracle.mesh_stage_jvp(Uc, dU, m, dt, dx, dy)

    want = _jacobian_transpose_w(oracle, U, w, m, sim, jvp)
    for rows_y, rows_x in ((1, 1), (2, 2), (3, 3), (1, 5), (2, 4), (6, 7)):
        sim.set_adjoint_tiling(rows_y, 1, rows_x, 2)
        got = sim.gather(gpu_stage_vjp(sim, U, sim.scatter(w), dt))
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        assert err <= TOL, (rows_y, rows_x, err)
    sim.close()


@pytest.mark.parametrize("case", CASES[:2])
def test_step_adjoint_full_transpose(oracle, case):
    nx, ny, npx, npy, bcx, bcy = case
    print_case("adjoint rk2 step full transpose", case)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy, mode="exact")
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.2 / max(nx, ny) / 3.0
    U = sim.scatter(rough_state(nx, ny, seed=5))
    w = np.random.default_rng(8).standard_normal((4, nx, ny))
    got = sim.gather(gpu_step_vjp(sim, U, sim.scatter(w), dt))

    def jvp(Uc, dU):
        oracle.mesh_rk2_jvp(Uc, dU, m, 1, dt, dx, dy)

    want = _jacobian_transpose_w(oracle, U, w, m, sim, jvp)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err <= TOL, err
    sim.close()


def _dot_check(a_vjp, dU, w, JdU):
    lhs = float(np.sum(a_vjp * dU))
    rhs = float(np.sum(w * JdU))
    scale = max(abs(lhs), abs(rhs))
    return abs(lhs - rhs) / scale


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("case", [
    (130, 190, 2, 3, "outflow", "outflow", "minmod"),
    (128, 128, 2, 2, "periodic", "periodic", "mc"),
    (96, 200, 3, 2, "periodic", "outflow", "minmod"),
])
def test_stage_and_step_dot_product(oracle, case, mode):
    nx, ny, npx, npy, bcx, bcy, lim = case
    print_case("adjoint dot product", case, mode)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy, limiter=lim, mode=mode)
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.25 / max(nx, ny) / 2.0
    U = sim.scatter(rough_state(nx, ny, seed=11))
    rng = np.random.default_rng(12)
    v = rng.standard_normal((4, nx, ny))
    w = rng.standard_normal((4, nx, ny))
    # stage
    Uc, dU = U.copy(), sim.scatter(v)
    oracle.mesh_stage_jvp(Uc, dU, m, dt, dx, dy, lim=lim)
    a = sim.gather(gpu_stage_vjp(sim, U, sim.scatter(w), dt))
    e1 = _dot_check(a, v, w, sim.gather(dU))
    # step
    Uc, dU = U.copy(), sim.scatter(v)
    oracle.mesh_rk2_jvp(Uc, dU, m, 1, dt, dx, dy, lim=lim)
    a = sim.gather(gpu_step_vjp(sim, U, sim.scatter(w), dt))
    e2 = _dot_check(a, v, w, sim.gather(dU))
    print(f"stage {e1:.2e} step {e2:.2e}")
    assert e1 <= TOL and e2 <= TOL
    sim.close()


@pytest.mark.parametrize("interval", [0, 1, 4, 6, 16, -1])
def test_rollout_checkpointed_adjoint(oracle, interval):
    """K = 11 fixed-dt steps with checkpoints every `interval` steps (0:
    ceil(sqrt K) = 4; 1: eleven intervals, each recomputed step by step
    during the reverse of the interval three later; 6: two intervals, two
    buffer sets; 16: one interval and one buffer set, recomputed whole
    before its reverse; -1: every state kept, nothing recomputed):
    the final state is bitwise the oracle's 11 steps, and the reverse sweep
    (recompute per interval) satisfies the dot-product identity against the
    dual-number JVP of the 11-step map."""
    _rollout_check(oracle, interval, 64, 48, 2, 2, "periodic", "periodic")


@pytest.mark.parametrize("interval", [4, -1])
@pytest.mark.parametrize("bcs", [("outflow", "outflow"), ("outflow", "periodic")])
def test_rollout_checkpointed_adjoint_outflow(oracle, interval, bcs):
    """The same on an uneven 3 x 2 mesh with outflow edges: the one-launch
    forward's stage-1 states at the edges (copied from the boundary cells for
    the second stage) are the ones the reverse reads."""
    _rollout_check(oracle, interval, 70, 53, 3, 2, *bcs)


def _rollout_check(oracle, interval, nx, ny, npx, npy, bcx, bcy):
    K = 11
    print_case("adjoint rollout", interval, nx, ny, npx, npy, bcx, bcy)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy, mode="exact")
    dx, dy = 1.0 / nx, 1.0 / ny
    dt = 0.3 * min(dx, dy) / 1.5
    U = sim.scatter(rough_state(nx, ny, seed=21, vmax=0.4))
    sim.upload(U)
    sim.rollout(K, dt, interval)
    k, mm, _ = sim.rollout_info()
    assert k == K and mm == {0: 4, -1: K}.get(interval, interval)
    final = sim.gather(sim.download())
    rng = np.random.default_rng(22)
    v = rng.standard_normal((4, nx, ny))
    w = rng.standard_normal((4, nx, ny))
    Uc, dU = U.copy(), sim.scatter(v)
    oracle.mesh_rk2_jvp(Uc, dU, m, K, dt, dx, dy)
    assert np.array_equal(final, sim.gather(Uc)), "rollout state differs from the oracle"
    Wd = dev(sim.scatter(w))
    torch.cuda.synchronize()
    sim.rollout_adjoint(Wd.data_ptr())
    sim.synchronize()
    a = sim.gather(host(Wd))
    err = _dot_check(a, v, w, sim.gather(dU))
    print(f"rollout K={K} interval {mm}: {err:.2e}")
    assert err <= TOL
    sim.close()


def test_pack_unpack_roundtrip():
    print_case("pack/unpack")
    nx, ny = 70, 45
    sim = Sim2D.mesh(nx, ny, 3, 2, mode="exact")
    G = rough_state(nx, ny, seed=31)
    U = sim.scatter(G)
    Ud = dev(U)
    g = torch.empty(sim.global_size(), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sim.pack(Ud.data_ptr(), g.data_ptr())
    sim.synchronize()
    assert np.array_equal(host(g).reshape(4, nx, ny), G)
    back = torch.zeros_like(Ud)
    torch.cuda.synchronize()  # the zero fill runs on torch's stream, unpack on the handle's
    sim.unpack(g.data_ptr(), back.data_ptr())
    sim.synchronize()
    assert np.array_equal(sim.gather(host(back)), G)
    sim.close()


def test_adjoint_rejects_reflecting_walls():
    """Reflecting walls are forward-only: the adjoint entry points return
    NEST_ERR_UNSUPPORTED and leave the output untouched."""
    from nest_amd._lib import NestError
    print_case("adjoint reflecting rejected")
    sim = Sim2D.mesh(32, 24, 2, 1, bcx="reflecting", bcy="periodic", mode="fast")
    U = sim.scatter(rough_state(32, 24, seed=5))
    with pytest.raises(NestError) as e:
        gpu_stage_vjp(sim, U, U, 1e-4)
    assert e.value.code == 5 and "reflecting" in str(e.value)


@pytest.mark.parametrize("interval", [1, 2, 3, -1])
def test_rollout_adjoint_twice(oracle, interval):
    """Two reverse sweeps of one rollout (as a second backward of the same
    forward makes) give bitwise the same cotangent. K = 11 with checkpoints
    every 1 or 2 steps is 11 or 6 intervals, more than the reverse sweep's
    four buffer sets: the first sweep's recomputes reuse the buffer sets the
    rollout kept its last four intervals in, so the second sweep must
    recompute every interval. Every 3 steps (4 intervals, all kept) and every
    state kept recompute nothing, in either sweep. Both equal the
    dual-number dot-product check."""
    nx, ny, npx, npy, K = 64, 48, 2, 2, 11
    print_case("adjoint rollout twice", interval)
    m = MeshSpec(nx, ny, npx, npy, bcx="periodic", bcy="periodic")
    sim = Sim2D.mesh(nx, ny, npx, npy, bcx="periodic", bcy="periodic", mode="exact")
    dx, dy = 1.0 / nx, 1.0 / ny
    dt = 0.3 * min(dx, dy) / 1.5
    U = sim.scatter(rough_state(nx, ny, seed=41, vmax=0.4))
    sim.upload(U)
    sim.rollout(K, dt, interval)
    rng = np.random.default_rng(42)
    v = rng.standard_normal((4, nx, ny))
    w = rng.standard_normal((4, nx, ny))
    Uc, dU = U.copy(), sim.scatter(v)
    oracle.mesh_rk2_jvp(Uc, dU, m, K, dt, dx, dy)
    outs = []
    for rep in range(3):
        Wd = dev(sim.scatter(w))
        torch.cuda.synchronize()
        sim.rollout_adjoint(Wd.data_ptr())
        sim.synchronize()
        outs.append(host(Wd))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    err = _dot_check(sim.gather(outs[0]), v, w, sim.gather(dU))
    assert err <= TOL
    sim.close()
