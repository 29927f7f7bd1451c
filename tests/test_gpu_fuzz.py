"""Brute-force comparison of the step kernels with the oracle on small random
problems (property-based: hypothesis draws the cases).

Each example draws a problem shape, a patch partition, the boundary
conditions, the limiter, the step path or 1D step kernel, the work
decomposition, the numerics mode and a seeded random state, runs it on the
GPU and compares with the oracle's restatement of the reference numerics (examples/sod1d/euler1d.hpp:27-242, examples/sod2d/euler2d.hpp:20-271,
stepped as examples/sod2d/sod2d.cpp:318-356 with the build-defined RK2):
  * 2D meshes: one RK2 step with the caller's dt (dt from compute_dt, itself
    asserted equal to the oracle's), then device-dt steps by step count;
  * 2D batches of independent problems of uneven sizes under a drawn tiling:
    device-dt steps;
  * 1D: device-dt steps of a drawn step kernel;
  * the discrete adjoint of a K-step rollout (checkpoint interval and
    adjoint kernel tiling drawn): its forward against the oracle, its
    reverse sweep by the dot-product identity against dual numbers.
Exact numerics must be bitwise equal, fast numerics within 1e-12 relative
L-infinity per variable (dt within 1e-13 relative).
The runs are derandomized: every run draws the same examples, so a failure
reproduces. The states are piecewise constant in four quadrants at drawn
split positions (a different Riemann problem per quadrant) with a seeded 5%
cell-by-cell perturbation, so every HLLE branch and limiter case occurs.
The unlimited scheme (limiter none) gets jumps of at most 10%: across a
strong jump it reconstructs negative face densities, whose sound speed is
not finite (the reference's too), and the run would end there.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from nest_amd import Sim2D, Solver1D, autotune
from oracle.oracle import MeshSpec
from tests.gpu_util import STEP_PATHS, configure, print_case, rel_linf_vars

pytestmark = pytest.mark.gpu
BCS = ["outflow", "periodic", "reflecting"]
LIMS = ["minmod", "mc", "none"]
GAMMA = 1.4
MODES = ["exact", "fast"]
TOL_FAST = 1e-12
SETTINGS = settings(max_examples=150, deadline=None, derandomize=True, database=None,
                    suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])


def _bounds(mild, nvar):
    """Ranges of the drawn primitives (rho, velocities..., p)."""
    lo, hi = (0.9, 1.1) if mild else (0.2, 2.0)
    vel = 0.1 if mild else 0.8
    return [lo] + [-vel] * (nvar - 2) + [lo], [hi] + [vel] * (nvar - 2) + [hi]


def _same(got, want, mode, what):
    if mode == "exact":
        assert np.array_equal(got, want), f"{what}: max diff {np.abs(got - want).max()}"
    else:
        err = rel_linf_vars(got, want)
        assert err <= TOL_FAST, f"{what}: {err}"


def random_state_2d(seed, nx, ny, mild=False):
    """Conserved [4, nx, ny]: four quadrants of drawn (rho, u, v, p) split at
    drawn positions, each cell perturbed by up to 5% (1% when mild)."""
    r = np.random.default_rng(seed)
    xs, ys = int(r.integers(0, nx + 1)), int(r.integers(0, ny + 1))
    q = (np.arange(nx)[:, None] >= xs).astype(int) * 2 + (np.arange(ny)[None, :] >= ys)
    lo, hi = _bounds(mild, 4)
    quad = r.uniform(lo, hi, (4, 4))  # [quadrant, (rho, u, v, p)]
    eps = 0.01 if mild else 0.05
    W = np.moveaxis(quad[q], -1, 0) * (1.0 + eps * r.uniform(-1.0, 1.0, (4, nx, ny)))
    rho, u, v, p = W
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E]))


def random_state_1d(seed, n, mild=False):
    """Extended 1D state [3 (n + 4)] (ghosts left at 0: the kernels read ghost
    cells through the boundary condition): two drawn states split at a drawn
    cell, each cell perturbed by up to 5% (1% when mild)."""
    r = np.random.default_rng(seed)
    xs = int(r.integers(0, n + 1))
    side = (np.arange(n) >= xs).astype(int)
    lo, hi = _bounds(mild, 3)
    lr = r.uniform(lo, hi, (2, 3))  # [side, (rho, u, p)]
    eps = 0.01 if mild else 0.05
    W = lr[side].T * (1.0 + eps * r.uniform(-1.0, 1.0, (3, n)))
    rho, u, p = W
    U = np.zeros((3, n + 4))
    U[:, 2:-2] = [rho, rho * u, p / (GAMMA - 1.0) + 0.5 * rho * u * u]
    return U.reshape(-1)


@SETTINGS
@given(nx=st.integers(2, 90), ny=st.integers(2, 90), px=st.integers(1, 4), py=st.integers(1, 4),
       bcx=st.sampled_from(BCS), bcy=st.sampled_from(BCS), lim=st.sampled_from(LIMS),
       path=st.sampled_from(STEP_PATHS), mode=st.sampled_from(MODES), seed=st.integers(0, 2**31 - 1))
def test_fuzz_2d_mesh(oracle, nx, ny, px, py, bcx, bcy, lim, path, mode, seed):
    npx, npy = max(1, min(px, nx // 2)), max(1, min(py, ny // 2))  # patches of at least 2 cells
    print_case("fuzz 2d mesh", (nx, ny, npx, npy), bcx, bcy, lim, path, mode, seed)
    dx, dy = 1.0 / nx, 1.0 / ny
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, dx, dy, bcx=bcx, bcy=bcy, limiter=lim, mode=mode)
    try:
        configure(sim, path)
        U = sim.scatter(random_state_2d(seed, nx, ny, mild=lim == "none"))
        oracle.mesh_fill(U, m)
        # one step with the caller's dt
        sim.upload(U)
        dt = oracle.mesh_dt(U, m, dx, dy, 0.4)
        dt_gpu = sim.compute_dt(0.4)[0]
        assert dt_gpu == dt if mode == "exact" else abs(dt_gpu - dt) <= 1e-13 * dt, (dt_gpu, dt)
        sim.rk2_step(dt)
        ref = U.copy()
        oracle.mesh_rk2_step(ref, m, dt, dx, dy, lim=lim)
        _same(sim.gather(sim.download()), m.to_global(ref), mode, "host-dt step")
        # then device-dt steps from the same state
        sim.upload(U)
        sim.start(cfl=0.4)
        sim.advance(3)
        ref = U.copy()
        k, _ = oracle.mesh_run(ref, m, 3, dx, dy, cfl=0.4, lim=lim)
        stt = sim.status()
        assert k == 3 and stt["steps"][0] == 3 and stt["done"][0] == 0, stt
        _same(sim.gather(sim.download()), m.to_global(ref), mode, "device-dt steps")
    finally:
        sim.close()


@SETTINGS
@given(sides=st.lists(st.tuples(st.integers(2, 70), st.integers(2, 70)), min_size=1, max_size=6),
       bc=st.sampled_from(BCS), lim=st.sampled_from(LIMS), rows=st.sampled_from(autotune.ROWS),
       waves=st.sampled_from(autotune.WAVES), persistent=st.integers(0, 1), path=st.sampled_from(STEP_PATHS),
       mode=st.sampled_from(MODES), seed=st.integers(0, 2**31 - 1))
def test_fuzz_2d_batch(oracle, sides, bc, lim, rows, waves, persistent, path, mode, seed):
    print_case("fuzz 2d batch", sides, bc, lim, (rows, waves, persistent), path, mode, seed)
    sim = Sim2D.batch(sides, bc=bc, limiter=lim, mode=mode)
    try:
        configure(sim, path)
        sim.set_tiling(rows, waves, persistent)
        U = np.zeros(sim.total)
        refs = []
        for k, (nx, ny) in enumerate(sides):
            m = MeshSpec(nx, ny, 1, 1, bcx=bc, bcy=bc)
            u = m.to_patches(random_state_2d(seed + k, nx, ny, mild=lim == "none"))
            oracle.mesh_fill(u, m)
            sim.patch_view(U, k)[:] = u.reshape(4, nx + 4, ny + 4)
            oracle.mesh_run(u, m, 3, 1.0 / nx, 1.0 / ny, cfl=0.4, lim=lim)
            refs.append(u.reshape(4, nx + 4, ny + 4)[:, 2:-2, 2:-2])
        sim.upload(U)
        sim.start(cfl=0.4)
        sim.advance(3)
        stt = sim.status()
        assert (stt["steps"] == 3).all() and (stt["done"] == 0).all(), stt
        got = sim.download()
        for k, want in enumerate(refs):
            _same(sim.patch_view(got, k)[:, 2:-2, 2:-2], want, mode, f"problem {k} {sides[k]}")
    finally:
        sim.close()


@SETTINGS
@given(n=st.integers(2, 1500), bc=st.sampled_from(BCS), lim=st.sampled_from(LIMS),
       kernel=st.sampled_from(["tile", "lanes", "pairs"]), nsteps=st.integers(1, 5),
       mode=st.sampled_from(MODES), seed=st.integers(0, 2**31 - 1))
def test_fuzz_1d(oracle, n, bc, lim, kernel, nsteps, mode, seed):
    print_case("fuzz 1d", n, bc, lim, kernel, nsteps, mode, seed)
    dx = 1.0 / n
    U = random_state_1d(seed, n, mild=lim == "none")
    oracle.fill_1d(U, n, bc=bc)
    ref = U.copy()
    for _ in range(nsteps):  # the device-dt loop of examples/sod1d/sod1d.cpp:219-272
        dt = 0.4 * dx / oracle.max_speed_1d(ref, n)
        oracle.rk2_step_1d(ref, n, dt, dx, bc=bc, lim=lim)
    oracle.fill_1d(ref, n, bc=bc)
    s = Solver1D(n, dx, bc=bc, mode=mode, limiter=lim)
    try:
        s.set_step_kernel(kernel)
        s.upload(U)
        s.start(0.4)
        s.advance(nsteps)
        stt = s.status()
        assert stt["steps"] == nsteps and stt["done"] == 0, stt
        got = s.download()
        if mode == "exact":
            assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"
        else:
            _same(got.reshape(3, -1)[:, 2:-2], ref.reshape(3, -1)[:, 2:-2], mode, "1d")
    finally:
        s.close()


@SETTINGS
@given(nx=st.integers(2, 80), ny=st.integers(2, 80), px=st.integers(1, 3), py=st.integers(1, 3),
       bcx=st.sampled_from(["periodic", "outflow"]), bcy=st.sampled_from(["periodic", "outflow"]),
       lim=st.sampled_from(LIMS), mode=st.sampled_from(MODES), K=st.integers(1, 8),
       interval=st.sampled_from([0, 1, 2, 3, -1, 100]),
       rows_y=st.sampled_from(autotune.ADJ_ROWS), waves_y=st.sampled_from(autotune.WAVES),
       rows_x=st.sampled_from(autotune.ADJ_ROWS), waves_x=st.sampled_from(autotune.WAVES),
       seed=st.integers(0, 2**31 - 1))
def test_fuzz_adjoint_rollout(oracle, nx, ny, px, py, bcx, bcy, lim, mode, K, interval, rows_y, waves_y, rows_x,
                              waves_x, seed):
    """The discrete adjoint of a K-step fixed-dt rollout (checkpoint interval,
    adjoint kernel tiling and the rest drawn): its final state equals the
    oracle's K steps (exact bitwise, fast within 1e-11), and its reverse
    sweep satisfies the dot-product identity <J^T w, v> = <w, J v> within
    1e-11 relative, J v from the oracle's dual-number JVP of the K-step map
    (the adjoint's own tests pin it entry by entry on fixed cases)."""
    torch = pytest.importorskip("torch")
    npx, npy = max(1, min(px, nx // 2)), max(1, min(py, ny // 2))
    print_case("fuzz adjoint", (nx, ny, npx, npy), bcx, bcy, lim, mode, K, interval,
               (rows_y, waves_y, rows_x, waves_x), seed)
    dx, dy = 1.0 / nx, 1.0 / ny
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    sim = Sim2D.mesh(nx, ny, npx, npy, dx, dy, bcx=bcx, bcy=bcy, limiter=lim, mode=mode)
    try:
        sim.set_adjoint_tiling(rows_y, waves_y, rows_x, waves_x)
        U = sim.scatter(random_state_2d(seed, nx, ny, mild=lim == "none"))
        oracle.mesh_fill(U, m)
        dt = 0.5 * oracle.mesh_dt(U, m, dx, dy, 0.4)
        sim.upload(U)
        sim.rollout(K, dt, interval)
        final = sim.gather(sim.download())
        r = np.random.default_rng(seed ^ 0x5EED)
        v = r.standard_normal((4, nx, ny))
        w = r.standard_normal((4, nx, ny))
        Uc, dU = U.copy(), sim.scatter(v)
        oracle.mesh_rk2_jvp(Uc, dU, m, K, dt, dx, dy, lim=lim)
        if mode == "exact":
            assert np.array_equal(final, sim.gather(Uc)), "rollout state differs from the oracle"
        else:
            assert rel_linf_vars(final, sim.gather(Uc)) <= 1e-11
        Wd = torch.from_numpy(sim.scatter(w)).to("cuda")
        torch.cuda.synchronize()
        sim.rollout_adjoint(Wd.data_ptr())
        sim.synchronize()
        a = sim.gather(Wd.cpu().numpy())
        lhs, rhs = float(np.sum(a * v)), float(np.sum(w * sim.gather(dU)))
        err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        assert err <= 1e-11, (lhs, rhs, err)
    finally:
        sim.close()
