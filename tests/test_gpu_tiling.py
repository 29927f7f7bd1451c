"""GPU parity of every work decomposition the step kernels can be given.

The autotuner (nest_amd.autotune) picks, per kernel (the one-launch step
kernel, or stage 1 and stage 2 of the two-launch step), the rows per work
item (autotune.ROWS), the waves per block (autotune.WAVES) and whether the
launch is persistent (grid capped at the resident waves, each wave looping
over the work list). Each of those choices changes where strips start and end, how
many rows a strip's X window warms up over, and which wave runs which item;
none may change a single bit of the result. Every choice runs here through
the device-dt path (start / advance, the path bench.py times) on an uneven
outflow mesh and on an uneven batch, against the oracle's own device-dt loop
(oracle.mesh_run, the reference numerics of examples/sod2d/euler2d.hpp:67-271
stepped as examples/sod2d/sod2d.cpp:318-356 with the build-defined RK2):
exact numerics bitwise, fast numerics within 1e-12.

The full-size 4096^2 planar Sod mesh then runs under the tiling the bench
reports (12-row strips, 4 waves per block, both stages), device dt, five
steps, each against the oracle's step from the same state with the same dt.
"""
import numpy as np
import pytest

from nest_amd import Sim2D, autotune, problems as P
from oracle.oracle import MeshSpec
from tests.gpu_util import STEP_PATHS, configure, print_case, rel_linf_vars

pytestmark = pytest.mark.gpu
TOL_FAST = 1e-12
STEPS = 3

MESH = (300, 260, 3, 2)  # uneven patches: 100 x (87, 87, 86)
BATCH = [(64, 64), (81, 120), (130, 97), (200, 33), (47, 260), (2, 9), (256, 256)]


@pytest.fixture(scope="module")
def mesh_case(oracle):
    nx, ny, npx, npy = MESH
    m = MeshSpec(nx, ny, npx, npy)
    U0 = m.to_patches(P.sod2d(nx, ny, jitter=1e-3, seed=2))
    G = P.quadrants(max(nx, ny), 9)[:, :nx, :ny]
    # a four-quadrant state on the left half, Sod on the right: every HLLE
    # branch and limiter case occurs within three steps
    S = m.to_global(U0)
    S[:, : nx // 2] = G[:, : nx // 2]
    U0 = m.to_patches(np.ascontiguousarray(S))
    oracle.mesh_fill(U0, m)
    ref = U0.copy()
    k, dts = oracle.mesh_run(ref, m, STEPS, 1.0 / nx, 1.0 / ny, cfl=0.4)
    assert k == STEPS
    return m, U0, m.to_global(ref)


@pytest.fixture(scope="module")
def batch_case(oracle):
    U = []
    refs = []
    for k, (nx, ny) in enumerate(BATCH):
        m = MeshSpec(nx, ny, 1, 1)
        u = m.to_patches(np.ascontiguousarray(P.quadrants(max(nx, ny), k + 5)[:, :nx, :ny]))
        oracle.mesh_fill(u, m)
        U.append(u.copy())
        oracle.mesh_run(u, m, STEPS, 1.0 / nx, 1.0 / ny, cfl=0.4)
        refs.append(u.reshape(4, nx + 4, ny + 4)[:, 2:-2, 2:-2].copy())
    return np.concatenate(U), refs


def _run_mesh(sim, U0):
    sim.upload(U0)
    sim.start(cfl=0.4)
    sim.advance(STEPS)
    st = sim.status()
    assert st["steps"][0] == STEPS and st["done"][0] == 0
    return sim.gather(sim.download())


def _run_batch(sim, U0):
    sim.upload(U0)
    sim.start(cfl=0.4)
    sim.advance(STEPS)
    st = sim.status()
    assert (st["steps"] == STEPS).all() and (st["done"] == 0).all()
    got = sim.download()
    return [sim.patch_view(got, k)[:, 2:-2, 2:-2] for k in range(len(BATCH))]


def _same(got, want, mode, what):
    if mode == "exact":
        assert np.array_equal(got, want), f"{what}: max diff {np.abs(got - want).max()}"
    else:
        err = rel_linf_vars(got, want)
        assert err <= TOL_FAST, f"{what}: {err}"


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("kernel", STEP_PATHS)
@pytest.mark.parametrize("persistent", [0, 1])
def test_mesh_every_tiling(mesh_case, mode, kernel, persistent):
    """Every rows x waves choice of set_tiling (every kernel alike), one wave
    per item and persistent, every step path, on the 300 x 260 mesh of
    3 x 2 uneven patches: 3 device-dt steps == the oracle."""
    m, U0, want = mesh_case
    sim = Sim2D.mesh(*MESH, mode=mode)
    configure(sim, kernel)
    for waves in autotune.WAVES:
        for rows in autotune.ROWS:
            print_case("mesh tiling", mode, kernel, persistent, waves, rows)
            sim.set_tiling(rows, waves, persistent)
            assert sim.get_tiling() == ((rows, rows), (waves, waves))
            assert sim.get_step_tiling() == (rows, waves)
            _same(_run_mesh(sim, U0), want, mode, f"rows {rows} waves {waves}")
    sim.close()


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_mesh_mixed_stage_tilings(mesh_case, mode):
    """Stage 1 and stage 2 with different rows and waves (the tuner sets them
    per stage), on both two-launch paths; the one-launch step with its own
    (stage 3) rows and waves."""
    m, U0, want = mesh_case
    rng = np.random.default_rng(11)
    sim = Sim2D.mesh(*MESH, mode=mode)
    for trial in range(12):
        r, wv = int(rng.choice(autotune.ROWS)), int(rng.choice(autotune.WAVES))
        pers = int(trial % 2)
        print_case("mesh step tiling", mode, (r, wv), pers)
        sim.set_tiling(0, 0, pers)
        sim.set_stage_tiling(3, r, wv)
        assert sim.get_step_tiling() == (r, wv)
        _same(_run_mesh(sim, U0), want, mode, f"step {(r, wv)} persistent {pers}")
    sim.close()
    for kernel in ("two_stage", "register"):
        sim = Sim2D.mesh(*MESH, mode=mode)
        configure(sim, kernel)
        for trial in range(12):
            r1, r2 = (int(x) for x in rng.choice(autotune.ROWS, 2))
            w1, w2 = (int(x) for x in rng.choice(autotune.WAVES, 2))
            pers = int(trial % 2)
            print_case("mesh stage tilings", mode, kernel, (r1, w1), (r2, w2), pers)
            sim.set_tiling(0, 0, pers)
            sim.set_stage_tiling(1, r1, w1)
            sim.set_stage_tiling(2, r2, w2)
            assert sim.get_tiling() == ((r1, r2), (w1, w2))
            _same(_run_mesh(sim, U0), want, mode, f"{(r1, w1)} {(r2, w2)} persistent {pers}")
        sim.close()


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("kernel", ["one_launch", "two_stage"])
def test_batch_every_tiling(batch_case, mode, persistent, kernel):
    """The uneven batch (2 x 9 up to 256 x 256, one patch per problem) under
    every rows x waves choice, then per-problem rows (set_problem_rows, as
    the tuner applies its per-size-class rows), for the one-launch step and
    the two-launch step."""
    U0, refs = batch_case
    sim = Sim2D.batch(BATCH, mode=mode)
    configure(sim, kernel)
    for waves in autotune.WAVES:
        for rows in autotune.ROWS:
            print_case("batch tiling", mode, persistent, waves, rows)
            sim.set_tiling(rows, waves, persistent)
            for k, (g, w) in enumerate(zip(_run_batch(sim, U0), refs)):
                _same(g, w, mode, f"problem {k}, rows {rows} waves {waves}")
    rng = np.random.default_rng(5)
    for trial in range(6):
        per = [[int(x) for x in rng.choice(autotune.ROWS, len(BATCH))] for _ in range(2)]
        print_case("batch per-problem rows", mode, persistent, kernel, per)
        sim.set_tiling(0, int(rng.choice(autotune.WAVES)), persistent)
        sim.set_problem_rows(1, per[0])
        sim.set_problem_rows(2, per[1])
        sim.set_problem_rows(3, per[0])
        for k, (g, w) in enumerate(zip(_run_batch(sim, U0), refs)):
            _same(g, w, mode, f"problem {k}, per-problem rows {per}")
    for st in (1, 2, 3):
        sim.set_problem_rows(st, None)
    sim.close()


def test_autotuned_choice_is_parity_checked(mesh_case):
    """The tuner's own choice on this mesh (fast numerics; whatever it picks
    on this device) reproduces the oracle within 1e-12."""
    m, U0, want = mesh_case
    import os
    import tempfile
    print_case("autotuned choice")
    sim = Sim2D.mesh(*MESH, mode="fast")
    sim.upload(U0)
    sim.start(cfl=0.4)
    with tempfile.TemporaryDirectory() as d:
        autotune.tune(sim, [MESH[:2]], cache=os.path.join(d, "t.json"))
    print("  tuned:", sim.get_tiling())
    _same(_run_mesh(sim, U0), want, "fast", "tuned tiling")
    sim.close()


@pytest.mark.parametrize("tiling", [("one_launch", (32, 4)), ("one_launch", (16, 1)), ("one_launch", (64, 2)),
                                    ("two_stage", ((12, 4), (12, 4))), ("two_stage", ((12, 4), (8, 1)))])
def test_sod4096_fast_device_dt_under_bench_tiling(oracle, tiling):
    """Full size, the workload bench.py times: 4096^2 planar Sod, 16 x 16
    patches of 256^2, outflow, fast numerics, device dt (start / advance:
    stage-2 speed slots, finalize, halo gathered in the kernels), under the
    tilings the bench runs with ((rows, waves) of the one-launch step, or of
    stage 1 and stage 2).
    Five steps; after each, the GPU's dt is within 1e-14 of the oracle's
    compute_dt_2d of the same state, and the state within 1e-12 of the
    oracle's step from that state with the GPU's dt (the GPU state is fed
    back each step: lockstep)."""
    n, npt = 4096, 16
    path, t = tiling
    print_case("sod 4096^2 fast device dt", tiling)
    m = MeshSpec(n, n, npt, npt)
    sim = Sim2D.mesh(n, n, npt, npt, mode="fast")
    configure(sim, path)
    if path == "one_launch":
        sim.set_stage_tiling(3, *t)
    else:
        sim.set_stage_tiling(1, *t[0])
        sim.set_stage_tiling(2, *t[1])
    U = m.to_patches(P.sod2d(n, n))
    oracle.mesh_fill(U, m)
    worst = 0.0
    for k in range(5):
        sim.upload(U)
        sim.start(cfl=0.4)
        sim.advance(1)
        st = sim.status()
        dt_gpu = st["dt"][0]
        dt_cpu = oracle.mesh_dt(U, m, 1.0 / n, 1.0 / n, 0.4)
        assert abs(dt_gpu - dt_cpu) <= 1e-14 * dt_cpu, (k, dt_gpu, dt_cpu)
        got = sim.download()
        ref = U.copy()
        oracle.mesh_rk2_step(ref, m, dt_gpu, 1.0 / n, 1.0 / n)
        err = rel_linf_vars(m.to_global(got), m.to_global(ref))
        print(f"  step {k}: dt {dt_gpu!r}, rel Linf {err:.2e}")
        worst = max(worst, err)
        U = got  # lockstep from the GPU's state
    sim.close()
    assert worst <= TOL_FAST, worst


@pytest.mark.parametrize("kernel", ["one_launch", "two_stage"])
def test_ensemble512_persistent_launch(oracle, kernel):
    """The 512-problem ensemble with the persistent launch (grid capped at
    the resident waves; each wave loops over the device work list), exact
    numerics, 3 device-dt steps from the captured step graph: bitwise the
    one-wave-per-item launch and, on a sample of problems, the oracle."""
    print_case("ensemble 512 persistent", kernel)
    a, sizes, U0 = P.ensemble_batch(512, seed=3, mode="exact")
    b = Sim2D.batch(sizes, mode="exact")
    configure(a, kernel)
    configure(b, kernel)
    a.set_tiling(0, 0, 1)
    b.upload(U0)
    for s in (a, b):
        s.start(cfl=0.4)
        s.advance(3, graph=True)
    ga, gb = a.download(), b.download()
    assert (a.status()["steps"] == 3).all()
    assert np.array_equal(ga, gb)
    for k in (0, 5, 300, 511):
        s = sizes[k][0]
        mk = MeshSpec(s, s, 1, 1)
        u = a.patch_view(U0, k).ravel().copy()
        oracle.mesh_run(u, mk, 3, 1.0 / s, 1.0 / s, cfl=0.4)
        assert np.array_equal(a.patch_view(ga, k)[:, 2:-2, 2:-2], u.reshape(4, s + 4, s + 4)[:, 2:-2, 2:-2]), k
    a.close()
    b.close()
