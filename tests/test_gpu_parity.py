"""GPU parity: HIP kernels (through the C ABI) against the CPU oracle.

Exact mode must be bitwise equal to the oracle (IEEE division and sqrt, no
FMA contraction on either side). Fast mode must agree to 1e-12 relative
L-infinity per step (the required bound); measured deviations are ~1e-15.
"""
import os

import numpy as np
import pytest

from nest_amd import Sim2D, Solver1D, problems as P
from oracle.oracle import MeshSpec
from tests.gpu_util import STEP_PATHS, configure, print_case, rel_linf_vars

pytestmark = pytest.mark.gpu
TOL_FAST = 1e-12


# ----------------------------------------------------------------------------
# 1D
# ----------------------------------------------------------------------------
LIMITERS = ["minmod", "mc", "none"]


@pytest.mark.parametrize("lim", LIMITERS)
@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("bc", ["outflow", "periodic", "reflecting"])
def test_1d_sod_lockstep(oracle, mode, bc, lim):
    """1D Sod, N=1000, CFL 0.4, each limiter (euler1d.hpp:106-137): same U^n
    and dt fed to both each step (the caller-chosen-dt stage kernels)."""
    n, cfl = 1000, 0.4
    dx = 1.0 / n
    print_case("1d lockstep", mode, bc, lim)
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc=bc)
    s = Solver1D(n, dx, bc=bc, mode=mode, limiter=lim)
    worst = 0.0
    for k in range(40):
        s.upload(U)
        dt_gpu = s.compute_dt(cfl)
        dt_cpu = cfl * dx / oracle.max_speed_1d(U, n)
        if mode == "exact":
            assert dt_gpu == dt_cpu
        else:
            assert abs(dt_gpu - dt_cpu) <= 1e-14 * dt_cpu
        s.rk2_step(dt_cpu)
        got = s.download().reshape(3, -1)[:, 2:-2]
        ref = U.copy()
        oracle.rk2_step_1d(ref, n, dt_cpu, dx, bc=bc, lim=lim)
        want = ref.reshape(3, -1)[:, 2:-2]
        if mode == "exact":
            assert np.array_equal(got, want), f"step {k}"
        else:
            worst = max(worst, rel_linf_vars(got, want))
        U = ref
    assert worst <= TOL_FAST


STEP1D = ["tile", "lanes", "pairs"]


@pytest.mark.parametrize("kernel", STEP1D)
@pytest.mark.parametrize("lim", LIMITERS)
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_1d_sod_run_to_t_end(oracle, mode, lim, kernel):
    """1D Sod, N=1000, to t = 0.2 with device-side dt (the fused one-launch
    step, both kernels), each limiter: step count and final state."""
    n, cfl, t_end = 1000, 0.4, 0.2
    print_case("1d run", mode, lim, kernel)
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc="outflow")
    ref = U.copy()
    k_ref, _ = oracle.run_1d(ref, n, t_end, cfl, 1.0 / n, bc="outflow", lim=lim)
    s = Solver1D(n, 1.0 / n, bc="outflow", mode=mode, limiter=lim)
    s.set_step_kernel(kernel)
    s.upload(U)
    k = s.run(t_end, cfl)
    got = s.download().reshape(3, -1)[:, 2:-2]
    want = ref.reshape(3, -1)[:, 2:-2]
    assert k == k_ref
    if mode == "exact":
        assert np.array_equal(got, want)
    else:
        assert rel_linf_vars(got, want) <= 1e-10  # free-running over ~1000 steps
    # conservation of mass with outflow: only boundary fluxes change it; the
    # Sod state is quiescent at both ends through t = 0.2
    assert abs(got[0].sum() - want[0].sum()) <= 1e-12 * want[0].sum()


@pytest.mark.parametrize("kernel", STEP1D)
@pytest.mark.parametrize("lim", LIMITERS)
@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_1d_run_fused_step_sizes(oracle, bc, lim, kernel):
    """Solver1D.run takes each device-dt step in one fused launch that keeps
    U1 in LDS, alternating the state between two buffers. Bitwise equal to
    the oracle's run (exact numerics) for sizes of 2 and 3 cells (one block
    whose halo wraps / mirrors / clamps onto itself), 257 (a last block of one
    cell) and 1000, each at two end times whose step counts differ in parity
    (the state ends in either buffer), for each limiter; 57 and 121 end
    one-cell-per-lane and two-cells-per-lane segments one cell past a whole
    segment."""
    # minmod on the Sod jump; MC and the unlimited scheme on a weak shock
    # tube (on the strong jump at 2-3 cells the unlimited scheme drives the
    # density negative, and the run ends with a non-finite dt on either side)
    left, right = ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1)) if lim == "minmod" else ((1.0, 0.0, 1.0), (0.8, 0.0, 0.75))
    for n, t_probe in [(2, 0.5), (3, 0.62), (57, 0.1), (121, 0.05), (257, 0.05), (1000, 0.05)]:
        U = P.sod1d(n, left=left, right=right)
        oracle.fill_1d(U, n, bc=bc)
        # end times half-way into steps j and j + 1 of the oracle's dt
        # sequence: runs of j + 1 and j + 2 steps, one of each parity
        k0, dts = oracle.run_1d(U.copy(), n, t_probe, 0.4, 1.0 / n, bc=bc, lim=lim)
        assert k0 >= 3 and np.isfinite(dts).all(), (n, k0)
        T = np.concatenate([[0.0], np.cumsum(dts)])
        j = k0 - 3
        parities = set()
        for t_end in (T[j] + 0.5 * dts[j], T[j + 1] + 0.5 * dts[j + 1]):
            print_case("1d run fused", n, bc, lim, kernel, t_end)
            ref = U.copy()
            k_ref, _ = oracle.run_1d(ref, n, t_end, 0.4, 1.0 / n, bc=bc, lim=lim)
            s = Solver1D(n, 1.0 / n, bc=bc, mode="exact", limiter=lim)
            s.set_step_kernel(kernel)
            s.upload(U)
            k = s.run(t_end, 0.4)
            got = s.download()
            s.close()
            assert k == k_ref, (n, t_end, k, k_ref)
            assert np.array_equal(got.reshape(3, -1)[:, 2:-2], ref.reshape(3, -1)[:, 2:-2]), (n, t_end)
            assert np.array_equal(got, ref), "ghost cells of the handed-out state"
            parities.add(k & 1)
        assert parities == {0, 1}, (n, bc)


def test_1d_reference_abi_periodic_sod(oracle):
    """Reference GPU-vs-CPU case (tests/test_euler_cuda.cpp:27-43): periodic
    Sod N=100, t=0.1, CFL 0.3 through the reference-compatible euler1d_gpu_*
    ABI. The reference's CPU loop (test_euler_cuda.cpp:82-106) takes 70 steps
    (the oracle, pinned by the reference benchmark's own step counts, agrees);
    docs/CUDA_BACKEND.md:160 prints 39, which its code does not produce."""
    from nest_amd import _lib
    import ctypes as C
    L = _lib.load()
    n = 100
    U = P.sod1d(n)
    h = L.euler1d_gpu_create(n, 2, 1.0 / n, 1.4)
    assert h
    try:
        L.euler1d_gpu_upload(h, U.ctypes.data_as(C.POINTER(C.c_double)))
        steps = L.euler1d_gpu_run(h, 0.1, 0.3)
        out = np.empty_like(U)
        L.euler1d_gpu_download(h, out.ctypes.data_as(C.POINTER(C.c_double)))
    finally:
        L.euler1d_gpu_destroy(h)
    ref = U.copy()
    oracle.fill_1d(ref, n, bc="periodic")
    k_ref, _ = oracle.run_1d(ref, n, 0.1, 0.3, 1.0 / n, bc="periodic")
    assert steps == k_ref == 70
    assert np.array_equal(out.reshape(3, -1)[:, 2:-2], ref.reshape(3, -1)[:, 2:-2])


def test_reference_gpu_test_program_links_and_passes():
    """The reference's own tests/test_euler_cuda.cpp, compiled from its source
    (oracle/Makefile target `drop-in`) and linked against libnest_hip.so
    through the euler1d_gpu_* ABI: its CPU half must accept our GPU result
    (its tolerance, 1e-8) — the drop-in boundary exercised by the reference's
    own harness."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "test_euler_cuda")
    if not os.path.exists(exe):
        pytest.skip("reference test program not built (needs the reference tree at build time)")
    print_case("reference test_euler_cuda")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout and "GPU: 70 steps" in r.stdout
    diffs = [float(l.split(":")[1]) for l in r.stdout.splitlines() if l.startswith("Max difference")]
    assert len(diffs) == 3 and max(diffs) == 0.0  # bitwise equal to the reference CPU code


def test_reference_benchmark_program_on_gpu():
    """The reference's own 1D benchmark (examples/benchmark/benchmark_euler1d.cpp,
    the program behind its published 1D GPU figure, N = 1e5), compiled from its
    source with its GPU section on and linked against libnest_hip.so through
    the euler1d_gpu_* ABI (oracle/Makefile target _ref/benchmark_euler1d_gpu).
    Its OpenMP CPU half and our GPU half must take the same number of steps;
    both rates are printed. t_final = 0.02 instead of the published 0.2 keeps
    the CPU half short; the per-step rates do not depend on it."""
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "benchmark_euler1d_gpu")
    if not os.path.exists(exe):
        pytest.skip("reference benchmark program not built (needs the reference tree at build time)")
    print_case("reference benchmark_euler1d on the GPU, N=1e5 t=0.02")
    r = subprocess.run([exe, "100000", "0.02"], capture_output=True, text=True, timeout=150)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    steps = [int(x) for x in re.findall(r"Steps:\s+(\d+)", r.stdout)]
    rates = [float(x) for x in re.findall(r"Updates/sec:\s+([0-9.eE+-]+)", r.stdout)]
    assert len(steps) == 2 and steps[0] == steps[1] > 0, steps
    assert len(rates) == 2 and rates[1] > rates[0], rates


# ----------------------------------------------------------------------------
# 2D mesh
# ----------------------------------------------------------------------------
CASES_2D = [
    # nx, ny, npx, npy, bcx, bcy, state
    (96, 80, 3, 2, "outflow", "outflow", "sod"),
    (64, 64, 2, 2, "periodic", "periodic", "quad"),
    (70, 130, 2, 3, "periodic", "outflow", "quad"),
    (40, 200, 1, 1, "outflow", "periodic", "sod"),
    # reflecting walls (normal momentum negated in the mirrored ghosts)
    (72, 56, 3, 2, "reflecting", "outflow", "quad"),
    (64, 96, 2, 3, "periodic", "reflecting", "quad"),
    (50, 90, 1, 1, "reflecting", "reflecting", "quad"),
]


def _state(kind, nx, ny):
    if kind == "sod":
        return P.sod2d(nx, ny, jitter=1e-3, seed=2)
    G = P.quadrants(max(nx, ny), 7)[:, :nx, :ny]
    return np.ascontiguousarray(G)


# every way a step can run (tests/gpu_util.py: STEP_PATHS)
STAGE_KERNELS = STEP_PATHS


@pytest.mark.parametrize("case", CASES_2D)
@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("kernel", STAGE_KERNELS)
def test_2d_mesh_lockstep(oracle, case, mode, kernel):
    nx, ny, npx, npy, bcx, bcy, kind = case
    print_case("2d lockstep", case, mode, kernel)
    dx, dy = 1.0 / nx, 1.0 / ny
    sim = Sim2D.mesh(nx, ny, npx, npy, dx, dy, bcx=bcx, bcy=bcy, mode=mode)
    configure(sim, kernel)
    m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
    U = sim.scatter(_state(kind, nx, ny))
    oracle.mesh_fill(U, m)
    worst = 0.0
    for k in range(6):
        sim.upload(U)
        dt_gpu = sim.compute_dt(0.4)[0]
        dt_cpu = oracle.mesh_dt(U, m, dx, dy, 0.4)
        if mode == "exact":
            assert dt_gpu == dt_cpu
        else:
            assert abs(dt_gpu - dt_cpu) <= 1e-14 * dt_cpu
        sim.rk2_step(dt_cpu)
        got = sim.download()
        ref = U.copy()
        oracle.mesh_rk2_step(ref, m, dt_cpu, dx, dy)
        if mode == "exact":
            assert np.array_equal(got, ref), f"step {k}: max diff {np.abs(got - ref).max()}"
        else:
            worst = max(worst, rel_linf_vars(sim.gather(got), sim.gather(ref)))
        U = ref
    assert worst <= TOL_FAST


BCS = ["periodic", "outflow", "reflecting"]


@pytest.mark.parametrize("kernel", STAGE_KERNELS)
def test_2d_minimum_size_patches(oracle, kernel):
    """Patches of 2-3 cells (the minimum is 2: the halo width) and tiny
    single-patch problems, every pair of boundary conditions, exact numerics:
    3 lockstep steps bitwise equal to the oracle. A 64-lane strip then spans
    many patches and runs past the problem's edge on both sides, so every
    lane's source comes through the wrap / clamp / mirror of its column."""
    shapes = [(9, 7, 4, 3), (2, 3, 1, 1), (5, 2, 2, 1), (3, 70, 1, 4)]
    for (nx, ny, npx, npy) in shapes:
        for bcx in BCS:
            for bcy in BCS:
                print_case("2d minimum patches", (nx, ny, npx, npy), bcx, bcy, kernel)
                dx, dy = 1.0 / nx, 1.0 / ny
                sim = Sim2D.mesh(nx, ny, npx, npy, dx, dy, bcx=bcx, bcy=bcy, mode="exact")
                configure(sim, kernel)
                m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
                G = P.quadrants(max(nx, ny), nx * 7 + ny)[:, :nx, :ny].copy()
                U = sim.scatter(G)
                oracle.mesh_fill(U, m)
                for k in range(3):
                    sim.upload(U)
                    dt = oracle.mesh_dt(U, m, dx, dy, 0.4)
                    assert sim.compute_dt(0.4)[0] == dt
                    sim.rk2_step(dt)
                    ref = U.copy()
                    oracle.mesh_rk2_step(ref, m, dt, dx, dy)
                    got = sim.download()
                    assert np.array_equal(got, ref), f"step {k}: max diff {np.abs(got - ref).max()}"
                    U = ref
                sim.close()


@pytest.mark.parametrize("bc", BCS)
def test_batch_of_tiny_problems(oracle, bc):
    """A batch mixing 2-cell-wide problems with larger ones (device dt, 4
    steps, exact numerics) == each problem alone in the oracle."""
    sizes = [(2, 2), (3, 5), (2, 9), (17, 2), (40, 33)]
    print_case("batch tiny", bc)
    sim = Sim2D.batch(sizes, bc=bc, mode="exact")
    U = np.zeros(sim.total)
    refs = []
    for k, (nx, ny) in enumerate(sizes):
        m = MeshSpec(nx, ny, 1, 1, bcx=bc, bcy=bc)
        u = m.to_patches(P.quadrants(max(nx, ny), k + 3)[:, :nx, :ny].copy())
        oracle.mesh_fill(u, m)
        sim.patch_view(U, k)[:] = u.reshape(4, nx + 4, ny + 4)
        refs.append((m, u))
    sim.upload(U)
    sim.start(cfl=0.4)
    sim.advance(4, graph=False)
    got = sim.download()
    st = sim.status()
    for k, (m, u) in enumerate(refs):
        oracle.mesh_run(u, m, 4, 1.0 / m.nx, 1.0 / m.ny, cfl=0.4)
        assert st["steps"][k] == 4
        g = sim.patch_view(got, k)[:, 2:-2, 2:-2]
        assert np.array_equal(g, u.reshape(4, m.nx + 4, m.ny + 4)[:, 2:-2, 2:-2]), k
    sim.close()


def test_2d_halo_fill_bitwise(oracle):
    """Device exchange plan (faces + corners; periodic, outflow, reflecting) == oracle."""
    for (nx, ny, npx, npy, bcx, bcy, _) in CASES_2D:
        print_case("halo", nx, ny, npx, npy, bcx, bcy)
        sim = Sim2D.mesh(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
        m = MeshSpec(nx, ny, npx, npy, bcx=bcx, bcy=bcy)
        rng = np.random.default_rng(nx * ny)
        U = sim.scatter(rng.uniform(-1, 1, (4, nx, ny)))
        sim.upload(U)
        sim.fill_halos()
        ref = U.copy()
        oracle.mesh_fill(ref, m)
        assert np.array_equal(sim.download(), ref)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("kernel", STAGE_KERNELS)
def test_2d_device_dt_run(oracle, graph, kernel):
    """Device-resident dt (reduction fused into stage 2 + finalize kernel):
    dt sequence and states equal the oracle's compute_dt_2d loop."""
    nx, ny, npx, npy = 64, 48, 2, 2
    print_case("2d device dt", graph, kernel)
    sim = Sim2D.mesh(nx, ny, npx, npy)
    configure(sim, kernel)
    m = MeshSpec(nx, ny, npx, npy)
    U = sim.scatter(P.sod2d(nx, ny))
    oracle.mesh_fill(U, m)
    sim.upload(U)
    sim.start(cfl=0.4, t_end=0.05)
    ref = U.copy()
    k_ref, dts = oracle.mesh_run(ref, m, 1000, 1.0 / nx, 1.0 / ny, cfl=0.4, t_end=0.05)
    sim.advance(k_ref + 3, graph=graph)  # extra steps must be no-ops once t_end is reached
    st = sim.status()
    assert st["steps"][0] == k_ref and st["done"][0] == 1
    assert st["t"][0] == np.cumsum(dts)[-1] or abs(st["t"][0] - 0.05) < 1e-15
    assert np.array_equal(sim.gather(sim.download()), sim.gather(ref))


def test_2d_device_dt_mixed_sequence(oracle):
    """One device-dt run of an uneven 3-problem batch (exact numerics) that
    switches between the one-launch step (dt computed in the step kernel)
    and the two-launch step (dt finalize + two stages), plain launches and
    graph replays starting in every phase of the speed-slot rotation
    (3 + 2 + 5 + 6 steps, then 40 more in graphs), and runs past t_end: each
    problem's states after 16 steps and at the end, its step count, done
    flag and time equal the oracle's compute_dt_2d loop for that problem
    alone. The problems reach t_end at different steps, so finished ones
    pass their state through while the others go on."""
    sizes = [(64, 64), (48, 80), (96, 40)]
    t_end = 0.1
    print_case("2d device dt mixed sequence", sizes)
    sim = Sim2D.batch(sizes, mode="exact")
    U = np.zeros(sim.total)
    refs = []
    for k, (nx, ny) in enumerate(sizes):
        m = MeshSpec(nx, ny, 1, 1)
        u = m.to_patches(P.sod2d(nx, ny, jitter=1e-3, seed=5 + k))
        oracle.mesh_fill(u, m)
        sim.patch_view(U, k)[:] = u.reshape(4, nx + 4, ny + 4)
        refs.append((m, u))
    sim.upload(U)
    sim.start(cfl=0.4, t_end=t_end)
    for kind, n, graph in (("auto", 3, False), ("two_stage", 2, True), ("two_stage", 5, True),
                           ("auto", 6, True)):
        sim.set_step_kernel(kind)
        sim.advance(n, graph=graph)
    mid = sim.download()
    st_mid = sim.status()
    sim.advance(40, graph=True)
    got = sim.download()
    st = sim.status()
    for k, (m, u) in enumerate(refs):
        ref16 = u.copy()
        k16, _ = oracle.mesh_run(ref16, m, 16, 1.0 / m.nx, 1.0 / m.ny, cfl=0.4, t_end=t_end)
        assert k16 == 16 and st_mid["steps"][k] == 16 and st_mid["done"][k] == 0, (k, k16, st_mid)
        shape = (4, m.nx + 4, m.ny + 4)
        assert np.array_equal(sim.patch_view(mid, k)[:, 2:-2, 2:-2], ref16.reshape(shape)[:, 2:-2, 2:-2]), k
        ref = u.copy()
        k_ref, dts = oracle.mesh_run(ref, m, 1000, 1.0 / m.nx, 1.0 / m.ny, cfl=0.4, t_end=t_end)
        assert k_ref < 56 and st["steps"][k] == k_ref and st["done"][k] == 1, (k, k_ref, st)
        assert st["t"][k] == np.cumsum(dts)[-1] or abs(st["t"][k] - t_end) < 1e-15
        assert np.array_equal(sim.patch_view(got, k)[:, 2:-2, 2:-2], ref.reshape(shape)[:, 2:-2, 2:-2]), k
    assert len(set(int(x) for x in st["steps"])) > 1  # the problems finished at different steps
    sim.close()


@pytest.mark.parametrize("kernel", STAGE_KERNELS)
def test_2d_host_dt_steps_between_device_dt_steps(oracle, kernel):
    """Device-dt steps, then an RK2 step with the caller's dt (compute_dt of
    the current state), then device-dt steps again, then another state
    uploaded and more device-dt steps (exact numerics, graph and plain
    launches): every dt and state equal the oracle's. The host-dt step
    leaves exactly the new state's speeds for the next device step; after
    the upload the next advance reduces them from the uploaded state."""
    nx, ny, npx, npy = 72, 56, 2, 2
    print_case("2d host dt between device dt", kernel)
    sim = Sim2D.mesh(nx, ny, npx, npy)
    configure(sim, kernel)
    m = MeshSpec(nx, ny, npx, npy)
    U = sim.scatter(P.sod2d(nx, ny, jitter=1e-3, seed=11))
    oracle.mesh_fill(U, m)
    sim.upload(U)
    sim.start(cfl=0.4)
    ref = U.copy()
    dx, dy = 1.0 / nx, 1.0 / ny
    sim.advance(3)
    oracle.mesh_run(ref, m, 3, dx, dy, cfl=0.4)
    dt = sim.compute_dt(0.4)[0]
    assert dt == oracle.mesh_dt(ref, m, dx, dy, 0.4)
    sim.rk2_step(dt)
    oracle.mesh_rk2_step(ref, m, dt, dx, dy)
    sim.advance(5, graph=True)
    _, dts = oracle.mesh_run(ref, m, 5, dx, dy, cfl=0.4)
    st = sim.status()
    assert st["steps"][0] == 8 and st["dt"][0] == dts[-1], (st, dts)
    assert np.array_equal(sim.gather(sim.download()), sim.gather(ref))
    # another state uploaded mid-run: the next steps' dt come from it
    V = sim.scatter(P.sod2d(nx, ny, jitter=1e-3, seed=12, x_disc=0.3))
    oracle.mesh_fill(V, m)
    sim.upload(V)
    sim.advance(6, graph=True)
    _, dts = oracle.mesh_run(V, m, 6, dx, dy, cfl=0.4)
    st = sim.status()
    assert st["steps"][0] == 14 and st["dt"][0] == dts[-1], (st, dts)
    assert np.array_equal(sim.gather(sim.download()), sim.gather(V))
    sim.close()


@pytest.mark.parametrize("kernel", STAGE_KERNELS)
def test_2d_limiters(oracle, kernel):
    nx, ny = 48, 40
    for lim in ["mc", "none"]:
        for mode in ["exact", "fast"]:
            print_case("limiter", lim, mode, kernel)
            sim = Sim2D.mesh(nx, ny, 2, 2, limiter=lim, mode=mode, bcx="periodic", bcy="periodic")
            configure(sim, kernel)
            m = MeshSpec(nx, ny, 2, 2, bcx="periodic", bcy="periodic")
            U = sim.scatter(P.smooth_periodic(48)[:, :nx, :ny].copy())
            oracle.mesh_fill(U, m)
            dt = oracle.mesh_dt(U, m, 1 / nx, 1 / ny, 0.4)
            sim.upload(U)
            sim.rk2_step(dt)
            ref = U.copy()
            oracle.mesh_rk2_step(ref, m, dt, 1 / nx, 1 / ny, lim=lim)
            got = sim.download()
            if mode == "exact":
                assert np.array_equal(got, ref)
            else:
                assert rel_linf_vars(sim.gather(got), sim.gather(ref)) <= TOL_FAST


def test_batch_matches_single_problems(oracle):
    """Ensemble of uneven problems stepped together == each problem alone."""
    sizes = [64, 81, 100, 130, 77]
    print_case("batch", sizes)
    sim = Sim2D.batch([(s, s) for s in sizes])
    U = np.zeros(sim.total)
    refs = []
    for k, s in enumerate(sizes):
        m = MeshSpec(s, s, 1, 1)
        G = P.quadrants(s, k)
        u = m.to_patches(G)
        oracle.mesh_fill(u, m)
        sim.patch_view(U, k)[:] = u.reshape(4, s + 4, s + 4)
        refs.append((m, u))
    sim.upload(U)
    sim.start(cfl=0.4)
    sim.advance(5, graph=True)
    got = sim.download()
    st = sim.status()
    for k, (m, u) in enumerate(refs):
        kk, dts = oracle.mesh_run(u, m, 5, 1.0 / m.nx, 1.0 / m.ny, cfl=0.4)
        assert st["steps"][k] == 5
        assert np.array_equal(sim.patch_view(got, k)[:, 2:-2, 2:-2], u.reshape(4, m.nx + 4, m.ny + 4)[:, 2:-2, 2:-2])


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_batch_reflecting_walls(oracle, mode):
    """Ensemble in closed boxes (all walls reflecting): single-patch problems
    read their ghosts straight from the interior, mirrored with the normal
    momentum negated. Device-dt run == the oracle's loop per problem; mass and
    energy of each box stay constant."""
    sizes = [(48, 64), (81, 81), (100, 52)]
    print_case("batch reflecting", sizes, mode)
    sim = Sim2D.batch(sizes, bc="reflecting", mode=mode)
    U = np.zeros(sim.total)
    refs = []
    for k, (nx, ny) in enumerate(sizes):
        m = MeshSpec(nx, ny, 1, 1, bcx="reflecting", bcy="reflecting")
        G = P.quadrants(max(nx, ny), k + 11)[:, :nx, :ny].copy()
        u = m.to_patches(G)
        oracle.mesh_fill(u, m)
        sim.patch_view(U, k)[:] = u.reshape(4, nx + 4, ny + 4)
        refs.append((m, u, G))
    sim.upload(U)
    sim.start(cfl=0.4)
    sim.advance(8, graph=False)
    got = sim.download()
    st = sim.status()
    for k, (m, u, G) in enumerate(refs):
        oracle.mesh_run(u, m, 8, 1.0 / m.nx, 1.0 / m.ny, cfl=0.4)
        assert st["steps"][k] == 8
        g = sim.patch_view(got, k)[:, 2:-2, 2:-2]
        w = u.reshape(4, m.nx + 4, m.ny + 4)[:, 2:-2, 2:-2]
        if mode == "exact":
            assert np.array_equal(g, w), k
        else:
            assert rel_linf_vars(g, w) <= 1e-11, k  # 8 free-running steps
        for var in (0, 3):
            assert abs(g[var].sum() - G[var].sum()) <= 1e-12 * G[var].sum()


def test_per_patch_path_equals_batched(oracle):
    nx, ny, npx, npy = 128, 96, 4, 3
    print_case("per-patch vs batched")
    a = Sim2D.mesh(nx, ny, npx, npy)
    b = Sim2D.mesh(nx, ny, npx, npy)
    U = a.scatter(P.sod2d(nx, ny))
    m = MeshSpec(nx, ny, npx, npy)
    oracle.mesh_fill(U, m)
    a.upload(U)
    b.upload(U)
    a.advance_per_patch(4, 0.4)
    b.start(cfl=0.4)
    b.advance(4)
    ga, gb = a.gather(a.download()), b.gather(b.download())
    assert np.array_equal(ga, gb)


def test_sod4096_full_size_one_step(oracle):
    """2D planar Sod at full size (4096^2, 16x16 patches of 256^2): one RK2 step
    bitwise vs the oracle, plus the seeded jitter's mass total."""
    n, npt = 4096, 16
    print_case("sod 4096^2 full size")
    sim = Sim2D.mesh(n, n, npt, npt)
    m = MeshSpec(n, n, npt, npt)
    U = sim.scatter(P.sod2d(n, n))
    oracle.mesh_fill(U, m)
    dt = oracle.mesh_dt(U, m, 1.0 / n, 1.0 / n, 0.4)
    sim.upload(U)
    assert sim.compute_dt(0.4)[0] == dt
    sim.rk2_step(dt)
    got = sim.download()
    oracle.mesh_rk2_step(U, m, dt, 1.0 / n, 1.0 / n)
    assert np.array_equal(got, U)


def test_state_over_4gib_runs_the_register_path(oracle):
    """A mesh whose state buffer is 4 GiB or more (11648^2 on 8x8 patches of
    1456^2: 545.7 M doubles per state). The one-launch step and the LDS-ring
    stage kernels address buffers with 32-bit byte offsets, so the handle
    must run the two-launch step with the register-path stage kernels
    (64-bit addressing). One device-dt step (exact numerics) equals the
    oracle's step, dt included, bitwise."""
    n, npt = 11648, 8
    print_case("state over 4 GiB", n)
    sim = Sim2D.mesh(n, n, npt, npt, mode="exact")
    assert sim.total * 8 >= 2**32 and not sim.one_launch
    m = MeshSpec(n, n, npt, npt)
    U = sim.scatter(P.sod2d(n, n))
    oracle.mesh_fill(U, m)
    sim.upload(U)
    sim.start(cfl=0.4)
    sim.advance(1)
    st = sim.status()
    got = m.to_global(sim.download())
    sim.close()
    k, dts = oracle.mesh_run(U, m, 1, 1.0 / n, 1.0 / n, cfl=0.4)
    assert k == 1 and st["steps"][0] == 1 and st["dt"][0] == dts[0], (st, dts)
    assert np.array_equal(got, m.to_global(U))


def test_kh8192_full_size_one_step(oracle):
    """The Kelvin-Helmholtz mesh at full size (8192^2, periodic, 64x64
    patches of 128^2): one RK2 step with the CFL dt, bitwise vs the oracle in
    exact numerics and within 1e-12 in fast numerics; the periodic totals of
    mass and energy are conserved to round-off."""
    n, npt = 8192, 64
    print_case("KH 8192^2 full size")
    m = MeshSpec(n, n, npt, npt, bcx="periodic", bcy="periodic")
    G = P.kelvin_helmholtz(n)
    U = m.to_patches(G)
    del G
    oracle.mesh_fill(U, m)
    dt = oracle.mesh_dt(U, m, 1.0 / n, 1.0 / n, 0.4)
    ref = U.copy()
    oracle.mesh_rk2_step(ref, m, dt, 1.0 / n, 1.0 / n)
    want = m.to_global(ref)
    del ref
    for mode in ("exact", "fast"):
        sim = Sim2D.mesh(n, n, npt, npt, bcx="periodic", bcy="periodic", mode=mode)
        sim.upload(U)
        sim.rk2_step(dt)
        got = m.to_global(sim.download())
        sim.close()
        if mode == "exact":
            assert np.array_equal(got, want)
        else:
            assert rel_linf_vars(got, want) <= TOL_FAST
        for k in (0, 3):
            tot0 = np.sum(m.to_global(U)[k], dtype=np.float64)
            assert abs(np.sum(got[k]) - tot0) <= 1e-12 * abs(tot0)


def test_ensemble512_batched_equals_per_patch_and_oracle(oracle):
    """The 512-problem ensemble at full size (64..512 cells a side): 3 batched
    device-dt steps equal 3 steps of the plain per-patch path bitwise, and a
    sample of problems equals the oracle run alone."""
    from nest_amd import problems as P
    print_case("ensemble of 512")
    a, sizes, U0 = P.ensemble_batch(512, seed=3, mode="exact")
    b = Sim2D.batch(sizes, mode="exact")
    b.upload(U0)
    a.start(cfl=0.4)
    a.advance(3, graph=True)
    b.advance_per_patch(3, 0.4)
    ga, gb = a.download(), b.download()
    st = a.status()
    assert (st["steps"] == 3).all() and (st["done"] == 0).all()
    for k in (0, 1, 77, 511):
        va, vb = a.patch_view(ga, k)[:, 2:-2, 2:-2], b.patch_view(gb, k)[:, 2:-2, 2:-2]
        assert np.array_equal(va, vb)
        s = sizes[k][0]
        m = MeshSpec(s, s, 1, 1)
        u = a.patch_view(U0, k).ravel().copy()
        oracle.mesh_run(u, m, 3, 1.0 / s, 1.0 / s, cfl=0.4)
        assert np.array_equal(va, u.reshape(4, s + 4, s + 4)[:, 2:-2, 2:-2]), k
    # every problem: identical between the two paths
    assert np.array_equal(ga, gb) or all(
        np.array_equal(a.patch_view(ga, k)[:, 2:-2, 2:-2], b.patch_view(gb, k)[:, 2:-2, 2:-2]) for k in range(512))


def test_device_copy_kernel():
    """The plain copy kernel (the measured bandwidth ceiling and counter
    calibration): exact copy of sizes around its grid cap and block size,
    including an empty copy and a size above 65536 * 256 elements (grid
    stride), and it rejects a negative size."""
    import ctypes as C

    import torch

    from nest_amd import _lib
    print_case("device copy")
    L = _lib.load()
    for n in (0, 1, 255, 257, 65536 * 256 + 3):
        src = torch.arange(n, dtype=torch.float64, device="cuda") * 0.5 - 7.0
        dst = torch.full((max(n, 1),), -1.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(L.nest_hip_device_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), n, None))
        torch.cuda.synchronize()
        assert torch.equal(dst[:n], src)
        if n == 0:
            assert dst[0].item() == -1.0
    with pytest.raises(_lib.NestError):
        _lib.check(L.nest_hip_device_copy(None, None, -1, None))


def test_dat_files_of_gpu_states(oracle, tmp_path):
    """The .dat outputs written from GPU-advanced states (Solver1D.write_dat,
    Sim2D.write_dat: device -> host -> file) are byte-identical to the same
    writers fed the oracle's state, and the writers themselves are pinned
    byte for byte against the reference's example writers
    (examples/sod1d/sod1d.cpp:50-91, examples/sod2d/sod2d.cpp:58-152;
    tests/test_dat_writers.py). Exact numerics: 1D Sod N=1000 to t=0.2 with
    device dt; 2D planar Sod 96 x 80 on 3 x 2 patches, 5 device-dt steps."""
    from nest_amd import dat
    print_case("dat files of GPU states")
    gdir, cdir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    cdir.mkdir()
    n = 1000
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc="outflow")
    s = Solver1D(n, 1.0 / n, bc="outflow", mode="exact")
    s.upload(U)
    k = s.run(0.2, 0.4)
    p_gpu = s.write_dat(str(gdir), 0.2, 4)
    s.close()
    ref = U.copy()
    k_ref, _ = oracle.run_1d(ref, n, 0.2, 0.4, 1.0 / n, bc="outflow")
    assert k == k_ref
    p_cpu = dat.write_sod1d(str(cdir), ref.reshape(3, -1)[:, 2:-2], 0.2, 4)
    assert open(p_gpu, "rb").read() == open(p_cpu, "rb").read()

    nx, ny, npx, npy = 96, 80, 3, 2
    m = MeshSpec(nx, ny, npx, npy)
    sim = Sim2D.mesh(nx, ny, npx, npy, mode="exact")
    U = sim.scatter(P.sod2d(nx, ny))
    oracle.mesh_fill(U, m)
    sim.upload(U)
    sim.start(cfl=0.4)
    sim.advance(5)
    t = float(sim.status()["t"][0])
    paths_gpu = sim.write_dat(str(gdir), t, 5)
    sim.close()
    ref = U.copy()
    oracle.mesh_run(ref, m, 5, 1.0 / nx, 1.0 / ny, cfl=0.4)
    G = m.to_global(ref)
    paths_cpu = [dat.write_sod2d(str(cdir), G, t, 5, npx, npy), dat.write_sod2d_xslice(str(cdir), G, t, 5)]
    assert len(paths_gpu) == 2
    for a, b in zip(paths_gpu, paths_cpu):
        assert os.path.basename(a) == os.path.basename(b)
        assert open(a, "rb").read() == open(b, "rb").read(), a


@pytest.mark.parametrize("kernel", STEP1D)
@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_1d_start_advance_and_stages(oracle, bc, kernel):
    """The 1D solver by step count (start / advance, device dt, one fused
    launch per step) in calls of 3, 4 and 1 steps: each call leaves the state
    in U whatever the parity of the steps taken, and equals the oracle's
    first 3, 7 and 8 steps bitwise; an advance that crosses t_end stops there
    (steps past it pass the state through); the per-stage entry points
    (stage 1, exchange of U1, stage 2) equal one oracle RK2 step."""
    n, cfl = 400, 0.4
    dx = 1.0 / n
    print_case("1d start/advance/stages", bc, kernel)
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc=bc)
    s = Solver1D(n, dx, bc=bc, mode="exact")
    s.set_step_kernel(kernel)
    s.upload(U)
    s.start(cfl)
    done, t, ref = 0, 0.0, U.copy()
    for calls in (3, 4, 1):
        s.advance(calls)
        done += calls
        for _ in range(calls):  # the oracle's loop (sod1d.cpp:219-272)
            dt = cfl * dx / oracle.max_speed_1d(ref, n)
            oracle.rk2_step_1d(ref, n, dt, dx, bc=bc)
            t += dt
        oracle.fill_1d(ref, n, bc=bc)
        st = s.status()
        assert st["steps"] == done and st["dt"] == dt and st["t"] == t, (st, dt, t)
        assert np.array_equal(s.download(), ref), (bc, done)
    # crossing the end time: 5 requested, the run stops after the oracle's count
    t_end = 0.004
    ref = U.copy()
    k_ref, _ = oracle.run_1d(ref, n, t_end, cfl, dx, bc=bc)
    s.upload(U)
    s.start(cfl, t_end)
    s.advance(k_ref + 5)
    st = s.status()
    assert st["steps"] == k_ref and st["done"] == 1
    assert np.array_equal(s.download(), ref)
    # per-stage entry points
    s.upload(U)
    dt = s.compute_dt(cfl)
    s.stage(1, dt)
    s.exchange(1)
    s.stage(2, dt)
    s.exchange(0)
    ref = U.copy()
    oracle.rk2_step_1d(ref, n, dt, dx, bc=bc)
    oracle.fill_1d(ref, n, bc=bc)
    assert np.array_equal(s.download(), ref)
    # a device-dt run continued after the state changed outside it (a
    # caller-dt step, then an upload): each next step's dt comes from the
    # current state
    s.upload(U)
    s.start(cfl)
    s.advance(2)
    ref = U.copy()
    for _ in range(2):
        oracle.rk2_step_1d(ref, n, cfl * dx / oracle.max_speed_1d(ref, n), dx, bc=bc)
    oracle.fill_1d(ref, n, bc=bc)
    dt = s.compute_dt(cfl)
    s.rk2_step(0.5 * dt)
    oracle.rk2_step_1d(ref, n, 0.5 * dt, dx, bc=bc)
    oracle.fill_1d(ref, n, bc=bc)
    s.advance(3)
    for _ in range(3):
        oracle.rk2_step_1d(ref, n, cfl * dx / oracle.max_speed_1d(ref, n), dx, bc=bc)
    oracle.fill_1d(ref, n, bc=bc)
    assert np.array_equal(s.download(), ref), bc
    V = P.sod1d(n, x_disc=0.3)  # another tube, uploaded mid-run
    oracle.fill_1d(V, n, bc=bc)
    s.upload(V)
    s.advance(2)
    for _ in range(2):
        oracle.rk2_step_1d(V, n, cfl * dx / oracle.max_speed_1d(V, n), dx, bc=bc)
    oracle.fill_1d(V, n, bc=bc)
    assert np.array_equal(s.download(), V), bc
    s.close()


@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflecting"])
def test_1d_large_lane_kernel_segments(oracle, bc):
    """The lane kernels with several segments per wave (n = 120 * 16384 + 78:
    3 segments of 56 cells or 2 of 120 cells per wave, a ragged last wave and
    block; n even, so the two-cells-per-lane kernel takes its 16-byte
    accesses, which the odd sizes of test_1d_run_fused_step_sizes do not):
    4 device-dt steps equal the oracle's bitwise (exact numerics) and within
    4e-12 (fast numerics)."""
    n = 120 * 16384 + 78
    print_case("1d lanes segments", bc)
    U = P.sod1d(n)
    oracle.fill_1d(U, n, bc=bc)
    ref = U.copy()
    dx = 1.0 / n
    for _ in range(4):
        dt = 0.4 * dx / oracle.max_speed_1d(ref, n)
        oracle.rk2_step_1d(ref, n, dt, dx, bc=bc)
    oracle.fill_1d(ref, n, bc=bc)
    outs = {}
    for mode in ("exact", "fast"):
        for kernel in ("lanes", "pairs", "tile"):
            s = Solver1D(n, dx, bc=bc, mode=mode)
            s.set_step_kernel(kernel)
            s.upload(U)
            s.start(0.4)
            s.advance(4)
            outs[(mode, kernel)] = s.download()
            s.close()
    for kernel in ("lanes", "pairs", "tile"):
        assert np.array_equal(outs[("exact", kernel)], ref), kernel
    for kernel in ("lanes", "pairs"):
        got, want = outs[("fast", kernel)].reshape(3, -1)[:, 2:-2], ref.reshape(3, -1)[:, 2:-2]
        assert rel_linf_vars(got, want) <= 4e-12, kernel


def test_1d_pair_kernel_size_limit():
    """The two-cells-per-lane kernel addresses the state with 32-bit byte
    offsets (3 (n + 4) doubles below 4 GiB): past that size an explicit
    request is refused (NEST_ERR_UNSUPPORTED) and the automatic choice takes
    the one-cell lane kernel, which one 2-step advance runs (no NaN, time
    advanced). Just below the limit the request is accepted."""
    from nest_amd import _lib
    n_max = (1 << 32) // 24 - 4  # largest n with 24 (n + 4) < 2^32
    print_case("1d pair kernel size limit", n_max)
    s = Solver1D(n_max, 1.0 / n_max, bc="outflow", mode="fast")
    s.set_step_kernel("pairs")
    s.close()
    n = n_max + 1
    s = Solver1D(n, 1.0 / n, bc="outflow", mode="fast")
    with pytest.raises(_lib.NestError) as e:
        s.set_step_kernel("pairs")
    assert e.value.code == 5
    s.set_step_kernel("auto")
    s.upload(P.sod1d(n))
    s.start(0.4)
    s.advance(2)
    st = s.status()
    assert st["steps"] == 2 and st["done"] == 0 and st["t"] > 0.0, st
    s.close()
