"""Benchmark of the gfx950 Euler step: 2D 4096^2 FP64 PLM+HLLE SSP-RK2.

Workload: planar Sod on [0,1]^2, 4096 x 4096 cells, seeded +-1e-3 density
jitter, outflow boundaries, decomposed into 16 x 16 patches of 256^2 with
2-cell halos exchanged on the device, CFL 0.4, device-resident dt. One step =
one SSP-RK2 step = two cell-updates per cell (the reference's "Updates/sec
(x2 for RK2)", examples/benchmark/benchmark_euler1d.cpp:217-218).

Timed region: K one-launch RK2 steps (the step kernel computes the step's
dt from the wave speeds the previous step left, runs both split stages in
one strip march, keeps U1 on chip, gathers its 4-cell halo from the
neighbour patches on the device and folds the wave speeds of the new state
for the next dt), launched one by one with no host
round trip, inputs resident in HBM, bracketed by stream synchronisation. It
starts after the W warm-up steps and, if those took less than WARM_S
seconds, after further untimed steps up to that time (the first launches
after an upload run slower: clocks and address translation warm up). The
same K steps replayed from the captured step graph give graph_value; a third
pass with HIP events on the stream the kernels run on gives each kernel's
duration for the roofline line.

Prints ONE JSON line (rank 0, N=1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "FP64 Mcell-updates/s, 2D 4096² PLM+HLLE RK2; % HBM peak; adjoint/fwd cost"
N, NP = 4096, 16
BYTES_STAGE1 = 64.0  # read U^n, write U1 (4 vars x 8 B each)
BYTES_STAGE2 = 96.0  # read U1, read U^n, write U^{n+1}
# algorithmic bytes of one RK2 step per cell with the stage state U1
# materialised (the metric's definition: stage 1 + stage 2 above); the
# one-launch step kernel's compulsory traffic is read U^n + write U^{n+1}
BYTES_STEP = BYTES_STAGE1 + BYTES_STAGE2
BYTES_STEP_COMPULSORY = 64.0
WARM_S = 1.5  # minimum warm-up time before the timed region


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(budget_s: float):
    """Oracle (CPU restatement of the reference numerics, OpenMP over patch
    rows) on the same 4096^2 workload, timed over whole RK2 steps until the
    time budget is spent. It runs in a child process so that its OpenMP
    runtime starts with OMP_PROC_BIND=close, OMP_PLACES=cores and one
    thread per CPU this process may run on (torch has already started this
    process's runtime)."""
    import subprocess

    threads = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", str(budget_s)],
                       capture_output=True, text=True, env=env, timeout=max(120.0, 6 * budget_s))
    if r.returncode != 0:
        log("bench: cpu baseline failed: " + r.stderr[-2000:])
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    steps, el, nthr = res["steps"], res["seconds"], res["threads"]
    val = N * N * 2.0 * steps / el / 1e6
    return {"value": round(val, 3), "unit": "Mcell-updates/s", "cores": nthr, "kind": "port",
            "sample": f"{steps} RK2 step(s) of the same 4096^2 16x16-patch Sod workload (dt from the "
                      f"state each step), {el:.1f} s wall on {nthr} OpenMP threads (OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores) of {os.cpu_count()} CPUs ({cpu_model()}); oracle built "
                      f"{res['build']}"}


ORACLE_FLAGS = ["-std=c++20", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"]


def _native_oracle(tmp: str):
    """The oracle compiled on this host with -march=native (a few seconds of
    g++), so the baseline uses the host's own instruction set; None where no
    compiler is available."""
    import subprocess

    src = os.path.join(ROOT, "oracle", "oracle_capi.cpp")
    out = os.path.join(tmp, "liboracle_native.so")
    try:
        r = subprocess.run(["g++", *ORACLE_FLAGS, "-march=native", "-shared", "-o", out, src],
                           capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    return out if r.returncode == 0 else None


def _cpu_baseline_child(budget_s: float) -> None:
    """The timed oracle loop of cpu_baseline (child process)."""
    import tempfile

    from nest_amd import problems as P
    from oracle.oracle import MeshSpec, Oracle

    with tempfile.TemporaryDirectory(prefix="nest_cpu_") as tmp:
        native = _native_oracle(tmp)
        o = Oracle(native) if native else Oracle()  # loaded: the file may go
    build = " ".join(ORACLE_FLAGS[1:2] + ORACLE_FLAGS[4:5]) + (
        " -march=native on this host" if native else " for the generic x86-64 target (no compiler here)")
    m = MeshSpec(N, N, NP, NP)
    U = m.to_patches(P.sod2d(N, N, jitter=1e-3, seed=2))
    o.mesh_fill(U, m)
    steps, t0 = 0, time.perf_counter()
    while True:
        dt = o.mesh_dt(U, m, 1.0 / N, 1.0 / N, 0.4)
        o.mesh_rk2_step(U, m, dt, 1.0 / N, 1.0 / N)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 20:
            break
    print(json.dumps({"steps": steps, "seconds": el, "threads": o.get_threads(), "build": build}), flush=True)


def copy_peak_gbs(gib: int = 2, reps: int = 10) -> dict:
    """Measured device-copy bandwidth (read + write bytes / time, HIP events
    on the launching stream), best of `reps` copies of `gib` GiB: the
    library's plain copy kernel (nest_hip_device_copy) and torch's copy."""
    import ctypes as C

    import torch

    from nest_amd import _lib

    L = _lib.load()
    n = gib * (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ts.append(a.elapsed_time(b) * 1e-3)
        return 2.0 * n * 8 / min(ts) / 1e9

    def lib_copy():
        _lib.check(L.nest_hip_device_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), n,
                                          C.c_void_p(stream.cuda_stream)))

    out = {"library_copy_kernel": round(best(lib_copy), 1), "torch_copy": round(best(lambda: dst.copy_(src)), 1)}
    del src, dst
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def lane_efficiency(sizes, width: int = 56) -> float:
    """Share of the step kernel's lane-rows that produce an output cell: a
    problem of nx x ny cells runs ceil(ny / width) 64-lane strips down its nx
    rows (the one-launch step kernel's strips produce 56 columns, the
    two-launch step's 60); warm-up rows not counted."""
    out = sum(a * b for a, b in sizes)
    lanes = sum(64 * (-(-b // width)) * a for a, b in sizes)
    return out / lanes


def ensemble_leg(steps: int, mode: str):
    """Ensemble: 512 uneven four-quadrant problems, 100 RK2 steps with
    per-problem device dt. Batched (one launch per stage over all problems,
    hipGraph-replayed step, autotuned tiles) against two per-patch paths:
    the plain one (host dt with a blocking read-back, one launch set per
    problem), and one handle per problem with its dt on the device (two
    steps per call, no host round trip), every handle on one stream and then
    each on a stream of its own."""
    import numpy as np
    import torch

    from nest_amd import Sim2D, autotune, problems as P

    sim, sizes, U0 = P.ensemble_batch(512, seed=3, mode=mode)
    sim_one = sim.one_launch
    cells = float(sum(a * b for a, b in sizes))
    sim.start(cfl=0.4)
    cfg = autotune.tune(sim, sizes, log=log)
    sim.upload(U0)
    sim.start(cfl=0.4)
    sim.advance(2, graph=True)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance(steps, graph=True)
    sim.synchronize()
    tb = time.perf_counter() - t0
    pp_steps = max(5, steps // 10)
    sim.upload(U0)
    sim.advance_per_patch(1, 0.4)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance_per_patch(pp_steps, 0.4)
    sim.synchronize()
    tp = time.perf_counter() - t0
    states = [np.ascontiguousarray(sim.patch_view(U0, k)).copy() for k in range(len(sizes))]
    sim.close()
    # one handle per problem (its single patch), device dt
    st = torch.cuda.Stream()  # default priority
    hs = []
    for (s, _), Uk in zip(sizes, states):
        h = Sim2D.mesh(s, s, 1, 1, mode=mode)
        h.set_stream(st.cuda_stream)
        h.upload(Uk)
        h.start(cfl=0.4)
        hs.append(h)
    for h in hs:
        h.advance(2)
    st.synchronize()
    dd_steps = 2 * max(3, steps // 20)
    t0 = time.perf_counter()
    for _ in range(dd_steps // 2):
        for h in hs:
            h.advance(2)
    st.synchronize()
    td = time.perf_counter() - t0
    # the same with every handle on a stream of its own (concurrent launches)
    for h in hs:
        h.set_stream(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(dd_steps // 2):
        for h in hs:
            h.advance(2)
    for h in hs:
        h.synchronize()
    tc = time.perf_counter() - t0
    for h in hs:
        h.close()
    batched = cells * 2.0 * steps / tb / 1e6
    per_patch = cells * 2.0 * pp_steps / tp / 1e6
    per_problem = cells * 2.0 * dd_steps / td / 1e6
    per_problem_streams = cells * 2.0 * dd_steps / tc / 1e6
    return {"problems": 512, "cells": int(cells), "steps": steps, "batched_mcups": round(batched, 1),
            "per_patch_mcups": round(per_patch, 1), "speedup": round(batched / per_patch, 2),
            "per_patch_steps_timed": pp_steps,
            "per_problem_device_dt_mcups": round(per_problem, 1),
            "speedup_vs_per_problem_device_dt": round(batched / per_problem, 2),
            "per_problem_device_dt_steps_timed": dd_steps,
            "per_problem_device_dt_own_streams_mcups": round(per_problem_streams, 1),
            "speedup_vs_per_problem_own_streams": round(batched / per_problem_streams, 2),
            "step_kernel": "one launch per step" if sim_one else "two stage launches",
            "lane_efficiency": round(lane_efficiency(sizes, 56 if sim_one else 60), 4),
            "tiling": {k: v for k, v in cfg.items() if k in ("1", "2", "3", "persistent")}}


def adjoint_leg(mode: str, K: int = 200, ndir: int = 16, eps: float = 1e-9, reps: int = 3):
    """Adjoint: 1024^2 periodic smooth state, K fixed-dt RK2 steps, loss =
    L2 misfit of the final density against the run from a seeded perturbed
    initial state. Times (median of `reps`): the plain forward rollout (one
    checkpoint); the forward and the whole gradient (forward + reverse sweep)
    under the rollout's default storage (every state kept when 2K states fit
    in half the free HBM, as at this size); the same with ceil(sqrt K)
    checkpoints, whose reverse recomputes each interval. cost = gradient /
    plain forward. Then checks the gradient against central differences along
    `ndir` seeded smooth directions of the full state at a generic point
    next to the initial state (its rho, u, v, p plus a seeded 1e-3 smooth
    field each; the initial state's exactly uniform u, v, p put every
    velocity / pressure limiter slope on its kink, where the loss has no
    two-sided derivative); and, for the unlimited scheme (limiter none,
    differentiable: every face stays subsonic, no floor is active), at the
    initial state itself, where central differences are accurate."""
    import numpy as np
    import torch

    from nest_amd import problems as P
    from nest_amd.adjoint import Rollout, density_misfit

    n = 1024
    U0, Ut, dt = P.adjoint_setup(n)
    ro = Rollout.mesh(n, n, 4, 4, nsteps=K, dt=dt, mode=mode)
    sizes = [(n, n)]
    ro.sim.upload(ro.sim.scatter(U0))
    fwd_cfg, adj_cfg = ro.tune(sizes, log=log)
    area = 1.0 / (n * n)
    with torch.no_grad():
        target = ro(torch.from_numpy(Ut).cuda())[0].clone()

    def loss(U):
        return density_misfit(ro(U), target, area)

    def timed(fn):
        out = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            out.append(time.perf_counter() - t0)
        return sorted(out[1:])[len(out[1:]) // 2]

    def gradient(U0d):
        U = U0d.clone().requires_grad_(True)
        loss(U).backward()
        return U.grad

    U0d = torch.from_numpy(U0).cuda()

    m_ck = int(np.ceil(np.sqrt(K)))

    def fwd(interval):
        def run():
            ro.interval = interval
            with torch.no_grad():
                loss(U0d)
        return run

    def grad(interval):
        def run():
            ro.interval = interval
            gradient(U0d)
        return run

    t_fwd = timed(fwd(K))
    # the same K fixed-dt steps with the one-launch step kernel (nothing kept:
    # the fastest forward of the map this library has)
    from nest_amd import Sim2D, autotune
    s1 = Sim2D.mesh(n, n, 4, 4, bcx="periodic", bcy="periodic", mode=mode)
    U0p = s1.scatter(U0)
    s1.upload(U0p)
    s1.start(cfl=0.4, fixed_dt=dt)
    autotune.tune(s1, sizes, log=log)

    def fwd_one():  # (each run continues from the last state: the same work at a fixed dt)
        s1.start(cfl=0.4, fixed_dt=dt)
        s1.advance(K, graph=True)
        s1.synchronize()
    t_fwd_one = timed(fwd_one)
    s1.close()
    t_store = timed(fwd(0))
    t_grad = timed(grad(0))
    kept = ro.rollout_interval() < 0
    t_ck = timed(fwd(m_ck))
    t_grad_ck = timed(grad(m_ck))
    ro.interval = 0
    scan = (1e-5, 1e-6, 1e-7, 1e-8, 1e-9)

    def fd_check(loss_fn, Ug, eps_pick):
        """Projected relative L2 error of the gradient against central
        differences along the ndir directions, per step in the scan; the
        error at eps_pick and the worst single direction there."""
        U = Ug.clone().requires_grad_(True)
        loss_fn(U).backward()
        g = U.grad
        d_adj, d_fd = [], {e: [] for e in scan}
        with torch.no_grad():
            for k in range(ndir):
                v = torch.from_numpy(P.smooth_direction(n, 200 + k)).cuda()
                d_adj.append(float(torch.sum(g * v)))
                for e in scan:
                    d_fd[e].append((float(loss_fn(Ug + e * v)) - float(loss_fn(Ug - e * v))) / (2 * e))
        a = np.array(d_adj)
        errs = {f"{e:.0e}": float(f"{np.linalg.norm(np.array(d_fd[e]) - a) / np.linalg.norm(a):.3e}")
                for e in scan}
        f = np.array(d_fd[eps_pick])
        return (errs, float(f"{np.linalg.norm(f - a) / np.linalg.norm(a):.3e}"),
                float(f"{np.max(np.abs(f - a) / np.abs(a)):.3e}"))

    scan_err, proj_err, worst, fd_none = None, None, None, None
    if ndir > 0:
        scan_err, proj_err, worst = fd_check(loss, torch.from_numpy(P.adjoint_generic_state(n)).cuda(), eps)
    ro.close()
    if ndir > 0:
        # the unlimited scheme at the initial state itself (its own target)
        rn = Rollout.mesh(n, n, 4, 4, nsteps=K, dt=dt, mode=mode, limiter="none")
        rn.interval = -1
        with torch.no_grad():
            tgt_n = rn(torch.from_numpy(Ut).cuda())[0].clone()
        errs, proj_n, worst_n = fd_check(lambda U: density_misfit(rn(U), tgt_n, area), U0d, 1e-6)
        rn.close()
        best_eps = min(errs, key=errs.get)
        fd_none = {"limiter": "none", "point": "the initial state", "directions": ndir, "eps_scan": errs,
                   "eps": 1e-6, "projected_rel_l2_err": proj_n, "max_rel_err_per_direction": worst_n,
                   "best_eps": best_eps}
    cells = n * n
    state_gb = ro.sim.total * 8 / 1e9
    storage = (f"every state and stage-1 state kept in HBM ({2 * K * state_gb:.1f} GB)" if kept
               else f"checkpoints every {m_ck} steps")
    return {"workload": f"1024^2 periodic smooth state (4x4 patches of 256^2), K={K} fixed-dt RK2 steps, "
                        f"density L2 misfit, {storage}",
            "forward_ms": round(t_fwd * 1e3, 2), "forward_stored_ms": round(t_store * 1e3, 2),
            "forward_one_launch_ms": round(t_fwd_one * 1e3, 2),
            "adjoint_cost_vs_one_launch_forward": round(t_grad / t_fwd_one, 3),
            "checkpointed_cost_vs_one_launch_forward": round(t_grad_ck / t_fwd_one, 3),
            "gradient_ms": round(t_grad * 1e3, 2), "adjoint_fwd_cost": round(t_grad / t_fwd, 3),
            "reverse_only_cost": round((t_grad - t_store) / t_fwd, 3),
            "checkpointed": {"interval": m_ck, "forward_ckpt_ms": round(t_ck * 1e3, 2),
                             "gradient_ms": round(t_grad_ck * 1e3, 2),
                             "adjoint_fwd_cost": round(t_grad_ck / t_fwd, 3),
                             "reverse_only_cost": round((t_grad_ck - t_ck) / t_fwd, 3)},
            "forward_mcups": round(cells * 2.0 * K / t_fwd / 1e6, 1),
            "tiling": {"forward": {k: v for k, v in fwd_cfg.items() if k in ("1", "2", "3", "persistent")},
                       "adjoint": {k: v for k, v in adj_cfg.items() if k != "ms_per_step"}},
            "fd_point": "initial state + seeded 1e-3 smooth perturbation of rho, u, v, p",
            "dt": float(dt).hex(),
            "fd_directions": ndir, "fd_eps": eps, "fd_projected_rel_l2_err": proj_err,
            "fd_eps_scan": scan_err,
            "fd_max_rel_err_per_direction": worst,
            "fd_unlimited": fd_none}


def oned_leg(large_steps: int = 200, sod_t_end: float = 0.2):
    """1D Sod with device dt through Solver1D.run (one launch of the lane step
    kernel per step, dt included, replayed 48 steps at a time from a captured
    graph):
    * the configuration of the reference's published 1D GPU figure
      (examples/benchmark/benchmark_euler1d.cpp: N = 1e5, periodic, CFL 0.4,
      t = 0.2; exact numerics, as the euler1d_gpu_* drop-in), median of 3 runs
      after one untimed run (`sod_t_end` shortens it for profiling runs);
    * N = 2^24 cells (outflow, fast numerics), about `large_steps` steps:
      hbm_frac counts the 48 B per cell-step the fused kernel moves (read
      Un, write Un+1, 3 doubles each); algorithmic_gbs_u1_materialised the
      120 B of the two stages with U1 materialised (stage 1 reads and writes
      3 doubles, stage 2 reads 6 and writes 3)."""
    from nest_amd import Solver1D, problems as P
    out = {}
    n = 100000
    U = P.sod1d(n)
    s = Solver1D(n, 1.0 / n, bc="periodic", mode="exact")
    times, steps = [], 0
    for r in range(4):
        s.upload(U)
        t0 = time.perf_counter()
        steps = s.run(sod_t_end, 0.4)
        times.append(time.perf_counter() - t0)
    s.close()
    t = sorted(times[1:])[1]
    out["sod_1e5"] = {"cells": n, "t_end": sod_t_end, "bc": "periodic", "numerics": "exact", "steps": steps,
                      "ms": round(t * 1e3, 2), "us_per_step": round(t * 1e6 / steps, 2),
                      "mcups": round(n * 2.0 * steps / t / 1e6, 1),
                      "reference_published_mcups": 5000.0}
    n = 1 << 24
    U = P.sod1d(n)
    s = Solver1D(n, 1.0 / n, bc="outflow", mode="fast")
    s.upload(U)
    t_end = large_steps * s.compute_dt(0.4)  # about large_steps steps (the state stays near the first dt)
    s.run(t_end, 0.4)  # untimed: captures the step graph
    s.upload(U)
    t0 = time.perf_counter()
    k = s.run(t_end, 0.4)
    t = time.perf_counter() - t0
    s.close()
    gbs = 48.0 * n * k / t / 1e9
    out["large"] = {"cells": n, "bc": "outflow", "numerics": "fast", "steps": k, "ms": round(t * 1e3, 2),
                    "mcups": round(n * 2.0 * k / t / 1e6, 1), "bytes_per_cell_step": 48,
                    "gbs": round(gbs, 1), "hbm_frac": round(gbs / 8000.0, 4),
                    "algorithmic_gbs_u1_materialised": round(gbs * 120.0 / 48.0, 1)}
    return out


def kh_leg(mode: str, steps: int = 100, nadj: int = 64, adj_steps: int = 200):
    """Kelvin-Helmholtz 8192^2 periodic on 64x64 patches of 128^2,
    `steps` RK2 steps with device dt (graph-replayed, autotuned tiling); then `nadj`
    concurrent adjoint gradients of 256^2 instances of the adjoint leg's setup
    (each its own seed) in one batch handle, K = adj_steps, against the same
    gradients one problem at a time; then both together (the KH steps and the
    batched gradients enqueued at once on their handles' own streams, which
    run side by side), against the two run alone."""
    import numpy as np
    import torch

    from nest_amd import Sim2D, problems as P
    from nest_amd.adjoint import Rollout, density_misfit

    n, npch = 8192, 64
    sim = Sim2D.mesh(n, n, npch, npch, bcx="periodic", bcy="periodic", mode=mode)
    U_kh = sim.scatter(P.kelvin_helmholtz(n))

    def kh_reset():
        sim.upload(U_kh)
        sim.fill_halos()
        sim.start(cfl=0.4)
        sim.synchronize()

    kh_reset()
    from nest_amd import autotune
    kh_cfg = autotune.tune(sim, [(n, n)], log=log)
    kh_reset()
    sim.advance(2, graph=True)
    kh_reset()
    t0 = time.perf_counter()
    sim.advance(steps, graph=True)
    sim.synchronize()
    t_kh = time.perf_counter() - t0
    st = sim.status()
    kh = {"cells": n * n, "patches": npch * npch, "steps": steps,
          "mcups": round(n * n * 2.0 * steps / t_kh / 1e6, 1), "ms_per_step": round(t_kh * 1e3 / steps, 3),
          "t_reached": float(f"{st['t'][0]:.6g}"),
          "tiling": {k: v for k, v in kh_cfg.items() if k in ("1", "2", "3", "persistent")}}

    m = 256
    setups = [P.adjoint_setup(m, seed=10 + 3 * b) for b in range(nadj)]
    dt = min(s[2] for s in setups)
    area = 1.0 / (m * m)
    rb = Rollout.batch(nadj, m, m, nsteps=adj_steps, dt=dt, mode=mode)
    fwd_cfg, adj_cfg = rb.tune([(m, m)] * nadj, log=log)
    U0 = torch.from_numpy(np.stack([s[0] for s in setups])).cuda()
    with torch.no_grad():
        tgt = rb(torch.from_numpy(np.stack([s[1] for s in setups])).cuda())[:, 0].clone()

    def batched():
        U = U0.clone().requires_grad_(True)
        density_misfit(rb(U), tgt, area).backward()
        return U.grad

    def forward():  # plain forward: no states kept
        rb.interval = adj_steps
        with torch.no_grad():
            density_misfit(rb(U0), tgt, area)
        rb.interval = 0

    def timed(fn, reps=3):  # median of `reps` runs after one untimed run
        fn()
        out = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            out.append(time.perf_counter() - t)
        return sorted(out)[reps // 2]

    t_b = timed(batched)
    t_f = timed(forward)
    kept = rb.rollout_interval() < 0
    kh_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.advance(steps, graph=True)  # enqueued on the KH handle's stream
    batched()  # the rollout's own stream
    torch.cuda.synchronize()
    sim.synchronize()
    t_both = time.perf_counter() - t0
    st_both = sim.status()
    sim.close()
    rb.close()
    rs = Rollout.mesh(m, m, nsteps=adj_steps, dt=dt, mode=mode)
    rs.tune([(m, m)], log=log)
    nseq = 8  # single-problem gradients timed (the rate is per gradient)

    def sequential():
        for b in range(nseq):
            u = U0[b].clone().requires_grad_(True)
            density_misfit(rs(u), tgt[b], area).backward()

    t_s = timed(sequential) / nseq
    rs.close()
    return {"kh": kh, "batched_adjoints": {
        "problems": nadj, "side": m, "steps": adj_steps,
        "storage": "every state kept" if kept else f"checkpoints every {int(np.ceil(np.sqrt(adj_steps)))} steps",
        "gradient_ms": round(t_b * 1e3, 2), "forward_ms": round(t_f * 1e3, 2),
        "adjoint_fwd_cost": round(t_b / t_f, 3),
        "gradients_per_s": round(nadj / t_b, 2),
        "single_problem_gradient_ms": round(t_s * 1e3, 2),
        "batching_speedup": round(t_s * nadj / t_b, 2)},
        "together": {"ms": round(t_both * 1e3, 2), "kh_alone_ms": round(t_kh * 1e3, 2),
                     "adjoints_alone_ms": round(t_b * 1e3, 2),
                     "sum_over_together": round((t_kh + t_b) / t_both, 3),
                     "kh_steps": int(st_both["steps"][0])}}


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--cpu-baseline-child":
        _cpu_baseline_child(float(sys.argv[2]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--two-stage", action="store_true",
                    help="time the two-launch step (stage 1 writes U1, stage 2 reads it) instead of the one-launch step")
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--per-patch-steps", type=int, default=3)
    ap.add_argument("--ensemble-steps", type=int, default=100, help="ensemble leg (512 uneven problems); 0 skips it")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="adjoint leg (1024^2, K steps); 0 skips it")
    ap.add_argument("--adjoint-fd-dirs", type=int, default=16,
                    help="central-difference directions of the adjoint gradient check; 0 skips the check")
    ap.add_argument("--kh-steps", type=int, default=100, help="KH 8192^2 + 64 batched 256^2 adjoints leg; 0 skips it")
    ap.add_argument("--oned-steps", type=int, default=200, help="1D leg (N=1e5 to t=0.2; N=2^24 this many steps); 0 skips it")
    ap.add_argument("--oned-sod-t-end", type=float, default=0.2, help="end time of the 1D N=1e5 run (profiling aid)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="profiling aid: one plain device copy of this many MiB after the legs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; sharding the path over several GPUs is "
              "designed but not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from nest_amd import Sim2D, problems as P

    log(f"bench: building 4096^2 Sod state ({args.mode})")
    G = P.sod2d(N, N, jitter=1e-3, seed=2)
    sim = Sim2D.mesh(N, N, NP, NP, mode=args.mode)
    if args.two_stage:
        sim.set_step_kernel("two_stage")
    U = sim.scatter(G)
    tiling = None
    if args.rows or args.wpb:
        sim.set_tiling(args.rows, args.wpb, 0)
    else:
        from nest_amd import autotune
        sim.upload(U)
        sim.fill_halos()
        sim.start(cfl=0.4)
        tiling = autotune.tune(sim, [(N, N)], log=log)
    sim.upload(U)
    sim.fill_halos()
    sim.start(cfl=0.4)
    log("bench: warmup")
    tw = time.perf_counter()
    sim.advance(args.warmup)
    sim.synchronize()
    extra_warm = 0
    while time.perf_counter() - tw < WARM_S:  # steady state before the timed region
        sim.advance(20)
        sim.synchronize()
        extra_warm += 20

    # the timed region: K device-dt steps launched one by one (the default
    # path; no graph, no events)
    log(f"bench: timed {args.steps} steps")
    cells = float(N * N)
    t0 = time.perf_counter()
    sim.advance(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = cells * 2.0 * args.steps / wall / 1e6

    # the same K steps replayed from the captured step graph; 2 more untimed
    # steps capture and warm the graph
    log(f"bench: {args.steps} steps (hipGraph replay)")
    sim.advance(2, graph=True)
    sim.synchronize()
    tg = time.perf_counter()
    sim.advance(args.steps, graph=True)
    sim.synchronize()
    graph_value = cells * 2.0 * args.steps / (time.perf_counter() - tg) / 1e6

    # per-kernel durations: the same steps with HIP events around each launch
    # on the handle's stream (the events add launch gaps, so this pass is not
    # the throughput figure)
    log("bench: per-kernel timing")
    comp_ms, ev_total_ms = sim.advance_timed(args.steps)
    st = sim.status()
    assert st["done"][0] == 0 and st["steps"][0] == args.warmup + extra_warm + 2 + 3 * args.steps
    one = sim.one_launch
    forced_tiling = sim.get_step_tiling() if (args.rows or args.wpb) else None
    k1_ms = comp_ms[1] / args.steps
    k2_ms = comp_ms[2] / args.steps
    if one:
        # dominant kernel: the one-launch step, 160 B per cell algorithmic
        # (U1 materialised, as the metric counts it), 64 B compulsory
        kernel_name = f"step2d_kernel<minmod> ({'fast' if args.mode == 'fast' else 'exact'})"
        kern_ms = k1_ms
        kern_bytes = BYTES_STEP
    else:
        # dominant kernel: stage 2 of the two-launch step (96 B per cell)
        kernel_name = "stage2d_dma_kernel<minmod, 2>"
        kern_ms = k2_ms
        kern_bytes = BYTES_STAGE2
    achieved = kern_bytes * cells / (kern_ms * 1e-3) / 1e9
    step_gbs = BYTES_STEP * cells / (ms_per_step * 1e-3) / 1e9
    stages_gbs = BYTES_STEP * cells / ((k1_ms + k2_ms) * 1e-3) / 1e9

    # plain per-patch drop-in path (host dt, one launch set per patch)
    pp = None
    if args.per_patch_steps > 0:
        log("bench: per-patch path")
        sim.advance_per_patch(1, 0.4)
        sim.synchronize()
        tp = time.perf_counter()
        sim.advance_per_patch(args.per_patch_steps, 0.4)
        sim.synchronize()
        pp_wall = time.perf_counter() - tp
        pp = round(cells * 2.0 * args.per_patch_steps / pp_wall / 1e6, 1)
    if args.copy_probe_mib > 0:
        # profiling aid for the HBM-traffic counters (outside every timed
        # region): the library's pack kernel reads the 4096^2 state's interior
        # (512 MiB, 8 B per lane, coalesced like the stage kernels) and writes
        # it densely (512 MiB)
        import torch
        packed = torch.empty(sim.global_size(), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        sim.pack(sim.device_ptrs()[0], packed.data_ptr())
        sim.synchronize()
        del packed
    sim.close()
    del U

    ens = None
    if args.ensemble_steps > 0:
        log("bench: ensemble (512 problems)")
        ens = ensemble_leg(args.ensemble_steps, args.mode)

    adj = None
    if args.adjoint_steps > 0:
        log("bench: adjoint (1024^2, K steps)")
        adj = adjoint_leg(args.mode, K=args.adjoint_steps, ndir=args.adjoint_fd_dirs)

    c5 = None
    if args.kh_steps > 0:
        log("bench: KH 8192^2, 64 batched 256^2 adjoints")
        c5 = kh_leg(args.mode, steps=args.kh_steps)

    oned = None
    if args.oned_steps > 0:
        log("bench: 1D Sod (N=1e5 to t=0.2; N=2^24)")
        oned = oned_leg(args.oned_steps, args.oned_sod_t_end)

    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(args.cpu_budget)

    if args.copy_probe_mib > 0:
        # a plain device copy of known size (the library's copy kernel), for
        # scaling the FETCH_SIZE / WRITE_SIZE counters of profiled runs (not
        # part of any timed region)
        import ctypes as C

        import torch

        from nest_amd import _lib
        n = args.copy_probe_mib * 131072
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        _lib.check(_lib.load().nest_hip_device_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), n,
                                                    C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        del src, dst

    copies = copy_peak_gbs()
    copy_gbs = copies["library_copy_kernel"]
    step_tiling = None
    if one:
        # measured HBM bytes per launch of the tiling this run used (rows per
        # work item x waves per block), from the rocprof counter passes
        if forced_tiling:
            step_tiling = forced_tiling
        else:
            from nest_amd import autotune
            cls = autotune.size_classes([(N, N)])[0]
            step_tiling = (tiling["3"]["rows"][cls], tiling["3"]["waves"])
        traffic = load_traffic("traffic_step2d_sod4096.json")
        entry = ((traffic or {}).get("tilings") or {}).get(f"{step_tiling[0]}x{step_tiling[1]}") or {}
        tr_bytes = entry.get("step_bytes_per_launch")
    else:
        traffic = load_traffic("traffic_stage2d_sod4096.json")
        tr_bytes = (traffic or {}).get("stage2_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded planar Sod with +-1e-3 density jitter)",
        "config": {"workload": "2D planar Sod 4096x4096 FP64, 16x16 patches of 256^2, halo 2, outflow, "
                               "PLM(minmod)+HLLE, SSP-RK2, CFL 0.4, device dt",
                   "numerics": args.mode, "cells": N * N, "patches": NP * NP},
        "mcell_steps_per_s": round(value / 2.0, 1),
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "kernels_hbm_frac": round(stages_gbs / HBM_PEAK_GBS, 4),
        "step_kernel": "one launch per step (U1 on chip)" if one else "two stage launches",
        "lane_efficiency": round(lane_efficiency([(N, N)], 56 if one else 60), 4),
        "kernel_ms": ({"step": round(k1_ms, 4)} if one else
                      {"finalize": round(comp_ms[0] / args.steps, 4), "stage1": round(k1_ms, 4),
                       "stage2": round(k2_ms, 4)}),
        "extra_warmup_steps": extra_warm,
        "tiling": tiling,
        "timed_path": "device-dt steps launched one by one (graph_value: the same steps replayed from the captured step graph)",
        "graph_value": round(graph_value, 1),
        "per_patch_value": pp,
        "ensemble": ens,
        "adjoint": adj,
        "kh_batched_adjoints": c5,
        "oned": oned,
        "roofline": {"bound": "hbm", "kernel": kernel_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": int(kern_bytes * cells),
                     "compulsory_bytes_per_launch": int((BYTES_STEP_COMPULSORY if one else kern_bytes) * cells),
                     "compulsory_gbs": round((BYTES_STEP_COMPULSORY if one else kern_bytes) * cells
                                             / (kern_ms * 1e-3) / 1e9, 1),
                     "measured_copy_peak": round(copy_gbs, 1), "frac_vs_copy_peak": round(achieved / copy_gbs, 4),
                     "copy_kernels": copies,
                     "traffic": tr_bytes,
                     "traffic_tiling": f"{step_tiling[0]} rows x {step_tiling[1]} waves" if step_tiling else None,
                     # measured bytes (rocprof counters) moved per second of the launch
                     "traffic_gbs": round(tr_bytes / (kern_ms * 1e-3) / 1e9, 1) if tr_bytes else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
