"""Autotuner of the step kernels' work decomposition.

Per kernel of the step that runs (the one-launch step kernel, or stage 1 and
stage 2 of the two-launch step) it picks, by measurement on the live problem
set:
  * rows per work item for each problem size class (the tile height of the
    strip march: short tiles give more waves and a finer tail, long tiles
    amortise the 4-row warm-up of the X window), and
  * waves per block of the launch, and
  * whether the launch is persistent (grid capped at the resident waves,
    each wave looping over the device work list) or one wave per item.
Search: coordinate descent from the defaults, each candidate timed as the
fastest of five runs of eight device-dt steps with HIP events on the
handle's stream (the same timing path as bench.py; shorter trials let
±3% noise pick tilings up to 7% slower at 4096^2). Results are cached in a
JSON file keyed by device name
and the problem set's size-class histogram, so a later run on the same kind
of input starts from the tuned point.

tune_adjoint() does the same for the two adjoint kernels (rows per work item
and waves per block of each), timing a short checkpointed rollout and its
reverse sweep (the path the differentiable rollouts take).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

ROWS = (4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128)
ADJ_ROWS = (6, 8, 10, 12, 16, 20, 24, 32, 48)
WAVES = (1, 2, 4)


def size_classes(sizes, edges=(128, 256)):
    """Class index of each problem from its largest side: < 128, < 256, >= 256."""
    return [int(np.searchsorted(edges, max(s), side="right")) for s in sizes]


def signature(sizes, classes, arch: str) -> str:
    hist = np.bincount(classes, minlength=3).tolist()
    cells = int(sum(a * b for a, b in sizes))
    return f"{arch}|classes={hist}|cells={cells}"


def _cache_path(path: str | None) -> str:
    if path:
        return path
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(base, "nest_amd", "tuning.json")


def _time(sim, steps, reps=5):
    """(stage 1, stage 2) kernel ms per step of the fastest of `reps` timed
    runs of `steps` device-dt steps."""
    sim.advance(1)
    best = None
    for _ in range(reps):
        ms, _ = sim.advance_timed(steps)
        t = (ms[1] / steps, ms[2] / steps)
        if best is None or sum(t) < sum(best):
            best = t
    return best


def apply(sim, classes, cfg) -> None:
    """Set a tuned configuration {kernel: {"rows": [per class], "waves": w},
    "persistent": 0|1}; kernel "1" / "2": the stages of the two-launch
    step, "3": the one-launch step kernel."""
    sim.set_tiling(0, 0, cfg.get("persistent", 0))
    for st in ("1", "2", "3"):
        if st not in cfg:
            continue
        c = cfg[st]
        sim.set_stage_tiling(int(st), 0, c["waves"])
        sim.set_problem_rows(int(st), [c["rows"][k] for k in classes])


def _load(path):
    if os.path.exists(path):
        try:
            return json.load(open(path))
        except (OSError, ValueError):
            return {}
    return {}


def _store(path, table):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        json.dump(table, open(path, "w"), indent=1, sort_keys=True)
    except OSError:
        pass


def tune_adjoint(sim, sizes, dt: float, arch: str = "gfx950", reps: int = 5, cache: str | None = None, log=None,
                 keep_all: bool = False):
    """Tune the adjoint kernels' tiling of `sim` in place: coordinate descent
    (two passes) over rows and waves per block of the Y and X kernels. One
    16-step rollout from the handle's own state is recorded first, with
    checkpoints every 2 steps: its reverse sweeps recompute intervals on the
    auxiliary stream while they run, as checkpointed gradients do. Each
    candidate is then timed as the fastest of `reps` reverse sweeps, with the
    handle's stage scratch as the cotangent: the tiling changes the reverse
    sweep only, and timing the forward with it hid differences of a few
    percent (an 8-step forward + reverse per trial picked 20 x 20 rows at
    1024^2, 8% slower per gradient than 10 x 16). A tiling that is fast
    alone can lose beside the recompute: one wave per block for the X kernel
    timed best on reverses with nothing to recompute, and the X kernel then
    took 36.7 against 33.2 us per launch in a checkpointed gradient (every
    state kept gained nothing from it). A candidate replaces the current
    choice only when it is 1% faster (timing noise). `keep_all` selects the
    cache entry of the caller's storage. Returns {"rows_y", "waves_y",
    "rows_x", "waves_x", "ms_per_step"} (ms per reversed step)."""
    classes = size_classes(sizes)
    key = ("adjoint-rollout-kept|" if keep_all else "adjoint-rollout|") + signature(sizes, classes, arch)
    path = _cache_path(cache)
    table = _load(path)
    if key in table:
        c = table[key]
        sim.set_adjoint_tiling(c["rows_y"], c["waves_y"], c["rows_x"], c["waves_x"])
        return c
    _, U1 = sim.device_ptrs()
    nsteps = 16
    sim.rollout(nsteps, dt, 2)

    def run(c):
        sim.set_adjoint_tiling(c["rows_y"], c["waves_y"], c["rows_x"], c["waves_x"])
        sim.rollout_adjoint(U1)
        sim.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            sim.rollout_adjoint(U1)
            sim.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best / nsteps

    cfg = {"rows_y": 12, "waves_y": 4, "rows_x": 16, "waves_x": 4}
    best = run(cfg)
    ntrials = 0
    for _ in range(2):
        moved = False
        for field, values in (("rows_y", ADJ_ROWS), ("rows_x", ADJ_ROWS), ("waves_y", WAVES), ("waves_x", WAVES)):
            for v in values:
                if v == cfg[field]:
                    continue
                trial = dict(cfg)
                trial[field] = v
                t = run(trial)
                ntrials += 1
                if t < 0.99 * best:  # a move must beat timing noise
                    best, cfg, moved = t, trial, True
        if not moved:
            break
    sim.set_adjoint_tiling(cfg["rows_y"], cfg["waves_y"], cfg["rows_x"], cfg["waves_x"])
    cfg["ms_per_step"] = round(best * 1e3, 4)
    if log:
        log(f"autotune {key}: {cfg} ({ntrials} trials)")
    table[key] = cfg
    _store(path, table)
    return cfg


def tune(sim, sizes, arch: str = "gfx950", steps: int = 8, cache: str | None = None, log=None,
         kernels: str = "step"):
    """Tune `sim` (already started: start() called) in place and return the
    configuration; `sizes` = [(nx, ny)] per problem. kernels = "step": the
    kernels the handle's device-dt step runs (the one-launch step kernel
    where it applies; fixed-dt rollouts and their recomputes run the same
    kernel); "stages": the two stage kernels, timed through the two-launch
    step."""
    classes = size_classes(sizes)
    if kernels == "stages":
        sim.set_step_kernel("two_stage")
        try:
            return tune(sim, sizes, arch, steps, cache, log, kernels="step")
        finally:
            sim.set_step_kernel("auto")
    one = sim.one_launch
    key = ("step|" if one else "") + signature(sizes, classes, arch)
    path = _cache_path(cache)
    table = _load(path)
    if key in table:
        apply(sim, classes, table[key])
        return table[key]
    ncls = max(classes) + 1
    if one:
        kernels = ("3",)
        cfg = {"3": {"rows": [48] * ncls, "waves": 4}}
    else:
        kernels = ("1", "2")
        cfg = {"1": {"rows": [32] * ncls, "waves": 4}, "2": {"rows": [48] * ncls, "waves": 4}}
    apply(sim, classes, cfg)
    best = sum(_time(sim, steps))
    trials = []
    for st in kernels:
        for k in range(ncls):
            if k not in classes:
                continue
            for r in ROWS:
                if r == cfg[st]["rows"][k]:
                    continue
                trial = json.loads(json.dumps(cfg))
                trial[st]["rows"][k] = r
                apply(sim, classes, trial)
                t = sum(_time(sim, steps))
                trials.append((st, k, r, t))
                if t < best:
                    best, cfg = t, trial
        for w in WAVES:
            if w == cfg[st]["waves"]:
                continue
            trial = json.loads(json.dumps(cfg))
            trial[st]["waves"] = w
            apply(sim, classes, trial)
            t = sum(_time(sim, steps))
            trials.append((st, "waves", w, t))
            if t < best:
                best, cfg = t, trial
    # persistent grid (capped at the resident waves, looping over the work
    # list) against one wave per work item
    trial = json.loads(json.dumps(cfg))
    trial["persistent"] = 1
    apply(sim, classes, trial)
    t = sum(_time(sim, steps))
    trials.append(("persistent", 1, t))
    if t < best:
        best, cfg = t, trial
    apply(sim, classes, cfg)
    cfg["ms_per_step"] = round(best, 5)
    if log:
        log(f"autotune {key}: {cfg} ({len(trials)} trials)")
    table[key] = cfg
    _store(path, table)
    return cfg
