"""Write the golden vectors under tests/golden/ from the reference's own code.

Run in a container where the reference tree is mounted and
`make -C oracle ref` has built oracle/_ref/ (the reference headers driven by
oracle/ref_harness.cpp, and the reference's 1D OpenMP benchmark program built
from its source). Inputs are seeded; outputs are whatever the reference
computes. Files: <name>.npz (float64 arrays, no pickles) and manifest.json
(shapes, parameters, sha256 of every file).

    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import REF_PATH, MeshSpec, RefHarness  # noqa: E402

GAMMA = 1.4


def rand_prims(rng, n, rho=(0.1, 2.0), u=(-2.0, 2.0), p=(0.1, 2.0)):
    return np.stack([rng.uniform(*rho, n), rng.uniform(*u, n), rng.uniform(*p, n)], 1)


def rand_cons2d(rng, shape, gentle=False):
    """Random positive 2D conserved states (rho, rho u, rho v, E)."""
    lo, hi = (0.8, 1.2) if gentle else (0.2, 1.5)
    rho = rng.uniform(lo, hi, shape)
    amp = 0.05 if gentle else 0.5
    u, v = rng.uniform(-amp, amp, shape), rng.uniform(-amp, amp, shape)
    p = rng.uniform(lo, hi, shape)
    E = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def main():
    ref = RefHarness(REF_PATH)
    rng = np.random.default_rng(20240601)
    files = {}

    # ---- scalar numerics: HLLE, PLM, limiters, EOS, passive flux ----
    n = 512
    WL, WR = rand_prims(rng, n), rand_prims(rng, n)
    WL[:64, 1] = rng.uniform(3.0, 12.0, 64)   # supersonic right-going: S_L >= 0 branch
    WR[:64, 1] = WL[:64, 1] + rng.uniform(-0.5, 0.5, 64)
    WL[64:128, 1] = rng.uniform(-12.0, -3.0, 64)  # supersonic left-going: S_R <= 0 branch
    WR[64:128, 1] = WL[64:128, 1] + rng.uniform(-0.5, 0.5, 64)
    F = ref.hlle(WL, WR, GAMMA)
    Wm, W0, Wp = rand_prims(rng, n), rand_prims(rng, n), rand_prims(rng, n)
    W0[:32] = Wm[:32]  # zero left differences
    plm = {}
    for lim in ("minmod", "mc", "none"):
        a, b = ref.plm(Wm, W0, Wp, lim)
        plm[f"plm_{lim}_WL"], plm[f"plm_{lim}_WR"] = a, b
    # limiter special values: zeros, opposite signs, equal magnitudes, underflowing products
    la = np.array([1.0, 2.0, -1.0, 1.0, -1.0, 0.0, 1.0, 3.0, -3.0, 1e-200, -1e-200, 2.5, -0.0, 1e-300])
    lb = np.array([2.0, 1.0, -2.0, -1.0, 1.0, 1.0, 0.0, 3.0, -3.0, 1e-200, -1e-200, -2.5, 4.0, 1e-300])
    lim_out = {name: np.array([getattr(ref, name)(a, b) for a, b in zip(la, lb)]) for name in ("minmod", "mc")}
    lim_none = np.array([ref.limiter(a, b, 2) for a, b in zip(la, lb)])
    eos_in = np.stack([rng.uniform(0.1, 2, 64), rng.uniform(-1, 1, 64), rng.uniform(0.5, 3, 64)], 1)
    eos_p = np.array([ref.eos_pressure(r, m, e, GAMMA) for r, m, e in eos_in])
    eos_E = np.array([ref.eos_total_energy(r, u, p, GAMMA) for r, u, p in eos_in])
    eos_c = np.array([ref.eos_sound_speed(r, abs(p), GAMMA) for r, _, p in eos_in])
    uic = np.stack([rng.uniform(-3, 3, 256), rng.uniform(-3, 3, 256), rng.uniform(0.5, 2, 256),
                    rng.uniform(0.5, 2, 256)], 1)
    uint = np.array([ref.interface_velocity(*row) for row in uic])
    qv = rng.uniform(-1, 1, (256, 2))
    pflux = np.array([ref.passive_flux(q[0], q[1], u) for q, u in zip(qv, uint)])
    files["numerics"] = dict(hlle_WL=WL, hlle_WR=WR, hlle_F=F, plm_Wm=Wm, plm_W0=W0, plm_Wp=Wp, **plm,
                             lim_a=la, lim_b=lb, lim_minmod=lim_out["minmod"], lim_mc=lim_out["mc"],
                             lim_none=lim_none, eos_in=eos_in, eos_p=eos_p, eos_E=eos_E, eos_c=eos_c,
                             uint_in=uic, uint=uint, pflux_q=qv, pflux=pflux)

    # ---- 2D sweeps (x_sweep_t / y_sweep_t) on one random patch ----
    nx, ny, h = 13, 11, 2
    sw = {}
    for lim, gentle in (("minmod", False), ("mc", False), ("none", True)):
        U = rand_cons2d(rng, (nx + 2 * h, ny + 2 * h), gentle).ravel().copy()
        sw[f"{lim}_in"] = U
        sw[f"{lim}_x"] = ref.sweep_2d(U, nx, ny, 7e-4, 1.0 / 64, True, lim=lim)
        sw[f"{lim}_y"] = ref.sweep_2d(U, nx, ny, 7e-4, 1.0 / 48, False, lim=lim)
    files["sweeps"] = sw

    # ---- 1D euler_rhs_t on an outflow-filled single patch ----
    n1 = 40
    r1 = {}
    for lim in ("minmod", "mc", "none"):
        rho = rng.uniform(0.5, 1.5, n1 + 4)
        u = rng.uniform(-0.3, 0.3, n1 + 4)
        p = rng.uniform(0.5, 1.5, n1 + 4)
        U = np.concatenate([rho, rho * u, p / (GAMMA - 1.0) + 0.5 * rho * u * u])
        Uv = U.reshape(3, -1)
        Uv[:, 0] = Uv[:, 1] = Uv[:, 2]
        Uv[:, -1] = Uv[:, -2] = Uv[:, -3]
        r1[f"{lim}_in"] = U.copy()
        r1[f"{lim}_out"] = ref.rhs_1d(U, n1, 5e-4, 1.0 / n1, lim=lim)
    files["rhs1d"] = r1

    # ---- reference periodic 2D mesh: exchange, Lie step, compute_dt_2d ----
    m = MeshSpec(16, 12, 2, 2, bcx="periodic", bcy="periodic")
    G = rand_cons2d(rng, (16, 12))
    U = m.to_patches(G)
    ex = U.copy()
    dt_ref = ref.periodic_mesh(ex, 16, 12, 2, 2, 0, cfl=0.4, dx=1 / 16, dy=1 / 12)
    lie = U.copy()
    ref.periodic_mesh(lie, 16, 12, 2, 2, 1, dt=2e-3, dx=1 / 16, dy=1 / 12)
    files["mesh_periodic"] = dict(U=U, exchanged=ex, lie_step=lie, dt=np.array([dt_ref]))

    # ---- step counts of the reference 1D OpenMP benchmark (periodic Sod, CFL 0.4) ----
    exe = os.path.join(ROOT, "oracle", "_ref", "benchmark_euler1d")
    counts = {}
    for nz, tf in ((100, 0.1), (400, 0.2), (1000, 0.2)):
        out = subprocess.run([exe, str(nz), str(tf)], capture_output=True, text=True, check=True).stdout
        counts[f"{nz}_{tf}"] = int(re.search(r"Steps:\s+(\d+)", out).group(1))

    manifest = {"generator": "tests/golden/gen_golden.py", "gamma": GAMMA,
                "source": "reference headers examples/sod1d/euler1d.hpp, examples/sod2d/euler2d.hpp, "
                          "include/nest/patch.hpp via oracle/ref_harness.cpp; "
                          "examples/benchmark/benchmark_euler1d.cpp built by oracle/Makefile",
                "params": {"sweeps": {"nx": nx, "ny": ny, "h": h, "dt": 7e-4, "dx": 1 / 64, "dy": 1 / 48},
                           "rhs1d": {"n": n1, "h": 2, "dt": 5e-4, "dx": 1 / n1},
                           "mesh_periodic": {"nx": 16, "ny": 12, "npx": 2, "npy": 2, "dt_lie": 2e-3,
                                             "cfl": 0.4}},
                "benchmark_steps_periodic_sod_cfl0.4": counts, "files": {}}
    for name, arrs in files.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez(path, **arrs)
        manifest["files"][f"{name}.npz"] = {
            "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(),
            "arrays": {k: list(np.asarray(v).shape) for k, v in arrs.items()}}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", sorted(manifest["files"]), counts)


if __name__ == "__main__":
    main()
