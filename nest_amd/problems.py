"""Initial states of the benchmark problems (synthetic inputs, host side).

Random numbers come from a counter-based SplitMix64 keyed by (seed, global
cell index), so an initial state does not depend on how the domain is split
into patches. Energy is formed as the reference drivers form it:
1D eos_t::total_energy (examples/sod1d/euler1d.hpp:47-51) in
examples/benchmark/benchmark_euler1d.cpp:53-77; 2D
E = p/(gamma-1) + 0.5*rho*(u*u + v*v) (examples/sod2d/sod2d.cpp:283).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """SplitMix64 output for counters idx under key seed (uint64 arrays)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, idx: np.ndarray) -> np.ndarray:
    """U[0,1) doubles, 53-bit."""
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def energy2d(rho, u, v, p, gamma=1.4):
    return p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)


def sod1d(n: int, gamma=1.4, x_disc=0.5, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), h=2):
    """Extended 1D Sod state [3*(n+4)], ghosts set from the same formula
    (benchmark_euler1d.cpp:53-77)."""
    dx = 1.0 / n
    i = np.arange(n + 2 * h)
    x = 0.0 + (i - h + 0.5) * dx
    L = x < x_disc
    rho = np.where(L, left[0], right[0])
    u = np.where(L, left[1], right[1])
    p = np.where(L, left[2], right[2])
    ke = 0.5 * rho * u * u
    E = p / (gamma - 1.0) + ke
    return np.concatenate([rho, rho * u, E]).astype(np.float64)


def sod2d(nx: int, ny: int, jitter=1e-3, seed=2, gamma=1.4, x_disc=0.5):
    """Global interior G[4, nx, ny]: planar Sod split along x = 0.5 (dim 0),
    density += jitter*(2*xi - 1) with xi ~ U[0,1) per cell."""
    dx = 1.0 / nx
    x = (np.arange(nx) + 0.5) * dx
    L = (x < x_disc)[:, None] * np.ones((1, ny), dtype=bool)
    rho = np.where(L, 1.0, 0.125)
    p = np.where(L, 1.0, 0.1)
    if jitter:
        idx = np.arange(nx * ny, dtype=np.uint64).reshape(nx, ny)
        rho = rho + jitter * (2.0 * uniform01(seed, idx) - 1.0)
    u = np.zeros_like(rho)
    v = np.zeros_like(rho)
    return np.stack([rho, rho * u, rho * v, energy2d(rho, u, v, p, gamma)])


def ensemble_sizes(nprob=512, seed=3, lo=64, hi=512):
    """Side lengths log-uniform in [lo, hi]: s = round(lo * (hi/lo)**xi)."""
    xi = uniform01(seed, np.arange(nprob, dtype=np.uint64))
    return [int(round(lo * (hi / lo) ** x)) for x in xi]


def quadrants(s: int, k: int, seed=3, gamma=1.4):
    """Four-quadrant Riemann problem number k on s x s cells: per quadrant
    rho, p ~ U[0.1, 1] and u, v ~ U[-0.5, 0.5], split at x = y = 0.5."""
    r = uniform01(seed + 1, np.arange(16 * k, 16 * k + 16, dtype=np.uint64)).reshape(4, 4)
    x = (np.arange(s) + 0.5) / s
    qx = (x >= 0.5).astype(int)
    q = qx[:, None] * 2 + qx[None, :]
    rho = 0.1 + 0.9 * r[q, 0]
    p = 0.1 + 0.9 * r[q, 1]
    u = -0.5 + r[q, 2]
    v = -0.5 + r[q, 3]
    return np.stack([rho, rho * u, rho * v, energy2d(rho, u, v, p, gamma)])


def smooth_periodic(n: int, seed=4, nbumps=8, amp=1e-2, width=0.05, gamma=1.4):
    """Adjoint test state: rho = 1 + 0.2 sin(2 pi x) sin(2 pi y) + Gaussian
    bumps (minimum-image periodic distance), u = v = 0.1, p = 1."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    c = uniform01(seed, np.arange(2 * nbumps, dtype=np.uint64)).reshape(nbumps, 2)
    for cx, cy in c:
        dxm = np.abs(X - cx)
        dxm = np.minimum(dxm, 1.0 - dxm)
        dym = np.abs(Y - cy)
        dym = np.minimum(dym, 1.0 - dym)
        rho = rho + amp * np.exp(-(dxm * dxm + dym * dym) / (2.0 * width * width))
    u = np.full_like(rho, 0.1)
    v = np.full_like(rho, 0.1)
    p = np.ones_like(rho)
    return np.stack([rho, rho * u, rho * v, energy2d(rho, u, v, p, gamma)])


def kelvin_helmholtz(n: int, seed=5, gamma=1.4):
    """KH shear layer: rho = 2, u = -0.5 for |y - 0.5| < 0.25, else rho = 1,
    u = +0.5; p = 2.5; v = 0.01 sin(4 pi x) + 1e-4 (2 xi - 1)."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, -0.5, 0.5)
    idx = np.arange(n * n, dtype=np.uint64).reshape(n, n)
    v = 0.01 * np.sin(4 * np.pi * X) + 1e-4 * (2.0 * uniform01(seed, idx) - 1.0)
    p = np.full_like(rho, 2.5)
    return np.stack([rho, rho * u, rho * v, energy2d(rho, u, v, p, gamma)])


def ensemble_batch(nprob=512, seed=3, mode="fast", limiter="minmod", Sim2D=None):
    """Ensemble problem set: `nprob` independent four-quadrant Riemann problems
    on s x s cells (s log-uniform in [64, 512]), outflow, all in one batch
    handle. Returns (sim, sizes, U) with U the uploaded halo-filled state."""
    if Sim2D is None:
        from .euler import Sim2D
    sides = ensemble_sizes(nprob, seed)
    sizes = [(s, s) for s in sides]
    sim = Sim2D.batch(sizes, mode=mode, limiter=limiter, bc="outflow")
    U = np.zeros(sim.total)
    for k, s in enumerate(sides):
        sim.patch_view(U, k)[:, 2:-2, 2:-2] = quadrants(s, k, seed=seed)
    sim.upload(U)
    sim.fill_halos()
    return sim, sizes, sim.download()


def adjoint_setup(n=1024, seed=4, cfl=0.4, perturb=1e-2, gamma=1.4):
    """Smooth-state adjoint setup on n x n periodic cells: the initial state
    (smooth_periodic), the initial state of the target run (density times
    1 + perturb*(2 xi - 1), xi ~ U[0,1) per cell from `seed` + 1, energy
    adjusted to keep p = 1), and the fixed dt = cfl*dx / max(|u| + c) of the
    initial state (the larger of the x and y speeds; dx = dy = 1/n)."""
    U0 = smooth_periodic(n, seed=seed, gamma=gamma)
    rho0 = U0[0]
    idx = np.arange(n * n, dtype=np.uint64).reshape(n, n)
    rho_t = rho0 * (1.0 + perturb * (2.0 * uniform01(seed + 1, idx) - 1.0))
    u = np.full_like(rho0, 0.1)
    Ut = np.stack([rho_t, rho_t * u, rho_t * u, energy2d(rho_t, u, u, np.ones_like(rho0), gamma)])
    c = np.sqrt(gamma * 1.0 / np.minimum(rho0, rho_t))
    dt = cfl * (1.0 / n) / float(np.max(0.1 + c))
    return U0, Ut, dt


def density_direction(n, seed, kmax=4, u=0.1, v=0.1):
    """A direction in conserved-state space that changes the density only
    (velocity and pressure held): drho = random smooth periodic field (the
    Fourier modes |kx|, |ky| <= kmax with seeded amplitudes and phases,
    scaled to max |drho| = 1); returns (drho, u drho, v drho,
    0.5 (u^2 + v^2) drho)."""
    r = uniform01(seed, np.arange(4 * (2 * kmax + 1) ** 2, dtype=np.uint64)).reshape(-1, 4)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    d = np.zeros((n, n))
    i = 0
    for kx in range(-kmax, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            a, ph = 2.0 * r[i, 0] - 1.0, 2.0 * np.pi * r[i, 1]
            d += a * np.cos(2.0 * np.pi * (kx * X + ky * Y) + ph)
            i += 1
    d /= np.max(np.abs(d))
    return np.stack([d, u * d, v * d, 0.5 * (u * u + v * v) * d])


def smooth_field(n, seed, kmax=4):
    """Seeded smooth periodic field on n x n cells (Fourier modes |kx|, |ky|
    <= kmax, random amplitudes and phases), scaled to max |f| = 1."""
    return density_direction(n, seed, kmax=kmax, u=0.0, v=0.0)[0]


def smooth_direction(n, seed, kmax=4):
    """Direction in conserved-state space: an independent seeded smooth
    field per conserved variable (seeds 4*seed .. 4*seed + 3)."""
    return np.stack([smooth_field(n, 4 * seed + k, kmax) for k in range(4)])


def adjoint_generic_state(n=1024, seed=4, amp=1e-3, gamma=1.4):
    """The adjoint setup's state made generic: rho, u, v, p of adjoint_setup's
    initial state each plus amp times a seeded smooth field. (The setup's
    state itself has exactly uniform u, v, p, so every velocity / pressure
    limiter slope sits on its kink and the loss has no two-sided derivative
    there along directions that move u, v or p.)"""
    U0, _, _ = adjoint_setup(n, seed=seed, gamma=gamma)
    r = U0[0]
    u = U0[1] / r
    v = U0[2] / r
    p = (gamma - 1.0) * (U0[3] - 0.5 * r * (u * u + v * v))
    base = 7919 + 10 * seed
    r = r + amp * smooth_field(n, base)
    u = u + amp * smooth_field(n, base + 1)
    v = v + amp * smooth_field(n, base + 2)
    p = p + amp * smooth_field(n, base + 3)
    return np.stack([r, r * u, r * v, energy2d(r, u, v, p, gamma)])
