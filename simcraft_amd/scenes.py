"""Seeded synthetic scenes of the FLIP workloads (host-side, numpy).

The reference's setInitialFluid draws unseeded random particles and ignores its mesh
(FluidSim/include/FluidSim/cpu/fluid-simulator.h:73-78); these generators replace it with a
reproducible initialiser: 8 particles per fluid cell, one per 2x2x2 sub-cell, each jittered
uniformly inside its sub-cell with a counter-based generator (splitmix64) keyed by
(seed, cell, sub-cell, axis). A cell is fluid when its centre satisfies the scene predicate.
"""
from __future__ import annotations

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from (seed, counter)."""
    with np.errstate(over="ignore"):
        key = splitmix64(np.uint64(seed) * _GOLD + counter.astype(np.uint64))
    return (key >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centres(n: int, h: float | None = None) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    h = 1.0 / n if h is None else h
    c = (np.arange(n) + 0.5) * h
    return np.meshgrid(c, c, c, indexing="ij")


def particles_in_cells(mask: np.ndarray, seed: int, h: float) -> np.ndarray:
    """8 jittered particles per True cell of mask (x-major cell order, sub-cell order q=4bx+2by+bz)."""
    cells = np.argwhere(mask)  # (C, 3), lexicographic = reference layout order
    ncell = len(cells)
    q = np.arange(8)
    sub = np.stack([(q >> 2) & 1, (q >> 1) & 1, q & 1], axis=1).astype(np.float64)  # (8,3)
    lin = (cells[:, 0].astype(np.uint64) * np.uint64(mask.shape[1]) + cells[:, 1].astype(np.uint64)) \
        * np.uint64(mask.shape[2]) + cells[:, 2].astype(np.uint64)
    ctr = (lin[:, None, None] * np.uint64(24) + (q[None, :, None] * 3 + np.arange(3)[None, None, :]).astype(np.uint64))
    u = uniform01(seed, ctr)  # (C, 8, 3)
    pos = (cells[:, None, :].astype(np.float64) + (sub[None, :, :] + u) * 0.5) * h
    return pos.reshape(ncell * 8, 3)


def dam_break_mask(n: int, x_max: float = 0.4, y_max: float = 0.6) -> np.ndarray:
    """Fluid block x < 0.4, y < 0.6 (all z) of the dam-break workloads, by cell centre."""
    X, Y, _ = cell_centres(n)
    return (X < x_max) & (Y < y_max)


def centred_block_mask(n: int, lo: float = 0.35, hi: float = 0.65) -> np.ndarray:
    X, Y, Z = cell_centres(n)
    return (X >= lo) & (X <= hi) & (Y >= lo) & (Y <= hi) & (Z >= lo) & (Z <= hi)


def sphere_sdf(n: int, centre=(0.7, 0.3, 0.5), radius: float = 0.15) -> np.ndarray:
    """Analytic collider SDF |x - c| - r sampled at cell centres (unit cube)."""
    X, Y, Z = cell_centres(n)
    return np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) - radius


def no_collider(n: int, value: float = 10.0) -> np.ndarray:
    """A collider SDF that is positive everywhere. An all-zero SDF (the reference's default)
    makes every face weight 0 and leaves no active cell."""
    return np.full((n, n, n), value)


def dam_break(n: int, seed: int = 2024, x_max: float = 0.4, y_max: float = 0.6):
    """Dam-break particles at rest: (positions (N,3), velocities (N,3))."""
    mask = dam_break_mask(n, x_max, y_max)
    pos = particles_in_cells(mask, seed, 1.0 / n)
    return pos, np.zeros_like(pos)


WORKLOADS = {
    # name: (n, dt, steps, alpha, precision, collider)
    "dam_break_32_fp64": dict(n=32, dt=0.004, steps=10, flip_alpha=0.95, precision="fp64", collider="none"),
    "dam_break_128_fp64": dict(n=128, dt=0.001, steps=50, flip_alpha=0.95, precision="fp64", collider="none"),
    "dam_break_sphere_256_mixed": dict(n=256, dt=5e-4, steps=20, flip_alpha=0.95, precision="mixed",
                                       collider="sphere"),
    "dam_break_sphere_256_fp64": dict(n=256, dt=5e-4, steps=20, flip_alpha=0.95, precision="fp64",
                                      collider="sphere"),
}


def collider_for(kind: str, n: int) -> np.ndarray:
    return sphere_sdf(n) if kind == "sphere" else no_collider(n)


def gaussian_potential_params(seed: int, count: int = 8, lo: float = 0.35, hi: float = 0.65,
                              amplitude: float = 0.05):
    """Centres (count, 3) uniform in [lo, hi]^3 and vector amplitudes (count, 3) uniform in
    [-amplitude, amplitude], drawn with the splitmix64 generator from `seed`."""
    u = uniform01(seed, np.arange(count * 6, dtype=np.uint64)).reshape(count, 6)
    return lo + (hi - lo) * u[:, :3], amplitude * (2.0 * u[:, 3:] - 1.0)


def curl_gaussian_velocity(n: int, seed: int = 4, sigma: float = 0.1, count: int = 8, lo: float = 0.35,
                           hi: float = 0.65, amplitude: float = 0.05):
    """Divergence-free face velocities (u, v, w) in the reference's face-grid layout: the discrete
    curl of a vector potential psi(x) = sum_m a_m exp(-|x - c_m|^2 / (2 sigma^2)) sampled on the
    cell edges (psi_x at (i+1/2, j, k) h, psi_y at (i, j+1/2, k) h, psi_z at (i, j, k+1/2) h). On
    the MAC grid the discrete divergence of a discrete curl telescopes to zero, so every cell's
    net face flux vanishes up to rounding.

    Shapes: u (n+1, n, n), v (n, n+1, n), w (n, n, n+1)."""
    h = 1.0 / n
    centres, amps = gaussian_potential_params(seed, count, lo, hi, amplitude)
    node = np.arange(n + 1) * h
    mid = (np.arange(n) + 0.5) * h

    def potential(xs, ys, zs, comp):
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        out = np.zeros(X.shape)
        for c, a in zip(centres, amps):
            r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
            out += a[comp] * np.exp(-r2 / (2.0 * sigma * sigma))
        return out

    px = potential(mid, node, node, 0)  # (n, n+1, n+1)
    py = potential(node, mid, node, 1)  # (n+1, n, n+1)
    pz = potential(node, node, mid, 2)  # (n+1, n+1, n)
    u = (pz[:, 1:, :] - pz[:, :-1, :]) / h - (py[:, :, 1:] - py[:, :, :-1]) / h
    v = (px[:, :, 1:] - px[:, :, :-1]) / h - (pz[1:, :, :] - pz[:-1, :, :]) / h
    w = (py[1:, :, :] - py[:-1, :, :]) / h - (px[:, 1:, :] - px[:, :-1, :]) / h
    return u, v, w


def centred_block(n: int, seed: int = 2024, lo: float = 0.35, hi: float = 0.65) -> np.ndarray:
    """Particles of the differentiable-rollout workload: 8 per cell of the centred block."""
    return particles_in_cells(centred_block_mask(n, lo, hi), seed, 1.0 / n)
