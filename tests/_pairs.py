"""Helpers that drive the device library and the oracle through identical calls."""
from __future__ import annotations

import hashlib

import numpy as np

from simcraft_amd import scenes

FACE = ["u", "v", "w"]
GRID_ARRAYS = ["u", "v", "w", "ubuf", "vbuf", "wbuf", "u_valid", "v_valid", "w_valid", "u_valid_buf", "v_valid_buf",
               "w_valid_buf", "uw", "vw", "ww", "u_last", "v_last", "w_last"]
CELL_ARRAYS = ["fluid_sdf", "fluid_sdf_buf", "sdf_valid", "sdf_valid_buf", "collider_sdf", "pg"]
SYSTEM_ARRAYS = ["u_weights", "v_weights", "w_weights", "adiag", "a_left", "a_right", "a_up", "a_down", "a_front",
                 "a_back", "rhs", "active"]
ALL_ARRAYS = GRID_ARRAYS + CELL_ARRAYS + SYSTEM_ARRAYS


def overlap_scene(n: int, seed: int = 7):
    """Fluid block overlapping the sphere collider, random initial velocities: exercises the
    collision push-out, the solid face fractions and applyCollider."""
    X, Y, Z = scenes.cell_centres(n)
    mask = (X > 0.5) & (X < 0.9) & (Y < 0.5) & (Z > 0.3) & (Z < 0.7)
    pos = scenes.particles_in_cells(mask, seed, 1.0 / n)
    rng = np.random.default_rng(seed)
    vel = rng.uniform(-0.5, 0.5, size=pos.shape)
    return pos, vel, scenes.sphere_sdf(n)


def dam_scene(n: int, seed: int = 2024):
    pos, vel = scenes.dam_break(n, seed)
    return pos, vel, scenes.no_collider(n)


def dense_scene(n: int, seed: int = 11, copies: int = 10):
    """Dam break with `copies` times the usual 8 particles per cell: a bin row holds more
    particle records than the P2G gather stages at once, so it is summed in several chunks."""
    mask = scenes.dam_break_mask(n)
    pos = np.concatenate([scenes.particles_in_cells(mask, seed + k, 1.0 / n) for k in range(copies)])
    vel = np.random.default_rng(seed).uniform(-0.5, 0.5, size=pos.shape)
    return pos, vel, scenes.no_collider(n)


def setup(sim, pos, vel, col):
    sim.set_collider_sdf(col)
    sim.set_particles(pos, vel)


def diff(a: np.ndarray, b: np.ndarray) -> str:
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    if a.dtype == np.uint8:
        bad = np.count_nonzero(a != b)
        return f"{bad} flags differ"
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    nan_mis = np.count_nonzero(np.isnan(a) != np.isnan(b))
    bad = np.count_nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    idx = np.unravel_index(np.nanargmax(d), d.shape) if d.size and np.any(~np.isnan(d)) else None
    return f"{bad} entries differ (nan mismatch {nan_mis}), max |d| {np.nanmax(d) if d.size else 0:g} at {idx}"


def assert_same(gpu, ora, names, where=""):
    bad = []
    for nm in names:
        a, b = gpu.get(nm), ora.get(nm)
        eq = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype != np.uint8 else (a == b)
        if not np.all(eq):
            bad.append(f"{nm}: {diff(a, b)}")
    ga, gv = gpu.particles()
    oa, ov = ora.particles()
    if not np.array_equal(ga, oa):
        bad.append(f"positions: {diff(ga, oa)}")
    if not np.array_equal(gv, ov):
        bad.append(f"velocities: {diff(gv, ov)}")
    assert not bad, f"{where}: " + "; ".join(bad)


def copy_state(src, dst, names=ALL_ARRAYS):
    for nm in names:
        dst.set(nm, src.get(nm))
    p, v = src.particles()
    dst.set_particles(p, v)


STAGES = [
    ("clear", lambda s, dt, h: s.clear()),
    ("advect", lambda s, dt, h: s.advect(dt)),
    ("reconstruct", lambda s, dt, h: s.reconstruct(0.8 * h)),
    ("extrapolate_fluid_sdf", lambda s, dt, h: s.extrapolate_fluid_sdf(10)),
    ("smooth_fluid_surface", lambda s, dt, h: s.smooth_fluid_surface(5)),
    ("p2g", lambda s, dt, h: s.p2g(dt)),
    ("apply_dirichlet_boundary", lambda s, dt, h: s.apply_dirichlet_boundary()),
    ("extrapolate_velocity", lambda s, dt, h: s.extrapolate_velocity(10)),
    ("apply_force", lambda s, dt, h: s.apply_force(dt)),
    ("build_system", lambda s, dt, h: s.build_system(dt)),
    ("solve_pressure", lambda s, dt, h: s.solve_pressure()),
    ("project", lambda s, dt, h: s.project(dt)),
    ("apply_collider", lambda s, dt, h: s.apply_collider()),
    ("g2p", lambda s, dt, h: s.g2p(dt)),
]


STRUCTURE_ARRAYS = ["active", "sdf_valid", "u_valid", "v_valid", "w_valid"]


def structure_hash(sim) -> str:
    """Digest of the discrete structure of the last substep. The substep is smooth in its inputs
    only while these stay the same; where one changes, the result jumps:
      * active cells and the valid flags of the fluid SDF and the face grids (after extrapolation);
      * the faces that P2G reached (weight sum > 0): a face a particle's stencil starts to reach
        takes that particle's velocity instead of an extrapolated one;
      * each particle's cell: the reconstruction skips the last particle of each cell's bin
        (ns-util.h:103), so a particle that changes cell can change which one is skipped."""
    h = hashlib.sha1()
    for nm in STRUCTURE_ARRAYS:
        h.update(np.ascontiguousarray(sim.get(nm)).tobytes())
    for nm in ("uw", "vw", "ww"):
        h.update(np.ascontiguousarray(sim.get(nm) > 0).tobytes())
    pos, _ = sim.particles()
    n = sim.shape[0]
    h.update(np.ascontiguousarray(np.floor(pos * n).astype(np.int32)).tobytes())
    return h.hexdigest()


def central_difference(run, x0, eps_list):
    """(f(x0 + e) - f(x0 - e)) / 2e for the first e of eps_list whose two runs pass through the
    same discrete structure; run(x) returns (f(x), signature). Returns (estimate, e), or
    (None, None) when every step size straddles a change of structure."""
    for eps in eps_list:
        fp, sp = run(x0 + eps)
        fm, sm = run(x0 - eps)
        if sp == sm:
            return (fp - fm) / (2 * eps), eps
    return None, None
