"""Mesh colliders: the OBJ and SDF text formats, the oracle's manifold2SDF
(FluidSim/src/cpu/util.cc:203-317) against closed forms and brute force, and (GPU) the device's
manifold2SDF bit for bit against the oracle, then substeps with a mesh collider."""
from __future__ import annotations

import numpy as np
import pytest

from simcraft_amd import meshio, scenes
from tests.conftest import announce


def _box_mesh(lo, hi):
    """Axis-aligned box, 12 outward-facing triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0], [x0, y0, z1], [x1, y0, z1], [x1, y1, z1],
                  [x0, y1, z1]], dtype=np.float64)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5),
         (0, 4, 7), (0, 7, 3)]
    return meshio.Mesh(v, np.array(f, dtype=np.uint32))


def _lumpy_sphere(seed=5):
    """A non-convex closed mesh: an icosphere with radially perturbed vertices."""
    m = meshio.icosphere((0.52, 0.47, 0.5), 0.3, 3)
    d = m.vertices - np.array([0.52, 0.47, 0.5])
    r = np.linalg.norm(d, axis=1, keepdims=True)
    bump = 1.0 + 0.25 * np.sin(7 * d[:, :1] / r) * np.cos(5 * d[:, 1:2] / r)
    return meshio.Mesh(np.array([0.52, 0.47, 0.5]) + d * bump, m.indices)


def _point_triangle_bruteforce(P, A, B, C, exact=True):
    """Distance from points P (N,3) to triangle ABC, from the closest point q found
    independently: the plane projection where it falls inside (barycentric test), else the
    nearest point of the three edge segments. exact=False applies the reference's measure
    instead (util.cc:101-125): |p + q - 2a| where q lies inside, or on edge AB or AC away from
    the vertices; |p - q| where q is a vertex or on edge BC."""
    n = np.cross(B - A, C - A)
    nn = n / np.linalg.norm(n)
    dist_plane = (P - A) @ nn
    Qp = P - dist_plane[:, None] * nn
    v0, v1, v2 = B - A, C - A, Qp - A
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    bv = (d11 * d20 - d01 * d21) / den
    bw = (d00 * d21 - d01 * d20) / den
    inside = (bv >= 0) & (bw >= 0) & (bv + bw <= 1)

    def seg(a, b):
        ab = b - a
        t = np.clip(((P - a) @ ab) / (ab @ ab), 0.0, 1.0)
        return a + t[:, None] * ab, t

    cands = [seg(A, B), seg(B, C), seg(C, A)]
    dists = np.stack([np.linalg.norm(P - q, axis=1) for q, _ in cands])
    e = np.argmin(dists, axis=0)
    Q = np.where(inside[:, None], Qp, np.choose(e[:, None], [q for q, _ in cands]))
    if exact:
        return np.linalg.norm(P - Q, axis=1)
    t = np.choose(e, [t for _, t in cands])
    on_ab_or_ca = ((e == 0) | (e == 2)) & (t > 0) & (t < 1)
    quirky = inside | on_ab_or_ca
    return np.where(quirky, np.linalg.norm(P + Q - 2 * A, axis=1), np.linalg.norm(P - Q, axis=1))


def _oracle_mesh_sdf(oracle, n, mesh, band=3, flags=0):
    o = oracle.OracleSim(oracle.default_config(nx=n, ny=n, nz=n))
    o.set_collider_mesh(mesh.vertices, mesh.indices, band, flags)
    s = o.get("collider_sdf")
    o.close()
    return s


def test_obj_round_trip_and_face_forms(tmp_path):
    m = meshio.icosphere(subdivisions=1)
    path = tmp_path / "sphere.obj"
    meshio.save_obj(path, m)
    r = meshio.load_obj(path)
    assert np.array_equal(r.vertices, m.vertices) and np.array_equal(r.indices, m.indices)
    assert np.allclose(np.linalg.norm(r.normals, axis=1), 1.0)
    plain = tmp_path / "plain.obj"
    plain.write_text("# comment\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 3 2\nf 1 2 4\nf 1 4 3\nf 2 3 4\n")
    p = meshio.load_obj(plain)
    assert p.triangle_count == 4 and p.indices[0].tolist() == [0, 2, 1]
    assert not p.normals.any()  # no "//" normals named: zero, as myLoadObj leaves them


def test_sdf_text_round_trip(tmp_path):
    s = np.random.default_rng(1).standard_normal((5, 4, 3))
    path = tmp_path / "a.sdf"
    meshio.save_sdf(path, s, spacing=(0.25, 0.25, 0.25))
    got, spacing, origin = meshio.load_sdf(path)
    assert np.array_equal(got, s) and np.array_equal(spacing, [0.25] * 3) and not origin.any()


@pytest.mark.parametrize("exact", [False, True], ids=["reference", "exact"])
def test_oracle_point_triangle_distance_single_triangle(oracle, exact):
    """A single triangle: the sweeps hand it to every cell, so every cell holds the (reference's
    or true) distance from its centre to that triangle, which pins every case of
    distancePointTriangle against the independent closest-point computation."""
    n = 16
    rng = np.random.default_rng(4)
    tri = rng.uniform(0.3, 0.7, (3, 3))
    mesh = meshio.Mesh(tri, np.array([[0, 1, 2]], dtype=np.uint32))
    s = _oracle_mesh_sdf(oracle, n, mesh, flags=1 if exact else 0)
    X, Y, Z = scenes.cell_centres(n)
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    want = _point_triangle_bruteforce(P, *tri, exact=exact).reshape(s.shape)
    np.testing.assert_allclose(np.abs(s), want, rtol=0, atol=1e-12)
    if not exact:  # the reference's measure differs from the distance for most cells
        true = _point_triangle_bruteforce(P, *tri).reshape(s.shape)
        assert np.mean(np.abs(np.abs(s) - true) > 1e-6) > 0.3


def test_oracle_mesh_sdf_sphere(oracle):
    """A finely triangulated sphere. With true distances: every value is the distance to some
    triangle of the mesh (the closest one near the surface, where it equals the brute-force
    minimum over all triangles; a neighbour's closest one elsewhere, so never below that
    minimum), and near the analytic |x - c| - r. In both modes: negative exactly inside."""
    n, c, r = 24, np.array([0.5, 0.5, 0.5]), 0.3
    m = meshio.icosphere(c, r, 3)
    X, Y, Z = scenes.cell_centres(n)
    ana = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r
    edge = max(np.linalg.norm(m.vertices[a] - m.vertices[b]) for a, b, _ in m.indices)
    sagitta = edge * edge / (8 * r)
    far = np.abs(ana) > sagitta
    sref = _oracle_mesh_sdf(oracle, n, m)
    assert np.array_equal(sref[far] < 0, ana[far] < 0)
    s = _oracle_mesh_sdf(oracle, n, m, flags=1)
    assert np.array_equal(s[far] < 0, ana[far] < 0)
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    best = np.full(len(P), np.inf)
    for t in m.indices:
        best = np.minimum(best, _point_triangle_bruteforce(P, *m.vertices[t]))
    best = best.reshape(s.shape)
    assert (np.abs(s) >= best - 1e-12).all()
    near = np.abs(ana) < 2.0 / n
    np.testing.assert_allclose(np.abs(s[near]), best[near], rtol=0, atol=1e-12)
    assert np.abs(np.abs(s) - best).max() <= 0.5 / n  # the sweeps' closest-triangle estimate
    assert np.abs(s - ana).max() <= 0.5 / n + sagitta


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["sphere", "lumpy", "box", "triangle"])
def test_device_mesh_sdf_matches_oracle(hip, oracle, shape):
    n = 32
    announce(f"mesh sdf {shape} n={n}")
    mesh = {"sphere": lambda: meshio.icosphere((0.7, 0.3, 0.5), 0.15, 3), "lumpy": _lumpy_sphere,
            "box": lambda: _box_mesh((0.25, 0.1, 0.3), (0.75, 0.5, 0.7)),
            "triangle": lambda: meshio.Mesh(np.random.default_rng(4).uniform(0.3, 0.7, (3, 3)),
                                            np.array([[0, 1, 2]], dtype=np.uint32))}[shape]()
    for flags in (0, hip.MESH_EXACT_DISTANCE):
        want = _oracle_mesh_sdf(oracle, n, mesh, flags=flags)
        for prec in (hip.PRECISION_FP64, hip.PRECISION_MIXED):
            sim = hip.FluidSim(hip.default_config(nx=n, ny=n, nz=n, precision=prec))
            sim.set_collider_mesh(mesh.vertices, mesh.indices, flags=flags)
            got = sim.get("collider_sdf")
            sim.close()
            ref = want if prec == hip.PRECISION_FP64 else want.astype(np.float32).astype(np.float64)
            assert np.array_equal(got, ref), (flags, prec, np.abs(got - ref).max())


@pytest.mark.gpu
def test_substeps_with_mesh_collider_bitexact(hip, oracle):
    """The dam break against the lumpy closed mesh as collider: three substeps bit-exact."""
    n, dt = 32, 0.004
    announce(f"substeps with a mesh collider n={n}")
    mesh = _lumpy_sphere()
    X, Y, Z = scenes.cell_centres(n)
    pos = scenes.particles_in_cells((X > 0.55) & (Y < 0.5), 9, 1.0 / n)
    vel = np.random.default_rng(9).uniform(-0.5, 0.5, pos.shape)
    cfg = hip.default_config(nx=n, ny=n, nz=n, flip_alpha=0.95)
    g = hip.FluidSim(cfg)
    o = oracle.OracleSim(cfg, dot_mode=oracle.DOT_TREE)
    for s in (g, o):
        s.set_collider_mesh(mesh.vertices, mesh.indices)
        s.set_particles(pos, vel)
    for step in range(3):
        sg, so = g.substep(dt), o.substep(dt)
        assert sg["cg_iterations"] == so["cg_iterations"], (step, sg, so)
    gp, gv = g.particles()
    op, ov = o.particles()
    assert np.array_equal(gp, op) and np.array_equal(gv, ov)
    assert np.array_equal(g.get("pg"), o.get("pg"))
