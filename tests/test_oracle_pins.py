"""Known-answer tests for the oracle functions that the basic set leaves open (CPU only).

Each test fixes an oracle function by something computed another way: a closed form of the
mathematics, a value worked by hand from the reference's branches, or a numpy brute force of a
small case. The inputs are chosen so that a plausible slip (a dropped term, a wrong sign,
index or spacing, a transposed operand, a copy the reference does not make) changes the answer.
Reference file:line citations name what is pinned.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
from simcraft_amd import scenes

G = 9.8


def _sim(n=6, **kw):
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n, **kw))
    sim.set_collider_sdf(scenes.no_collider(n))
    return sim


# ---- extrapolateFluidSdf (FluidSim/include/FluidSim/cpu/fluid-simulator.h:123-166) -----------
def test_sdf_extrapolation_plane_and_stale_buffer(oracle):
    """A valid plane i = 0 grows one plane per pass. Valid cells are not copied into the buffer
    that is swapped in, so after two passes plane 1 shows its original (never extrapolated)
    value again and plane 2 carries plane 0's values; cells out of reach keep the buffer."""
    n = 6
    sim = _sim(n)
    rng = np.random.default_rng(1)
    cur0 = rng.uniform(-1, 1, (n, n, n))
    valid = np.zeros((n, n, n), np.uint8)
    valid[0] = 1
    stale = 777.0
    sim.set("fluid_sdf", cur0)
    sim.set("fluid_sdf_buf", np.full((n, n, n), stale))
    sim.set("sdf_valid", valid)
    sim.extrapolate_fluid_sdf(1)
    a1, v1 = sim.get("fluid_sdf"), sim.get("sdf_valid")
    assert np.array_equal(a1[1], cur0[0])      # one valid neighbour: its value
    assert np.all(a1[0] == stale)              # the valid plane is not copied
    assert np.all(a1[2:] == stale)             # no valid neighbour: the buffer is left as it was
    assert np.all(v1[:2] == 1) and np.all(v1[2:] == 0)
    assert np.array_equal(sim.get("fluid_sdf_buf"), cur0)
    sim.extrapolate_fluid_sdf(1)
    a2, v2 = sim.get("fluid_sdf"), sim.get("sdf_valid")
    assert np.array_equal(a2[0], cur0[0])
    assert np.array_equal(a2[1], cur0[1])      # the stale odd ring (Q3)
    assert np.array_equal(a2[2], cur0[0])
    assert np.array_equal(a2[3:], cur0[3:])
    assert np.all(v2[:3] == 1) and np.all(v2[3:] == 0)
    sim.close()


def test_sdf_extrapolation_mean_of_valid_neighbours(oracle):
    n = 6
    sim = _sim(n)
    cur = np.zeros((n, n, n))
    a, b = (2, 2, 1), (2, 2, 3)
    cur[a], cur[b] = 0.375, -1.25
    cur[3, 2, 2] = 55.0  # an invalid neighbour: must not enter the mean
    valid = np.zeros((n, n, n), np.uint8)
    valid[a] = valid[b] = 1
    sim.set("fluid_sdf", cur)
    sim.set("fluid_sdf_buf", np.zeros((n, n, n)))
    sim.set("sdf_valid", valid)
    sim.extrapolate_fluid_sdf(1)
    out = sim.get("fluid_sdf")
    assert out[2, 2, 2] == (0.375 + -1.25) / 2
    assert out[1, 2, 1] == 0.375 and out[2, 3, 3] == -1.25
    assert out[2, 2, 0] == 0.375 and out[2, 2, 4] == -1.25
    sim.close()


# ---- smoothFluidSurface (FluidSim/src/cpu/fluid-simulator.cc:39-74): buf = min(avg6, cur) ------
def test_smoothing_closed_forms(oracle):
    """For f = -(y^2 + z^2) (constant along x) the mean of the in-range face neighbours is
    f - 2h^2/3 inside (6 neighbours) and f - 4h^2/5 on an x-wall (5 neighbours): both below f,
    so the minimum takes them. For +(y^2 + z^2) the inner means lie above f and f is kept."""
    n = 8
    h = 1.0 / n
    X, Y, Z = scenes.cell_centres(n)
    for sign in (-1.0, 1.0):
        sim = _sim(n)
        f = sign * (Y ** 2 + Z ** 2)
        sim.set("fluid_sdf", f)
        sim.set("fluid_sdf_buf", np.full((n, n, n), 3.0))
        sim.smooth_fluid_surface(1)
        out = sim.get("fluid_sdf")
        assert np.array_equal(sim.get("fluid_sdf_buf"), f)  # the arrays swapped
        inner = (slice(1, -1),) * 3
        wall = (0, slice(1, -1), slice(1, -1))
        if sign < 0:
            np.testing.assert_allclose(out[inner], f[inner] - 2 * h * h / 3, rtol=0, atol=1e-15)
            np.testing.assert_allclose(out[wall], f[wall] - 4 * h * h / 5, rtol=0, atol=1e-15)
        else:
            assert np.array_equal(out[inner], f[inner])
        sim.close()


# ---- extrapolate<GridType> (cpu/fluid-simulator.h:204-253) --------------------------------------
def test_velocity_extrapolation_zero_fill_and_stale_buffer(oracle):
    """Like the SDF version, but an invalid face without valid neighbours is written 0."""
    n = 6
    sim = _sim(n)
    shape, _ = sim.array_info("u")
    rng = np.random.default_rng(2)
    u0 = rng.uniform(-1, 1, shape)
    valid = np.zeros(shape, np.uint8)
    valid[2] = 1
    sim.set("u", u0)
    sim.set("ubuf", np.full(shape, 777.0))
    sim.set("u_valid", valid)
    sim.extrapolate_velocity(1)
    u1, v1 = sim.get("u"), sim.get("u_valid")
    assert np.array_equal(u1[1], u0[2]) and np.array_equal(u1[3], u0[2])
    assert np.all(u1[2] == 777.0)               # valid faces are not copied
    assert np.all(u1[0] == 0.0) and np.all(u1[4:] == 0.0)
    assert np.array_equal(v1[1:4], np.ones_like(v1[1:4])) and not v1[0].any() and not v1[4:].any()
    assert np.array_equal(sim.get("ubuf"), u0)
    sim.extrapolate_velocity(1)
    u2 = sim.get("u")
    assert np.array_equal(u2[2], u0[2]) and np.array_equal(u2[1], u0[1]) and np.array_equal(u2[3], u0[3])
    assert np.array_equal(u2[0], u0[2]) and np.array_equal(u2[4], u0[2])
    assert np.all(u2[5:] == 0.0)
    sim.close()


# ---- applyDirichletBoundary (fluid-simulator.cc:113-138), applyForce (:13-18) ------------------
def test_dirichlet_walls_and_gravity_faces(oracle):
    n = 6
    dt = 0.004
    sim = _sim(n)
    rng = np.random.default_rng(3)
    g0, v0 = {}, {}
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        g0[nm] = rng.uniform(-1, 1, shape)
        v0[nm] = (rng.uniform(size=shape) < 0.5).astype(np.uint8)
        sim.set(nm, g0[nm])
        sim.set(nm + "_valid", v0[nm])
    sim.apply_dirichlet_boundary()
    for ax, nm in enumerate("uvw"):
        g, v = sim.get(nm), sim.get(nm + "_valid")
        wall = np.zeros(g.shape, bool)
        idx = [slice(None)] * 3
        idx[ax] = 0
        wall[tuple(idx)] = True
        idx[ax] = -1
        wall[tuple(idx)] = True
        assert np.all(g[wall] == 0.0) and np.all(v[wall] == 1)
        assert np.array_equal(g[~wall], g0[nm][~wall]) and np.array_equal(v[~wall], v0[nm][~wall])
    vv, vvalid = sim.get("v"), sim.get("v_valid").astype(bool)
    uu, ww = sim.get("u"), sim.get("w")
    sim.apply_force(dt)
    v2 = sim.get("v")
    changed = v2 != vv
    interior = np.zeros(vv.shape, bool)
    interior[:, 1:-1] = True
    # exactly the valid v-faces off the y-walls move, each by -g dt; u and w are untouched
    assert np.array_equal(changed, vvalid & interior)
    np.testing.assert_array_equal(v2[changed], vv[changed] - G * dt)
    assert changed.sum() == int((vvalid & interior).sum()) > 0
    assert np.array_equal(sim.get("u"), uu) and np.array_equal(sim.get("w"), ww)
    sim.close()


# ---- SDF::grad (FluidSim/include/FluidSim/cpu/sdf.h:62-118) ------------------------------------
def test_sdf_grad_linear_and_quadratic_fields(oracle):
    n = 8
    h = 1.0 / n
    sim = _sim(n)
    X, Y, Z = scenes.cell_centres(n)
    a = np.array([0.3, -1.7, 2.2])
    sim.set_collider_sdf(a[0] * X + a[1] * Y + a[2] * Z)
    rng = np.random.default_rng(4)
    for p in rng.uniform(0.5 * h, 1 - 0.5 * h, (60, 3)):
        np.testing.assert_allclose(sim.sdf_grad(p), a, rtol=1e-12)
    # f = x^2: node gradients 2x inside (central, spacing 2h) and 2x0 + h at node 0 (one-sided,
    # spacing h); the blend weights the lower node by t (mirrored like eval)
    sim.set_collider_sdf(X ** 2)
    x0, x1 = 0.5 * h, 1.5 * h
    for t in (0.1, 0.3, 0.8):
        got = sim.sdf_grad([x0 + t * h, 0.5, 0.5])[0]
        assert got == pytest.approx(t * (2 * x0 + h) + (1 - t) * (2 * x1), rel=1e-12)
        xi = 3.5 * h
        got = sim.sdf_grad([xi + t * h, 0.5, 0.5])[0]
        assert got == pytest.approx(t * 2 * xi + (1 - t) * 2 * (xi + h), rel=1e-12)
    sim.close()


# ---- applyCollider (fluid-simulator.cc:76-111): u -= u n_x inside the collider -----------------
def test_apply_collider_planar_collider(oracle):
    n = 8
    h = 1.0 / n
    sim = _sim(n)
    X, Y, Z = scenes.cell_centres(n)
    a = np.array([0.6, 0.8, 0.0])  # unit normal; the solid is a . x < 0.5
    c = 0.5
    sim.set_collider_sdf(a[0] * X + a[1] * Y + a[2] * Z - c)
    rng = np.random.default_rng(5)
    g0 = {}
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        g0[nm] = rng.uniform(-1, 1, shape)
        sim.set(nm, g0[nm])
    sim.apply_collider()
    for ax, nm in enumerate("uvw"):
        g = sim.get(nm)
        idx = np.indices(g.shape).astype(float)
        pts = [(idx[d] + (0.0 if d == ax else 0.5)) * h for d in range(3)]
        f = a[0] * pts[0] + a[1] * pts[1] + a[2] * pts[2] - c
        wall = (idx[ax] == 0) | (idx[ax] == g.shape[ax] - 1)
        assert np.all(g[wall] == 0.0)
        inside = (f < -1.5 * h) & ~wall
        outside = (f > 1.5 * h) & ~wall
        assert inside.sum() > 10 and outside.sum() > 10
        np.testing.assert_allclose(g[inside], g0[nm][inside] * (1.0 - a[ax]), rtol=1e-13, atol=1e-15)
        assert np.array_equal(g[outside], g0[nm][outside])
    sim.close()


# ---- advect (advect-solver.cc:170-187) and handleCollision (:100-137) --------------------------
def test_advect_on_cleared_grid_is_two_ninths_step(oracle):
    """clear() before advect (Q2): k2 = k3 = 0, so p' = p + (2/9) dt v exactly."""
    n = 8
    dt = 0.01
    sim = _sim(n)
    rng = np.random.default_rng(6)
    pos = rng.uniform(0.2, 0.8, (40, 3))
    vel = rng.uniform(-1, 1, (40, 3))
    sim.set_particles(pos, vel)
    sim.clear()
    sim.advect(dt)
    p, v = sim.particles()
    np.testing.assert_array_equal(p, pos + (2.0 / 9.0 * dt) * vel)
    np.testing.assert_array_equal(v, vel)
    sim.close()


def test_advect_rk3_stages_on_a_varying_grid(oracle):
    """On a grid that is not cleared, Ralston's RK3 (advect-solver.cc:170-187):
    k2 = S(p + dt/2 k1), k3 = S(p + 3dt/4 k2), p' = p + dt (2/9 k1 + 1/3 k2 + 4/9 k3), with S the
    grid sampling (pinned by the brute-force transfer test) and k1 the particle velocity."""
    n = 8
    dt = 0.05
    sim = _sim(n)
    rng = np.random.default_rng(7)
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        sim.set(nm, rng.uniform(-1, 1, shape))
    pos = rng.uniform(0.3, 0.7, (30, 3))
    vel = rng.uniform(-1, 1, (30, 3))
    want = []
    for p, v in zip(pos, vel):
        k2 = sim.sample_velocity(p + 0.5 * dt * v)
        k3 = sim.sample_velocity(p + 0.75 * dt * k2)
        want.append(p + dt * (2.0 / 9.0 * v + 1.0 / 3.0 * k2 + 4.0 / 9.0 * k3))
    sim.set_particles(pos, vel)
    sim.advect(dt)
    p, _ = sim.particles()
    np.testing.assert_allclose(p, np.array(want), rtol=0, atol=1e-15)
    sim.close()


def test_collision_push_out_and_walls(oracle):
    """Planar collider f = y - y0: a particle that ends inside is moved by -f n, where f is read
    with the mirrored weights, i.e. at y_c + (1 - t) h for p_y = y_c + t h; its inward velocity
    is removed. A particle leaving through x = 0 is clamped there with v_x zeroed."""
    n = 8
    h = 1.0 / n
    dt = 0.01
    y0 = 0.55
    sim = _sim(n)
    X, Y, Z = scenes.cell_centres(n)
    sim.set_collider_sdf(Y - y0)
    pos = np.array([[0.41, 0.40, 0.52], [0.001, 0.7, 0.5]])
    vel = np.array([[0.2, -0.5, 0.1], [-1.0, 0.3, 0.2]])
    sim.set_particles(pos, vel)
    sim.clear()
    sim.advect(dt)
    p, v = sim.particles()
    q = pos[0] + (2.0 / 9.0 * dt) * vel[0]
    yc = (np.floor(q[1] / h - 0.5) + 0.5) * h
    t = (q[1] - yc) / h
    d = yc + (1.0 - t) * h - y0
    assert d < 0
    assert p[0][0] == q[0] and p[0][2] == q[2]
    assert p[0][1] == pytest.approx(q[1] - d, abs=1e-15)
    np.testing.assert_allclose(v[0], [0.2, 0.0, 0.1], atol=1e-16)
    assert p[1][0] == 0.0 and v[1][0] == 0.0
    np.testing.assert_array_equal(v[1][1:], vel[1][1:])
    sim.close()


# ---- buildSystem (FluidSim/src/cpu/project-solver.cc:77-226) ------------------------------------
def test_assembly_ghost_fluid_theta_and_clamps(oracle):
    """Cell c = (1,1,1) with SDF -0.3 and six chosen neighbours, worked by hand:
      left  air +1000 -> theta = min(1000/1000.3, 0.99) = 0.99 (clamp),  Adiag += f/(1 - 0.99)
      right air +0.1  -> theta = max(-0.3/-0.4, 0.01) = 0.75,            Adiag += f/0.75
      down  fluid     -> Adiag += f, A_down = -f
      up    air +100  -> theta = max(-0.3/-100.3, 0.01) = 0.01 (clamp),  Adiag += f/0.01
      back  air +0.7  -> theta = min(0.7/1.0, 0.99) = 0.7,               Adiag += f/(1 - 0.7)
      front fluid     -> Adiag += f, A_front = -f
    with f = dt/h and all face weights 1 (no solid)."""
    n = 4
    h = 1.0 / n
    dt = 0.01
    f = dt / h
    sim = _sim(n)
    sd = np.full((n, n, n), 1.0)
    sd[1, 1, 1] = -0.3
    sd[0, 1, 1], sd[2, 1, 1], sd[1, 0, 1], sd[1, 2, 1], sd[1, 1, 0], sd[1, 1, 2] = 1000.0, 0.1, -0.5, 100.0, 0.7, -0.2
    sim.set("fluid_sdf", sd)
    rng = np.random.default_rng(8)
    vel = {}
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        vel[nm] = rng.uniform(-1, 1, shape)
        sim.set(nm, vel[nm])
    sim.build_system(dt)
    adiag = sim.get("adiag")
    want = f / (1.0 - 0.99) + f / 0.75 + f + f / 0.01 + f / (1.0 - 0.7) + f
    assert adiag[1, 1, 1] == pytest.approx(want, rel=1e-14)
    assert sim.get("a_down")[1, 1, 1] == -f and sim.get("a_front")[1, 1, 1] == -f
    for nm in ("a_left", "a_right", "a_up", "a_back"):
        assert sim.get(nm)[1, 1, 1] == 0.0, nm
    u, v, w = vel["u"], vel["v"], vel["w"]
    rhs = u[1, 1, 1] - u[2, 1, 1] + v[1, 1, 1] - v[1, 2, 1] + w[1, 1, 1] - w[1, 1, 2]
    assert sim.get("rhs")[1, 1, 1] == pytest.approx(rhs, abs=1e-15)
    act = sim.get("active")
    assert act[1, 1, 1] == act[1, 0, 1] == act[1, 1, 2] == 1 and act.sum() == 3
    uw = sim.get("u_weights")
    assert uw[1, 1, 1] == 1.0 and uw[2, 1, 1] == 1.0   # faces of the fluid cell
    assert uw[3, 1, 1] == 0.0                          # both sides air: skipped
    assert uw[0].max() == 0.0 and uw[-1].max() == 0.0  # domain walls
    sim.close()


def test_face_weights_sample_four_corners(oracle):
    """With the whole domain fluid and a planar collider f = 0.9x + 2y + 3.1z - c, each interior
    face weight is 1 - fractionInside of f at the face's four corners (+-h/2 along the two
    other axes), in the reference's corner order (project-solver.cc:85-122)."""
    n = 8
    h = 1.0 / n
    sim = _sim(n)
    X, Y, Z = scenes.cell_centres(n)
    a, c = np.array([0.9, 2.0, 3.1]), 3.0137
    fn = lambda p: a[0] * p[0] + a[1] * p[1] + a[2] * p[2] - c  # noqa: E731
    sim.set_collider_sdf(fn((X, Y, Z)))
    sim.set("fluid_sdf", np.full((n, n, n), -1.0))
    sim.build_system(0.01)
    corners = {  # (offsets along (x, y, z) in units of h/2), the reference's order per grid
        0: [(0, 1, -1), (0, -1, -1), (0, -1, 1), (0, 1, 1)],
        1: [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)],
        2: [(-1, -1, 0), (-1, 1, 0), (1, 1, 0), (1, -1, 0)],
    }
    checked = 0
    for ax, nm in enumerate(("u_weights", "v_weights", "w_weights")):
        wts = sim.get(nm)
        for idx in np.argwhere(np.ones(wts.shape, bool)):
            p = (idx + np.where(np.arange(3) == ax, 0.0, 0.5)) * h
            cs = [p + 0.5 * h * np.array(o) for o in corners[ax]]
            if idx[ax] in (0, wts.shape[ax] - 1):
                assert wts[tuple(idx)] == 0.0
                continue
            if any(np.any(q < 0.5 * h) or np.any(q > 1 - 0.5 * h) for q in cs):
                continue  # eval clamps outside the cell-centre range: not a planar value there
            vals = [fn(q) for q in cs]
            if min(abs(x) for x in vals) < 1e-9:
                continue  # a corner on the plane: its sign is decided by rounding
            want = 1.0 - oracle.fraction_inside(*vals)
            assert wts[tuple(idx)] == pytest.approx(want, abs=1e-12), (nm, idx)
            checked += 1
    assert checked > 300
    sim.close()


# ---- MIC(0) tau terms of the Down and Back blocks (project-solver.cc:252-269) -----------------
def _mic_case(n, cells, coeffs, tau, sigma=0.0):
    act = np.zeros((n, n, n), np.uint8)
    for c in cells:
        act[c] = 1
    arrs = {k: np.zeros((n, n, n)) for k in ("a_left", "a_down", "a_back")}
    for (nm, c), val in coeffs.items():
        arrs[nm][c] = val
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n, mic_tau=tau, mic_sigma=sigma))
    for nm, arr in (("active", act), ("adiag", np.where(act == 1, 4.0, 0.0)), *arrs.items()):
        sim.set(nm, arr)
    sim.mic_apply(np.zeros((n, n, n)))
    E = sim.get("mic_e")
    sim.close()
    return E


_BLOCK = {0: "a_left", 1: "a_down", 2: "a_back"}


@pytest.mark.parametrize("axis,other", [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)],
                         ids=["left+y", "left+z", "down+x", "down+z", "back+x", "back+y"])
def test_mic_each_tau_term_hand_computed(oracle, axis, other):
    """Each of the six tau terms alone (project-solver.cc:245-269). Cells d = (1,1,1), c = d + e_axis,
    side s = d + e_other and s - e_axis (so that s's coefficient A_axis(s) couples two cells).
    d has no active lower neighbour, so E(d) = 1/sqrt(4) = 1/2, and with A_axis(c) = -1,
    A_axis(s) = -0.5, tau = 0.5, worked by hand:
      e(c) = 4 - (-1 * 1/2)^2 - 0.5 * (-1)(-0.5) / (1/2)^2 = 4 - 0.25 - 1 = 2.75."""
    d = np.array([1, 1, 1])
    ea, eo = np.eye(3, dtype=int)[axis], np.eye(3, dtype=int)[other]
    c, side, below = d + ea, d + eo, d + eo - ea
    nm = _BLOCK[axis]
    E = _mic_case(4, [tuple(d), tuple(c), tuple(side), tuple(below)],
                  {(nm, tuple(c)): -1.0, (nm, tuple(side)): -0.5}, 0.5)
    assert E[tuple(d)] == 0.5
    assert E[tuple(c)] == pytest.approx(1 / np.sqrt(2.75), rel=1e-15)
    # without the side cell the term vanishes: e(c) = 4 - 0.25
    E2 = _mic_case(4, [tuple(d), tuple(c), tuple(below)], {(nm, tuple(c)): -1.0, (nm, tuple(side)): -0.5}, 0.5)
    assert E2[tuple(c)] == pytest.approx(1 / np.sqrt(3.75), rel=1e-15)


def test_mic_sigma_safeguard(oracle):
    """tau 0.97, AD(c) = -1.9 for c = (1,2,1) over d = (1,1,1), side (2,1,1):
    e = 4 - 0.9025 - 0.97 * 1.9 * 0.5 * 4 = -0.5885 < sigma * 4, so the pivot falls back to
    Adiag = 4 and E = 1/2 exactly."""
    E = _mic_case(4, [(1, 1, 1), (1, 2, 1), (2, 1, 1), (2, 0, 1)],
                  {("a_down", (1, 2, 1)): -1.9, ("a_down", (2, 1, 1)): -0.5}, 0.97, sigma=0.25)
    assert E[1, 2, 1] == 0.5


# ---- transfers: numpy brute force (advect-solver.cc:188-271, util.cc:40-85) --------------------
def _bspline(r):
    r = np.abs(r)
    return np.where(r < 1, 2.0 / 3.0 - r * r + 0.5 * r ** 3, np.where(r < 2, (2 - r) ** 3 / 6.0, 0.0))


def _stencil(p, ax, h, hk):
    """Nodes nearest-1 .. nearest+2 of face grid `ax` per axis and their weights N(|d|/hk)."""
    out = []
    for d in range(3):
        off = 0.0 if d == ax else 0.5
        nb = int(np.floor((p[d] - off * h[d]) / h[d] + 0.5))
        idx = np.arange(nb - 1, nb + 3)
        out.append((idx, _bspline((p[d] - (idx + off) * h[d]) / hk[d])))
    return out


def test_p2g_g2p_brute_force(oracle):
    """20 particles on an 8 x 8 x 6 grid with h_z != h_x: P2G weighs each axis by its own
    spacing, sampleVelocity uses the scalar h.x on every axis (util.cc:47)."""
    nx, nz = 8, 6
    size_z = 0.9
    h = np.array([1 / nx, 1 / nx, size_z / nz])
    cfg = B.default_config(nx=nx, ny=nx, nz=nz, size_z=size_z)
    sim = B.OracleSim(cfg)
    sim.set_collider_sdf(np.full((nx, nx, nz), 10.0))
    rng = np.random.default_rng(9)
    pos = rng.uniform(0.02, 0.98, (20, 3)) * np.array([1.0, 1.0, size_z])
    vel = rng.uniform(-1, 1, (20, 3))
    sim.set_particles(pos, vel)
    sim.clear()
    sim.p2g(0.0)
    for ax, nm in enumerate("uvw"):
        shape, _ = sim.array_info(nm)
        acc, wsum = np.zeros(shape), np.zeros(shape)
        for p, v in zip(pos, vel):
            (ix, wx), (iy, wy), (iz, wz) = _stencil(p, ax, h, h)
            for a, x in zip(wx, ix):
                for b, y in zip(wy, iy):
                    for c, z in zip(wz, iz):
                        if 0 <= x < shape[0] and 0 <= y < shape[1] and 0 <= z < shape[2]:
                            acc[x, y, z] += v[ax] * a * b * c
                            wsum[x, y, z] += a * b * c
        want = np.where(wsum > 0, acc / np.where(wsum > 0, wsum, 1), 0.0)
        np.testing.assert_allclose(sim.get(nm), want, rtol=1e-13, atol=1e-15)
        np.testing.assert_array_equal(sim.get(nm + "_valid"), (wsum > 0).astype(np.uint8))
    grids = {}
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        grids[nm] = rng.uniform(-1, 1, shape)
        sim.set(nm, grids[nm])
    hx = np.full(3, h[0])
    for p in pos:
        want = []
        for ax, nm in enumerate("uvw"):
            g = grids[nm]
            (ix, wx), (iy, wy), (iz, wz) = _stencil(p, ax, h, hx)
            num = den = 0.0
            for a, x in zip(wx, ix):
                for b, y in zip(wy, iy):
                    for c, z in zip(wz, iz):
                        if 0 <= x < g.shape[0] and 0 <= y < g.shape[1] and 0 <= z < g.shape[2]:
                            num += g[x, y, z] * a * b * c
                            den += a * b * c
            want.append(num / den)
        np.testing.assert_allclose(sim.sample_velocity(p), want, rtol=1e-13, atol=1e-15)
    sim.close()


# ---- CFL (fluid-simulator.cc:184-203): 5 max(1e-3, min h/|u|) -----------------------------------
def test_cfl_closed_form(oracle):
    n = 8
    h = 1.0 / n
    sim = _sim(n)
    assert sim.cfl() == pytest.approx(5 * h / 1e-6, rel=1e-15)  # all zero: h / 1e-6
    rng = np.random.default_rng(10)
    for nm in "uvw":
        shape, _ = sim.array_info(nm)
        sim.set(nm, rng.uniform(-1, 1, shape))
    w = sim.get("w")
    w[3, 4, 5] = -3.0
    sim.set("w", w)
    assert sim.cfl() == pytest.approx(5 * h / 3.0, rel=1e-15)
    w[3, 4, 5] = 1e3  # h/|u| = 1.25e-4 < 1e-3: the floor
    sim.set("w", w)
    assert sim.cfl() == pytest.approx(5e-3, rel=1e-15)
    sim.close()
