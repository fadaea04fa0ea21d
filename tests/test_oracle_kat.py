"""Known-answer tests that pin the oracle (CPU only).

Every test checks an oracle function against something computed another way: a hand-worked
value from the reference's formula, a brute-force search, a textbook algorithm run on a dense
matrix, or an invariant of the mathematics. Reference file:line citations name what is pinned.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
from simcraft_amd import scenes


# ---- fractionInside (FluidSim/src/cpu/util.cc:333-378) -----------------------------------------
# Worked by hand from the reference's branches. The reference's edge fractions are
# positive/(positive - negative), i.e. the share of the edge on the positive side, so
# asymmetric corner values do not give the textbook areas; values with |d| = 1 do.
@pytest.mark.parametrize("corners,expect", [
    ((1.0, 1.0, 1.0, 1.0), 0.0),
    ((-1.0, -1.0, -1.0, -1.0), 1.0),
    ((0.0, 0.0, 0.0, 0.0), 1.0),          # 0 counts as negative (d <= 0)
    ((-1.0, 1.0, 1.0, 1.0), 0.125),
    ((-1.0, 3.0, 3.0, 3.0), 0.28125),     # 0.5 * (3/4) * (3/4)
    ((3.0, 3.0, -1.0, 3.0), 0.28125),     # same, after cycling
    ((-1.0, -1.0, 1.0, 1.0), 0.5),
    ((-1.0, -1.0, 3.0, 3.0), 0.75),       # ru <= 0 branch: 0.5 * (3/4 + 3/4)
    ((-1.0, 3.0, 3.0, -1.0), 0.75),       # ld <= 0 branch
    ((-1.0, 3.0, -1.0, 3.0), 0.5625),     # diagonal, sum > 0: 0.5 * (9/16 + 9/16)
    ((-3.0, 1.0, -3.0, 1.0), 0.4375),     # diagonal, sum <= 0: 1 - 0.5 * (9/16 + 9/16)
    ((-1.0, -1.0, -1.0, 1.0), 0.875),
    ((-1.0, -1.0, -1.0, 3.0), 0.96875),   # 1 - 0.5 * (1/4) * (1/4)
])
def test_fraction_inside_known_answers(oracle, corners, expect):
    assert oracle.fraction_inside(*corners) == pytest.approx(expect, abs=1e-15)


def test_fraction_inside_rotation_invariant(oracle):
    rng = np.random.default_rng(5)
    for _ in range(500):
        c = list(rng.uniform(-1, 1, 4))
        vals = [oracle.fraction_inside(*(c[k:] + c[:k])) for k in range(4)]
        assert max(vals) - min(vals) < 1e-14, c
        assert 0.0 <= vals[0] <= 1.0


# ---- Morton keys (Spatify/include/Spatify/ns-util.h:36-45) ------------------------------------
def test_morton_matches_bit_loop(oracle):
    rng = np.random.default_rng(11)
    for x, y, z in rng.integers(0, 1024, size=(300, 3)):
        want = 0
        for bit in range(10):
            want |= ((int(x) >> bit) & 1) << (3 * bit)
            want |= ((int(y) >> bit) & 1) << (3 * bit + 1)
            want |= ((int(z) >> bit) & 1) << (3 * bit + 2)
        assert oracle.morton3d(int(x), int(y), int(z)) == want


# ---- transfer kernel (assumed cubic B-spline; CubicKernel is absent from the reference) -------
def test_bspline_values_and_partition_of_unity(oracle):
    assert oracle.bspline(0.0) == pytest.approx(2.0 / 3.0, abs=1e-16)
    assert oracle.bspline(0.5) == pytest.approx(23.0 / 48.0, abs=1e-16)
    assert oracle.bspline(1.0) == pytest.approx(1.0 / 6.0, abs=1e-16)
    assert oracle.bspline(1.5) == pytest.approx(1.0 / 48.0, abs=1e-16)
    assert oracle.bspline(2.0) == 0.0
    assert oracle.bspline(2.5) == 0.0
    for t in np.linspace(0.0, 0.999, 37):
        total = sum(oracle.bspline(abs(t - k)) for k in (-1, 0, 1, 2))
        assert total == pytest.approx(1.0, abs=1e-15)


# ---- nearest node (Spatify/include/Spatify/grids.h:287-292, face offsets :479-485) ------------
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_nearest_node_brute_force(oracle, axis):
    n = 8
    h = 1.0 / n
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    rng = np.random.default_rng(axis)
    for p in rng.uniform(0.01, 0.99, size=(200, 3)):
        got = sim.nearest(axis, p)
        for c in range(3):
            off = 0.0 if c == axis else 0.5
            nodes = (np.arange(-2, n + 3) + off) * h
            want = int(np.arange(-2, n + 3)[np.argmin(np.abs(p[c] - nodes))])
            assert got[c] == want, (axis, c, p)
    sim.close()


# ---- SDF::eval with the reference's mirrored weights (FluidSim/include/FluidSim/cpu/sdf.h:35-61)
def test_sdf_eval_linear_field_mirrored_weights(oracle):
    n = 8
    h = 1.0 / n
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    a = np.array([0.3, -1.7, 2.2])
    X, Y, Z = scenes.cell_centres(n)
    sim.set_collider_sdf(a[0] * X + a[1] * Y + a[2] * Z)
    rng = np.random.default_rng(3)
    for _ in range(100):
        i0 = rng.integers(0, n - 1, 3)
        t = rng.uniform(0.0, 1.0, 3)
        x0 = (i0 + 0.5) * h
        p = x0 + t * h
        # weight t on the lower node and 1-t on the upper one: the field is read at x0 + (1-t) h
        want = float(np.dot(a, x0 + (1.0 - t) * h))
        assert sim.sdf_eval(p) == pytest.approx(want, abs=1e-13)
    sim.close()


# ---- sampleVelocity (FluidSim/src/cpu/util.cc:40-85) -------------------------------------------
def test_sample_velocity_reproduces_constant_field(oracle):
    n = 8
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    const = (0.3, -0.7, 1.1)
    for nm, c in zip(("u", "v", "w"), const):
        shape, _ = sim.array_info(nm)
        sim.set(nm, np.full(shape, c))
    rng = np.random.default_rng(9)
    for p in rng.uniform(0.2, 0.8, size=(50, 3)):
        np.testing.assert_allclose(sim.sample_velocity(p), const, rtol=1e-14)
    sim.close()


# ---- P2G / G2P (FluidSim/src/cpu/advect-solver.cc:157-271) -------------------------------------
def _dam(n, seed=1):
    pos, _ = scenes.dam_break(n, seed)
    return pos


def test_p2g_reproduces_constant_velocity(oracle):
    n = 16
    pos = _dam(n)
    vel = np.tile([0.3, -0.7, 1.1], (len(pos), 1))
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    sim.set_collider_sdf(scenes.no_collider(n))
    sim.set_particles(pos, vel)
    sim.clear()
    sim.p2g(0.0)
    for nm, c in zip(("u", "v", "w"), (0.3, -0.7, 1.1)):
        g, valid = sim.get(nm), sim.get(nm + "_valid").astype(bool)
        assert valid.sum() > 0
        np.testing.assert_allclose(g[valid], c, rtol=1e-14)
        assert np.all(g[~valid] == 0.0)
        np.testing.assert_array_equal(sim.get(nm + "_last"), g)
    sim.close()


def test_flip_g2p_is_blend_of_particle_and_pic_velocity(oracle):
    """With u_last == ug (right after P2G) FLIP gives alpha*v + (1-alpha)*pic."""
    n = 16
    alpha = 0.95
    pos = _dam(n, seed=4)
    vel = np.random.default_rng(4).uniform(-1, 1, pos.shape)
    out = {}
    for adv, a in (("pic", 0.0), ("flip", alpha)):
        sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n, flip_alpha=a))
        sim.set_collider_sdf(scenes.no_collider(n))
        sim.set_particles(pos, vel)
        sim.clear()
        sim.p2g(0.0)
        sim.g2p(0.0)
        out[adv] = sim.particles()[1]
        sim.close()
    np.testing.assert_allclose(out["flip"], alpha * vel + (1 - alpha) * out["pic"], rtol=1e-12, atol=1e-14)


# ---- assembly (FluidSim/src/cpu/project-solver.cc:77-226) -------------------------------------
def test_poisson_all_fluid_box(oracle):
    """Every cell fluid, no solid: Adiag = open faces * dt/h, neighbours -dt/h, rhs = -div u."""
    n = 6
    h = 1.0 / n
    dt = 0.01
    f = dt / h
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    sim.set_collider_sdf(scenes.no_collider(n))
    sim.set("fluid_sdf", np.full((n, n, n), -1.0))
    rng = np.random.default_rng(0)
    vel = {}
    for nm in ("u", "v", "w"):
        shape, _ = sim.array_info(nm)
        vel[nm] = rng.uniform(-1, 1, shape)
        sim.set(nm, vel[nm])
    sim.build_system(dt)
    assert np.all(sim.get("active") == 1)
    interior = np.zeros((n, n, n))  # number of faces shared with another cell
    for ax in range(3):
        lo = np.moveaxis(np.zeros((n, n, n)), ax, 0)
        lo[1:] += 1   # a face towards the lower neighbour exists
        lo[:-1] += 1  # a face towards the upper neighbour exists
        interior += np.moveaxis(lo, 0, ax)
    np.testing.assert_allclose(sim.get("adiag"), interior * f, rtol=1e-15)
    al = sim.get("a_left")
    assert np.all(al[1:] == -f) and np.all(al[0] == 0.0)
    ad = sim.get("a_down")
    assert np.all(ad[:, 1:] == -f) and np.all(ad[:, 0] == 0.0)
    ab = sim.get("a_back")
    assert np.all(ab[:, :, 1:] == -f) and np.all(ab[:, :, 0] == 0.0)
    # walls carry weight 0, interior faces 1
    u, v, w = vel["u"].copy(), vel["v"].copy(), vel["w"].copy()
    u[0], u[-1], v[:, 0], v[:, -1], w[:, :, 0], w[:, :, -1] = 0, 0, 0, 0, 0, 0
    div = (u[1:] - u[:-1]) + (v[:, 1:] - v[:, :-1]) + (w[:, :, 1:] - w[:, :, :-1])
    np.testing.assert_allclose(sim.get("rhs"), -div, atol=1e-14)
    # symmetry of the stored stencil (project-solver.cc:41-57)
    np.testing.assert_array_equal(sim.get("a_right")[:-1], al[1:])
    np.testing.assert_array_equal(sim.get("a_up")[:, :-1], ad[:, 1:])
    np.testing.assert_array_equal(sim.get("a_front")[:, :, :-1], ab[:, :, 1:])
    sim.close()


# ---- SpMV, MIC(0) and PCG (FluidSim/src/cpu/project-solver.cc:16-39, 230-346) -----------------
def _random_system(n, seed, frac=0.7):
    rng = np.random.default_rng(seed)
    act = (rng.uniform(size=(n, n, n)) < frac).astype(np.uint8)
    a = act.astype(bool)
    al = np.zeros((n, n, n))
    ad = np.zeros((n, n, n))
    ab = np.zeros((n, n, n))
    al[1:] = np.where(a[1:] & a[:-1], -rng.uniform(0.5, 1.0, (n - 1, n, n)), 0.0)
    ad[:, 1:] = np.where(a[:, 1:] & a[:, :-1], -rng.uniform(0.5, 1.0, (n, n - 1, n)), 0.0)
    ab[:, :, 1:] = np.where(a[:, :, 1:] & a[:, :, :-1], -rng.uniform(0.5, 1.0, (n, n, n - 1)), 0.0)
    adiag = np.where(a, 6.5 + rng.uniform(0, 0.5, (n, n, n)), 0.0)
    return act, adiag, al, ad, ab


def _dense(act, adiag, al, ad, ab):
    n = act.shape[0]
    cells = [tuple(c) for c in np.argwhere(act)]
    pos = {c: i for i, c in enumerate(cells)}
    A = np.zeros((len(cells), len(cells)))
    for c, i in pos.items():
        A[i, i] = adiag[c]
        for ax, arr in ((0, al), (1, ad), (2, ab)):
            lo = list(c)
            lo[ax] -= 1
            lo = tuple(lo)
            if lo in pos:
                A[i, pos[lo]] = A[pos[lo], i] = arr[c]
    return cells, A


def _load(sim, act, adiag, al, ad, ab):
    ar, au, af = np.zeros_like(al), np.zeros_like(ad), np.zeros_like(ab)
    ar[:-1], au[:, :-1], af[:, :, :-1] = al[1:], ad[:, 1:], ab[:, :, 1:]
    for nm, arr in (("active", act), ("adiag", adiag), ("a_left", al), ("a_down", ad), ("a_back", ab),
                    ("a_right", ar), ("a_up", au), ("a_front", af)):
        sim.set(nm, arr)


def test_spmv_matches_dense(oracle):
    n = 6
    act, adiag, al, ad, ab = _random_system(n, 1)
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    _load(sim, act, adiag, al, ad, ab)
    x = np.random.default_rng(2).standard_normal((n, n, n))
    y0 = np.full((n, n, n), 7.0)
    y = sim.spmv(x, y0)
    cells, A = _dense(act, adiag, al, ad, ab)
    idx = tuple(np.array(cells).T)
    np.testing.assert_allclose(y[idx], A @ x[idx], rtol=1e-13, atol=1e-13)
    assert np.all(y[act == 0] == 7.0)  # inactive rows are not written
    sim.close()


def test_mic_tau0_is_textbook_ic0(oracle):
    """With tau = 0 (and sigma = 0) the reference's factor is IC(0) of the 7-point matrix."""
    n = 5
    act, adiag, al, ad, ab = _random_system(n, 3)
    cells, A = _dense(act, adiag, al, ad, ab)
    m = len(cells)
    # textbook IC(0): Cholesky restricted to the sparsity pattern of A (lexicographic order)
    L = np.zeros_like(A)
    pattern = A != 0
    for j in range(m):
        L[j, j] = np.sqrt(A[j, j] - np.sum(L[j, :j] ** 2))
        for i in range(j + 1, m):
            if pattern[i, j]:
                L[i, j] = (A[i, j] - np.sum(L[i, :j] * L[j, :j])) / L[j, j]
    r = np.random.default_rng(4).standard_normal((n, n, n))
    idx = tuple(np.array(cells).T)
    want = np.linalg.solve(L.T, np.linalg.solve(L, r[idx]))
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n, mic_tau=0.0, mic_sigma=0.0))
    _load(sim, act, adiag, al, ad, ab)
    z = sim.mic_apply(r)
    np.testing.assert_allclose(z[idx], want, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(sim.get("mic_e")[idx], 1.0 / np.diag(L), rtol=1e-13)
    sim.close()


def test_mic_tau_term_hand_computed(oracle):
    """Five cells where exactly one tau term is nonzero (project-solver.cc:245-250), worked by
    hand: the term of cell (2,0,0) through its left neighbour (1,0,0) and side cell (1,1,0)."""
    n = 4
    act = np.zeros((n, n, n), np.uint8)
    for c in ((0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 0)):
        act[c] = 1
    adiag = np.where(act == 1, 4.0, 0.0)
    al = np.zeros((n, n, n))
    ad = np.zeros((n, n, n))
    ab = np.zeros((n, n, n))
    al[1, 0, 0], al[2, 0, 0], al[1, 1, 0] = -1.0, -1.0, -0.5
    ad[0, 1, 0], ad[1, 1, 0] = -1.0, -1.0
    tau = 0.5
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n, mic_tau=tau, mic_sigma=0.0))
    _load(sim, act, adiag, al, ad, ab)
    sim.mic_apply(np.zeros((n, n, n)))
    E = sim.get("mic_e")
    e000 = 4.0
    E000 = 1 / np.sqrt(e000)
    # (1,0,0): left (0,0,0) with side (0,1,0): AL(1,0,0) * AL(0,1,0) = 0, no tau contribution
    e100 = 4.0 - (al[1, 0, 0] * E000) ** 2
    E100 = 1 / np.sqrt(e100)
    # (0,1,0): down (0,0,0); side (1,0,0) along +x: tau * AD(0,1,0) * AD(1,0,0) / E000^2 = 0 (AD(1,0,0) = 0)
    e010 = 4.0 - (ad[0, 1, 0] * E000) ** 2
    E010 = 1 / np.sqrt(e010)
    # (2,0,0): left (1,0,0), side (1,1,0): tau * AL(2,0,0) * AL(1,1,0) / E100^2
    e200 = 4.0 - (al[2, 0, 0] * E100) ** 2 - tau * al[2, 0, 0] * al[1, 1, 0] / E100 ** 2
    E200 = 1 / np.sqrt(e200)
    np.testing.assert_allclose([E[0, 0, 0], E[1, 0, 0], E[0, 1, 0], E[2, 0, 0]], [E000, E100, E010, E200],
                               rtol=1e-15)
    sim.close()


def test_pcg_tridiagonal_converges_in_one_iteration(oracle):
    """A single active x-line: IC(0) is the exact Cholesky factor, so PCG stops after one step."""
    n = 16
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n), dot_mode=B.DOT_SERIAL)
    rng = np.random.default_rng(1)
    act = np.zeros((n, n, n), np.uint8)
    act[:, 3, 5] = 1
    adiag = np.zeros((n, n, n))
    adiag[:, 3, 5] = rng.uniform(2.5, 3.0, n)
    al = np.zeros((n, n, n))
    al[1:, 3, 5] = -rng.uniform(0.5, 1.0, n - 1)
    _load(sim, act, adiag, al, np.zeros_like(al), np.zeros_like(al))
    rhs = np.zeros((n, n, n))
    rhs[:, 3, 5] = rng.standard_normal(n)
    sim.set("rhs", rhs)
    st = sim.solve_pressure()
    assert st["cg_iterations"] == 1
    A = np.diag(adiag[:, 3, 5]) + np.diag(al[1:, 3, 5], -1) + np.diag(al[1:, 3, 5], 1)
    np.testing.assert_allclose(sim.get("pg")[:, 3, 5], np.linalg.solve(A, rhs[:, 3, 5]), atol=1e-14)
    sim.close()


@pytest.mark.parametrize("dot_mode", [B.DOT_SERIAL, B.DOT_TREE])
def test_pcg_solves_random_system(oracle, dot_mode):
    n = 8
    act, adiag, al, ad, ab = _random_system(n, 7, frac=0.8)
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n), dot_mode=dot_mode)
    _load(sim, act, adiag, al, ad, ab)
    rhs = np.where(act == 1, np.random.default_rng(8).standard_normal((n, n, n)), 0.0)
    sim.set("rhs", rhs)
    st = sim.solve_pressure()
    cells, A = _dense(act, adiag, al, ad, ab)
    idx = tuple(np.array(cells).T)
    res = A @ sim.get("pg")[idx] - rhs[idx]
    assert st["status"] == 0 and st["cg_iterations"] < 300
    assert np.abs(res).max() < 1e-6
    assert st["residual_inf"] == pytest.approx(np.abs(res).max(), rel=1e-6, abs=1e-15)
    sim.close()


# ---- projection (FluidSim/src/cpu/project-solver.cc:348-435) -----------------------------------
def test_projection_leaves_active_cells_divergence_free(oracle):
    n = 16
    dt = 0.004
    h = 1.0 / n
    pos, vel = scenes.dam_break(n, 5)
    vel = np.random.default_rng(5).uniform(-0.5, 0.5, pos.shape)
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    sim.set_collider_sdf(scenes.sphere_sdf(n, centre=(0.3, 0.3, 0.5), radius=0.12))
    sim.set_particles(pos, vel)
    for name, fn in (("clear", lambda: sim.clear()), ("advect", lambda: sim.advect(dt)),
                     ("reconstruct", lambda: sim.reconstruct(0.8 * h)),
                     ("esdf", lambda: sim.extrapolate_fluid_sdf(10)), ("smooth", lambda: sim.smooth_fluid_surface(5)),
                     ("p2g", lambda: sim.p2g(dt)), ("dirichlet", lambda: sim.apply_dirichlet_boundary()),
                     ("extrap", lambda: sim.extrapolate_velocity(10)), ("force", lambda: sim.apply_force(dt)),
                     ("build", lambda: sim.build_system(dt))):
        fn()
    st = sim.solve_pressure()
    assert st["status"] == 0
    sim.project(dt)
    u, v, w = sim.get("u"), sim.get("v"), sim.get("w")
    uw, vw, ww = sim.get("u_weights"), sim.get("v_weights"), sim.get("w_weights")
    act = sim.get("active").astype(bool)
    div = (u[:-1] * uw[:-1] - u[1:] * uw[1:]) + (v[:, :-1] * vw[:, :-1] - v[:, 1:] * vw[:, 1:]) + \
          (w[:, :, :-1] * ww[:, :, :-1] - w[:, :, 1:] * ww[:, :, 1:])
    assert act.sum() > 100
    # the weighted divergence is the pressure residual scaled by h/dt
    assert np.abs(div[act]).max() < 1e-6 * h / dt * 1.01
    sim.close()


# ---- binning and reconstruction (ns-util.h:46-109, sdf.h:210-225) -----------------------------
def test_reconstruct_matches_brute_force_with_skipped_last_particle(oracle):
    """Each Morton bin's last particle (highest index, stable order) is never visited; cells
    whose bins hold nothing else stay invalid at 1e9."""
    n = 8
    h = 1.0 / n
    radius = 0.8 * h
    rng = np.random.default_rng(12)
    pos = rng.uniform(0.05, 0.95, size=(150, 3))
    sim = B.OracleSim(B.default_config(nx=n, ny=n, nz=n))
    sim.set_collider_sdf(scenes.no_collider(n))
    sim.set_particles(pos, np.zeros_like(pos))
    sim.reconstruct(radius)
    sdf, valid = sim.get("fluid_sdf"), sim.get("sdf_valid")
    bins = np.clip(np.floor(pos / h).astype(int), 0, n - 1)
    last = {}
    for idx, b in enumerate(map(tuple, bins)):
        last[b] = idx  # ascending index: the final assignment is the bin's last particle
    skipped = set(last.values())
    r2 = 2 * radius
    want = np.full((n, n, n), 1e9)
    want_valid = np.zeros((n, n, n), np.uint8)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = (np.array([i, j, k]) + 0.5) * h
                lo = np.maximum(0, np.floor((c - r2) / h).astype(int))
                hi = np.minimum(n - 1, np.ceil((c + r2) / h).astype(int))
                for p_idx in range(len(pos)):
                    if p_idx in skipped or np.any(bins[p_idx] < lo) or np.any(bins[p_idx] > hi):
                        continue
                    d = pos[p_idx] - c
                    dist = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
                    if dist <= r2:
                        want[i, j, k] = min(want[i, j, k], dist - radius)
                        want_valid[i, j, k] = 1
    np.testing.assert_array_equal(valid, want_valid)
    np.testing.assert_array_equal(sdf, want)
    sim.close()


# ---- synthetic inputs (the workloads' initialiser) -------------------------------------------
def test_splitmix64_published_value():
    # the splitmix64 generator seeded with 0 (Steele, Lea and Flood 2014) outputs
    # mix(k * gamma) for k = 1, 2, 3, ...; scenes.splitmix64(x) is mix(x + gamma)
    gamma = 0x9E3779B97F4A7C15
    x = np.array([0, gamma, (2 * gamma) % 2**64], dtype=np.uint64)
    assert [int(v) for v in scenes.splitmix64(x)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("n,cells", [(32, 13 * 19 * 32), (128, 51 * 77 * 128)])
def test_dam_break_counts(n, cells):
    assert int(scenes.dam_break_mask(n).sum()) == cells
    if n == 32:
        pos, vel = scenes.dam_break(n)
        assert pos.shape == (8 * cells, 3) and np.all(vel == 0)
        h = 1.0 / n
        # 8 particles per cell, one per 2x2x2 sub-cell
        sub = np.floor(pos / (h / 2)).astype(int)
        assert len({tuple(s) for s in sub}) == 8 * cells
