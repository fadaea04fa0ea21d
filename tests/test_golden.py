"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

The oracle must reproduce them bit for bit (a regression pin on the oracle itself); the device
reproduces them within the north-star tolerances, since the fixtures use the reference's
serial PCG dot products while the device uses a fixed reduction tree.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import binding as B
from tests.conftest import announce

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def run(sim, g):
    sim.set_collider_sdf(g["collider"])
    sim.set_particles(g["pos0"], g["vel0"])
    iters = [sim.substep(float(g["dt"]))["cg_iterations"] for _ in range(int(g["steps"]))]
    p, v = sim.particles()
    return p, v, sim.get("pg"), sim.get("active"), np.array(iters)


def config(mod, g):
    n = int(g["n"])
    return mod.default_config(nx=n, ny=n, nz=n, flip_alpha=float(g["flip_alpha"]))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(oracle, name):
    g = load(name)
    sim = B.OracleSim(config(B, g), dot_mode=B.DOT_SERIAL)
    p, v, pg, act, iters = run(sim, g)
    sim.close()
    np.testing.assert_array_equal(p, g["pos"])
    np.testing.assert_array_equal(v, g["vel"])
    np.testing.assert_array_equal(pg, g["pg"])
    np.testing.assert_array_equal(act, g["active"])
    np.testing.assert_array_equal(iters, g["iters"])


@pytest.mark.parametrize("name", CASES)
def test_tree_dots_stay_within_tolerance_of_serial(oracle, name):
    g = load(name)
    sim = B.OracleSim(config(B, g), dot_mode=B.DOT_TREE)
    p, v, pg, act, iters = run(sim, g)
    sim.close()
    np.testing.assert_array_equal(act, g["active"])
    a = act.astype(bool)
    assert np.linalg.norm(pg[a] - g["pg"][a]) <= 1e-5 * np.linalg.norm(g["pg"][a])
    np.testing.assert_allclose(p, g["pos"], atol=1e-9)
    assert np.all(np.abs(iters - g["iters"]) <= 2)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_device_reproduces_golden(hip, name):
    announce(f"golden {name}")
    g = load(name)
    sim = hip.FluidSim(config(hip, g))
    p, v, pg, act, iters = run(sim, g)
    sim.close()
    np.testing.assert_array_equal(act, g["active"])
    a = act.astype(bool)
    assert np.linalg.norm(pg[a] - g["pg"][a]) <= 1e-5 * np.linalg.norm(g["pg"][a])
    np.testing.assert_allclose(p, g["pos"], atol=1e-9)
    np.testing.assert_allclose(v, g["vel"], atol=1e-6)
    assert np.all(np.abs(iters - g["iters"]) <= 2)
