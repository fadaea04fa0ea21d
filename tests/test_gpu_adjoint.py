"""Reverse mode on the device: transpose identities, central finite differences of the FP64
CPU oracle's forward, checkpoint invariance and the torch.autograd binding."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as B
from simcraft_amd import scenes
from tests.conftest import announce

N = 16
DT = 0.004


def _scene(seed=3, kind="dam"):
    if kind == "dam":
        pos, _ = scenes.dam_break(N, seed)
    else:  # fluid fills the box: no free surface
        pos = scenes.particles_in_cells(np.ones((N, N, N), bool), seed, 1.0 / N)
    vel = np.random.default_rng(seed).uniform(-0.3, 0.3, pos.shape)
    return pos, vel


def _oracle_loss(pos0, vel0, cx, cv, steps, gravity=9.8, alpha=0.95):
    o = B.OracleSim(B.default_config(nx=N, ny=N, nz=N, flip_alpha=alpha, gravity=gravity), dot_mode=B.DOT_TREE)
    o.set_collider_sdf(scenes.no_collider(N))
    o.set_particles(pos0, vel0)
    for _ in range(steps):
        o.substep(DT)
    p, v = o.particles()
    o.close()
    return float(np.sum(cx * p) + np.sum(cv * v))


def _device(hip, pos, vel, alpha=0.95, precision=None, gravity=9.8):
    cfg = hip.default_config(nx=N, ny=N, nz=N, flip_alpha=alpha, gravity=gravity)
    if precision is not None:
        cfg.precision = precision
    sim = hip.FluidSim(cfg)
    sim.set_collider_sdf(scenes.no_collider(N))
    sim.set_particles(pos, vel)
    return sim


def _vjp(hip, pos, vel, cx, cv, steps, every, **kw):
    sim = _device(hip, pos, vel, **kw)
    sim.rollout_begin(every)
    for _ in range(steps):
        sim.substep(DT)
    out = sim.rollout_vjp(cx, cv)
    sim.close()
    return out


@pytest.mark.gpu
def test_sample_grid_transpose(hip):
    announce("adjoint sample_grid transpose")
    pos, vel = _scene()
    sim = _device(hip, pos, vel)
    rng = np.random.default_rng(1)
    g = [rng.standard_normal(sim.array_info(nm)[0]) for nm in ("u", "v", "w")]
    d = rng.standard_normal((len(pos), 3))
    v = sim.sample_grid(*g)
    gb = sim.sample_grid_vjp(d)
    lhs = float(np.sum(v * d))
    rhs = float(sum(np.sum(a * b) for a, b in zip(g, gb)))
    sim.close()
    assert lhs == pytest.approx(rhs, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,alpha", [("full", 0.95), ("dam", 0.95), ("dam", 0.0)])
def test_rollout_vjp_matches_oracle_finite_differences(hip, kind, alpha):
    """Directional derivatives against central differences of the oracle's FP64 forward. The
    full box has no free surface; the dam break exercises the fluid-SDF path (theta)."""
    announce(f"adjoint FD {kind} alpha={alpha}")
    steps = 3
    pos, vel = _scene(kind=kind)
    rng = np.random.default_rng(7)
    cx, cv = rng.standard_normal(pos.shape), rng.standard_normal(pos.shape)
    dp0, dv0, dg = _vjp(hip, pos, vel, cx, cv, steps, every=2, alpha=alpha)
    # directional derivatives along random velocity and position perturbations, and gravity
    # small steps: the discrete structure (marking, masks, branches) must not change under them
    for which, grad, eps in (("vel", dv0, 1e-6), ("pos", dp0, 1e-8)):
        d = rng.standard_normal(pos.shape)
        if which == "pos":
            d *= (pos > 2e-6) & (pos < 1 - 2e-6)  # stay inside the domain
        fp = _oracle_loss(pos + eps * d if which == "pos" else pos, vel + eps * d if which == "vel" else vel,
                          cx, cv, steps, alpha=alpha)
        fm = _oracle_loss(pos - eps * d if which == "pos" else pos, vel - eps * d if which == "vel" else vel,
                          cx, cv, steps, alpha=alpha)
        fd = (fp - fm) / (2 * eps)
        an = float(np.sum(grad * d))
        assert an == pytest.approx(fd, rel=1e-4, abs=1e-6 * max(1.0, abs(fd))), (which, an, fd)
    if kind == "full":
        return  # in a closed full box the projection cancels gravity: dL/dg is solver-tolerance noise
    eps = 1e-6
    fd = (_oracle_loss(pos, vel, cx, cv, steps, gravity=9.8 + eps, alpha=alpha) -
          _oracle_loss(pos, vel, cx, cv, steps, gravity=9.8 - eps, alpha=alpha)) / (2 * eps)
    # G = (0, -g, 0): dL/dg = -dL/dG_y
    assert -dg[1] == pytest.approx(fd, rel=1e-4, abs=1e-8), (dg, fd)


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_the_gradient(hip):
    announce("adjoint checkpoint invariance")
    pos, vel = _scene(5)
    rng = np.random.default_rng(2)
    cx, cv = rng.standard_normal(pos.shape), rng.standard_normal(pos.shape)
    a = _vjp(hip, pos, vel, cx, cv, 4, every=1)
    b = _vjp(hip, pos, vel, cx, cv, 4, every=3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
def test_mixed_precision_gradient_close_to_fp64(hip):
    announce("adjoint mixed vs fp64")
    pos, vel = _scene(6)
    rng = np.random.default_rng(4)
    cx, cv = rng.standard_normal(pos.shape), rng.standard_normal(pos.shape)
    a = _vjp(hip, pos, vel, cx, cv, 3, every=2)
    b = _vjp(hip, pos, vel, cx, cv, 3, every=2, precision=hip.PRECISION_MIXED)
    for x, y in zip(a[:2], b[:2]):
        assert np.linalg.norm(x - y) <= 1e-3 * np.linalg.norm(x)
    assert b[2][1] == pytest.approx(a[2][1], rel=1e-3)


@pytest.mark.gpu
def test_torch_autograd_binding(hip):
    announce("adjoint torch autograd")
    import torch

    from simcraft_amd.autograd import flip_rollout, grid_to_particles

    pos, vel = _scene(8)
    sim = _device(hip, pos, vel)
    rng = np.random.default_rng(9)
    grids = [torch.tensor(rng.standard_normal(sim.array_info(nm)[0]) * 0.1, requires_grad=True) for nm in "uvw"]
    v0 = grid_to_particles(sim, *grids)
    grav = torch.tensor([0.0, -9.8, 0.0], dtype=torch.float64, requires_grad=True)
    target = torch.tensor([0.5, 0.2, 0.5], dtype=torch.float64)
    p, v = flip_rollout(sim, torch.from_numpy(pos), v0, grav, [DT, DT], checkpoint_every=1)
    loss = ((p - target) ** 2).sum(dim=1).mean()
    loss.backward()
    # the same gradient through the ABI directly
    dpos = (2.0 * (p.detach().numpy() - target.numpy())) / len(pos)
    sim2 = _device(hip, pos, v0.detach().numpy())
    sim2.rollout_begin(1)
    sim2.substep(DT)
    sim2.substep(DT)
    _, dv0, dg = sim2.rollout_vjp(dpos, None)
    gu = sim2.sample_grid_vjp(dv0, pos=pos)
    sim2.close()
    sim.close()
    np.testing.assert_allclose(grav.grad.numpy(), dg, rtol=1e-12, atol=1e-15)
    for g, ref in zip(grids, gu):
        np.testing.assert_allclose(g.grad.numpy(), ref, rtol=1e-12, atol=1e-15)
