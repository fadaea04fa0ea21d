"""torch.autograd bindings of the reverse mode (fs_rollout_vjp, fs_sample_grid_vjp).

The simulation state lives on the device inside a FluidSim; these Functions move particle and
grid data through host float64 tensors, so they suit moderate particle counts and act as the
differentiable interface, not as the fast path of a forward-only run.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

from . import FluidSim

_ROLLOUT_IDS = itertools.count(1)


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class FlipRollout(torch.autograd.Function):
    """(positions0, velocities0, gravity) -> (positions, velocities) after `len(dts)` substeps.

    gravity is the vector G = (0, -g, 0) the simulator was configured with (the reference applies
    gravity to v only); it is an input so that dL/dG is returned. Checkpoints are taken every
    `checkpoint_every` substeps and the backward pass recomputes one segment at a time.

    Recording stops when forward returns, so substeps run on `sim` between forward and backward
    are not part of the differentiated rollout. One sim holds one tape: a later flip_rollout on
    the same sim replaces it, and backward of the earlier one then raises instead of returning
    the gradient of a different rollout.
    """

    @staticmethod
    def forward(ctx, pos0, vel0, gravity, sim: FluidSim, dts, checkpoint_every: int = 8):
        g = _np(gravity).reshape(3)
        if g[0] != 0.0 or g[2] != 0.0 or g[1] != -sim.cfg.gravity:
            raise ValueError(f"gravity must be (0, {-sim.cfg.gravity}, 0), the simulator's configuration")
        sim.set_particles(_np(pos0), _np(vel0))
        sim.rollout_begin(checkpoint_every)
        for dt in dts:
            sim.substep(float(dt))
        sim.rollout_stop()
        pos, vel = sim.particles()
        ctx.sim = sim
        ctx.nsteps = len(dts)
        ctx.rollout_id = next(_ROLLOUT_IDS)
        sim._rollout_id = ctx.rollout_id
        return torch.from_numpy(pos), torch.from_numpy(vel)

    @staticmethod
    def backward(ctx, gpos, gvel):
        sim = ctx.sim
        if getattr(sim, "_rollout_id", None) != ctx.rollout_id or sim.rollout_steps() != ctx.nsteps:
            raise RuntimeError("the simulator's recorded rollout is no longer the one this output came from "
                               "(another rollout was recorded on it since)")
        dp0, dv0, dg = sim.rollout_vjp(None if gpos is None else _np(gpos), None if gvel is None else _np(gvel))
        return torch.from_numpy(dp0), torch.from_numpy(dv0), torch.from_numpy(dg), None, None, None


class GridToParticles(torch.autograd.Function):
    """PIC transfer of face grids (u, v, w) to particles at `pos` (default: the simulator's
    current particles, read at forward time)."""

    @staticmethod
    def forward(ctx, u, v, w, sim: FluidSim, pos=None):
        ctx.sim = sim
        ctx.pos = sim.positions() if pos is None else _np(pos)
        return torch.from_numpy(sim.sample_grid(_np(u), _np(v), _np(w), pos=ctx.pos))

    @staticmethod
    def backward(ctx, gvel):
        du, dv, dw = ctx.sim.sample_grid_vjp(_np(gvel), pos=ctx.pos)
        return torch.from_numpy(du), torch.from_numpy(dv), torch.from_numpy(dw), None, None


def flip_rollout(sim: FluidSim, pos0, vel0, gravity, dts, checkpoint_every: int = 8):
    return FlipRollout.apply(pos0, vel0, gravity, sim, list(dts), checkpoint_every)


def grid_to_particles(sim: FluidSim, u, v, w, pos=None):
    return GridToParticles.apply(u, v, w, sim, pos)
