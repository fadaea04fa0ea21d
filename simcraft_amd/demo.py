"""Headless run of the reference's demo application (FluidSim/disabled/fluid-sim.cc:175-250) on the
device: a triangle-mesh collider from an OBJ file (or a built-in sphere), scaled into the domain as
relocateMesh does (:67-84), the dam-break fluid block, CFL-limited frames (FluidSimulator::step),
and the fluid surface exported as OBJ files for a range of frames (reconstruct, smooth 5 times,
rebuildSurface, exportObj; the app's window and particle rendering are left out).

    python -m simcraft_amd.demo [-r 64] [-c collider.obj] [--frames 140] [--export-from 80]
                                [--out DIR] [--precision mixed|fp64] [--precond mic|mg]
                                [--advection pic|flip|apic]
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

from . import (ADVECT_APIC, ADVECT_FLIP, ADVECT_PIC, PRECISION_FP64, PRECISION_MIXED, PRECOND_MIC,
               PRECOND_MULTIGRID, FluidSim, default_config, meshio, scenes)


def relocate_mesh(mesh: meshio.Mesh, size: float) -> meshio.Mesh:
    """relocateMesh (fluid-sim.cc:67-84): shift the mesh to the origin by its smallest coordinate,
    scale its largest extent to the domain size, then halve it."""
    v = np.asarray(mesh.vertices, dtype=np.float64)
    lo, hi = float(v.min()), float(v.max())
    scale = size / (hi - lo)
    v = (v - lo) * scale * 0.5
    return meshio.Mesh(vertices=v, indices=np.asarray(mesh.indices, dtype=np.int64), normals=mesh.normals)


def run(n: int = 64, frames: int = 140, export_from: int = 80, export_to: int | None = None,
        collider: str | None = None, out: str | Path = ".", precision: str = "mixed", precond: str = "mic",
        advection: str = "pic", frame_dt: float = 1.0 / 60.0, log=print) -> list[Path]:
    """Runs `frames` frames at n^3 and writes fluid_<frame>.obj for frames in [export_from, export_to);
    returns the paths written."""
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    export_to = frames if export_to is None else export_to
    mesh = meshio.load_obj(collider) if collider else meshio.icosphere(subdivisions=3)
    mesh = relocate_mesh(mesh, 1.0)
    adv = {"pic": ADVECT_PIC, "flip": ADVECT_FLIP, "apic": ADVECT_APIC}[advection]
    cfg = default_config(nx=n, ny=n, nz=n, advection=adv,
                         precision=PRECISION_MIXED if precision == "mixed" else PRECISION_FP64,
                         preconditioner=PRECOND_MULTIGRID if precond == "mg" else PRECOND_MIC)
    written = []
    with FluidSim(cfg) as sim:
        sim.set_collider_mesh(mesh.vertices, mesh.indices)
        pos, vel = scenes.dam_break(n)
        sim.set_particles(pos, vel)
        h = 1.0 / n
        for frame in range(frames):
            t0 = time.perf_counter()
            st = sim.step_frame(frame_dt)
            msg = f"frame {frame}: {st['substeps']} substeps, {time.perf_counter() - t0:.3f}s"
            if export_from <= frame < export_to:
                sim.reconstruct(0.8 * h)
                sim.smooth_fluid_surface(5)
                v, nr, tri = sim.rebuild_surface()
                path = out / f"fluid_{frame:04d}.obj"
                meshio.save_obj(path, meshio.Mesh(vertices=v, indices=tri, normals=nr))
                written.append(path)
                msg += f", surface {len(v)} vertices / {len(tri)} triangles -> {path}"
            log(msg)
    return written


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("-r", "--resolution", type=int, default=64)
    ap.add_argument("-c", "--collider", default=None, help="OBJ file of the collider (default: a sphere)")
    ap.add_argument("--frames", type=int, default=140)
    ap.add_argument("--export-from", type=int, default=80)
    ap.add_argument("--export-to", type=int, default=None)
    ap.add_argument("--out", default=".")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--precond", choices=["mic", "mg"], default="mic")
    ap.add_argument("--advection", choices=["pic", "flip", "apic"], default="pic")
    a = ap.parse_args(argv)
    run(a.resolution, a.frames, a.export_from, a.export_to, a.collider, a.out, a.precision, a.precond, a.advection)
    return 0


if __name__ == "__main__":
    sys.exit(main())
