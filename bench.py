"""FLIP step throughput on one MI355X.

Metric: FLIP steps/sec at 256^3 (MIC-PCG to 1e-6). One step = one fs_substep of the dam break
with the static sphere collider (256^3 cells, 8 particles per fluid cell, dt 5e-4), inputs
resident in HBM before the timed region. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_sim(simcraft_amd, scenes, wl: dict, tol_mode: int):
    n = wl["n"]
    prec = simcraft_amd.PRECISION_MIXED if wl["precision"] == "mixed" else simcraft_amd.PRECISION_FP64
    cfg = simcraft_amd.default_config(nx=n, ny=n, nz=n, flip_alpha=wl["flip_alpha"], precision=prec,
                                      tol_mode=tol_mode, cg_max_iterations=wl.get("max_iters", 300))
    sim = simcraft_amd.FluidSim(cfg)
    pos, vel = scenes.dam_break(n)
    sim.set_collider_sdf(scenes.collider_for(wl["collider"], n))
    sim.set_particles(pos, vel)
    return sim, len(pos)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dam_break_sphere_256_fp64")
    ap.add_argument("--tol-mode", choices=["abs", "rel"], default="abs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is described in DESIGN.md"}))
        return 2

    import simcraft_amd
    from simcraft_amd import scenes

    wl = dict(scenes.WORKLOADS[args.workload])
    tol_mode = simcraft_amd.TOL_REL_L2 if args.tol_mode == "rel" else simcraft_amd.TOL_ABS_LINF
    log(f"workload {args.workload}: {wl}")
    t0 = time.perf_counter()
    sim, npart = make_sim(simcraft_amd, scenes, wl, tol_mode)
    log(f"setup {time.perf_counter() - t0:.1f}s, {npart} particles")
    dt = wl["dt"]
    for w in range(args.warmup):
        st = sim.substep(dt)
        log(f"warmup {w}: {st['cg_iterations']} PCG iterations, residual {st['residual_inf']:.3g}")
    sim.synchronize()
    iters = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        st = sim.substep(dt)
        iters.append(st["cg_iterations"])
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f}s; PCG iterations {iters}")
    sps = args.steps / elapsed
    out = {
        "metric": "FLIP steps/sec at 256^3 (MIC-PCG to 1e-6)",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if wl["precision"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "grid": wl["n"],
            "particles": npart,
            "dt": dt,
            "collider": wl["collider"],
            "pcg_tolerance": 1e-6,
            "tol_mode": args.tol_mode,
            "pcg_iterations_mean": float(np.mean(iters)),
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
