"""FLIP step throughput on one MI355X.

Metric: FLIP steps/sec at 256^3 (MIC-PCG to 1e-6). One step = one fs_substep of the dam break
(fluid block x < 0.4, y < 0.6, 8 jittered particles per fluid cell, at rest) with the static
sphere collider (r 0.15 at (0.7, 0.3, 0.5)), 256^3 cells, dt 5e-4. Inputs are resident in HBM
before the timed region; the timed region is K substeps between two device synchronisations,
run without per-launch timing events.

The JSON line carries:
  roofline      the kernel class with the largest device time over T further substeps
                (--timing-steps) run with every launch bracketed by HIP events on the library's
                stream: algorithmic bytes per launch / mean launch duration, against 8 TB/s HBM.
                (The events, and the synchronisations that drain them, cost several percent of
                a step, so they stay out of the timed region.)
  cpu_baseline  the FP64 CPU oracle (a restatement of the reference, which does not build) on
                a bounded sample: one substep of the same scene at 160^3, scaled by the
                cell/particle count and the device's PCG iteration count.
  config.pcg_*  how many timed solves stopped without meeting the stopping rule (FS_NOT_CONVERGED
                at the iteration cap) and the largest final FP64 residuals.
  other_stopping_rule  the same workload timed under the other PCG stopping rule: the
                reference's ||r||_inf < 1e-6 with its cap of 300 (--tol-mode abs, the default
                headline), or ||r||_2 <= 1e-6 ||rhs||_2 with a cap of 1000 (--tol-mode rel).
  fp64_step     the same workload with FP64 arithmetic throughout, and the mixed/FP64 ratio.
  other_preconditioner  the same workload with the other preconditioner (multigrid for the
                MIC(0) headline, --precond mg for the reverse) and its speedup over the timed run.

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

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_config(simcraft_amd, wl: dict, tol_mode: int, profile: int):
    n = wl["n"]
    prec = simcraft_amd.PRECISION_MIXED if wl["precision"] == "mixed" else simcraft_amd.PRECISION_FP64
    pre = simcraft_amd.PRECOND_MULTIGRID if wl.get("precond", "mic") == "mg" else simcraft_amd.PRECOND_MIC
    return simcraft_amd.default_config(nx=n, ny=n, nz=n, flip_alpha=wl["flip_alpha"], precision=prec,
                                       tol_mode=tol_mode, cg_max_iterations=wl.get("max_iters", 300),
                                       profile_stages=profile, preconditioner=pre)


REL_MAX_ITERS = 1000


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def solve_summary(stats: list) -> dict:
    """Convergence record of the timed substeps: how many solves stopped without meeting the
    stopping rule, and the largest final residuals (FP64, both norms)."""
    return {
        "pcg_not_converged_steps": int(sum(1 for s in stats if s["status"] != 0)),
        "pcg_max_residual_inf": float(max(s["residual_inf"] for s in stats)),
        "pcg_max_residual_rel_l2": float(max(s["residual_rel"] for s in stats)),
    }


def timed_run(simcraft_amd, scenes, wl: dict, tol_mode: str, steps: int, warmup: int) -> dict:
    """Steps/s of `steps` substeps of the workload (after `warmup`) under one stopping rule:
    "abs" = the reference's ||r||_inf < 1e-6 with its cap of 300 iterations, "rel" =
    ||r||_2 <= 1e-6 ||rhs||_2 with a cap of REL_MAX_ITERS."""
    wl = dict(wl, max_iters=300 if tol_mode == "abs" else REL_MAX_ITERS)
    tm = simcraft_amd.TOL_REL_L2 if tol_mode == "rel" else simcraft_amd.TOL_ABS_LINF
    sim = simcraft_amd.FluidSim(workload_config(simcraft_amd, wl, tm, 0))
    pos, vel = scenes.dam_break(wl["n"])
    sim.set_collider_sdf(scenes.collider_for(wl["collider"], wl["n"]))
    sim.set_particles(pos, vel)
    del pos, vel
    for _ in range(warmup):
        sim.substep(wl["dt"])
    sim.synchronize()
    t0 = time.perf_counter()
    stats = [sim.substep(wl["dt"]) for _ in range(steps)]
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    sim.close()
    return {"value": steps / elapsed, "unit": "steps/s", "steps": steps, "warmup": warmup,
            "ms_per_step": 1000.0 * elapsed / steps, "precision": wl["precision"],
            "preconditioner": "multigrid" if wl.get("precond", "mic") == "mg" else "mic",
            "tol_mode": "abs_linf" if tol_mode == "abs" else "rel_l2", "pcg_max_iterations": wl["max_iters"],
            "pcg_iterations_mean": float(np.mean([s["cg_iterations"] for s in stats])), **solve_summary(stats)}


def cpu_baseline(simcraft_amd, scenes, wl: dict, tol_mode: int, gpu_iters: float, cap: int, sample_n: int) -> dict:
    """Time the oracle on one substep of the same scene at sample_n^3 (PCG capped at `cap`
    iterations) and scale by the cell and particle count to the workload's n^3."""
    from oracle.binding import DOT_SERIAL, P2G_PARALLEL, OracleSim

    n_full = wl["n"]
    n = min(sample_n, n_full)
    scale = (n_full / n) ** 3
    cfg = workload_config(simcraft_amd, dict(wl, n=n), tol_mode, 0)
    cfg.cg_max_iterations = cap + 1  # `cap` iterations of the loop body
    pos, vel = scenes.dam_break(n)
    o = OracleSim(cfg, dot_mode=DOT_SERIAL, p2g_order=P2G_PARALLEL)
    o.set_collider_sdf(scenes.collider_for(wl["collider"], n))
    o.set_particles(pos, vel)
    del pos, vel
    dt, h = wl["dt"], 1.0 / n
    stages = [("clear", lambda: o.clear()), ("advect", lambda: o.advect(dt)),
              ("reconstruct", lambda: o.reconstruct(0.8 * h)),
              ("extrapolate_fluid_sdf", lambda: o.extrapolate_fluid_sdf(10)),
              ("smooth_fluid_surface", lambda: o.smooth_fluid_surface(5)), ("p2g", lambda: o.p2g(dt)),
              ("dirichlet", lambda: o.apply_dirichlet_boundary()),
              ("extrapolate_velocity", lambda: o.extrapolate_velocity(10)), ("force", lambda: o.apply_force(dt)),
              ("build_system", lambda: o.build_system(dt)), ("solve_pressure", None),
              ("project", lambda: o.project(dt)), ("collider", lambda: o.apply_collider()),
              ("g2p", lambda: o.g2p(dt))]
    times = {}
    pcg_iters = 0
    for name, fn in stages:
        t0 = time.perf_counter()
        if fn is None:
            st = o.solve_pressure()
            pcg_iters = st["cg_iterations"]
        else:
            fn()
        times[name] = time.perf_counter() - t0
    o.close()
    t_pcg = times.pop("solve_pressure")
    per_iter = t_pcg / max(pcg_iters, 1)
    step_s = scale * (sum(times.values()) + per_iter * gpu_iters)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / step_s,
        "unit": "steps/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"one substep of the same scene at {n}^3 (cells and particles 1/{scale:g} of {n_full}^3), "
                   f"timed per stage and scaled x{scale:g}; its PCG ran {pcg_iters} iterations (cap {cap}, "
                   f"{t_pcg:.2f}s) and is scaled to the device's mean {gpu_iters:.1f} iterations at {n_full}^3; "
                   f"P2G with the reference's parallel particle loop"),
        "seconds_per_step": step_s,
        "stage_seconds_sample": {k: round(v, 3) for k, v in times.items()},
    }


def adjoint_ratio(simcraft_amd, scenes, n: int, steps: int, every: int, precision: str, x_min: float) -> dict:
    """Forward rollout vs its reverse pass: dam break at n^3, `steps` substeps of dt 1e-3,
    loss = kinetic energy of the particles with x > x_min at the end; gradients with respect to
    the initial particle velocities and the gravity vector. The reverse time includes the
    recomputation of every checkpoint segment. (After 40 substeps of 1e-3 s the front of the
    x < 0.4 block has not reached x = 0.7, so a region starting there would hold no particle
    and give a zero gradient and a trivial adjoint solve; the default region starts at 0.3.)"""
    prec = simcraft_amd.PRECISION_MIXED if precision == "mixed" else simcraft_amd.PRECISION_FP64
    cfg = simcraft_amd.default_config(nx=n, ny=n, nz=n, flip_alpha=0.95, precision=prec)
    sim = simcraft_amd.FluidSim(cfg)
    sim.set_collider_sdf(scenes.no_collider(n))
    pos, vel = scenes.dam_break(n)
    sim.set_particles(pos, vel)
    dt = 1e-3
    sim.substep(dt)  # warm-up outside the rollout
    p_start, v_start = sim.particles()
    sim.set_particles(p_start, v_start)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.rollout_begin(every)
    for _ in range(steps):
        sim.substep(dt)
    sim.synchronize()
    t_fwd = time.perf_counter() - t0
    p, v = sim.particles()
    mask = (p[:, 0] > x_min)[:, None]
    dvel = np.where(mask, v, 0.0)  # d(0.5 |v|^2)/dv on the region
    t0 = time.perf_counter()
    _, dv0, dg = sim.rollout_vjp(None, dvel)
    sim.synchronize()
    t_bwd = time.perf_counter() - t0
    sim.close()
    return {"workload": f"dam_break_{n}_{precision}_rollout", "steps": steps, "checkpoint_every": every,
            "loss": f"kinetic energy of particles with x > {x_min}", "particles_in_region": int(mask.sum()),
            "particles": int(len(pos)), "forward_s": round(t_fwd, 3), "reverse_s": round(t_bwd, 3),
            "ratio": round(t_bwd / t_fwd, 3), "dL_dgravity": [float(x) for x in dg],
            "dL_dv0_norm": float(np.linalg.norm(dv0))}


def copy_calibration() -> int:
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatches,
    divided by the bytes it must move, scale the same counters read for the sweep kernels."""
    import torch

    n = 1 << 27  # 128 Mi floats = 512 MiB each way
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": 4 * n, "dispatches": 5}), flush=True)
    return 0


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary (two rocprofv3 --pmc runs,
    FETCH_SIZE and WRITE_SIZE, each scaled by the same counter for a plain copy), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(workload, {}).get(kernel)
    return None if e is None else e.get("bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dam_break_sphere_256_mixed")
    ap.add_argument("--tol-mode", choices=["abs", "rel"], default="abs")
    ap.add_argument("--precond", choices=["mic", "mg"], default="mic",
                    help="preconditioner of the timed run: the reference's MIC(0) (the headline) or multigrid")
    ap.add_argument("--max-iters", type=int, default=None,
                    help="PCG iteration cap (default: the reference's 300 for abs, 1000 for rel)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the secondary lines (the other stopping rule, the FP64 step)")
    ap.add_argument("--secondary-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-pcg-iters", type=int, default=300)
    ap.add_argument("--cpu-sample-n", type=int, default=160)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-n", type=int, default=128)
    ap.add_argument("--adjoint-steps", type=int, default=40)
    ap.add_argument("--adjoint-every", type=int, default=8)
    ap.add_argument("--adjoint-precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--adjoint-x-min", type=float, default=0.3)
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=3,
                    help="substeps after the timed region with per-launch event timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 512 MiB five times (scale for the PMC byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the path's sharding is designed, not built)"}))
        return 2

    import simcraft_amd
    from simcraft_amd import scenes

    wl = dict(scenes.WORKLOADS[args.workload], precond=args.precond)
    tol_mode = simcraft_amd.TOL_REL_L2 if args.tol_mode == "rel" else simcraft_amd.TOL_ABS_LINF
    log(f"workload {args.workload}: {wl}")
    t0 = time.perf_counter()
    if args.max_iters is not None:
        wl["max_iters"] = args.max_iters
    elif args.tol_mode == "rel":
        wl["max_iters"] = REL_MAX_ITERS
    cfg = workload_config(simcraft_amd, wl, tol_mode, 0 if args.no_kernel_timing else 2)
    sim = simcraft_amd.FluidSim(cfg)
    pos, vel = scenes.dam_break(wl["n"])
    npart = len(pos)
    sim.set_collider_sdf(scenes.collider_for(wl["collider"], wl["n"]))
    sim.set_particles(pos, vel)
    del pos, vel
    log(f"setup {time.perf_counter() - t0:.1f}s, {npart} particles")
    dt = wl["dt"]
    for w in range(args.warmup):
        st = sim.substep(dt)
        log(f"warmup {w}: {st['cg_iterations']} PCG iterations, residual {st['residual_inf']:.3g}")
    sim.synchronize()
    if not args.no_kernel_timing:
        sim.set_kernel_timing(False)
    stats = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        stats.append(sim.substep(dt))
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    iters = [st["cg_iterations"] for st in stats]
    active = [st["n_active"] for st in stats]
    log(f"timed {args.steps} steps in {elapsed:.3f}s; PCG iterations {iters}")
    sps = args.steps / elapsed

    roofline = None
    kernels = {}
    if not args.no_kernel_timing and args.timing_steps > 0:
        sim.set_kernel_timing(True)
        sim.kernel_profile_reset()
        for s in range(args.timing_steps):
            sim.substep(dt)
        sim.synchronize()
        prof = sim.kernel_profile()
        for name, p in prof.items():
            if p["launches"]:
                avg_ms = p["ms"] / p["launches"]
                gbs = (p["bytes"] / p["launches"]) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
                kernels[name] = {"launches": p["launches"], "ms_total": round(p["ms"], 3),
                                 "us_per_launch": round(1e3 * avg_ms, 2),
                                 "bytes_per_launch": p["bytes"] / p["launches"], "GBps": round(gbs, 1)}
        if kernels:
            dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
            kd = kernels[dom]
            roofline = {"bound": "hbm", "kernel": dom, "achieved": kd["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(kd["GBps"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(args.workload, dom),
                        "us_per_launch": kd["us_per_launch"], "bytes_per_launch": kd["bytes_per_launch"]}
            log("kernels: " + json.dumps(kernels))
    out = {
        "metric": ("FLIP steps/sec at 256^3 (MIC-PCG to 1e-6) + adjoint/forward time ratio" if args.precond == "mic"
                   else "FLIP steps/sec at 256^3 (multigrid-PCG to 1e-6)"),
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "kernel_timing_steps": 0 if args.no_kernel_timing else args.timing_steps,
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
            "active_cells_mean": float(np.mean(active)),
            "dt": dt,
            "collider": wl["collider"],
            "precision": wl["precision"],
            "pcg_tolerance": 1e-6,
            "tol_mode": "abs_linf" if args.tol_mode == "abs" else "rel_l2",
            "pcg_max_iterations": wl.get("max_iters", 300),
            "preconditioner": "mic" if args.precond == "mic" else "multigrid",
            "pcg_iterations_mean": float(np.mean(iters)),
            **solve_summary(stats),
        },
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": None,
        "adjoint": None,
        "other_stopping_rule": None,
        "fp64_step": None,
        "other_preconditioner": None,
    }
    sim.close()
    del sim
    if not args.no_secondary:
        other = "rel" if args.tol_mode == "abs" else "abs"
        log(f"secondary: the same workload with the {other} stopping rule ...")
        out["other_stopping_rule"] = timed_run(simcraft_amd, scenes, wl, other, args.secondary_steps, args.warmup)
        log(f"secondary: {out['other_stopping_rule']}")
        if wl["precision"] == "mixed":
            wl64 = dict(wl, precision="fp64")
            log("secondary: the FP64 step of the same workload ...")
            r64 = timed_run(simcraft_amd, scenes, wl64, args.tol_mode, args.secondary_steps, args.warmup)
            r64["mixed_over_fp64"] = round(sps / r64["value"], 3)
            out["fp64_step"] = r64
            log(f"secondary: {r64}")
        other_pc = "mg" if args.precond == "mic" else "mic"
        log(f"secondary: the same workload with the {other_pc} preconditioner ...")
        rpc = timed_run(simcraft_amd, scenes, dict(wl, precond=other_pc), args.tol_mode, args.secondary_steps,
                        args.warmup)
        rpc["speedup"] = round(rpc["value"] / sps, 3)
        out["other_preconditioner"] = rpc
        log(f"secondary: {rpc}")
    if not args.no_adjoint:
        log("adjoint rollout ...")
        out["adjoint"] = adjoint_ratio(simcraft_amd, scenes, args.adjoint_n, args.adjoint_steps, args.adjoint_every,
                                       args.adjoint_precision, args.adjoint_x_min)
        log(f"adjoint: {out['adjoint']}")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(simcraft_amd, scenes, wl, tol_mode, float(np.mean(iters)),
                                           args.cpu_pcg_iters, args.cpu_sample_n)
        log(f"cpu baseline took {time.perf_counter() - t0:.1f}s: {out['cpu_baseline']['value']:.5f} steps/s")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
