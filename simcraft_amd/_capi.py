"""ctypes binding of libsimcraft_fluid.so (include/simcraft_fluid.h).

The structures below mirror fs_config / fs_stats field by field; natural C alignment is the
same under ctypes.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsimcraft_fluid.so"

FS_OK, FS_NOT_CONVERGED, FS_BREAKDOWN, FS_NONFINITE, FS_BAD_ARG, FS_DEVICE_ERROR, FS_UNSUPPORTED, FS_NO_DEVICE = range(8)
STATUS_NAMES = {
    0: "FS_OK", 1: "FS_NOT_CONVERGED", 2: "FS_BREAKDOWN", 3: "FS_NONFINITE", 4: "FS_BAD_ARG",
    5: "FS_DEVICE_ERROR", 6: "FS_UNSUPPORTED", 7: "FS_NO_DEVICE",
}
ADVECT_PIC, ADVECT_FLIP, ADVECT_APIC = 0, 1, 2
PROJECT_FVM, PROJECT_FDM = 0, 1
PRECOND_MULTIGRID, PRECOND_MIC = 0, 1
SOLVER_CG = 0
RECON_NAIVE = 0
PRECISION_FP64, PRECISION_MIXED = 0, 1
TOL_ABS_LINF, TOL_REL_L2 = 0, 1
MESH_EXACT_DISTANCE = 1


class fs_config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("size_x", C.c_double), ("size_y", C.c_double), ("size_z", C.c_double),
        ("precision", C.c_int32), ("advection", C.c_int32),
        ("flip_alpha", C.c_double), ("mic_tau", C.c_double), ("mic_sigma", C.c_double),
        ("cg_tolerance", C.c_double), ("cg_max_iterations", C.c_int32), ("tol_mode", C.c_int32),
        ("gravity", C.c_double),
        ("sdf_extrapolate_iters", C.c_int32), ("sdf_smooth_iters", C.c_int32), ("vel_extrapolate_iters", C.c_int32),
        ("recon_radius_factor", C.c_double),
        ("advect_on_cleared_grid", C.c_int32),
        ("cfl_factor", C.c_double),
        ("inner_tolerance", C.c_double),
        ("max_refinements", C.c_int32), ("check_interval", C.c_int32), ("profile_stages", C.c_int32),
        ("preconditioner", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class fs_stats(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("cg_iterations", C.c_int32), ("refinements", C.c_int32), ("n_active", C.c_int32),
        ("residual_inf", C.c_double), ("residual_rel", C.c_double), ("rhs_norm2", C.c_double),
        ("substeps", C.c_int32), ("pad0", C.c_int32),
        ("ms_advect", C.c_double), ("ms_reconstruct", C.c_double), ("ms_p2g", C.c_double),
        ("ms_extrapolate", C.c_double), ("ms_build", C.c_double), ("ms_pcg", C.c_double),
        ("ms_project", C.c_double), ("ms_g2p", C.c_double), ("ms_total", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad0"}


_P = C.c_void_p
_D = C.POINTER(C.c_double)
# name -> (restype, argtypes), exactly the declarations of include/simcraft_fluid.h
SIGNATURES = {
    "fs_config_default": (None, [C.POINTER(fs_config)]),
    "fs_create": (C.c_int, [C.POINTER(fs_config), C.POINTER(_P)]),
    "fs_destroy": (None, [_P]),
    "fs_last_error": (C.c_char_p, [_P]),
    "fs_device_count": (C.c_int, []),
    "fs_set_collider_sdf": (C.c_int, [_P, _D]),
    "fs_set_particles": (C.c_int, [_P, _D, _D, C.c_int64]),
    "fs_get_particles": (C.c_int, [_P, _D, _D]),
    "fs_num_particles": (C.c_int64, [_P]),
    "fs_set_advector": (C.c_int, [_P, C.c_int32]),
    "fs_set_projector": (C.c_int, [_P, C.c_int32]),
    "fs_set_compressed_solver": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "fs_set_reconstructor": (C.c_int, [_P, C.c_int32]),
    "fs_substep": (C.c_int, [_P, C.c_double, C.POINTER(fs_stats)]),
    "fs_step_frame": (C.c_int, [_P, C.c_double, C.POINTER(fs_stats)]),
    "fs_cfl": (C.c_int, [_P, _D]),
    "fs_get_particle_affine": (C.c_int, [_P, _D]),
    "fs_set_particle_affine": (C.c_int, [_P, _D]),
    "fs_rebuild_surface": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fs_get_surface": (C.c_int, [_P, _D, _D, C.POINTER(C.c_uint32)]),
    "fs_set_collider_mesh": (C.c_int, [_P, _D, C.c_int64, C.POINTER(C.c_uint32), C.c_int64, C.c_int32, C.c_int32]),
    "fs_clear": (C.c_int, [_P]),
    "fs_advect": (C.c_int, [_P, C.c_double]),
    "fs_reconstruct": (C.c_int, [_P, C.c_double]),
    "fs_extrapolate_fluid_sdf": (C.c_int, [_P, C.c_int32]),
    "fs_smooth_fluid_surface": (C.c_int, [_P, C.c_int32]),
    "fs_p2g": (C.c_int, [_P, C.c_double]),
    "fs_apply_dirichlet_boundary": (C.c_int, [_P]),
    "fs_extrapolate_velocity": (C.c_int, [_P, C.c_int32]),
    "fs_apply_force": (C.c_int, [_P, C.c_double]),
    "fs_build_system": (C.c_int, [_P, C.c_double]),
    "fs_solve_pressure": (C.c_int, [_P, C.POINTER(fs_stats)]),
    "fs_project": (C.c_int, [_P, C.c_double]),
    "fs_apply_collider": (C.c_int, [_P]),
    "fs_g2p": (C.c_int, [_P, C.c_double]),
    "fs_spmv": (C.c_int, [_P, _D, _D]),
    "fs_mic_apply": (C.c_int, [_P, _D, _D]),
    "fs_mg_apply": (C.c_int, [_P, _D, _D]),
    "fs_array_info": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "fs_get_array": (C.c_int, [_P, C.c_char_p, _P]),
    "fs_set_array": (C.c_int, [_P, C.c_char_p, _P]),
    "fs_synchronize": (C.c_int, [_P]),
    "fs_kernel_profile": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "fs_kernel_profile_reset": (C.c_int, [_P]),
    "fs_set_kernel_timing": (C.c_int, [_P, C.c_int32]),
    "fs_rollout_begin": (C.c_int, [_P, C.c_int32]),
    "fs_rollout_stop": (C.c_int, [_P]),
    "fs_rollout_end": (C.c_int, [_P]),
    "fs_rollout_steps": (C.c_int, [_P]),
    "fs_rollout_vjp": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "fs_sample_grid": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "fs_sample_grid_vjp": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "fs_cfl_device_noalloc": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, _P,
                                        _P]),
}

KERNEL_NAMES = ["spmv", "mic_fwd", "mic_bwd", "pcg_update", "mic_factor", "p2g", "g2p", "advect", "reconstruct",
                "mg_level0", "mg_coarse"]
# HIP kernel symbols behind each class (for matching rocprofv3 traces)
KERNEL_SYMBOLS = {
    "spmv": ["k_spmv", "k_mg_spmv"], "mic_fwd": ["k_fwd"], "mic_bwd": ["k_bwd"], "pcg_update": ["k_update", "k_mg_update"],
    "mic_factor": ["k_factor_sweep"], "p2g": ["k_p2g_records", "k_p2g_gather", "k_keys_nearest", "k_count",
                                             "k_scatter", "k_sort_bins"],
    "g2p": ["k_g2p"], "advect": ["k_advect"], "reconstruct": ["k_keys_morton", "k_bin_last", "k_recon_splat",
                                                               "k_recon_final"],
    "mg_level0": ["k_mg_pre", "k_mg_restrict", "k_mg_post1", "k_mg_post2"], "mg_coarse": [],
}

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load the HIP library; raises if it has not been built (there is no CPU fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def default_config(**overrides) -> fs_config:
    cfg = fs_config()
    load().fs_config_default(C.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise AttributeError(f"fs_config has no field {k}")
        setattr(cfg, k, v)
    return cfg
