"""ctypes binding of the oracle (oracle/libfs_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg. The methods mirror simcraft_amd.FluidSim so that a parity test
drives both through the same calls.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

# ctypes mirrors of the C structs only (no call into the device library)
from simcraft_amd._capi import fs_config, fs_stats

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfs_oracle.so"
DOT_SERIAL, DOT_TREE = 0, 1
P2G_SORTED, P2G_INDEX, P2G_PARALLEL = 0, 1, 2

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_SIGS = {
    "fso_config_default": (None, [C.POINTER(fs_config)]),
    "fso_create": (_P, [C.POINTER(fs_config)]),
    "fso_destroy": (None, [_P]),
    "fso_set_dot_mode": (C.c_int, [_P, C.c_int]),
    "fso_set_p2g_order": (C.c_int, [_P, C.c_int]),
    "fso_set_particles": (C.c_int, [_P, _D, _D, C.c_int64]),
    "fso_get_particles": (C.c_int, [_P, _D, _D]),
    "fso_num_particles": (C.c_int64, [_P]),
    "fso_set_collider_sdf": (C.c_int, [_P, _D]),
    "fso_clear": (C.c_int, [_P]),
    "fso_advect": (C.c_int, [_P, C.c_double]),
    "fso_reconstruct": (C.c_int, [_P, C.c_double]),
    "fso_extrapolate_fluid_sdf": (C.c_int, [_P, C.c_int]),
    "fso_smooth_fluid_surface": (C.c_int, [_P, C.c_int]),
    "fso_p2g": (C.c_int, [_P, C.c_double]),
    "fso_apply_dirichlet_boundary": (C.c_int, [_P]),
    "fso_extrapolate_velocity": (C.c_int, [_P, C.c_int]),
    "fso_apply_force": (C.c_int, [_P, C.c_double]),
    "fso_build_system": (C.c_int, [_P, C.c_double]),
    "fso_solve_pressure": (C.c_int, [_P, C.POINTER(fs_stats)]),
    "fso_project": (C.c_int, [_P, C.c_double]),
    "fso_apply_collider": (C.c_int, [_P]),
    "fso_g2p": (C.c_int, [_P, C.c_double]),
    "fso_substep": (C.c_int, [_P, C.c_double, C.POINTER(fs_stats)]),
    "fso_step_frame": (C.c_int, [_P, C.c_double, C.POINTER(fs_stats)]),
    "fso_cfl": (C.c_int, [_P, _D]),
    "fso_spmv": (C.c_int, [_P, _D, _D]),
    "fso_mic_apply": (C.c_int, [_P, _D, _D]),
    "fso_mg_apply": (C.c_int, [_P, _D, _D]),
    "fso_array_info": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "fso_get_array": (C.c_int, [_P, C.c_char_p, _P]),
    "fso_set_array": (C.c_int, [_P, C.c_char_p, _P]),
    "fso_fraction_inside": (C.c_double, [C.c_double] * 4),
    "fso_morton3d": (C.c_int, [C.c_int] * 3),
    "fso_bspline": (C.c_double, [C.c_double]),
    "fso_sdf_eval": (C.c_double, [_P, _D]),
    "fso_sdf_grad": (None, [_P, _D, _D]),
    "fso_sample_velocity": (None, [_P, _D, _D]),
    "fso_nearest": (None, [_P, C.c_int, _D, C.POINTER(C.c_int)]),
    "fso_last_iterations": (C.c_int, [_P]),
    "fso_set_collider_mesh": (C.c_int, [_P, _D, C.c_int64, C.POINTER(C.c_uint32), C.c_int64, C.c_int, C.c_int]),
    "fso_rebuild_surface": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fso_get_surface": (C.c_int, [_P, _D, _D, C.POINTER(C.c_uint32)]),
    "fso_mc_table": (C.c_int, [C.POINTER(C.c_int8), C.POINTER(C.c_int8)]),
    "fso_get_apic": (C.c_int, [_P, _D]),
    "fso_set_apic": (C.c_int, [_P, _D]),
    "fso_rollout_begin": (C.c_int, [_P]),
    "fso_rollout_stop": (C.c_int, [_P]),
    "fso_rollout_steps": (C.c_int, [_P]),
    "fso_rollout_vjp": (C.c_int, [_P, _D, _D, _D, _D, _D]),
}
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
        for n, (res, args) in _SIGS.items():
            f = getattr(lib, n)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def default_config(**overrides) -> fs_config:
    """The reference's parameters (fso_config_default) with the given overrides."""
    cfg = fs_config()
    load().fso_config_default(C.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise AttributeError(f"fs_config has no field {k}")
        setattr(cfg, k, v)
    return cfg


def fraction_inside(lu, ru, rd, ld) -> float:
    return load().fso_fraction_inside(lu, ru, rd, ld)


def morton3d(x, y, z) -> int:
    return load().fso_morton3d(x, y, z)


def bspline(r) -> float:
    return load().fso_bspline(r)


def mc_table():
    """The marching-cubes case table: triangle counts (256,) and edge triples (256, 12, 3)."""
    ntri = np.zeros(256, np.int8)
    tri = np.zeros((256, 36), np.int8)
    load().fso_mc_table(ntri.ctypes.data_as(C.POINTER(C.c_int8)), tri.ctypes.data_as(C.POINTER(C.c_int8)))
    return ntri, tri.reshape(256, 12, 3)


class OracleSim:
    """CPU FP64 restatement of cpu::FluidSimulator with the same methods as FluidSim."""

    def __init__(self, cfg: fs_config | None = None, dot_mode: int = DOT_TREE, p2g_order: int = P2G_SORTED,
                 **overrides):
        self._lib = load()
        self.cfg = cfg if cfg is not None else default_config()
        for k, v in overrides.items():
            setattr(self.cfg, k, v)
        self._ctx = self._lib.fso_create(C.byref(self.cfg))
        if not self._ctx:
            raise ValueError("bad oracle config")
        self._lib.fso_set_dot_mode(self._ctx, dot_mode)
        self._lib.fso_set_p2g_order(self._ctx, p2g_order)
        self.shape = (self.cfg.nx, self.cfg.ny, self.cfg.nz)

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.fso_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_dot_mode(self, mode: int):
        self._lib.fso_set_dot_mode(self._ctx, mode)

    def set_collider_sdf(self, sdf):
        a = np.ascontiguousarray(sdf, dtype=np.float64)
        assert a.shape == self.shape
        self._lib.fso_set_collider_sdf(self._ctx, _dp(a))

    def set_particles(self, pos, vel=None):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        self._lib.fso_set_particles(self._ctx, _dp(p), _dp(v) if v is not None else None, len(p))

    def num_particles(self):
        return int(self._lib.fso_num_particles(self._ctx))

    def particles(self):
        n = self.num_particles()
        p, v = np.empty((n, 3)), np.empty((n, 3))
        self._lib.fso_get_particles(self._ctx, _dp(p), _dp(v))
        return p, v

    def positions(self):
        return self.particles()[0]

    def substep(self, dt):
        st = fs_stats()
        self._lib.fso_substep(self._ctx, dt, C.byref(st))
        return st.as_dict()

    def step_frame(self, frame_dt=1.0 / 60.0):
        st = fs_stats()
        self._lib.fso_step_frame(self._ctx, frame_dt, C.byref(st))
        return st.as_dict()

    def cfl(self):
        out = C.c_double()
        self._lib.fso_cfl(self._ctx, C.byref(out))
        return out.value

    def clear(self):
        self._lib.fso_clear(self._ctx)

    def advect(self, dt):
        self._lib.fso_advect(self._ctx, dt)

    def reconstruct(self, radius):
        self._lib.fso_reconstruct(self._ctx, radius)

    def extrapolate_fluid_sdf(self, iters):
        self._lib.fso_extrapolate_fluid_sdf(self._ctx, iters)

    def smooth_fluid_surface(self, iters):
        self._lib.fso_smooth_fluid_surface(self._ctx, iters)

    def p2g(self, dt=0.0):
        self._lib.fso_p2g(self._ctx, dt)

    def apply_dirichlet_boundary(self):
        self._lib.fso_apply_dirichlet_boundary(self._ctx)

    def extrapolate_velocity(self, iters):
        self._lib.fso_extrapolate_velocity(self._ctx, iters)

    def apply_force(self, dt):
        self._lib.fso_apply_force(self._ctx, dt)

    def build_system(self, dt):
        self._lib.fso_build_system(self._ctx, dt)

    def solve_pressure(self):
        st = fs_stats()
        self._lib.fso_solve_pressure(self._ctx, C.byref(st))
        return st.as_dict()

    def project(self, dt):
        self._lib.fso_project(self._ctx, dt)

    def apply_collider(self):
        self._lib.fso_apply_collider(self._ctx)

    def g2p(self, dt=0.0):
        self._lib.fso_g2p(self._ctx, dt)

    def spmv(self, x, y=None):
        xa = np.ascontiguousarray(x, dtype=np.float64)
        ya = np.zeros(self.shape) if y is None else np.array(y, dtype=np.float64, copy=True)
        self._lib.fso_spmv(self._ctx, _dp(xa), _dp(ya))
        return ya

    def mic_apply(self, r, z=None):
        ra = np.ascontiguousarray(r, dtype=np.float64)
        za = np.zeros(self.shape) if z is None else np.array(z, dtype=np.float64, copy=True)
        self._lib.fso_mic_apply(self._ctx, _dp(ra), _dp(za))
        return za

    def mg_apply(self, r):
        ra = np.ascontiguousarray(r, dtype=np.float64)
        za = np.zeros(self.shape)
        self._lib.fso_mg_apply(self._ctx, _dp(ra), _dp(za))
        return za

    def array_info(self, name):
        dims = (C.c_int64 * 3)()
        dt = C.c_int32()
        if self._lib.fso_array_info(self._ctx, name.encode(), dims, C.byref(dt)) != 0:
            raise KeyError(name)
        return (dims[0], dims[1], dims[2]), (np.dtype(np.uint8) if dt.value == 1 else np.dtype(np.float64))

    def get(self, name):
        shape, dtype = self.array_info(name)
        out = np.empty(shape, dtype=dtype)
        self._lib.fso_get_array(self._ctx, name.encode(), out.ctypes.data_as(C.c_void_p))
        return out

    def set(self, name, value):
        shape, dtype = self.array_info(name)
        a = np.ascontiguousarray(value, dtype=dtype)
        assert a.shape == shape, (name, a.shape, shape)
        self._lib.fso_set_array(self._ctx, name.encode(), a.ctypes.data_as(C.c_void_p))

    # point queries used by known-answer tests
    def sdf_eval(self, p):
        pa = np.ascontiguousarray(p, dtype=np.float64)
        return self._lib.fso_sdf_eval(self._ctx, _dp(pa))

    def sdf_grad(self, p):
        pa = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty(3)
        self._lib.fso_sdf_grad(self._ctx, _dp(pa), _dp(out))
        return out

    def sample_velocity(self, p):
        pa = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty(3)
        self._lib.fso_sample_velocity(self._ctx, _dp(pa), _dp(out))
        return out

    def set_collider_mesh(self, vertices, indices, exact_band: int = 3, flags: int = 0):
        """Collider SDF of a closed triangle mesh (manifold2SDF, exact band of 3 as setCollider uses;
        flags 1 = true point-triangle distances instead of the reference's)."""
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        if self._lib.fso_set_collider_mesh(self._ctx, _dp(v), len(v), ix.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           len(ix) // 3, exact_band, flags) != 0:
            raise ValueError("bad mesh")

    def rebuild_surface(self):
        """Marching-cubes surface of the fluid SDF: (vertices (V,3), normals (V,3), triangles (T,3))."""
        nv, nt = C.c_int64(), C.c_int64()
        self._lib.fso_rebuild_surface(self._ctx, C.byref(nv), C.byref(nt))
        v, n = np.zeros((nv.value, 3)), np.zeros((nv.value, 3))
        ix = np.zeros((nt.value, 3), np.uint32)
        self._lib.fso_get_surface(self._ctx, _dp(v), _dp(n), ix.ctypes.data_as(C.POINTER(C.c_uint32)))
        return v, n, ix

    def apic_matrices(self):
        """APIC affine rows (N, 3, 3): [particle, component u/v/w, x/y/z]."""
        c = np.zeros((self.num_particles(), 3, 3))
        self._lib.fso_get_apic(self._ctx, _dp(c))
        return c

    def set_apic_matrices(self, c):
        a = np.ascontiguousarray(c, np.float64)
        self._lib.fso_set_apic(self._ctx, _dp(a))

    def rollout_begin(self):
        """Record the following substeps for rollout_vjp (all of them are kept in memory)."""
        if self._lib.fso_rollout_begin(self._ctx) != 0:
            raise ValueError("the reverse pass needs advect_on_cleared_grid = 1 and smoothing")

    def rollout_stop(self):
        self._lib.fso_rollout_stop(self._ctx)

    def rollout_steps(self) -> int:
        return self._lib.fso_rollout_steps(self._ctx)

    def rollout_vjp(self, dpos, dvel):
        """dL/dpos0, dL/dvel0 and dL/dG of the recorded substeps from dL/dpos, dL/dvel at their end
        (the FP64 CPU reverse pass)."""
        n = self.num_particles()
        dp0, dv0, dg = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(3)
        a = None if dpos is None else np.ascontiguousarray(dpos, np.float64)
        b = None if dvel is None else np.ascontiguousarray(dvel, np.float64)
        st = self._lib.fso_rollout_vjp(self._ctx, None if a is None else _dp(a), None if b is None else _dp(b),
                                       _dp(dp0), _dp(dv0), _dp(dg))
        if st != 0:
            raise RuntimeError(f"oracle reverse pass failed (status {st})")
        return dp0, dv0, dg

    def nearest(self, axis, p):
        pa = np.ascontiguousarray(p, dtype=np.float64)
        out = (C.c_int * 3)()
        self._lib.fso_nearest(self._ctx, axis, _dp(pa), out)
        return tuple(out)
