"""simcraft_amd - SimCraft FluidSim's FLIP/PIC substep on MI355X (HIP, gfx950).

Python host mirror of the reference's simulator interface (fluid::cpu::FluidSimulator and its
solver plug-ins, FluidSim/include/FluidSim/cpu/fluid-simulator.h:25-272) over the C ABI of
libsimcraft_fluid.so. All compute runs in HIP kernels; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _capi
from ._capi import (ADVECT_APIC, ADVECT_FLIP, ADVECT_PIC, MESH_EXACT_DISTANCE, PRECISION_FP64, PRECISION_MIXED, PRECOND_MIC,
                    PRECOND_MULTIGRID,
                    PROJECT_FVM, RECON_NAIVE, SOLVER_CG, TOL_ABS_LINF, TOL_REL_L2, fs_config, fs_stats)

__all__ = ["FluidSim", "FluidSimError", "default_config", "device_count", "fs_config", "fs_stats"]

_DP = C.POINTER(C.c_double)


class FluidSimError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_capi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def default_config(**overrides) -> fs_config:
    """fs_config with the reference defaults (fs_config_default) and the given overrides."""
    return _capi.default_config(**overrides)


def device_count() -> int:
    return int(_capi.load().fs_device_count())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class FluidSim:
    """One simulation context on the current HIP device.

    Mirrors cpu::FluidSimulator: construct with a resolution/size config, set the collider
    and the initial fluid, choose the solvers, then step. Operator-level methods run single
    stages of substep() on the device-resident state (used for parity tests).
    """

    def __init__(self, cfg: fs_config | None = None, **overrides):
        self._lib = _capi.load()
        self.cfg = cfg if cfg is not None else default_config()
        for k, v in overrides.items():
            setattr(self.cfg, k, v)
        ptr = C.c_void_p()
        st = self._lib.fs_create(C.byref(self.cfg), C.byref(ptr))
        if st != _capi.FS_OK:
            raise FluidSimError(st, (self._lib.fs_last_error(None) or b"").decode())
        self._ctx = ptr
        self.shape = (self.cfg.nx, self.cfg.ny, self.cfg.nz)

    # -- lifetime -------------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.fs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, allow=(_capi.FS_OK,)) -> int:
        if st in allow:
            return st
        msg = (self._lib.fs_last_error(self._ctx) or b"").decode()
        raise FluidSimError(st, msg)

    # -- backend level (FluidComputeBackend) --------------------------------------------------
    def set_collider_sdf(self, sdf: np.ndarray) -> None:
        a = np.ascontiguousarray(sdf, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"collider SDF must have shape {self.shape}")
        self._check(self._lib.fs_set_collider_sdf(self._ctx, _dptr(a)))

    def set_collider_mesh(self, vertices: np.ndarray, indices: np.ndarray, exact_band: int = 3,
                          flags: int = 0) -> None:
        """setCollider(mesh): the collider SDF of a closed triangle mesh (manifold2SDF on the
        device; exact band of 3 cells as the reference's setCollider uses). flags =
        MESH_EXACT_DISTANCE replaces the reference's point-triangle distance by the true one."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        ix = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        if len(ix) % 3:
            raise ValueError("indices must hold three vertex indices per triangle")
        self._check(self._lib.fs_set_collider_mesh(self._ctx, _dptr(v), len(v),
                                                   ix.ctypes.data_as(C.POINTER(C.c_uint32)), len(ix) // 3,
                                                   exact_band, flags))

    def particle_affine(self) -> np.ndarray:
        """APIC rows (N, 3, 3): [particle, velocity component u/v/w, x/y/z]."""
        c = np.zeros((self.num_particles(), 3, 3))
        self._check(self._lib.fs_get_particle_affine(self._ctx, _dptr(c)))
        return c

    def set_particle_affine(self, c9: np.ndarray) -> None:
        a = np.ascontiguousarray(c9, dtype=np.float64).reshape(self.num_particles(), 3, 3)
        self._check(self._lib.fs_set_particle_affine(self._ctx, _dptr(a)))

    def rebuild_surface(self):
        """rebuildSurface: marching-cubes mesh of the fluid SDF on the device; returns
        (vertices (V, 3), normals (V, 3), triangles (T, 3) uint32)."""
        nv, nt = C.c_int64(), C.c_int64()
        self._check(self._lib.fs_rebuild_surface(self._ctx, C.byref(nv), C.byref(nt)))
        v, n = np.zeros((nv.value, 3)), np.zeros((nv.value, 3))
        ix = np.zeros((nt.value, 3), dtype=np.uint32)
        self._check(self._lib.fs_get_surface(self._ctx, _dptr(v), _dptr(n), ix.ctypes.data_as(C.POINTER(C.c_uint32))))
        return v, n, ix

    def set_particles(self, pos: np.ndarray, vel: np.ndarray | None = None) -> None:
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        if v is not None and v.shape != p.shape:
            raise ValueError("velocities must match positions")
        self._check(self._lib.fs_set_particles(self._ctx, _dptr(p), _dptr(v) if v is not None else None, len(p)))

    def num_particles(self) -> int:
        return int(self._lib.fs_num_particles(self._ctx))

    def particles(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.num_particles()
        p = np.empty((n, 3))
        v = np.empty((n, 3))
        self._check(self._lib.fs_get_particles(self._ctx, _dptr(p), _dptr(v)))
        return p, v

    def positions(self) -> np.ndarray:
        return self.particles()[0]

    def set_advector(self, method: int = ADVECT_FLIP) -> None:
        self._check(self._lib.fs_set_advector(self._ctx, method))

    def set_projector(self, method: int = PROJECT_FVM) -> None:
        self._check(self._lib.fs_set_projector(self._ctx, method))

    def set_compressed_solver(self, solver: int = SOLVER_CG, precond: int = PRECOND_MIC) -> None:
        self._check(self._lib.fs_set_compressed_solver(self._ctx, solver, precond))

    def set_reconstructor(self, method: int = RECON_NAIVE) -> None:
        self._check(self._lib.fs_set_reconstructor(self._ctx, method))

    def substep(self, dt: float) -> dict:
        st = fs_stats()
        s = self._check(self._lib.fs_substep(self._ctx, dt, C.byref(st)),
                        allow=(_capi.FS_OK, _capi.FS_NOT_CONVERGED, _capi.FS_BREAKDOWN, _capi.FS_NONFINITE))
        if s != _capi.FS_OK:
            warnings.warn(f"pressure solve ended with {_capi.STATUS_NAMES[s]} (residual {st.residual_inf:g})")
        return st.as_dict()

    def step_frame(self, frame_dt: float = 1.0 / 60.0) -> dict:
        st = fs_stats()
        self._check(self._lib.fs_step_frame(self._ctx, frame_dt, C.byref(st)),
                    allow=(_capi.FS_OK, _capi.FS_NOT_CONVERGED))
        return st.as_dict()

    def cfl(self) -> float:
        out = C.c_double()
        self._check(self._lib.fs_cfl(self._ctx, C.byref(out)))
        return out.value

    # -- operator level ------------------------------------------------------------------------
    def clear(self):
        self._check(self._lib.fs_clear(self._ctx))

    def advect(self, dt: float):
        self._check(self._lib.fs_advect(self._ctx, dt))

    def reconstruct(self, radius: float):
        self._check(self._lib.fs_reconstruct(self._ctx, radius))

    def extrapolate_fluid_sdf(self, iters: int):
        self._check(self._lib.fs_extrapolate_fluid_sdf(self._ctx, iters))

    def smooth_fluid_surface(self, iters: int):
        self._check(self._lib.fs_smooth_fluid_surface(self._ctx, iters))

    def p2g(self, dt: float = 0.0):
        self._check(self._lib.fs_p2g(self._ctx, dt))

    def apply_dirichlet_boundary(self):
        self._check(self._lib.fs_apply_dirichlet_boundary(self._ctx))

    def extrapolate_velocity(self, iters: int):
        self._check(self._lib.fs_extrapolate_velocity(self._ctx, iters))

    def apply_force(self, dt: float):
        self._check(self._lib.fs_apply_force(self._ctx, dt))

    def build_system(self, dt: float):
        self._check(self._lib.fs_build_system(self._ctx, dt))

    def solve_pressure(self) -> dict:
        st = fs_stats()
        self._check(self._lib.fs_solve_pressure(self._ctx, C.byref(st)),
                    allow=(_capi.FS_OK, _capi.FS_NOT_CONVERGED, _capi.FS_BREAKDOWN, _capi.FS_NONFINITE))
        return st.as_dict()

    def project(self, dt: float):
        self._check(self._lib.fs_project(self._ctx, dt))

    def apply_collider(self):
        self._check(self._lib.fs_apply_collider(self._ctx))

    def g2p(self, dt: float = 0.0):
        self._check(self._lib.fs_g2p(self._ctx, dt))

    def spmv(self, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
        xa = np.ascontiguousarray(x, dtype=np.float64)
        ya = np.zeros(self.shape) if y is None else np.array(y, dtype=np.float64, copy=True)
        self._check(self._lib.fs_spmv(self._ctx, _dptr(xa), _dptr(ya)))
        return ya

    def mic_apply(self, r: np.ndarray, z: np.ndarray | None = None) -> np.ndarray:
        ra = np.ascontiguousarray(r, dtype=np.float64)
        za = np.zeros(self.shape) if z is None else np.array(z, dtype=np.float64, copy=True)
        self._check(self._lib.fs_mic_apply(self._ctx, _dptr(ra), _dptr(za)))
        return za

    def mg_apply(self, r: np.ndarray) -> np.ndarray:
        """One multigrid V-cycle of the assembled system (FS_PRECOND_MULTIGRID's preconditioner)."""
        ra = np.ascontiguousarray(r, dtype=np.float64)
        za = np.zeros(self.shape)
        self._check(self._lib.fs_mg_apply(self._ctx, _dptr(ra), _dptr(za)))
        return za

    # -- named state ---------------------------------------------------------------------------
    def array_info(self, name: str) -> tuple[tuple[int, int, int], np.dtype]:
        dims = (C.c_int64 * 3)()
        dt = C.c_int32()
        self._check(self._lib.fs_array_info(self._ctx, name.encode(), dims, C.byref(dt)))
        return (dims[0], dims[1], dims[2]), (np.dtype(np.uint8) if dt.value == 1 else np.dtype(np.float64))

    def get(self, name: str) -> np.ndarray:
        shape, dtype = self.array_info(name)
        out = np.empty(shape, dtype=dtype)
        self._check(self._lib.fs_get_array(self._ctx, name.encode(), out.ctypes.data_as(C.c_void_p)))
        return out

    def set(self, name: str, value: np.ndarray) -> None:
        shape, dtype = self.array_info(name)
        a = np.ascontiguousarray(value, dtype=dtype)
        if a.shape != shape:
            raise ValueError(f"{name} must have shape {shape}")
        self._check(self._lib.fs_set_array(self._ctx, name.encode(), a.ctypes.data_as(C.c_void_p)))

    def synchronize(self) -> None:
        self._check(self._lib.fs_synchronize(self._ctx))

    # -- kernel timing (fs_config.profile_stages >= 2) ------------------------------------------
    def kernel_profile(self) -> dict:
        """Per kernel class: launches, device ms and algorithmic bytes since the last reset."""
        out = {}
        for k, name in enumerate(_capi.KERNEL_NAMES):
            n, ms, by = C.c_int64(), C.c_double(), C.c_double()
            self._check(self._lib.fs_kernel_profile(self._ctx, k, C.byref(n), C.byref(ms), C.byref(by)))
            out[name] = {"launches": n.value, "ms": ms.value, "bytes": by.value}
        return out

    def kernel_profile_reset(self) -> None:
        self._check(self._lib.fs_kernel_profile_reset(self._ctx))

    def set_kernel_timing(self, on: bool) -> None:
        """Switch the per-launch event bracketing on or off (profile_stages >= 2)."""
        self._check(self._lib.fs_set_kernel_timing(self._ctx, 1 if on else 0))

    # -- reverse mode (no counterpart in the reference) ------------------------------------------
    def rollout_begin(self, checkpoint_every: int = 8) -> None:
        """Record the following substeps for rollout_vjp (checkpoints every `checkpoint_every`)."""
        self._check(self._lib.fs_rollout_begin(self._ctx, checkpoint_every))

    def rollout_stop(self) -> None:
        """Stop appending substeps to the recorded rollout; the tape stays for rollout_vjp."""
        self._check(self._lib.fs_rollout_stop(self._ctx))

    def rollout_end(self) -> None:
        self._check(self._lib.fs_rollout_end(self._ctx))

    def rollout_steps(self) -> int:
        return int(self._lib.fs_rollout_steps(self._ctx))

    def rollout_vjp(self, dpos: np.ndarray | None, dvel: np.ndarray | None):
        """dL/d(initial positions), dL/d(initial velocities), dL/dG (gravity vector G = (0, -g, 0))
        from dL/d(final positions) and dL/d(final velocities) of the recorded rollout."""
        n = self.num_particles()
        a = None if dpos is None else np.ascontiguousarray(dpos, dtype=np.float64).reshape(n, 3)
        b = None if dvel is None else np.ascontiguousarray(dvel, dtype=np.float64).reshape(n, 3)
        dp0, dv0, dg = np.empty((n, 3)), np.empty((n, 3)), np.empty(3)
        self._check(self._lib.fs_rollout_vjp(self._ctx, _dptr(a) if a is not None else None,
                                             _dptr(b) if b is not None else None, _dptr(dp0), _dptr(dv0),
                                             _dptr(dg)))
        return dp0, dv0, dg

    def _pos_arg(self, pos):
        if pos is None:
            return None, None
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.num_particles(), 3)
        return p, _dptr(p)

    def sample_grid(self, u: np.ndarray, v: np.ndarray, w: np.ndarray, pos: np.ndarray | None = None) -> np.ndarray:
        """PIC transfer of the face grids u, v, w to particles at `pos` (default: the current
        particles) as velocities (N, 3); the simulation state is not changed."""
        arrs = [np.ascontiguousarray(g, dtype=np.float64) for g in (u, v, w)]
        for nm, g in zip(("u", "v", "w"), arrs):
            if g.shape != self.array_info(nm)[0]:
                raise ValueError(f"{nm} must have shape {self.array_info(nm)[0]}")
        p, pp = self._pos_arg(pos)
        out = np.empty((self.num_particles(), 3))
        self._check(self._lib.fs_sample_grid(self._ctx, pp, *[_dptr(g) for g in arrs], _dptr(out)))
        return out

    def sample_grid_vjp(self, dvel: np.ndarray, pos: np.ndarray | None = None):
        """Transpose of sample_grid: face-grid gradients from particle-velocity gradients."""
        d = np.ascontiguousarray(dvel, dtype=np.float64).reshape(self.num_particles(), 3)
        p, pp = self._pos_arg(pos)
        outs = [np.empty(self.array_info(nm)[0]) for nm in ("u", "v", "w")]
        self._check(self._lib.fs_sample_grid_vjp(self._ctx, pp, _dptr(d), *[_dptr(o) for o in outs]))
        return tuple(outs)
