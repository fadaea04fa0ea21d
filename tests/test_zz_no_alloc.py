"""fs_cfl_device_noalloc: the CFL step of device face grids with no allocation (its partial
maxima live in a file-scope device array), against fs_cfl on a context holding the same grids
and against the oracle's CFL (fluid-simulator.cc:184-203). Two calls in a row, the second on a
smaller grid, must both agree bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from simcraft_amd import _capi
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _grids(n, seed, scale):
    rng = np.random.default_rng(seed)
    shapes = [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    return [rng.uniform(-scale, scale, s) for s in shapes]


def _cfl_via_context(hip, n, grids):
    sim = hip.FluidSim(hip.default_config(nx=n, ny=n, nz=n))
    for nm, g in zip("uvw", grids):
        sim.set(nm, g)
    dt = sim.cfl()
    sim.close()
    return dt


def _cfl_oracle(oracle, n, grids):
    o = oracle.OracleSim(oracle.default_config(nx=n, ny=n, nz=n))
    for nm, g in zip("uvw", grids):
        o.set(nm, g)
    dt = o.cfl()
    o.close()
    return dt


class _Hip:
    """The HIP runtime the library itself loaded (libamdhip64, found by its soname), for the
    test's own device buffers: the entry point under test allocates nothing, its caller does."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def upload(self, a: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p

    def download(self, p: C.c_void_p, out: np.ndarray) -> np.ndarray:
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out


def test_cfl_noalloc_matches_fs_cfl_twice(hip, oracle):
    lib = _capi.load()
    fn = lib.fs_cfl_device_noalloc
    rt = _Hip()
    out = rt.upload(np.full(1, -1.0))
    bufs = []
    try:
        for n, seed, scale in ((32, 1, 3.0), (12, 2, 0.5)):  # the second call on a smaller grid
            announce(f"cfl noalloc n={n}")
            grids = _grids(n, seed, scale)
            want = _cfl_via_context(hip, n, grids)
            assert want == _cfl_oracle(oracle, n, grids)
            dg = [rt.upload(np.ascontiguousarray(g)) for g in grids]
            bufs += dg
            st = fn(dg[0], dg[1], dg[2], n, n, n, C.c_double(1.0 / n), C.c_double(5.0), out, None)
            assert st == _capi.FS_OK
            assert rt.lib.hipDeviceSynchronize() == 0
            got = float(rt.download(out, np.zeros(1))[0])
            print(f"[cfl noalloc n={n}] {got!r} vs fs_cfl {want!r}")
            assert got == want
        # sizes beyond the 32-bit face count are refused before any launch
        big = 1 << 11
        st = fn(bufs[0], bufs[1], bufs[2], big, big, big, C.c_double(1.0 / big), C.c_double(5.0), out, None)
        assert st == _capi.FS_BAD_ARG
    finally:
        for p in bufs + [out]:
            rt.lib.hipFree(p)
