"""The C ABI (include/simcraft_fluid.h) on the CPU: declarations, exports, struct layout and
the loud failure without a device. No kernel is launched here."""
from __future__ import annotations

import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from simcraft_amd import _capi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "simcraft_fluid.h"


def declared_functions() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text))


def test_header_matches_binding():
    assert declared_functions() == set(_capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    nm = shutil.which("nm")
    if nm:
        out = subprocess.run([nm, "-D", "--defined-only", str(_capi.LIB_PATH)], capture_output=True, text=True,
                             check=True).stdout
        exported = {ln.split()[-1] for ln in out.splitlines() if re.search(r" T fs_", ln)}
        assert exported == declared_functions()


def test_manual_build_line_lists_every_source():
    """The hand-written hipcc line in INTEGRATION.md builds the same library as build.py."""
    from simcraft_amd import build

    text = (ROOT / "INTEGRATION.md").read_text()
    start = text.index("```sh")
    block = text[start:text.index("```", start + 5)]
    assert [s for s in build.SOURCES if f"simcraft_amd/csrc/{s}" not in block] == []


def test_struct_layout_matches_c(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for struct in (_capi.fs_config, _capi.fs_stats):
        cname = struct.__name__
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in struct._fields_:
            lines.append(f'  printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c99", "-o", str(exe), str(src)], check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                            check=True).stdout.splitlines())
    for struct in (_capi.fs_config, _capi.fs_stats):
        cname = struct.__name__
        assert int(got[cname]) == C.sizeof(struct), cname
        for field, _ in struct._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(struct, field).offset, (cname, field)


def test_device_defaults_equal_oracle_defaults(oracle):
    dev, ora = _capi.default_config(), oracle.default_config()
    for field, _ in _capi.fs_config._fields_:
        a, b = getattr(dev, field), getattr(ora, field)
        if field == "reserved":
            assert list(a) == list(b)
        else:
            assert a == b, field


def test_no_device_fails_loudly():
    import simcraft_amd

    if simcraft_amd.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(simcraft_amd.FluidSimError) as err:
        simcraft_amd.FluidSim(simcraft_amd.default_config(nx=8, ny=8, nz=8))
    assert err.value.status == _capi.FS_NO_DEVICE


def test_create_rejects_bad_config():
    lib = _capi.load()
    ctx = C.c_void_p()
    assert lib.fs_create(None, C.byref(ctx)) == _capi.FS_BAD_ARG
    cfg = _capi.default_config(nx=8, ny=4, nz=8)  # the reference's "up" bound assumes nx == ny
    st = lib.fs_create(C.byref(cfg), C.byref(ctx))
    assert st in (_capi.FS_UNSUPPORTED, _capi.FS_NO_DEVICE)
    assert not ctx.value
    # the configuration is checked before any device is looked for, so these hold without one
    bad = [(dict(nx=0, ny=0, nz=8), "positive"), (dict(nx=8, ny=8, nz=-1), "positive"),
           (dict(nx=8, ny=8, nz=8, size_y=0.0), "positive"),
           (dict(nx=1290, ny=1290, nz=1290), "32-bit"),  # 1291^3 nodes >= 2^31
           (dict(nx=8, ny=8, nz=8, precision=7), "precision"),
           (dict(nx=8, ny=8, nz=8, preconditioner=9), "preconditioner")]
    for kw, msg in bad:
        cfg = _capi.default_config(**kw)
        st = lib.fs_create(C.byref(cfg), C.byref(ctx))
        assert st == _capi.FS_BAD_ARG, kw
        assert not ctx.value
        assert msg in lib.fs_last_error(None).decode(), (kw, lib.fs_last_error(None))


CALLER = ROOT / "tests" / "capi_caller.c"


def build_c_caller(out_dir: Path) -> Path:
    """Compile tests/capi_caller.c with gcc against the header and link the library."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    exe = out_dir / "capi_caller"
    lib_dir = _capi.LIB_PATH.parent
    subprocess.run([gcc, "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", str(HEADER.parent), str(CALLER),
                    "-L", str(lib_dir), "-lsimcraft_fluid", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    exe = build_c_caller(tmp_path)
    out = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert "libsimcraft_fluid.so" in out and "not found" not in out


@pytest.mark.gpu
@pytest.mark.parametrize("advection", [1, 0], ids=["flip", "pic"])
def test_c_caller_two_substeps_match_oracle(tmp_path, oracle, advection):
    """The C program's two 16^3 substeps give the oracle's particles and pressure bit for bit."""
    import numpy as np

    from simcraft_amd import scenes
    from tests.conftest import announce

    n, dt, steps, alpha = 16, 0.004, 2, 0.95
    announce(f"C caller n={n} advection={advection}")
    pos, _ = scenes.dam_break(n)
    vel = np.random.default_rng(12).uniform(-0.3, 0.3, pos.shape)
    col = scenes.sphere_sdf(n)
    src = tmp_path / "in.bin"
    with open(src, "wb") as f:
        np.array([n, advection, steps, 0], np.int32).tofile(f)
        np.array([dt, alpha], np.float64).tofile(f)
        np.array([len(pos)], np.int64).tofile(f)
        for a in (col, pos, vel):
            np.ascontiguousarray(a, np.float64).tofile(f)
    exe = build_c_caller(tmp_path)
    dst = tmp_path / "out.bin"
    r = subprocess.run([str(exe), str(src), str(dst)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    raw = dst.read_bytes()
    assert int(np.frombuffer(raw, np.int32, 1)[0]) == steps
    rec = np.frombuffer(raw, np.dtype([("status", np.int32), ("iters", np.int32), ("res", np.float64)]), steps, 4)
    off = 4 + rec.nbytes
    body = np.frombuffer(raw, np.float64, offset=off)
    npart = len(pos)
    gp, gv = body[:3 * npart].reshape(-1, 3), body[3 * npart:6 * npart].reshape(-1, 3)
    gpg = body[6 * npart:].reshape(n, n, n)

    cfg = oracle.default_config(nx=n, ny=n, nz=n, flip_alpha=alpha, advection=advection)
    o = oracle.OracleSim(cfg, dot_mode=oracle.DOT_TREE)
    o.set_collider_sdf(col)
    o.set_particles(pos, vel)
    for s in range(steps):
        st = o.substep(dt)
        assert rec["iters"][s] == st["cg_iterations"] and rec["res"][s] == st["residual_inf"], (s, rec[s], st)
    op, ov = o.particles()
    assert np.array_equal(gp, op) and np.array_equal(gv, ov)
    assert np.array_equal(gpg, o.get("pg"))
