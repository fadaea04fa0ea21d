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
