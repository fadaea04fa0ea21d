"""Build libsimcraft_fluid.so (HIP, gfx950) in-tree.

The library is compiled with hipcc straight from simcraft_amd/csrc/*.hip. Flags:
  * --offload-arch=gfx950 only;
  * -ffp-contract=off so FP64 arithmetic keeps the reference's rounding (no fused multiply-add);
  * packed FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJDIR = PKG / "_obj"
LIB = PKG / "libsimcraft_fluid.so"
SOURCES = ["fs_grid.hip", "fs_pcg.hip", "fs_capi.hip", "fs_adjoint.hip", "fs_noalloc.hip", "fs_mesh.hip", "fs_surface.hip", "fs_mg.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")


def build_library(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    include = ["-I", str(CSRC), "-I", str(PKG.parent / "include")]

    def compile_one(src: str) -> Path:
        obj = OBJDIR / (Path(src).stem + ".o")
        cmd = [HIPCC, *COMMON, *include, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
