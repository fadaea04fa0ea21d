"""Host-side file formats around the collider: triangle meshes (OBJ) and SDF text files.

* load_obj / save_obj: the subset of OBJ that the reference reads and writes
  (core::myLoadObj / core::exportObj, Core/src/obj-io.cc:13-91): "v x y z" vertices, "vn" normals,
  and triangular faces "f a b c" or "f a//na b//nb c//nc" (1-based). As in myLoadObj, a vertex's
  normal is the normalised mean of the face-normal entries that name it; faces with more than three
  corners are read as their first three (myLoadObj reads three).
* save_sdf / load_sdf: SDF<3>::saveSDF (FluidSim/include/FluidSim/cpu/sdf.h:160-173) and loadSDF
  (FluidSim/src/cpu/util.cc:412-429): three text lines - the origin, the cell counts
  (TGrid::size(), Spatify/include/Spatify/grids.h:259) and the spacing - followed by every value,
  x slowest and z fastest, separated by spaces.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


@dataclass
class Mesh:
    """core::Mesh (Core/include/Core/mesh.h:9-14): vertices (V, 3), per-vertex normals (V, 3) and
    three 0-based vertex indices per triangle (T, 3)."""
    vertices: np.ndarray
    indices: np.ndarray
    normals: np.ndarray = field(default=None)

    @property
    def triangle_count(self) -> int:
        return len(self.indices)


def load_obj(path: str | Path) -> Mesh:
    verts, vnormals, idx = [], [], []
    acc, cnt = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            tag = parts[0]
            if tag == "v":
                verts.append([float(x) for x in parts[1:4]])
                acc.append(np.zeros(3))
                cnt.append(0)
            elif tag == "vn":
                vnormals.append(np.array([float(x) for x in parts[1:4]]))
            elif tag == "f":
                if len(parts) < 4:
                    raise ValueError(f"face with fewer than three corners: {line.strip()}")
                for corner in parts[1:4]:
                    if "//" in corner:
                        a, b = corner.split("//", 1)
                        vi, ni = int(a) - 1, int(b) - 1
                        acc[vi] = acc[vi] + vnormals[ni]
                        cnt[vi] += 1
                    else:
                        vi = int(corner.split("/", 1)[0]) - 1
                    idx.append(vi)
    v = np.array(verts, dtype=np.float64).reshape(-1, 3)
    ix = np.array(idx, dtype=np.uint32).reshape(-1, 3)
    if len(ix) and int(ix.max()) >= len(v):
        raise ValueError("a face names a missing vertex")
    normals = np.zeros_like(v)
    for i, (a, c) in enumerate(zip(acc, cnt)):
        if c:
            m = a / c
            normals[i] = m / np.sqrt(np.dot(m, m))
    return Mesh(v, ix, normals)


def save_obj(path: str | Path, mesh: Mesh) -> None:
    """core::exportObj: vertices, normals, then faces "f a//a b//b c//c" (1-based)."""
    normals = mesh.normals if mesh.normals is not None else np.zeros_like(mesh.vertices)
    with open(path, "w") as f:
        for v in np.asarray(mesh.vertices, dtype=np.float64).tolist():
            f.write(f"v {v[0]!r} {v[1]!r} {v[2]!r}\n")
        for n in np.asarray(normals, dtype=np.float64).tolist():
            f.write(f"vn {n[0]!r} {n[1]!r} {n[2]!r}\n")
        for t in mesh.indices:
            a, b, c = (int(x) + 1 for x in t)
            f.write(f"f {a}//{a} {b}//{b} {c}//{c} \n")


def save_sdf(path: str | Path, sdf: np.ndarray, spacing=(None, None, None), origin=(0.0, 0.0, 0.0)) -> None:
    """Writes a cell-centred SDF (nx, ny, nz) in the reference's text format."""
    sdf = np.asarray(sdf, dtype=np.float64)
    nx, ny, nz = sdf.shape
    h = [1.0 / n if s is None else s for s, n in zip(spacing, sdf.shape)]
    with open(path, "w") as f:
        f.write(" ".join(repr(float(x)) for x in origin) + "\n")
        f.write(f"{nx} {ny} {nz}\n")
        f.write(" ".join(repr(float(x)) for x in h) + "\n")
        f.write(" ".join(repr(x) for x in sdf.reshape(-1).tolist()))
        f.write(" ")


def load_sdf(path: str | Path):
    """Reads the reference's SDF text format: (values (nx, ny, nz), spacing (3,), origin (3,))."""
    tok = Path(path).read_text().split()
    origin = np.array([float(x) for x in tok[0:3]])
    dims = [int(float(x)) for x in tok[3:6]]
    spacing = np.array([float(x) for x in tok[6:9]])
    n = dims[0] * dims[1] * dims[2]
    vals = np.array([float(x) for x in tok[9:9 + n]])
    if len(vals) != n:
        raise ValueError(f"SDF file holds {len(vals)} values, expected {n}")
    return vals.reshape(dims), spacing, origin


def icosphere(centre=(0.5, 0.5, 0.5), radius: float = 0.25, subdivisions: int = 2) -> Mesh:
    """A closed triangulated sphere (outward-facing triangles): the icosahedron, each triangle
    split in four `subdivisions` times, vertices projected onto the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                mid[key] = len(v) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    vv = np.array(v) * radius + np.asarray(centre, dtype=np.float64)
    return Mesh(vv, np.array(faces, dtype=np.uint32), np.array(v))
