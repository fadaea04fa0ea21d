"""The headless demo (simcraft_amd.demo, after FluidSim/disabled/fluid-sim.cc): the mesh placement
of relocateMesh on the CPU, and on the GPU a few frames at 16^3 with the surface exported as OBJ
files that read back as closed, consistently oriented meshes inside the domain."""
from __future__ import annotations

import numpy as np
import pytest

from simcraft_amd import meshio
from tests.conftest import announce


def test_relocate_mesh_scales_into_half_the_domain():
    from simcraft_amd.demo import relocate_mesh

    m = meshio.Mesh(vertices=np.array([[-2.0, 0.0, 1.0], [2.0, 1.0, -1.0], [0.0, 3.0, 0.0]]),
                    indices=np.array([[0, 1, 2]]))
    r = relocate_mesh(m, 1.0)
    # shifted by the smallest coordinate (-2), scaled by 1 / (3 - (-2)), halved
    np.testing.assert_allclose(r.vertices, (m.vertices + 2.0) / 5.0 * 0.5, rtol=0, atol=1e-15)
    assert r.vertices.min() == 0.0 and r.vertices.max() == pytest.approx(0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("advection,precond", [("pic", "mic"), ("apic", "mg")])
def test_demo_frames_export_closed_surfaces(hip, tmp_path, advection, precond):
    from simcraft_amd.demo import run

    announce(f"demo 16^3 advection={advection} precond={precond}")
    paths = run(n=16, frames=3, export_from=1, out=tmp_path, advection=advection, precond=precond,
                log=lambda m: print(m, flush=True))
    assert [p.name for p in paths] == ["fluid_0001.obj", "fluid_0002.obj"]
    for p in paths:
        m = meshio.load_obj(p)
        tri = m.indices
        assert len(tri) > 0
        assert np.all(m.vertices >= -1.0 / 16) and np.all(m.vertices <= 1.0 + 1.0 / 16)
        e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]).astype(np.int64)
        key, rev = e[:, 0] * (1 << 32) + e[:, 1], e[:, 1] * (1 << 32) + e[:, 0]
        assert len(np.unique(key)) == len(key) and np.array_equal(np.sort(key), np.sort(rev))
