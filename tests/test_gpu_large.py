"""GPU: a frame large enough that the extraction's scans take their widest path.

The counts the host reads back (leaves, records, points, dual cells, triangles, vertices) are
published by the scans that produce them. A scan of more than 2048 tiles of 2048 items scans its
tile totals recursively and publishes from its last kernel; the small test frames never get
there. The crown splash with a shallow pool keeps the full surface (about 3 M triangles, so
the vertex scan runs over more than 4.2 M corners) at 0.6 M particles. Its device extraction is
checked against the extraction in the reference's traversal order (on the host, from the same
tree, with none of those scans): same counts, vertices equal up to a few units in the last place
(the bisection orientation can differ in degenerate dual cells), and the same triangles by the
positions of their corners.
"""
import numpy as np
import pytest

from particleisosurface_amd import capi, synth
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


def _tri_rows(t):
    s = np.sort(t, axis=1)
    return s[np.lexsort((s[:, 2], s[:, 1], s[:, 0]))]


def test_large_frame_device_extraction_matches_host_order():
    from scipy.spatial import cKDTree
    xyz, r = synth.crown_splash(pool_layers=4)
    dev = gpu_context(xyz, r, precision=capi.FP32_KAHAN, mesh_order=0)
    host = gpu_context(xyz, r, precision=capi.FP32_KAHAN, mesh_order=1)
    dev.run_frame()
    host.run_frame()
    v, t = dev.mesh()
    v2, t2 = host.mesh()
    # wide enough: more than 2048 scan tiles of 2048 corners
    assert 3 * len(t) > 2048 * 2048, len(t)
    assert v.shape == v2.shape and t.shape == t2.shape
    st = dev.stats()
    assert st["n_tris"] == len(t) and st["n_verts"] == len(v)
    assert t.min() == 1 and t.max() == len(v) and len(np.unique(t)) == len(v)
    d, idx = cKDTree(v2).query(v)
    ulp = float(np.spacing(np.float32(np.abs(v2).max())))
    assert d.max() <= 4 * ulp, d.max()
    assert np.mean(d == 0) > 0.99
    # every device vertex replaced by its host counterpart: the same triangles. Vertices snapped to
    # one endpoint coincide, so triangles are compared by the positions of their corners.
    _, pos = np.unique(v2, axis=0, return_inverse=True)
    pos = pos.ravel()
    assert np.array_equal(_tri_rows(pos[idx[t - 1]]), _tri_rows(pos[t2 - 1]))
