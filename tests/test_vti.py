"""VTK image-data writer / reader of the dense-grid output (host logic, no GPU)."""
import numpy as np

from particleisosurface_amd.vti import read_vti, write_vti


def test_vti_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 4, 3)).astype(np.float32)
    p = tmp_path / "g.vti"
    write_vti(str(p), a, (0.25, -1.0, 3.5), 0.005)
    b, origin, spacing = read_vti(str(p))
    assert b.shape == (5, 4, 3) and np.array_equal(a, b)
    assert np.allclose(origin, [0.25, -1.0, 3.5]) and np.allclose(spacing, 0.005)
    head = p.read_bytes()[:400].decode(errors="replace")
    # x fastest: WholeExtent lists x, y, z extents = shape reversed minus one
    assert 'WholeExtent="0 2 0 3 0 4"' in head and 'type="ImageData"' in head


def test_vti_single_point(tmp_path):
    a = np.array([[[1.5]]], np.float32)
    p = tmp_path / "one.vti"
    write_vti(str(p), a, (0, 0, 0), 1.0)
    b, _, _ = read_vti(str(p))
    assert np.array_equal(a, b)
