"""Input front-end of the frame driver (controlData.json, CSV, ShonDy HDF5, raw float32), run on
the host through `pisurf_run --load-only` (no GPU)."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from particleisosurface_amd import capi

RUN = os.path.join(os.path.dirname(capi.LIB_PATH), "pisurf_run")
H5_CANDIDATES = [os.environ.get("PISURF_HDF5_LIB", ""), "/opt/conda/lib/libhdf5.so", "libhdf5.so"]

pytestmark = pytest.mark.skipif(not os.path.exists(RUN), reason="pisurf_run not built")


def _hdf5():
    for p in H5_CANDIDATES:
        if not p:
            continue
        try:
            L = C.CDLL(p)
            if L.H5open() >= 0:
                return L, p
        except OSError:
            pass
    return None, None


def write_h5(path, position, radius):
    """ShonDy layout: flat float64 'position' [3N] and float64 'particleRadius' [N]."""
    L, _ = _hdf5()
    for fn in ("H5Fcreate", "H5Screate_simple", "H5Dcreate2"):
        getattr(L, fn).restype = C.c_int64
    L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, C.c_int64, C.c_int64]
    L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]
    L.H5Dcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
    L.H5Dwrite.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    for fn in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(L, fn).argtypes = [C.c_int64]
    dbl = C.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    f = L.H5Fcreate(str(path).encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    assert f >= 0
    for name, arr in (("position", np.ascontiguousarray(position, np.float64).ravel()),
                      ("particleRadius", np.ascontiguousarray(radius, np.float64).ravel())):
        dims = (C.c_uint64 * 1)(len(arr))
        sp = L.H5Screate_simple(1, dims, None)
        ds = L.H5Dcreate2(f, name.encode(), dbl, sp, 0, 0, 0)
        assert ds >= 0 and L.H5Dwrite(ds, dbl, 0, 0, 0, arr.ctypes.data) >= 0
        L.H5Dclose(ds)
        L.H5Sclose(sp)
    L.H5Fclose(f)


def run_load_only(d, env=None):
    res = subprocess.run([RUN, str(d), "--load-only"], capture_output=True, text=True, timeout=60,
                         env=dict(os.environ, **(env or {})))
    return res


def _frames(n=50, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, 3)).astype(np.float32)


def test_csv_and_raw_frames(tmp_path):
    a = _frames(40, 1)
    b = _frames(25, 2)
    # a ParaView-style export: quoted "Points:0.." titles, a radius column, extra columns
    with open(tmp_path / "f1.csv", "w") as f:
        f.write('"id","Points:0","Points:1","Points:2","radius"\n')
        for i, p in enumerate(a):
            f.write("%d,%r,%r,%r,0.01\n" % (i, float(p[0]), float(p[1]), float(p[2])))
    b.tofile(tmp_path / "f2.f32")
    (tmp_path / "controlData.json").write_text(json.dumps(
        {"DATA_FILE": "f1.csv,f2.f32", "RADIUS": 0.01, "OUTPUT_TYPE": "obj", "USE_ANI": 1, "ISO_FACTOR": 1.5}))
    res = run_load_only(tmp_path)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout)
    assert [f["name"] for f in out["frames"]] == ["f1.csv", "f2.f32"]
    assert [f["n"] for f in out["frames"]] == [40, 25]
    assert np.allclose(out["frames"][0]["sum"], a.astype(np.float64).sum(0), rtol=1e-6, atol=1e-6)
    assert np.allclose(out["frames"][1]["sum"], b.astype(np.float64).sum(0), rtol=1e-6, atol=1e-6)
    assert out["frames"][0]["radius"] == pytest.approx(0.01)
    assert out["output_type"] == "obj" and out["iso_factor"] == pytest.approx(1.5)
    assert out["smooth_factor"] == pytest.approx(2.0) and out["use_ani"] == 1


def test_csv_axis_titles_and_variable_radius(tmp_path):
    a = _frames(10, 3)
    # plain x / y / z titles; no RADIUS in the config: the radius column is read, and equal radii
    # make the frame constant-radius (src/main.cpp:358-366)
    with open(tmp_path / "p.csv", "w") as f:
        f.write("X,y,Z,R\n")
        for p in a:
            f.write("%r,%r,%r,0.004\n" % tuple(float(v) for v in p))
    (tmp_path / "controlData.json").write_text('{"SUFFIX": ".csv", "USE_POLY6": false}')
    out = json.loads(run_load_only(tmp_path).stdout)
    assert out["frames"][0]["n"] == 10 and out["frames"][0]["radius"] == pytest.approx(0.004)
    assert out["output_type"] == "ply" and out["use_poly6"] == 0
    # radii that really vary are refused
    with open(tmp_path / "p.csv", "a") as f:
        f.write("0,0,0,0.5\n")
    res = run_load_only(tmp_path)
    assert res.returncode == 1 and "radii vary" in res.stderr


def test_errors_are_reported(tmp_path):
    res = run_load_only(tmp_path)
    assert res.returncode == 1 and "controlData.json" in res.stderr
    (tmp_path / "controlData.json").write_text('{"DATA_FILE": "a.csv", "RADIUS": 0.01}')
    (tmp_path / "a.csv").write_text("u,v,w\n1,2,3\n")
    res = run_load_only(tmp_path)
    assert res.returncode == 1 and "axis column" in res.stderr
    (tmp_path / "controlData.json").write_text('{"DATA_FILE": "a.csv", "RADIUS": [1]}')
    res = run_load_only(tmp_path)
    assert res.returncode == 1


def test_hdf5_frame(tmp_path):
    L, path = _hdf5()
    if L is None:
        pytest.skip("libhdf5 not available")
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 0.3, (123, 3))
    write_h5(tmp_path / "frame_0001.h5", pos, np.full(123, 0.005))
    (tmp_path / "controlData.json").write_text('{"DATA_FILE": "frame_0001.h5", "OUTPUT_TYPE": "obj"}')
    res = run_load_only(tmp_path, {"PISURF_HDF5_LIB": path})
    assert res.returncode == 0, res.stderr
    fr = json.loads(res.stdout)["frames"][0]
    assert fr["n"] == 123 and fr["radius"] == pytest.approx(0.005)
    assert np.allclose(fr["sum"], pos.astype(np.float32).astype(np.float64).sum(0), rtol=1e-7)
