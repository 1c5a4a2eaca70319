"""C ABI checks that need no GPU: exports, argument validation, and the product/oracle boundary."""
import ctypes as C
import os
import re

import pytest

from particleisosurface_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "pisurf.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void\*)\s+(pisurf_\w+)\s*\(", hdr, re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.EXPORTS)


def test_null_and_default_handling():
    lib = capi.load()
    assert lib.pisurf_destroy(None) == -1
    assert lib.pisurf_last_error(None) == b"null context"
    assert lib.pisurf_default_params(None) == -1
    p = capi.Params()
    assert lib.pisurf_default_params(C.byref(p)) == 0
    assert (p.smooth_factor, p.use_ani, p.use_poly6, p.calc_normals, p.oversample) == (2.0, 1, 0, 1, 2)
    assert abs(p.iso_factor - 1.9) < 1e-6 and p.qef_mode == 0 and p.precision == 0
    out = C.c_void_p()
    assert lib.pisurf_create(None, 0) == -1
    assert lib.pisurf_create(C.byref(out), -1) != 0 and not out.value
    for fn in ("pisurf_anisotropy", "pisurf_extract", "pisurf_reset_stats"):
        assert getattr(lib, fn)(None) == -1


def test_no_gpu_fails_loudly():
    lib = capi.load()
    n = C.c_int()
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.PisurfError):
        capi.Context(0)


def test_product_never_uses_the_oracle():
    pkg = os.path.join(ROOT, "particleisosurface_amd")
    bad = []
    for dp, _, fns in os.walk(pkg):
        for fn in fns:
            if fn.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dp, fn)).read()
                if re.search(r"import oracle|from oracle|pisurf_oracle|orc_\w+\(|pyoracle", src):
                    bad.append(fn)
    for fn in ("bench.py",):
        src = open(os.path.join(ROOT, fn)).read()
        # bench.py may use the oracle only inside its CPU-baseline leg
        body = src.split("def cpu_baseline", 1)[0]
        assert "oracle" not in body.replace("the oracle", "")
    assert not bad, bad
