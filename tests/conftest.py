import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    lib = os.path.join(ROOT, "oracle", "libpisurf_oracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def cfg1():
    from particleisosurface_amd import synth
    return synth.dam_break()


@pytest.fixture(scope="session")
def sphere():
    from particleisosurface_amd import synth
    return synth.lattice_sphere(radius=0.05, r=0.005)


def gpu_context(xyz, r, **kw):
    from particleisosurface_amd import capi
    ctx = capi.Context(0)
    ctx.set_params(r, **kw)
    ctx.upload(xyz)
    return ctx


def field_tol_ok(S, S_ref, rel, floor):
    """|S - S_ref| <= rel * S_ref + floor (S = density sum, positive)."""
    return np.abs(S - S_ref) <= rel * np.abs(S_ref) + floor
