"""Build libpisurf.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python -m particleisosurface_amd.build   (or call build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libpisurf.so")
SOURCES = ["k_grid.hip", "k_aniso.hip", "k_field.hip", "k_octree.hip", "k_vertex.hip", "k_extract.hip", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PISURF_ARCH", "gfx950")
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, msg))
    return out, msg


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        res = list(ex.map(_compile, SOURCES))
    for _, msg in res:
        if verbose and msg.strip():
            print(msg)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + [o for o, _ in res]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(LIB + ".tmp", LIB)
    # the C++ mirror of the reference classes, exercised by a small driver program
    demo = os.path.join(HERE, "pisurf_facade_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "facade_demo.cpp"), "-o", demo + ".tmp",
           "-L" + HERE, "-lpisurf", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade demo build failed:\n" + r.stdout + r.stderr)
    os.replace(demo + ".tmp", demo)
    # the reference's extraction stage (calc_vertices on OpenMP threads, generate_mesh) driven
    # through the C++ mirror; float operations are not contracted, as in the reference's build
    ext = os.path.join(HERE, "pisurf_extract_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fopenmp", "-ffp-contract=off", os.path.join(CSRC, "facade_extract.cpp"),
           "-o", ext + ".tmp", "-L" + HERE, "-lpisurf", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("extraction driver build failed:\n" + r.stdout + r.stderr)
    os.replace(ext + ".tmp", ext)
    # the frame driver over a data directory (controlData.json, CSV / HDF5 / raw frames)
    run = os.path.join(HERE, "pisurf_run")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "pisurf_run.cpp"), "-o", run + ".tmp",
           "-L" + HERE, "-lpisurf", "-ldl", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("frame driver build failed:\n" + r.stdout + r.stderr)
    os.replace(run + ".tmp", run)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
