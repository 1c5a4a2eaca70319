"""ctypes binding of the pisurf C ABI (include/pisurf.h).

This is how Python (tests, bench) reaches the HIP path. There is no fallback: if
libpisurf.so is missing or no GPU is visible, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpisurf.so")

FP32, FP32_KAHAN, FP64 = 0, 1, 2
QEF_OFF, QEF_REFERENCE, QEF_SVD64 = 0, 1, 2
EMPTY, INTERNAL, LEAF = 0, 1, 2

EXPORTS = [
    "pisurf_create", "pisurf_destroy", "pisurf_last_error", "pisurf_default_params", "pisurf_set_params",
    "pisurf_upload_particles", "pisurf_hashgrid_build", "pisurf_hashgrid_download", "pisurf_neighbors",
    "pisurf_neighbors_noalloc",
    "pisurf_box_particles", "pisurf_anisotropy", "pisurf_download_aniso", "pisurf_iso_constants",
    "pisurf_particle_normals", "pisurf_eval_points", "pisurf_uniform_grid", "pisurf_sweep_level", "pisurf_build_tree",
    "pisurf_tree_download", "pisurf_qef_solve", "pisurf_bisect_edges", "pisurf_dual_cells", "pisurf_extract",
    "pisurf_mesh_download",
    "pisurf_run_frame", "pisurf_run_frames", "pisurf_run_batch", "pisurf_get_stats", "pisurf_reset_stats", "pisurf_stream",
    "pisurf_run_tuning",
]


class Params(C.Structure):
    _fields_ = [
        ("radius", C.c_float), ("smooth_factor", C.c_float), ("iso_factor", C.c_float),
        ("use_ani", C.c_int32), ("use_poly6", C.c_int32), ("calc_normals", C.c_int32), ("oversample", C.c_int32),
        ("precision", C.c_int32), ("qef_mode", C.c_int32), ("qef_svd_tol", C.c_float),
        ("depth_max_override", C.c_int32), ("mesh_order", C.c_int32), ("sample_dedupe", C.c_int32),
        ("frames_in_flight", C.c_int32), ("level_graph", C.c_int32), ("reserved", C.c_int32 * 1),
    ]


class GridInfo(C.Structure):
    _fields_ = [("bounding", C.c_float * 6), ("xyz_cell_num", C.c_uint32 * 3), ("cell_num", C.c_int64),
                ("cell_size", C.c_float), ("n_invalid", C.c_int64)]


class TreeInfo(C.Structure):
    _fields_ = [("depth_max", C.c_int32), ("depth_min", C.c_int32), ("root_center", C.c_float * 3),
                ("root_half", C.c_float), ("n_nodes", C.c_int64), ("n_levels", C.c_int32), ("n_leaves", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("field_evals", C.c_int64), ("grad_evals", C.c_int64), ("qef_solves", C.c_int64),
                ("candidates", C.c_int64), ("sample_evals", C.c_int64), ("sample_candidates", C.c_int64),
                ("sample_launches", C.c_int64), ("fp64_eigensolves", C.c_int64),
                ("ms_grid", C.c_double), ("ms_aniso", C.c_double),
                ("ms_normals", C.c_double), ("ms_tree", C.c_double), ("ms_extract_host", C.c_double),
                ("ms_bisect", C.c_double), ("ms_total", C.c_double), ("ms_sample_kernel", C.c_double),
                ("n_dual_cells", C.c_int64), ("n_verts", C.c_int64), ("n_tris", C.c_int64), ("n_nodes", C.c_int64),
                ("sample_unique", C.c_int64), ("run_evals", C.c_int64), ("run_candidates", C.c_int64),
                ("run_launches", C.c_int64), ("ms_run_kernel", C.c_double)]


_lib = None

# void (*)(void* user, int32 frame, const pisurf_tree_info*, const double* v, int64 nv, const int32* t, int64 nt)
MESH_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                      C.POINTER(C.c_int32), C.c_int64)


def load():
    """Load libpisurf.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpisurf.so is not built: run `python -m particleisosurface_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.pisurf_last_error.restype = C.c_char_p
        L.pisurf_last_error.argtypes = [C.c_void_p]
        L.pisurf_stream.restype = C.c_void_p
        L.pisurf_stream.argtypes = [C.c_void_p]
        L.pisurf_run_frames.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_float), C.c_int32, MESH_CB, C.c_void_p, C.c_void_p]
        L.pisurf_run_batch.argtypes = L.pisurf_run_frames.argtypes
        L.pisurf_run_tuning.argtypes = [C.c_int32, C.c_int32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PisurfError(RuntimeError):
    pass


class Context:
    """One pisurf context on one GPU."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.pisurf_create(C.byref(h), device)
        if rc != 0:
            raise PisurfError("pisurf_create failed (%d): no usable GPU %d" % (rc, device))
        self.h = h
        self.n = 0
        self.params = Params()
        self.L.pisurf_default_params(C.byref(self.params))

    def close(self):
        if getattr(self, "h", None):
            self.L.pisurf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc, what):
        if rc != 0:
            raise PisurfError("%s failed (%d): %s" % (what, rc, self.L.pisurf_last_error(self.h).decode()))

    def set_params(self, radius, precision=FP32, use_ani=True, use_poly6=False, calc_normals=True, qef_mode=QEF_OFF,
                   qef_svd_tol=1e-3, depth_max_override=0, smooth_factor=2.0, iso_factor=1.9, mesh_order=0, sample_dedupe=True,
                   frames_in_flight=0, level_graph=False):
        p = Params()
        self.L.pisurf_default_params(C.byref(p))
        p.radius = radius
        p.mesh_order = mesh_order
        p.sample_dedupe = int(sample_dedupe)
        p.frames_in_flight = int(frames_in_flight)
        p.level_graph = int(level_graph)
        p.precision = precision
        p.use_ani = int(use_ani)
        p.use_poly6 = int(use_poly6)
        p.calc_normals = int(calc_normals)
        p.qef_mode = qef_mode
        p.qef_svd_tol = qef_svd_tol
        p.depth_max_override = depth_max_override
        p.smooth_factor = smooth_factor
        p.iso_factor = iso_factor
        self._c(self.L.pisurf_set_params(self.h, C.byref(p)), "set_params")
        self.params = p

    def upload(self, xyz, device_ptr=None, n=None):
        if device_ptr is not None:
            self._c(self.L.pisurf_upload_particles(self.h, C.c_void_p(device_ptr), C.c_int64(n), 1), "upload")
            self.n = n
            return
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self._c(self.L.pisurf_upload_particles(self.h, _p(x), C.c_int64(len(x)), 0), "upload")
        self.n = len(x)

    def hashgrid_build(self, bbox=None, inf_factor=4.0):
        gi = GridInfo()
        bb = None if bbox is None else np.ascontiguousarray(bbox, np.float32)
        self._c(self.L.pisurf_hashgrid_build(self.h, _p(bb), C.c_float(inf_factor), C.byref(gi)), "hashgrid_build")
        return dict(bounding=np.array(gi.bounding[:], np.float32), cellnum3=np.array(gi.xyz_cell_num[:], np.uint32),
                    cellnum=gi.cell_num, cellsize=gi.cell_size, n_invalid=gi.n_invalid)

    def hashgrid_download(self, ncell):
        il = np.zeros(self.n, np.int32)
        s = np.zeros(ncell, np.int32)
        e = np.zeros(ncell, np.int32)
        self._c(self.L.pisurf_hashgrid_download(self.h, _p(il), _p(s), _p(e)), "hashgrid_download")
        return il, s, e

    def neighbors(self, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        off = np.zeros(len(q) + 1, np.int64)
        tot = C.c_int64()
        self._c(self.L.pisurf_neighbors(self.h, _p(q), C.c_int64(len(q)), _p(off), None, C.c_int64(0),
                                        C.byref(tot)), "neighbors")
        ids = np.zeros(max(tot.value, 1), np.int32)
        self._c(self.L.pisurf_neighbors(self.h, _p(q), C.c_int64(len(q)), _p(off), _p(ids), C.c_int64(tot.value),
                                        C.byref(tot)), "neighbors")
        return off, ids[:tot.value]

    def box_particles(self, b1, b2):
        b1 = np.ascontiguousarray(b1, np.float32)
        b2 = np.ascontiguousarray(b2, np.float32)
        tot = C.c_int64()
        self._c(self.L.pisurf_box_particles(self.h, _p(b1), _p(b2), None, C.c_int64(0), C.byref(tot)), "box")
        ids = np.zeros(max(tot.value, 1), np.int32)
        self._c(self.L.pisurf_box_particles(self.h, _p(b1), _p(b2), _p(ids), C.c_int64(tot.value), C.byref(tot)), "box")
        return ids[:tot.value]

    def anisotropy(self):
        self._c(self.L.pisurf_anisotropy(self.h), "anisotropy")

    def aniso(self):
        xm = np.zeros((self.n, 3))
        G = np.zeros((self.n, 9))
        d = np.zeros(self.n)
        self._c(self.L.pisurf_download_aniso(self.h, _p(xm), _p(G), _p(d)), "download_aniso")
        return xm, G, d

    def iso_constants(self):
        a, b = C.c_double(), C.c_double()
        self._c(self.L.pisurf_iso_constants(self.h, C.byref(a), C.byref(b)), "iso_constants")
        return a.value, b.value

    def particle_normals(self):
        n = np.zeros((self.n, 3))
        self._c(self.L.pisurf_particle_normals(self.h, _p(n)), "particle_normals")
        return n

    def eval_points(self, q, grad=False):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        f = np.zeros(len(q))
        g = np.zeros((len(q), 3)) if grad else None
        self._c(self.L.pisurf_eval_points(self.h, _p(q), C.c_int64(len(q)), _p(f), _p(g)), "eval_points")
        return (f, g) if grad else f

    def sweep_level(self, centers, halfs, depths, want_samples=False):
        centers = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        halfs = np.ascontiguousarray(halfs, np.float32)
        depths = np.ascontiguousarray(depths, np.int32)
        n = len(halfs)
        t = np.zeros(n, np.uint8)
        f = np.zeros(n)
        cv = np.zeros(n, np.float32)
        sc = np.zeros(n, np.int32)
        s = np.zeros((n, 27, 4)) if want_samples else None
        g = np.zeros((n, 27, 3)) if want_samples else None
        self._c(self.L.pisurf_sweep_level(self.h, _p(centers), _p(halfs), _p(depths), C.c_int64(n), _p(t), _p(f),
                                          _p(cv), _p(sc), _p(s), _p(g)), "sweep_level")
        return (t, f, cv, sc, s, g) if want_samples else (t, f, cv, sc)

    def build_tree(self):
        ti = TreeInfo()
        self._c(self.L.pisurf_build_tree(self.h, C.byref(ti)), "build_tree")
        return dict(depth_max=ti.depth_max, depth_min=ti.depth_min, root_center=np.array(ti.root_center[:]),
                    root_half=ti.root_half, n_nodes=ti.n_nodes, n_levels=ti.n_levels)

    def tree_nodes(self, n):
        d = dict(center=np.zeros((n, 3), np.float32), half=np.zeros(n, np.float32), depth=np.zeros(n, np.int32),
                 type=np.zeros(n, np.uint8), id=np.zeros(n, np.uint64), node=np.zeros((n, 4)),
                 child=np.zeros(n, np.int32))
        self._c(self.L.pisurf_tree_download(self.h, _p(d["center"]), _p(d["half"]), _p(d["depth"]), _p(d["type"]),
                                            _p(d["id"]), _p(d["node"]), _p(d["child"])), "tree_download")
        return d

    def qef_solve(self, centers, halfs, samples, grads, mode):
        centers = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        halfs = np.ascontiguousarray(halfs, np.float32)
        s = np.ascontiguousarray(samples, np.float64)
        g = np.ascontiguousarray(grads, np.float64)
        out = np.zeros((len(halfs), 4))
        self._c(self.L.pisurf_qef_solve(self.h, _p(centers), _p(halfs), _p(s), _p(g), C.c_int64(len(halfs)), mode,
                                        _p(out)), "qef_solve")
        return out

    def bisect_edges(self, ends):
        e = np.ascontiguousarray(ends, np.float64).reshape(-1, 2, 4)
        out = np.zeros((len(e), 3))
        nev = np.zeros(len(e), np.int32)
        self._c(self.L.pisurf_bisect_edges(self.h, _p(e), C.c_int64(len(e)), _p(out), _p(nev)), "bisect_edges")
        return out, nev

    def extract(self):
        nv, nt = C.c_int64(), C.c_int64()
        self._c(self.L.pisurf_extract(self.h, C.byref(nv), C.byref(nt)), "extract")
        return nv.value, nt.value

    def mesh(self):
        st = self.stats()
        v = np.zeros((st["n_verts"], 3))
        t = np.zeros((st["n_tris"], 3), np.int32)
        self._c(self.L.pisurf_mesh_download(self.h, _p(v), _p(t)), "mesh_download")
        return v, t

    def run_frame(self):
        ti = TreeInfo()
        self._c(self.L.pisurf_run_frame(self.h, C.byref(ti)), "run_frame")
        return dict(depth_max=ti.depth_max, depth_min=ti.depth_min, n_nodes=ti.n_nodes, n_levels=ti.n_levels)

    def uniform_grid(self, values=True):
        """UniformGrid sampling: (scalars (dims2, dims1, dims0) float32 or None, box[6], spacing)."""
        dims = (C.c_int32 * 3)()
        box = (C.c_float * 6)()
        sp = C.c_float()
        self._c(self.L.pisurf_uniform_grid(self.h, dims, box, C.byref(sp), None), "uniform_grid")
        d = [dims[0], dims[1], dims[2]]
        out = None
        if values:
            out = np.empty((d[2], d[1], d[0]), np.float32)
            self._c(self.L.pisurf_uniform_grid(self.h, None, None, None, _p(out)), "uniform_grid")
        return out, np.array(list(box), np.float32), sp.value

    def run_batch(self, frames, radii, on_mesh=None):
        """Mesh frames together (pisurf_run_batch): lockstep levels, one run-queue launch per level
        for every frame. Arguments and result as run_frames."""
        return self.run_frames(frames, radii, on_mesh, entry="pisurf_run_batch")

    def run_frames(self, frames, radii, on_mesh=None, entry="pisurf_run_frames", frames_in_flight=None):
        """Mesh a batch of frames in one call. frames: list of (n, 3) float32 arrays;
        on_mesh(k, verts (nv, 3) float64, tris (nt, 3) int32, tree dict) is called after each frame,
        in frame order. frames_in_flight (None: the context's parameter) sets how many frames are
        meshed at once. Returns the per-frame statistics."""
        if frames_in_flight is not None:
            self.params.frames_in_flight = int(frames_in_flight)
            self._c(self.L.pisurf_set_params(self.h, C.byref(self.params)), "set_params")
        nf = len(frames)
        frames = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1, 3) for f in frames]
        ptrs = (C.c_void_p * max(nf, 1))(*[f.ctypes.data for f in frames])
        ns = (C.c_int64 * max(nf, 1))(*[len(f) for f in frames])
        rs = (C.c_float * max(nf, 1))(*[float(r) for r in radii])
        per = (Stats * max(nf, 1))()
        err = []

        def cb(user, k, ti, v, nv, t, nt):
            try:
                if on_mesh is not None:
                    tip = C.cast(ti, C.POINTER(TreeInfo)).contents
                    tree = dict(depth_max=tip.depth_max, depth_min=tip.depth_min, n_nodes=tip.n_nodes,
                                n_levels=tip.n_levels)
                    va = np.ctypeslib.as_array(v, (nv, 3)).copy() if nv else np.zeros((0, 3))
                    ta = np.ctypeslib.as_array(t, (nt, 3)).copy() if nt else np.zeros((0, 3), np.int32)
                    on_mesh(k, va, ta, tree)
            except Exception as e:  # an exception must not unwind through the C frames
                err.append(e)

        fn = MESH_CB(cb)
        self._c(getattr(self.L, entry)(self.h, nf, ptrs, ns, rs, 0, fn, None, C.cast(per, C.c_void_p)), entry)
        if err:
            raise err[0]
        if frames:
            self.params.radius = float(radii[-1])
        return [{k: getattr(per[i], k) for k, _ in Stats._fields_} for i in range(nf)]

    def stats(self):
        s = Stats()
        self._c(self.L.pisurf_get_stats(self.h, C.byref(s)), "get_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def reset_stats(self):
        self._c(self.L.pisurf_reset_stats(self.h), "reset_stats")

    def stream(self):
        return self.L.pisurf_stream(self.h)
