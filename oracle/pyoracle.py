"""ctypes view of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
and only as the checker / CPU baseline; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpisurf_oracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("radius", C.c_float), ("smooth_factor", C.c_float), ("iso_factor", C.c_float),
        ("use_ani", C.c_int), ("use_poly6", C.c_int), ("calc_normals", C.c_int), ("oversample", C.c_int),
        ("fp64", C.c_int), ("qef_mode", C.c_int), ("qef_svd_tol", C.c_float),
        ("depth_max_override", C.c_int), ("threads", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L = _lib
        L.orc_create.restype = P
        L.orc_create.argtypes = [P, C.c_int64, C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        for name in ["orc_hashgrid_cells", "orc_neighbors", "orc_box_particles", "orc_eval_count"]:
            getattr(L, name).restype = C.c_int64
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, xyz, radius, fp64=False, use_poly6=False, use_ani=True, calc_normals=True,
                 qef_mode=0, qef_svd_tol=1e-3, depth_max_override=0, threads=0, smooth_factor=2.0,
                 iso_factor=1.9, oversample=2):
        self.L = lib()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.n = len(self.xyz)
        pr = OrcParams(radius, smooth_factor, iso_factor, int(use_ani), int(use_poly6), int(calc_normals),
                       oversample, int(fp64), qef_mode, qef_svd_tol, depth_max_override, threads)
        self.params = pr
        self.ns = (oversample + 1) ** 3
        self.h = self.L.orc_create(_p(self.xyz), self.n, C.byref(pr))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(C.c_void_p(self.h))
            self.h = None

    def _c(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d)" % (what, rc))

    # H1-H5
    def hashgrid_build(self, bbox=None, inf_factor=4.0):
        b = np.zeros(6, np.float32)
        cn = np.zeros(3, np.uint32)
        tot = C.c_int64()
        cs = C.c_float()
        ninv = C.c_int64()
        bb = None if bbox is None else np.ascontiguousarray(bbox, np.float32)
        self._c(self.L.orc_hashgrid_build(C.c_void_p(self.h), _p(bb), C.c_float(inf_factor), _p(b), _p(cn),
                                          C.byref(tot), C.byref(cs), C.byref(ninv)), "hashgrid_build")
        return dict(bounding=b, cellnum3=cn, cellnum=tot.value, cellsize=cs.value, n_invalid=ninv.value)

    def index_list(self):
        a = np.zeros(self.n, np.int32)
        self._c(self.L.orc_hashgrid_index_list(C.c_void_p(self.h), _p(a)), "index_list")
        return a

    def cells(self):
        k = self.L.orc_hashgrid_cells(C.c_void_p(self.h), None, None, None)
        keys = np.zeros(k, np.int64)
        s = np.zeros(k, np.int32)
        e = np.zeros(k, np.int32)
        self.L.orc_hashgrid_cells(C.c_void_p(self.h), _p(keys), _p(s), _p(e))
        return keys, s, e

    def neighbors(self, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        off = np.zeros(len(q) + 1, np.int64)
        tot = self.L.orc_neighbors(C.c_void_p(self.h), _p(q), len(q), _p(off), None, 0)
        ids = np.zeros(max(tot, 1), np.int32)
        self.L.orc_neighbors(C.c_void_p(self.h), _p(q), len(q), _p(off), _p(ids), tot)
        return off, ids[:tot]

    def box_particles(self, b1, b2):
        b1 = np.ascontiguousarray(b1, np.float32)
        b2 = np.ascontiguousarray(b2, np.float32)
        n = self.L.orc_box_particles(C.c_void_p(self.h), _p(b1), _p(b2), None, 0)
        ids = np.zeros(max(n, 1), np.int32)
        self.L.orc_box_particles(C.c_void_p(self.h), _p(b1), _p(b2), _p(ids), n)
        return ids[:n]

    # H6-H12
    def anisotropy(self):
        self._c(self.L.orc_anisotropy(C.c_void_p(self.h)), "anisotropy")

    def aniso(self):
        xm = np.zeros((self.n, 3))
        G = np.zeros((self.n, 9))
        d = np.zeros(self.n)
        self._c(self.L.orc_get_aniso(C.c_void_p(self.h), _p(xm), _p(G), _p(d)), "get_aniso")
        return xm, G, d

    def set_aniso(self, xm, G, d):
        xm = np.ascontiguousarray(xm, np.float64)
        G = np.ascontiguousarray(G, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        self._c(self.L.orc_set_aniso(C.c_void_p(self.h), _p(xm), _p(G), _p(d)), "set_aniso")

    def iso_constants(self):
        ms, iso = C.c_double(), C.c_double()
        self._c(self.L.orc_iso_constants(C.c_void_p(self.h), C.byref(ms), C.byref(iso)), "iso_constants")
        return ms.value, iso.value

    def constants(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_get_constants(C.c_void_p(self.h), C.byref(a), C.byref(b))
        return a.value, b.value

    def particle_normals(self):
        n = np.zeros((self.n, 3))
        self._c(self.L.orc_particle_normals(C.c_void_p(self.h), _p(n)), "particle_normals")
        return n

    def eval_points(self, q, grad=False):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        f = np.zeros(len(q))
        g = np.zeros((len(q), 3)) if grad else None
        self._c(self.L.orc_eval_points(C.c_void_p(self.h), _p(q), len(q), _p(f), _p(g)), "eval_points")
        return (f, g) if grad else f

    def grid_eval(self, center, half):
        c = np.ascontiguousarray(center, np.float32)
        s = np.zeros((self.ns, 4))
        g = np.zeros((self.ns, 3))
        sc = C.c_int()
        self._c(self.L.orc_grid_eval(C.c_void_p(self.h), _p(c), C.c_float(half), _p(s), _p(g), C.byref(sc)),
                "grid_eval")
        return s, g, bool(sc.value)

    def classify_nodes(self, centers, halfs, depths):
        centers = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        halfs = np.ascontiguousarray(halfs, np.float32)
        depths = np.ascontiguousarray(depths, np.int32)
        n = len(halfs)
        t = np.zeros(n, np.uint8)
        f = np.zeros(n)
        cv = np.zeros(n, np.float32)
        sc = np.zeros(n, np.int32)
        self._c(self.L.orc_classify_nodes(C.c_void_p(self.h), _p(centers), _p(halfs), _p(depths), n, _p(t), _p(f),
                                          _p(cv), _p(sc)), "classify_nodes")
        return t, f, cv, sc

    def build_tree(self):
        self._c(self.L.orc_build_tree(C.c_void_p(self.h)), "build_tree")

    def tree_info(self):
        dmax, dmin = C.c_int(), C.c_int()
        rc = np.zeros(3, np.float32)
        rh = C.c_float()
        nn = C.c_int64()
        self.L.orc_tree_info(C.c_void_p(self.h), C.byref(dmax), C.byref(dmin), _p(rc), C.byref(rh), C.byref(nn))
        return dict(depth_max=dmax.value, depth_min=dmin.value, root_center=rc, root_half=rh.value, n_nodes=nn.value)

    def tree_nodes(self):
        n = self.tree_info()["n_nodes"]
        d = dict(center=np.zeros((n, 3), np.float32), half=np.zeros(n, np.float32), depth=np.zeros(n, np.int32),
                 type=np.zeros(n, np.uint8), id=np.zeros(n, np.uint64), node=np.zeros((n, 4)),
                 child=np.zeros(n, np.int32))
        self.L.orc_tree_nodes(C.c_void_p(self.h), _p(d["center"]), _p(d["half"]), _p(d["depth"]), _p(d["type"]),
                              _p(d["id"]), _p(d["node"]), _p(d["child"]))
        return d

    def qef(self, center, half, samples, grads, mode):
        c = np.ascontiguousarray(center, np.float32)
        s = np.ascontiguousarray(samples, np.float64)
        g = np.ascontiguousarray(grads, np.float64)
        out = np.zeros(4)
        self._c(self.L.orc_qef(C.c_void_p(self.h), _p(c), C.c_float(half), _p(s), _p(g), mode, _p(out)), "qef")
        return out

    def bisect_edges(self, ends):
        e = np.ascontiguousarray(ends, np.float64).reshape(-1, 2, 4)
        out = np.zeros((len(e), 3))
        nev = np.zeros(len(e), np.int32)
        self._c(self.L.orc_bisect_edges(C.c_void_p(self.h), _p(e), len(e), _p(out), _p(nev)), "bisect")
        return out, nev

    def extract(self):
        self._c(self.L.orc_extract(C.c_void_p(self.h)), "extract")

    def mesh(self):
        nv, nt, nd = C.c_int64(), C.c_int64(), C.c_int64()
        self.L.orc_mesh_counts(C.c_void_p(self.h), C.byref(nv), C.byref(nt), C.byref(nd))
        v = np.zeros((nv.value, 3))
        t = np.zeros((nt.value, 3), np.int32)
        self.L.orc_mesh_get(C.c_void_p(self.h), _p(v), _p(t))
        return v, t

    def dual_cells(self):
        nv, nt, nd = C.c_int64(), C.c_int64(), C.c_int64()
        self.L.orc_mesh_counts(C.c_void_p(self.h), C.byref(nv), C.byref(nt), C.byref(nd))
        n8 = np.zeros((nd.value, 8), np.int32)
        si = np.zeros(nd.value, np.uint8)
        self.L.orc_dual_cells(C.c_void_p(self.h), _p(n8), _p(si))
        return n8, si

    def write_obj(self, path):
        self._c(self.L.orc_write_obj(C.c_void_p(self.h), path.encode()), "write_obj")

    def run(self):
        t = np.zeros(7)
        self._c(self.L.orc_run(C.c_void_p(self.h), _p(t)), "run")
        return dict(zip(["hashgrid", "anisotropy", "constants", "normals", "tree", "extract", "total"], t))


def num_threads():
    return lib().orc_num_threads()


def eigensolve(a, fp64=True):
    """The oracle's Jacobi eigensolver on a symmetric 3x3 or 4x4 matrix."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    lam = np.zeros(n)
    Q = np.zeros((n, n))
    rc = lib().orc_test_eig(_p(a), n, int(fp64), _p(lam), _p(Q))
    if rc != 0:
        raise RuntimeError("orc_test_eig failed")
    return lam, Q


def g_from_cov(cov6, fp64=True):
    """compute_G_Yus tail (clamped spectrum -> G) on a given covariance."""
    c = np.ascontiguousarray(cov6, np.float64)
    G = np.zeros(9)
    det = C.c_double()
    lib().orc_test_g_from_cov(_p(c), int(fp64), _p(G), C.byref(det))
    return G.reshape(3, 3, order="F"), det.value
