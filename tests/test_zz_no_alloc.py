"""pisurf_neighbors_noalloc against pisurf_neighbors, through the C ABI.

The no-allocation entry point takes device buffers and keeps its working memory in the library's
fixed device array, which holds the previous call's contents: two calls in a row, the second
with fewer queries, must each equal the allocating entry point's result; a call whose total
exceeds the ids capacity writes nothing and reports the total; a call above the query limit is
refused. The test's own device buffers come from the HIP runtime the library itself links (the
same process-wide runtime), allocated before and freed after the calls under test.
"""
import ctypes as C

import numpy as np
import pytest

from particleisosurface_amd import capi, synth
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu

MAX_Q = 1 << 20
H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, n, dtype):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(ctx, q, nq, off, ids, cap):
    tot = C.c_int64(-1)
    rc = capi.load().pisurf_neighbors_noalloc(ctx.h, q.p, C.c_int64(nq), off.p, ids.p, C.c_int64(cap), C.byref(tot))
    return rc, tot.value


def test_neighbors_noalloc_matches_allocating_entry_point():
    capi.load()
    hip = C.CDLL("libamdhip64.so.7")  # the runtime libpisurf.so is linked against, already loaded
    xyz, r = synth.dam_break()
    ctx = gpu_context(xyz, r)
    ctx.hashgrid_build()
    rng = np.random.default_rng(17)
    cap = 12_000_000
    q_d = DevBuf(hip, 3 * 4 * 60000)
    off_d = DevBuf(hip, 8 * (MAX_Q + 1))
    ids_d = DevBuf(hip, 4 * cap)
    try:
        for nq in (60000, 7001):  # the second call is smaller and must not see the first's counts
            q = (xyz[rng.choice(len(xyz), nq)] + rng.uniform(-6 * r, 6 * r, (nq, 3))).astype(np.float32)
            q_d.put(q)
            rc, tot = _noalloc(ctx, q_d, nq, off_d, ids_d, cap)
            assert rc == 0, capi.load().pisurf_last_error(ctx.h)
            off_ref, ids_ref = ctx.neighbors(q)
            assert tot == off_ref[-1] > 0
            assert np.array_equal(off_d.get(nq + 1, np.int64), off_ref)
            assert np.array_equal(ids_d.get(tot, np.int32), ids_ref)
        # capacity below the total: nothing written, the total reported
        rc, tot = _noalloc(ctx, q_d, 7001, off_d, ids_d, 10)
        assert rc == -1 and tot == off_ref[-1]  # PISURF_E_INVALID
        # above the query limit: refused before any launch
        rc, _ = _noalloc(ctx, q_d, MAX_Q + 1, off_d, ids_d, cap)
        assert rc == -4  # PISURF_E_UNSUPPORTED
    finally:
        for b in (q_d, off_d, ids_d):
            b.free()
