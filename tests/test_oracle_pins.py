"""Pins of the CPU oracle against what the mathematics fixes (no GPU needed).

The reference ships no tests, fixtures or golden vectors, so every oracle function is checked
here against closed forms, brute force over all particles, a library eigensolver, or structural
invariants, each chosen so that a plausible mistake (wrong multiplicity, sign, index, order,
truncation rule) fails.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle, eigensolve, g_from_cov
from particleisosurface_amd import synth


def m4(q):
    """Cubic B-spline of Monaghan & Lattanzio (1985) normalised to W(0)=1, support q<2."""
    q = np.asarray(q, np.float64)
    return np.where(q <= 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def w_cubic(d, h):
    """Cubic spline with support h normalised to w(0)=1 (1 - 6u^2 + 6u^3, 2(1-u)^3), u = d/h."""
    u = np.asarray(d, np.float64) / h
    return np.where(u < 0.5, 1 - 6 * u ** 2 + 6 * u ** 3, np.where(u < 1, 2 * (1 - u) ** 3, 0.0))


# ---------------------------------------------------------------- H6 / H7 / H10 closed forms
@pytest.mark.parametrize("r", [0.005, 0.002, 0.001, 0.0005])
def test_iso_constants_closed_form(r):
    xyz = np.array([[0, 0, 0]], np.float32)
    mscale = 1 + 6 * 0.25 + 12 * 0.25 * (2 - math.sqrt(2)) ** 3 + 8 * 0.25 * (2 - math.sqrt(3)) ** 3
    assert abs(mscale - 3.141506) < 1e-6
    for fp64 in (False, True):
        o = Oracle(xyz, r, fp64=fp64)
        ms, iso = o.iso_constants()
        h2, sigma = o.constants()
        h = 2 * np.float64(np.float32(r) * np.float32(2.0)) / 2
        tol = 1e-12 if fp64 else 2e-7
        assert abs(h2 / h ** 2 - 1) < tol
        assert abs(sigma * h ** 6 - 1) < tol
        # the reference's sqrt2 / sqrt3 are float constants (include/evaluator.h:16-17)
        m32 = 1 + 6 * m4(1.0) + 12 * m4(float(np.float32(math.sqrt(2)))) + 8 * m4(float(np.float32(math.sqrt(3))))
        assert abs(ms / sigma / m32 - 1) < (1e-12 if fp64 else 1e-6)
        assert abs(ms / sigma / mscale - 1) < 1e-6
        assert abs(iso / sigma / float(m4(np.float32(0.01))) - 1) < (1e-12 if fp64 else 1e-6)
        assert abs(iso / sigma - 0.99985075) < (1e-10 if fp64 else 2e-7)  # ratio 0.01 is a float too


def test_isolated_particles_closed_form():
    """Far-apart particles: xMean = x, G = I, det = 1 and f(q) = iso - sigma B(|q-x|/2r)."""
    r = 0.01
    g = np.arange(4) * 30 * r
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    for fp64 in (True, False):
        o = Oracle(xyz, r, fp64=fp64)
        o.anisotropy()
        xm, G, det = o.aniso()
        assert np.allclose(xm, xyz, atol=0, rtol=0)
        assert np.allclose(G, np.eye(3).ravel(), atol=1e-12 if fp64 else 1e-6)
        assert np.allclose(det, 1, atol=1e-12 if fp64 else 1e-6)
        ms, iso = o.iso_constants()
        h2, sigma = o.constants()
        rho = np.array([0.0, 0.01, 0.02, 0.5, 1.0, 1.5, 2.0, 3.0, 3.9, 4.1]) * r
        q = (xyz[21] + np.stack([rho, np.zeros_like(rho), np.zeros_like(rho)], 1)).astype(np.float32)
        f = o.eval_points(q)
        d = np.linalg.norm(q.astype(np.float64) - xyz[21], axis=1)
        expect = iso - sigma * m4(d / math.sqrt(h2))
        assert np.allclose(f, expect, rtol=1e-10 if fp64 else 2e-6, atol=(1e-9 if fp64 else 1e-6) * iso)
        # the iso-surface of one particle is the sphere |q - x| = 0.01 * 2r
        ff = o.eval_points(np.array([xyz[21] + [0.0199 * r, 0, 0], xyz[21] + [0.0201 * r, 0, 0]], np.float32))
        assert ff[0] < 0 < ff[1]


# ---------------------------------------------------------------- H8 / H9 anisotropy pins
def test_lattice_interior_is_isotropic():
    """Perfect cubic lattice: interior xMean = x, covariance isotropic, G = I, det = 1."""
    r = 0.01
    n = 14
    g = np.arange(n) * 2 * r
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    o = Oracle(xyz, r, fp64=True)
    o.anisotropy()
    xm, G, det = o.aniso()
    idx = np.arange(n)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    interior = ((I >= 4) & (I < n - 4) & (J >= 4) & (J < n - 4) & (K >= 4) & (K < n - 4)).ravel()
    assert interior.sum() > 100
    assert np.allclose(xm[interior], xyz[interior], atol=1e-12)
    assert np.allclose(G[interior], np.eye(3).ravel(), atol=1e-9)
    assert np.allclose(det[interior], 1, atol=1e-9)
    # boundary particles are anisotropic (their xMean moves inward)
    assert np.max(np.abs(xm[~interior] - xyz[~interior])) > 0.1 * r


def test_planar_layer_closed_form():
    """Single square layer: cov = diag(l, l, h^2/W) with l = (h^2 + sum w dx^2)/W; the normal
    eigenvalue is clamped to l/4; G = diag(s/l, s/l, s/l') with s = cbrt(l^2 l')."""
    r = 1.0 / 128  # lattice exactly representable, so the neighbourhood is exactly symmetric
    n = 16  # extent 7.5 cells: an exact multiple would trigger the max-face -1 shift
    g = np.arange(n) * 2 * r
    X, Y = np.meshgrid(g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1).astype(np.float32)
    c = (n // 2) * n + n // 2
    # brute force over the lattice neighbours of the centre particle (all within 4r are interior)
    d = xyz.astype(np.float64) - xyz[c]
    dist = np.linalg.norm(d, axis=1)
    w = w_cubic(dist, 4 * r)
    W = w.sum()
    h2 = (2 * r) ** 2
    lam_par = (h2 + np.sum(w * d[:, 0] ** 2)) / W
    lam_perp = h2 / W
    assert abs(np.sum(w * d[:, 0] * d[:, 1])) == 0
    lp = max(lam_perp, lam_par / 4)
    s = (lam_par * lam_par * lp) ** (1 / 3)
    G_expect = np.diag([s / lam_par, s / lam_par, s / lp])
    for fp64, tol in ((True, 1e-9), (False, 2e-5)):
        o = Oracle(xyz, r, fp64=fp64)
        o.anisotropy()
        xm, G, det = o.aniso()
        assert np.allclose(xm[c], xyz[c], atol=1e-9)
        assert np.allclose(G[c].reshape(3, 3, order="F"), G_expect, atol=tol)
        assert abs(det[c] - 1) < tol


def test_eigensolver_against_lapack():
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = 3 if trial % 2 == 0 else 4
        A = rng.normal(size=(n, n))
        A = A @ A.T + n * np.eye(n) * rng.uniform(0, 2)
        if trial % 7 == 0:  # repeated eigenvalues
            Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
            ev = np.array([2.0, 2.0] + list(rng.uniform(0.1, 5, n - 2)))
            A = Q @ np.diag(ev) @ Q.T
        A = 0.5 * (A + A.T)
        lam, Q = eigensolve(A, fp64=True)
        ref = np.linalg.eigvalsh(A)
        assert np.allclose(np.sort(lam), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        assert np.allclose(Q @ np.diag(lam) @ Q.T, A, atol=1e-12 * np.abs(A).max())
        assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-12)
        if n == 3:
            lam32, Q32 = eigensolve(A, fp64=False)
            assert np.allclose(np.sort(lam32), ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("diag,expect", [
    ((3.0, 3.0, 3.0), (1.0, 1.0, 1.0)),
    ((4.0, 1.0, 1.0), (4 ** (1 / 3) / 4, 4 ** (1 / 3), 4 ** (1 / 3))),
    ((16.0, 1.0, 1.0), (256 ** (1 / 3) / 16, 256 ** (1 / 3) / 4, 256 ** (1 / 3) / 4)),  # clamped at max/4
    ((16.0, 16.0, 0.5), (1024 ** (1 / 3) / 16, 1024 ** (1 / 3) / 16, 1024 ** (1 / 3) / 4)),
])
def test_g_from_covariance_closed_form(diag, expect):
    rng = np.random.default_rng(int(sum(diag) * 10))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    cov = Q @ np.diag(diag) @ Q.T
    c6 = np.array([cov[0, 0], cov[0, 1], cov[0, 2], cov[1, 1], cov[1, 2], cov[2, 2]])
    for fp64, tol in ((True, 1e-11), (False, 3e-6)):
        G, det = g_from_cov(c6, fp64)
        assert np.allclose(G, Q @ np.diag(expect) @ Q.T, atol=tol * max(expect))
        assert abs(det - 1) < 10 * tol


# ---------------------------------------------------------------- H1-H5 hash grid by brute force
def _grid_numpy(xyz, bbox, r, inf=4.0):
    cell = np.float32(r) * np.float32(inf)
    b = np.zeros(6, np.float32)
    for i in range(3):
        length = np.ceil((bbox[2 * i + 1] - bbox[2 * i]) / cell) * cell
        center = (bbox[2 * i] + bbox[2 * i + 1]) / np.float32(2)
        b[2 * i] = center - length / np.float32(2)
        b[2 * i + 1] = center + length / np.float32(2)
    n3 = np.array([max(int(np.ceil((b[2 * i + 1] - b[2 * i]) / cell)), 1) for i in range(3)])
    idx = ((xyz - b[[0, 2, 4]]) / cell).astype(np.int64)  # float32 ops, truncation toward zero
    ok = np.all((idx >= 0) & (idx < n3), axis=1)
    key = np.where(ok, idx[:, 2] * n3[0] * n3[1] + idx[:, 1] * n3[0] + idx[:, 0], -1)
    return b, n3, key


def _bbox(xyz):
    return np.array([xyz[:, 0].min(), xyz[:, 0].max(), xyz[:, 1].min(), xyz[:, 1].max(), xyz[:, 2].min(),
                     xyz[:, 2].max()], np.float32)


@pytest.mark.parametrize("which", ["cfg1", "cloud", "sphere"])
def test_hashgrid_brute_force(which):
    if which == "cfg1":
        xyz, r = synth.dam_break()
    elif which == "cloud":
        xyz, r = synth.random_cloud(20000, r=0.004, extent=0.2)
    else:
        xyz, r = synth.lattice_sphere()
    o = Oracle(xyz, r)
    gi = o.hashgrid_build()
    b, n3, key = _grid_numpy(xyz, _bbox(xyz), r)
    assert np.array_equal(gi["bounding"], b)
    assert np.array_equal(gi["cellnum3"], n3)
    order = np.argsort(key, kind="stable")
    assert np.array_equal(o.index_list(), order)
    assert gi["n_invalid"] == int(np.sum(key < 0))
    keys, s, e = o.cells()
    valid = key[key >= 0]
    uk, cnt = np.unique(valid, return_counts=True)
    assert np.array_equal(keys, uk)
    assert np.array_equal(e - s, cnt)
    assert np.array_equal(s, np.concatenate([[0], np.cumsum(cnt)[:-1]]))


def test_hashgrid_max_face_shift():
    """Extent an exact multiple of the cell: particles on the max faces hash to -1, sort first, and
    shift every cell window by their number (src/hash_grid.cpp:103-122)."""
    r = 0.25  # cell = 1.0 exactly
    g = np.arange(4, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    o = Oracle(xyz, r)
    gi = o.hashgrid_build()
    b, n3, key = _grid_numpy(xyz, _bbox(xyz), r)
    assert list(n3) == [3, 3, 3]
    n_inv = int(np.sum(key < 0))
    assert n_inv == 4 ** 3 - 3 ** 3 and gi["n_invalid"] == n_inv
    order = np.argsort(key, kind="stable")
    il = o.index_list()
    assert np.array_equal(il, order)
    # the window of a cell indexes the full sorted list at "valid-count" positions
    q = np.array([[1.5, 1.5, 1.5]], np.float32)
    off, ids = o.neighbors(q)
    valid_sorted = np.sort(key[key >= 0], kind="stable")
    start = {k: int(np.searchsorted(valid_sorted, k)) for k in np.unique(valid_sorted)}
    cnt = {k: int(np.sum(valid_sorted == k)) for k in np.unique(valid_sorted)}
    expect = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                c = (1 + dz) * 9 + (1 + dy) * 3 + (1 + dx)
                expect += list(il[start[c]:start[c] + cnt[c]])
    assert list(ids) == expect
    # ... so the returned ids are not the particles of those cells
    assert set(ids) != set(np.nonzero(key >= 0)[0])


def test_truncation_below_box_maps_to_cell_zero():
    xyz, r = synth.random_cloud(500, r=0.01, extent=0.1)
    o = Oracle(xyz, r)
    gi = o.hashgrid_build()
    b = gi["bounding"]
    q = np.array([[b[0] - 0.5 * gi["cellsize"], (b[2] + b[3]) / 2, (b[4] + b[5]) / 2]], np.float32)
    off, ids = o.neighbors(q)
    q2 = q.copy()
    q2[0, 0] = b[0] + 0.5 * gi["cellsize"]
    off2, ids2 = o.neighbors(q2)
    assert list(ids) == list(ids2) and len(ids) > 0


def test_neighbors_and_box_brute_force(cfg1):
    xyz, r = cfg1
    o = Oracle(xyz, r)
    gi = o.hashgrid_build()
    b, n3, key = _grid_numpy(xyz, _bbox(xyz), r)
    cell = np.float32(gi["cellsize"])
    ijk = np.stack([key % n3[0], (key // n3[0]) % n3[1], key // (n3[0] * n3[1])], 1)
    rng = np.random.default_rng(2)
    q = (xyz[rng.choice(len(xyz), 300)] + rng.uniform(-0.02, 0.02, (300, 3))).astype(np.float32)
    off, ids = o.neighbors(q)
    qi = ((q - b[[0, 2, 4]]) / cell).astype(np.int64)
    for i in range(len(q)):
        near = np.all(np.abs(ijk - qi[i]) <= 1, axis=1) & (key >= 0)
        assert sorted(ids[off[i]:off[i + 1]]) == sorted(np.nonzero(near)[0])
    for _ in range(20):
        lo = (xyz[rng.integers(len(xyz))] - 0.03).astype(np.float32)
        hi = (lo + rng.uniform(0.0, 0.08, 3)).astype(np.float32)
        got = o.box_particles(lo, hi)
        lo_c = np.maximum(lo, b[[0, 2, 4]])
        hi_c = np.minimum(hi, b[[1, 3, 5]])
        l3 = ((lo_c - b[[0, 2, 4]]) / cell).astype(np.int64) - 1
        h3 = ((hi_c - b[[0, 2, 4]]) / cell).astype(np.int64) + 1
        inb = np.all((ijk >= l3) & (ijk <= h3), axis=1) & (key >= 0)
        assert sorted(got) == sorted(np.nonzero(inb)[0])


# ---------------------------------------------------------------- H11 / H13 / H14 / H15
def test_field_fp64_brute_force(cfg1):
    """SingleEval in FP64 = direct sum over the same 27-cell candidates (found by brute force)."""
    xyz, r = cfg1
    o = Oracle(xyz, r, fp64=True)
    gi = o.hashgrid_build()
    o.anisotropy()
    xm, G, det = o.aniso()
    ms, iso = o.iso_constants()
    h2, sigma = o.constants()
    b, n3, key = _grid_numpy(xyz, _bbox(xyz), r)
    ijk = np.stack([key % n3[0], (key // n3[0]) % n3[1], key // (n3[0] * n3[1])], 1)
    rng = np.random.default_rng(4)
    q = (xyz[rng.choice(len(xyz), 200)] + rng.uniform(-3 * r, 3 * r, (200, 3))).astype(np.float32)
    f = o.eval_points(q)
    qi = ((q - b[[0, 2, 4]]) / np.float32(gi["cellsize"])).astype(np.int64)
    Gm = G.reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major storage
    for i in range(len(q)):
        near = np.nonzero(np.all(np.abs(ijk - qi[i]) <= 1, axis=1) & (key >= 0))[0]
        d = q[i].astype(np.float64) - xm[near]
        gd = np.einsum("nab,nb->na", Gm[near], d)
        S = np.sum(det[near] * m4(np.sqrt(np.sum(gd ** 2, 1) / h2)))
        assert abs((iso - f[i]) - sigma * S) <= 1e-11 * sigma * max(S, 1e-3)


def test_depth_closed_form(cfg1, sphere):
    for xyz, r, dmax in ((cfg1[0], cfg1[1], 7), (sphere[0], sphere[1], 5)):
        o = Oracle(xyz, r)
        info = o.tree_info()
        bb = _bbox(xyz)
        max_len = max(bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4])
        D = math.ceil(math.log2(math.ceil(max_len / r)))
        while 2 ** D * r - max_len < 8 * r:
            D += 1
        assert info["depth_max"] == D == dmax
        assert info["depth_min"] == D - 3
        assert abs(info["root_half"] - 2 ** (D - 1) * r) < 1e-7


def test_grid_eval_consistency(cfg1):
    xyz, r = cfg1
    o = Oracle(xyz, r)
    o.anisotropy()
    o.iso_constants()
    c = np.array([0.15, 0.30, 0.2], np.float32)  # on the droplet's top surface
    half = np.float32(2 * r)
    s, g, sc = o.grid_eval(c, half)
    mn, mx = c - half, c + half
    t = np.array([[x, y, z] for z in range(3) for y in range(3) for x in range(3)], np.float32) / np.float32(2)
    pts = (np.float32(1) - t) * mn + t * mx
    assert np.array_equal(s[:, :3].astype(np.float32), pts)
    assert np.array_equal(s[:, 3], o.eval_points(pts))
    assert sc == bool(np.any((s[:, 3] >= 0) != (s[0, 3] >= 0))) and sc
    f = s[:, 3].reshape(3, 3, 3)  # [z, y, x]
    step = float(half)
    assert np.allclose(g[:, 0].reshape(3, 3, 3)[:, :, 1], (f[:, :, 0] - f[:, :, 2]) / (2 * step), rtol=1e-5)
    assert np.allclose(g[:, 2].reshape(3, 3, 3)[0], (f[0] - f[1]) / step, rtol=1e-5)


def test_emptiness_and_curvature_brute_force(cfg1):
    xyz, r = cfg1
    o = Oracle(xyz, r)
    o.anisotropy()
    o.iso_constants()
    nrm = o.particle_normals()
    info = o.tree_info()
    rng = np.random.default_rng(9)
    c = (xyz[rng.choice(len(xyz), 150)] + rng.uniform(-6 * r, 6 * r, (150, 3))).astype(np.float32)
    half = np.full(150, 2 * r, np.float32)
    depth = np.full(150, info["depth_min"] + 1, np.int32)
    t, f, curv, sc = o.classify_nodes(c, half, depth)
    m = np.float32(r) * np.float32(2.0)
    for i in range(150):
        lo, hi = c[i] - half[i] - m, c[i] + half[i] + m
        inside = np.all((xyz > lo) & (xyz < hi), axis=1)
        assert (t[i] == 0) == (not inside.any())
        if inside.any():
            n = nrm[inside].astype(np.float32)
            area = np.sum(np.linalg.norm(n, axis=1))
            expect = 1.0 if area == 0 else np.linalg.norm(n.sum(0)) / area
            assert abs(curv[i] - expect) < 1e-5


def test_qef_modes_agree_on_well_posed_systems(cfg1):
    xyz, r = cfg1
    o = Oracle(xyz, r, qef_svd_tol=1e-12)
    o.anisotropy()
    o.iso_constants()
    rng = np.random.default_rng(3)
    c = np.array([0.1, 0.1, 0.1], np.float32)
    half = np.float32(0.01)
    t = np.array([[x, y, z] for z in range(3) for y in range(3) for x in range(3)], np.float64) / 2
    pts = (1 - t) * (c - half) + t * (c + half)
    for _ in range(10):
        gr = rng.normal(size=(27, 3))
        f = rng.normal(size=27) * 0.01
        samples = np.concatenate([pts, f[:, None]], 1)
        rows = np.concatenate([gr, -np.ones((27, 1)), (f - np.sum(pts * gr, 1))[:, None]], 1)
        A = rows[:, :4].T @ rows[:, :4]
        x = np.linalg.solve(A, -rows[:, :4].T @ rows[:, 4])
        n1 = o.qef(c, half, samples, gr, 1)
        n2 = o.qef(c, half, samples, gr, 2)
        inside = np.all(np.abs(x[:3] - c) <= half)
        if inside:
            assert np.allclose(n1[:3], x[:3], atol=1e-4 * half)
            assert np.allclose(n2[:3], x[:3], atol=1e-9)
        else:
            assert np.array_equal(n1[:3], c.astype(np.float64))
    # rank-deficient system: the float inverse fails (centre kept), the truncated SVD stays finite
    gr = np.tile([0.0, 0.0, 1.0], (27, 1))
    samples = np.concatenate([pts, (pts[:, 2:3] - c[2])], 1)
    n1 = o.qef(c, half, samples, gr, 1)
    n2 = o.qef(c, half, samples, gr, 2)
    assert np.array_equal(n1[:3], c.astype(np.float64))
    assert np.all(np.isfinite(n2))


def test_bisection_brackets_the_zero_crossing():
    """One isolated particle: f < 0 at the centre, zero at rho0 = 0.02 r. The bisection halves the
    bracket until it is <= r/2, so the vertex lies within r/2 of the crossing and takes
    ceil(log2(d / (r/2))) evaluations."""
    r = 0.01
    xyz = np.array([[0.5, 0.5, 0.5], [5.0, 5.0, 5.0]], np.float32)
    o = Oracle(xyz, r, fp64=True)
    o.anisotropy()
    o.iso_constants()
    for L in (0.4 * r, 0.9 * r, 1.7 * r, 3.1 * r):
        a = np.array([0.5, 0.5, 0.5])
        b = a + [L, 0, 0]
        fa, fb = o.eval_points(np.array([a, b], np.float32))
        assert fa < 0 < fb or L < 0.02 * r
        out, nev = o.bisect_edges(np.array([[list(a) + [fa], list(b) + [fb]]]))
        steps = max(0, math.ceil(math.log2(L / (r / 2))))
        assert nev[0] == steps
        assert abs(out[0, 0] - (0.5 + 0.02 * r)) <= max(r / 2, 0.1 * L) + 1e-12


# ---------------------------------------------------------------- extraction tables and mesh
def _dmc_tables():
    import re
    import os
    src = open(os.path.join(os.path.dirname(__file__), "..", "particleisosurface_amd", "csrc", "dmc_tables.h")).read()

    def arr(name):
        body = re.search(name + r"\)?\[[^\]]*\](?:\[[^\]]*\])*\s*=\s*\{(.*?)\};", src, re.S).group(1)
        return [int(v, 0) for v in re.findall(r"0x[0-9a-f]+|\d+", body)]
    perm = arr("kDmcPerm")
    case = arr("kDmcCase")
    mirror = arr("kDmcMirror")
    canon = arr("kDmcCanon")
    ntri = arr("kDmcNTri")
    tri = np.array(arr("kDmcTri")).reshape(len(canon), -1, 3)
    edge = np.array(arr("kDmcEdge")).reshape(12, 2)
    return perm, case, mirror, canon, ntri, tri, edge


def test_dmc_table_invariants():
    perm, case, mirror, canon, ntri, tri, edge = _dmc_tables()
    corner = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)])
    cube_edges = {tuple(sorted(e)) for e in edge}
    assert len(cube_edges) == 12
    for s in range(256):
        p = [(perm[s] >> (4 * k)) & 15 for k in range(8)]
        assert sorted(p) == list(range(8))
        # a cube symmetry maps edges to edges
        assert {tuple(sorted((p[a], p[b]))) for a, b in edge} == cube_edges
        # canonical configuration = permuted sign bits
        assert case[s] == sum(((s >> p[k]) & 1) << k for k in range(8))
        # mirror flag = improper symmetry (orientation reversing)
        M = np.linalg.lstsq(np.c_[corner, np.ones(8)], corner[p], rcond=None)[0][:3]
        assert mirror[s] == (1 if np.linalg.det(M) < 0 else 0)
    for k, code in enumerate(canon):
        used = {int(e) for t in tri[k][:ntri[k]] for e in t}
        change = {e for e in range(12) if ((code >> edge[e][0]) & 1) != ((code >> edge[e][1]) & 1)}
        assert used == change


def _edges_of(t):
    from collections import Counter
    directed = Counter()
    for a, b, c in t:
        directed.update([(a, b), (b, c), (c, a)])
    return directed


def test_sphere_mesh_is_closed_and_oriented(sphere):
    xyz, r = sphere
    o = Oracle(xyz, r)
    o.run()
    v, t = o.mesh()
    assert len(t) > 1000
    directed = _edges_of(t)
    undirected = {}
    for (a, b), k in directed.items():
        assert k == 1  # consistent orientation: each directed edge once
        undirected[tuple(sorted((a, b)))] = undirected.get(tuple(sorted((a, b))), 0) + 1
    assert all(k == 2 for k in undirected.values())
    assert len(v) - len(undirected) + len(t) == 2  # sphere
    d = np.linalg.norm(v - xyz.mean(0), axis=1)
    assert d.min() > 0.04 and d.max() < 0.07


def test_mesh_vertices_keyed_by_leaf_id_pairs(cfg1):
    xyz, r = cfg1
    o = Oracle(xyz, r)
    o.run()
    v, t = o.mesh()
    nd = o.tree_nodes()
    n8, sidx = o.dual_cells()
    perm, case, mirror, canon, ntri, tri, edge = _dmc_tables()
    pairs = set()
    for cell, s in zip(n8, sidx):
        k = canon.index(case[s])
        for e in {int(x) for tt in tri[k][:ntri[k]] for x in tt}:
            a, b = nd["id"][cell[edge[e][0]]], nd["id"][cell[edge[e][1]]]
            pairs.add((min(a, b), max(a, b)))
    assert len(pairs) == len(v)
    assert t.min() == 1 and t.max() == len(v)
