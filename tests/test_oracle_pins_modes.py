"""Pins of the oracle functions that the default mode does not exercise, and of the anisotropy
and gradient formulas on inputs where the answer is known without the oracle's own code.

Each case is built so that a plausible restatement mistake fails it:
- particle normals / gradients (src/evaluator.cpp:89-125,408-415,455-463): a closed form for one
  isolated particle with an anisotropic G of determinant != 1 (catches d instead of G d, a dropped
  det or sigma, a wrong factor or sign), and a central difference of the isotropic poly6 field
  (whose poly6-form gradient is the true gradient);
- xMean (src/evaluator.cpp:465-512,656-718): collinear clusters with hand-computed weights
  (catches self inclusion, the 0.9 blend, the (4r)^2 cut, wij's branches and the <= 2 candidate
  rule) and a brute force over a half-lattice surface;
- compute_G_Yus on an asymmetric pair, compute_G_ours (USE_POLY6) on a lattice interior and on a
  planar layer (src/evaluator.cpp:514-654);
- poly6 constants and the isotropic field (src/evaluator.cpp:272-291,323-335,377-381,417-435).
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle


def m4(q):
    q = np.asarray(q, np.float64)
    return np.where(q <= 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def rot(seed):
    Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return Q


def set_records(o, xm, Gs, dets):
    """Install xMean / G / det records (G as 3x3 matrices) in the oracle."""
    G9 = np.stack([np.asarray(G, np.float64).ravel(order="F") for G in Gs])
    o.set_aniso(np.asarray(xm, np.float64), G9, np.asarray(dets, np.float64))


# ------------------------------------------------------------------ H12 / H10 gradients
@pytest.mark.parametrize("fp64", [True, False])
def test_gradient_closed_form_one_particle(fp64):
    """One isolated particle with G = Q diag(1.6, 0.8, 0.5) Q^T (det 0.64): at q = x + d the
    value is det sigma B(|G d| / h) and the gradient sum is det sigma (-6 G d) (h^2 - |G d|^2)^2
    for |G d|^2 <= h^2, else 0. A gradient formed with d (or G^T G d, the true gradient) instead
    of G d, or without det, differs in direction or size."""
    r = 0.01
    xyz = np.array([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]], np.float32)
    o = Oracle(xyz, r, fp64=fp64)
    o.hashgrid_build()
    Q = rot(3)
    G = Q @ np.diag([1.6, 0.8, 0.5]) @ Q.T
    G = np.float32(G).astype(np.float64)  # representable in both modes
    det = float(np.float32(np.linalg.det(G)))
    set_records(o, xyz.astype(np.float64), [G, G], [det, det])
    ms, iso = o.iso_constants()
    h2, sigma = o.constants()
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(400, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    rho = rng.uniform(0, 4.5 * r, 400)
    q = (xyz[0] + dirs * rho[:, None]).astype(np.float32)
    f, grad = o.eval_points(q, grad=True)
    d = q.astype(np.float64) - xyz[0]
    gd = d @ G.T
    n2 = np.sum(gd ** 2, 1)
    S = det * sigma * m4(np.sqrt(n2 / h2))
    g_exp = det * sigma * (-6 * gd) * np.where(n2 <= h2, (h2 - n2) ** 2, 0.0)[:, None]
    rel = 1e-12 if fp64 else 2e-5
    scale = np.abs(g_exp).max()
    assert np.max(np.abs(grad - g_exp)) <= rel * scale
    assert np.max(np.abs((iso - f) - S)) <= rel * sigma
    inside = n2 < 0.9 * h2
    assert inside.sum() > 50 and (~(n2 <= h2)).sum() > 50
    # the reference's form is not the true gradient of the anisotropic poly6 kernel
    true = det * sigma * (-6 * (gd @ G)) * np.where(n2 <= h2, (h2 - n2) ** 2, 0.0)[:, None]
    assert np.max(np.abs(grad[inside] - true[inside])) > 0.1 * scale


@pytest.mark.parametrize("fp64", [True, False])
def test_particle_normals_sum_of_closed_forms(fp64):
    """Three particles within one another's support, records installed by hand (xMeans moved off
    the positions): each normal is the sum of the closed-form gradients of all three records at
    the particle's own position, self included."""
    r = 0.01
    xyz = np.array([[0.5, 0.5, 0.5], [0.51, 0.505, 0.5], [0.5, 0.512, 0.507]], np.float32)
    o = Oracle(xyz, r, fp64=fp64)
    o.hashgrid_build()
    xm = xyz.astype(np.float64) + np.array([[0.001, 0, 0], [0, -0.002, 0.001], [0.0005, 0.0005, -0.001]])
    xm = np.float32(xm).astype(np.float64)
    Gs = [np.float32(rot(s) @ np.diag(dg) @ rot(s).T).astype(np.float64)
          for s, dg in ((4, (1.2, 1.0, 0.7)), (5, (1.5, 0.6, 0.9)), (6, (1.0, 1.0, 1.0)))]
    dets = [float(np.float32(np.linalg.det(G))) for G in Gs]
    set_records(o, xm, Gs, dets)
    o.iso_constants()
    h2, sigma = o.constants()
    n = o.particle_normals()
    for i in range(3):
        e = np.zeros(3)
        for j in range(3):
            gd = Gs[j] @ (xyz[i].astype(np.float64) - xm[j])
            n2 = gd @ gd
            if n2 <= h2:
                e += dets[j] * sigma * (-6 * gd) * (h2 - n2) ** 2
        assert np.max(np.abs(n[i] - e)) <= (1e-12 if fp64 else 2e-5) * np.abs(e).max()
        assert np.abs(e).max() > 0


def test_isotropic_gradient_is_central_difference():
    """USE_ANI = 0: the field is iso - sum sigma (h^2 - |q - x|^2)^3, whose gradient is exactly
    the poly6-form sum; a central difference of the value (FP64) must agree. (Poly6 mode, so that
    iso is of the order of the sums; in B-spline mode the isotropic sums are poly6 as well but
    iso = sigma B(0.01) is 1e10 times larger, so the difference would drown in its rounding.)"""
    r = 0.01
    rng = np.random.default_rng(8)
    xyz = (0.5 + rng.uniform(-3 * r, 3 * r, (40, 3))).astype(np.float32)
    o = Oracle(xyz, r, fp64=True, use_ani=False, use_poly6=True)
    o.anisotropy()
    ms, iso = o.iso_constants()
    q = (xyz[rng.choice(40, 60)] + rng.uniform(-r, r, (60, 3))).astype(np.float32)
    f, g = o.eval_points(q, grad=True)
    # q must be float: step in float and difference with the exact float offsets
    eps = np.float32(1e-5)
    for a in range(3):
        qp, qm = q.copy(), q.copy()
        qp[:, a] += eps
        qm[:, a] -= eps
        hstep = (qp[:, a].astype(np.float64) - qm[:, a].astype(np.float64))
        fd = -(o.eval_points(qp) - o.eval_points(qm)) / hstep  # S = iso - f
        assert np.allclose(g[:, a], fd, rtol=1e-4, atol=1e-6 * np.abs(g).max())


# ------------------------------------------------------------------ H6 / H10 poly6 and isotropic
@pytest.mark.parametrize("r", [0.005, 0.001])
def test_poly6_constants_closed_form(r):
    """USE_POLY6: MaxScalar sums poly6 at d^2 = 0, 4r^2, 8r^2, 12r^2 with no multiplicities; with
    h = 2r only d = 0 is inside the support, so MaxScalar = sigma h^6 = 1. iso = poly6(1.9 r^2) =
    sigma (h^2 - 1.9 r^2)^3 = (2.1 / 4)^3."""
    xyz = np.array([[0, 0, 0]], np.float32)
    for fp64 in (True, False):
        o = Oracle(xyz, r, fp64=fp64, use_poly6=True)
        ms, iso = o.iso_constants()
        tol = 1e-12 if fp64 else 1e-6
        assert abs(ms - 1) < tol
        assert abs(iso - (2.1 / 4) ** 3) < (1e-6 if fp64 else 2e-6)  # 1.9 and r^2 are float products


@pytest.mark.parametrize("use_poly6", [False, True])
def test_isotropic_single_particle_field(use_poly6):
    """USE_ANI = 0 uses the poly6 kernel in both kernel modes (IsotropicInterpolate):
    f(q) = iso - sigma (h^2 - rho^2)^3 for rho^2 <= h^2."""
    r = 0.01
    xyz = np.array([[0.5, 0.5, 0.5], [3, 3, 3]], np.float32)
    o = Oracle(xyz, r, fp64=True, use_ani=False, use_poly6=use_poly6)
    o.anisotropy()
    ms, iso = o.iso_constants()
    h2, sigma = o.constants()
    rho = np.array([0, 0.3, 0.9, 1.5, 1.99, 2.01, 3.0]) * r
    q = (xyz[0] + np.stack([rho, 0 * rho, 0 * rho], 1)).astype(np.float32)
    f = o.eval_points(q)
    d2 = (q[:, 0].astype(np.float64) - 0.5) ** 2
    expect = iso - sigma * np.where(d2 <= h2, (h2 - d2) ** 3, 0.0)
    assert np.allclose(f, expect, rtol=1e-12, atol=1e-12)
    if use_poly6:
        # poly6 iso-surface of one particle: sigma (h^2 - rho^2)^3 = (2.1/4)^3, rho^2 = 1.9 r^2
        rho0 = math.sqrt(1.9) * r
        ff = o.eval_points(np.array([[0.5 + 0.99 * rho0, 0.5, 0.5], [0.5 + 1.01 * rho0, 0.5, 0.5]], np.float32))
        assert ff[0] < 0 < ff[1]


@pytest.mark.parametrize("fp64", [True, False])
def test_aniso_poly6_value_closed_form(fp64):
    """USE_ANI with USE_POLY6: value det sigma (h^2 - |G d|^2)^3."""
    r = 0.01
    xyz = np.array([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]], np.float32)
    o = Oracle(xyz, r, fp64=fp64, use_poly6=True)
    o.hashgrid_build()
    G = np.float32(rot(9) @ np.diag([0.6, 0.45, 0.3]) @ rot(9).T).astype(np.float64)
    det = float(np.float32(np.linalg.det(G)))
    set_records(o, xyz.astype(np.float64), [G, G], [det, det])
    ms, iso = o.iso_constants()
    h2, sigma = o.constants()
    rng = np.random.default_rng(2)
    q = (xyz[0] + rng.uniform(-4 * r, 4 * r, (300, 3))).astype(np.float32)
    f = o.eval_points(q)
    gd = (q.astype(np.float64) - xyz[0]) @ G.T
    n2 = np.sum(gd ** 2, 1)
    S = det * sigma * np.where(n2 <= h2, (h2 - n2) ** 3, 0.0)
    assert np.max(np.abs((iso - f) - S)) <= (1e-12 if fp64 else 2e-6) * sigma * h2 ** 3
    assert np.sum(S > 0) > 30


# ------------------------------------------------------------------ H8 xMean
W_R, W_2R, W_3R = 0.71875, 0.25, 1.0 / 32  # wij at d = r, 2r, 3r with h = 4r
# the blend weights are float members (_XMEAN_DELTA = 0.9f, include/evaluator.h:30), so
# 1 - delta = 0.100000024 and delta = 0.899999976 in every mode
DELTA = float(np.float32(0.9))
KEEP = float(np.float32(1) - np.float32(0.9))


def blend(x, mean):
    return x * KEEP + mean * DELTA


def _line(xs, r):
    return np.array([[0.25 + x * r, 0.25, 0.25] for x in xs], np.float32)


@pytest.mark.parametrize("fp64", [True, False])
def test_xmean_collinear_closed_form(fp64):
    """Three particles at 0, r, 2r: the middle one is symmetric (xMean = x); an end particle gets
    x + 0.9 (sum w x_j / sum w - x) over the other two (self excluded), w(r) = 0.71875 (the
    d < h/2 branch) and w(2r) = 0.25 (d = h/2 takes the outer branch)."""
    r = 1.0 / 64
    xyz = _line([0, 1, 2], r)
    o = Oracle(xyz, r, fp64=fp64)
    o.anisotropy()
    xm, G, det = o.aniso()
    m0 = (W_R * 1 + W_2R * 2) / (W_R + W_2R)
    x = 0.25 + np.array([0, 1, 2]) * r
    expect = np.array([blend(x[0], 0.25 + m0 * r), x[1], blend(x[2], 0.25 + (2 - m0) * r)])
    tol = 1e-15 if fp64 else 4e-8
    assert np.allclose(xm[:, 0], expect, rtol=0, atol=tol)
    assert np.all(xm[:, 1:] == 0.25)


@pytest.mark.parametrize("fp64", [True, False])
def test_xmean_cut_and_outer_branch(fp64):
    """Particles at 0, r and 4r: the pair at distance exactly 4r is cut ((4r)^2 >= nI^2), the pair
    at 3r weighs 2 (1/4)^3 = 1/32. (A fourth particle at 11r, 7r beyond the others, keeps the box
    from being an exact multiple of the cell, whose max face would hash to -1.)"""
    r = 1.0 / 64
    xyz = _line([0, 1, 4, 11], r)
    o = Oracle(xyz, r, fp64=fp64)
    o.anisotropy()
    xm, _, _ = o.aniso()
    m1 = (W_R * 0 + W_3R * 4) / (W_R + W_3R)
    x = 0.25 + np.array([0, 1, 4]) * r
    expect = np.array([blend(x[0], x[1]), blend(x[1], 0.25 + m1 * r), blend(x[2], x[1])])
    assert xm[3, 0] == xyz[3, 0]  # two candidates: itself and the particle at 4r
    assert np.allclose(xm[:3, 0], expect, rtol=0, atol=1e-15 if fp64 else 4e-8)


def test_xmean_two_candidates_keep_position():
    """A pair alone has two candidates (self included): xMean = x and G = I in pass 1 (the
    smoothed value would be 0.1 x + 0.9 x_other). In poly6 mode xMean = x always."""
    r = 1.0 / 64
    xyz = _line([0, 1], r)
    for poly6 in (False, True):
        o = Oracle(xyz, r, fp64=True, use_poly6=poly6, use_ani=False)
        o.anisotropy()
        xm, _, _ = o.aniso()
        assert np.array_equal(xm, xyz.astype(np.float64))
    o = Oracle(_line([0, 1, 2], r), r, fp64=True, use_poly6=True)
    o.anisotropy()
    xm, _, _ = o.aniso()
    assert np.array_equal(xm, _line([0, 1, 2], r).astype(np.float64))


def test_xmean_half_lattice_brute_force():
    """A lattice slab with a free top face: xMean of every particle by brute force over all
    particles closer than 4r (self excluded; the 27 cells of 4r contain all of them)."""
    r = 0.01
    g = np.arange(8) * 2 * r
    X, Y, Z = np.meshgrid(g, g, np.arange(4) * 2 * r, indexing="ij")
    rng = np.random.default_rng(3)
    xyz = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + rng.uniform(-0.1 * r, 0.1 * r, (256, 3))).astype(
        np.float32)
    o = Oracle(xyz, r, fp64=True)
    o.anisotropy()
    xm, _, _ = o.aniso()
    P = xyz.astype(np.float64)
    h = float(np.float32(r) * np.float32(4))  # nI = Radius * 4 and nI^2 are floats
    h2 = float(np.float32(h) * np.float32(h))
    for i in range(len(P)):
        d2 = np.sum((P - P[i]) ** 2, 1)
        d = np.sqrt(d2)
        m = (d2 < h2) & (np.arange(len(P)) != i)
        u = d[m] / h
        w = np.where(u < 0.5, 1 - 6 * u ** 2 * (1 - u), 2 * (1 - u) ** 3)
        mean = (w[:, None] * P[m]).sum(0) / w.sum()
        assert np.allclose(xm[i], blend(P[i], mean), rtol=0, atol=1e-14)
    top = xyz[:, 2] > 5 * r
    assert np.all(xm[top, 2] < xyz[top, 2])  # the free surface is pulled inward


# ------------------------------------------------------------------ H9 covariance -> G
@pytest.mark.parametrize("fp64", [True, False])
def test_g_yu_pair_closed_form(fp64):
    """Pair at distance r (xMean = x by the two-candidate rule): the G pass includes self (weight
    1, offset 0), so cov = (h^2 I + w(r) r^2 e e^T) / (1 + w(r)); unclamped, G = diag(s / l_e,
    s / l_p, s / l_p) with s = cbrt(l_e l_p^2)."""
    r = 1.0 / 64
    xyz = _line([0, 1], r)
    o = Oracle(xyz, r, fp64=fp64)
    o.anisotropy()
    xm, G, det = o.aniso()
    h2 = (2 * r) ** 2
    le = (h2 + W_R * r * r) / (1 + W_R)
    lp = h2 / (1 + W_R)
    s = (le * lp * lp) ** (1 / 3)
    Ge = np.diag([s / le, s / lp, s / lp])
    for i in range(2):
        assert np.allclose(G[i].reshape(3, 3, order="F"), Ge, atol=1e-12 if fp64 else 1e-6)
        assert abs(det[i] - 1) < (1e-12 if fp64 else 1e-6)


def test_g_ours_lattice_interior_and_planar():
    """compute_G_ours (USE_POLY6 with USE_ANI): cov = (1.5 I + sum w (xm_j - xm_i)(xm_j - xm_i)^T)
    / sum w, weights from |x_i - xm_j|, G scaled by 0.5. On a lattice interior cov is isotropic, so
    G = 0.5 I (det 1/8). On a planar layer cov = diag(l, l, 1.5 / W) with l = (1.5 + sum w dx^2) /
    W, clamped at l / 4: the regulariser 1.5 does not scale with r, so with r = 2 the in-plane spread
    dominates and the clamp is active."""
    r = 0.01
    n = 14
    g = np.arange(n) * 2 * r
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    o = Oracle(xyz, r, fp64=True, use_poly6=True)
    o.anisotropy()
    xm, G, det = o.aniso()
    idx = np.arange(n)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    inner = ((I >= 4) & (I < n - 4) & (J >= 4) & (J < n - 4) & (K >= 4) & (K < n - 4)).ravel()
    assert np.array_equal(xm, xyz.astype(np.float64))
    assert np.allclose(G[inner], 0.5 * np.eye(3).ravel(), atol=1e-12)
    assert np.allclose(det[inner], 0.125, atol=1e-12)
    # planar layer
    r = 2.0
    m = 16
    gp = np.arange(m) * 2 * r
    X, Y = np.meshgrid(gp, gp, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(m * m)], 1).astype(np.float32)
    c = (m // 2) * m + m // 2
    d = xyz.astype(np.float64) - xyz[c]
    dist = np.linalg.norm(d, axis=1)
    u = dist / (4 * r)
    w = np.where(u < 0.5, 1 - 6 * u ** 2 * (1 - u), np.where(u < 1, 2 * (1 - u) ** 3, 0.0))
    Wt = w.sum()
    l = (1.5 + np.sum(w * d[:, 0] ** 2)) / Wt
    lz = max(1.5 / Wt, l / 4)
    assert 1.5 / Wt < l / 4  # the clamp is active
    s = (l * l * lz) ** (1 / 3)
    o = Oracle(xyz, r, fp64=True, use_poly6=True)
    o.anisotropy()
    xm, G, det = o.aniso()
    assert np.allclose(G[c].reshape(3, 3, order="F"), 0.5 * np.diag([s / l, s / l, s / lz]), atol=1e-12)
    assert abs(det[c] - 0.125) < 1e-12
