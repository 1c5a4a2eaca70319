"""GPU parity for the reference's other kernel modes and for the full-size workloads.

Modes (use_ani, use_poly6, calc_normals), each against the oracle on the dam-break frame (14 015
particles, r = 0.005):
- (1, 1, 1): anisotropic poly6 with compute_G_ours (src/evaluator.cpp:514-583);
- (0, 0, 1): isotropic kernel (IsotropicInterpolate, :417-435) with the B-spline iso constants:
  the isotropic sums are poly6 (of order 1) while iso = sigma B(0.01) is of order 1e10, so the
  frame has no surface, as in the reference;
- (0, 1, 1): isotropic kernel with the poly6 iso constants (a frame with a surface);
- (1, 0, 0): default kernel without particle normals (curvature 1 everywhere).

Workloads at full size:
- the 24-frame batch (sloshing frames of 50 k - 3 M particles, 19.6 M in all) through
  pisurf_run_frames, every frame bit-identical to a fresh context, the smallest against the oracle;
- the 12.1 M multi-body pour with the maximum depth forced to 12: cell table bit-exact and FP64
  field values against the FP64 oracle, compensated FP32 against FP64 on the same GPU;
- the 4.02 M crown splash: cell table bit-exact and field values against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from particleisosurface_amd import capi, synth
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu

MODES = {"ani_poly6": (1, 1, 1), "iso": (0, 0, 1), "iso_poly6": (0, 1, 1), "no_normals": (1, 0, 0)}
PREC = {"fp32": capi.FP32, "fp32k": capi.FP32_KAHAN, "fp64": capi.FP64}
FIELD_TOL = {"fp32": (3e-5, 5e-5), "fp32k": (1e-5, 2e-5), "fp64": (1e-10, 1e-9)}


def _mode_kw(mode):
    a, p, n = MODES[mode]
    return dict(use_ani=bool(a), use_poly6=bool(p), calc_normals=bool(n))


def _queries(xyz, r, n, seed=1):
    rng = np.random.default_rng(seed)
    return (xyz[rng.choice(len(xyz), n)] + rng.uniform(-3 * r, 3 * r, (n, 3))).astype(np.float32)


def _field_check(f, f2, iso, prec, floor=None):
    """|dS| / max(S, floor) within the precision's bound (floor = iso by default), and signs equal
    outside |f| <= eps iso"""
    rel, eps = FIELD_TOL[prec]
    S, S2 = iso - f, iso - f2
    err = np.abs(S - S2) / np.maximum(np.abs(S2), abs(iso) if floor is None else floor)
    assert err.max() <= rel, float(err.max())
    tie = np.abs(f2) <= eps * abs(iso)
    assert np.array_equal((f > 0)[~tie], (f2 > 0)[~tie])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("mode", sorted(MODES))
def test_mode_records_field_normals(mode, prec):
    xyz, r = synth.dam_break()
    kw = _mode_kw(mode)
    o = Oracle(xyz, r, fp64=(prec == "fp64"), **kw)
    ctx = gpu_context(xyz, r, precision=PREC[prec], **kw)
    o.anisotropy()
    ctx.anisotropy()
    xm, G, d = ctx.aniso()
    xm2, G2, d2 = o.aniso()
    if prec == "fp64":
        assert np.max(np.abs(xm - xm2)) <= 1e-12
    else:
        assert np.max(np.abs(xm - xm2)) <= 4 * np.spacing(np.float32(np.abs(xyz).max())) + 1e-6 * r
    if kw["use_ani"]:
        tol = 1e-9 if prec == "fp64" else 5e-5
        assert np.max(np.abs(G - G2)) <= tol * np.abs(G2).max()
        assert np.max(np.abs(d - d2)) <= tol * np.abs(d2).max()
        if kw["use_poly6"]:  # compute_G_ours: G ~ 0.5 I, det ~ 1/8 with the r-independent 1.5 regulariser
            assert np.allclose(d2, 0.125, rtol=1e-3)
    ms, iso = ctx.iso_constants()
    ms2, iso2 = o.iso_constants()
    assert abs(ms / ms2 - 1) < 1e-15 and abs(iso / iso2 - 1) < 1e-15
    q = _queries(xyz, r, 20000, seed=3)
    f, g = ctx.eval_points(q, grad=True)
    f2, g2 = o.eval_points(q, grad=True)
    if mode != "iso":
        _field_check(f, f2, iso, prec)
    elif prec == "fp64":
        # S is of order 1 beside iso ~ 1e10: compare S itself (a float f cannot hold it)
        S2 = iso - f2
        _field_check(f, f2, iso, prec, floor=1e-2 * S2.max())
        assert np.all(f > 0) and np.all(f2 > 0)
    sc = np.abs(g2).max()
    assert np.max(np.abs(g - g2)) <= (1e-9 if prec == "fp64" else 1e-4) * sc
    if kw["calc_normals"]:
        n, n2 = ctx.particle_normals(), o.particle_normals()
        assert np.max(np.abs(n - n2)) <= (1e-7 if prec == "fp64" else 5e-5) * np.abs(n2).max()


@pytest.mark.parametrize("mode", sorted(MODES))
def test_mode_full_tree_and_mesh(mode):
    """FP64 frame in each mode against the FP64 oracle: identical octree (types, ids, depths,
    children), leaf values, triangles (reference traversal order) and vertices."""
    xyz, r = synth.dam_break()
    kw = _mode_kw(mode)
    o = Oracle(xyz, r, fp64=True, **kw)
    o.run()
    ctx = gpu_context(xyz, r, precision=capi.FP64, mesh_order=1, **kw)
    info = ctx.run_frame()
    oi = o.tree_info()
    assert (info["depth_max"], info["depth_min"], info["n_nodes"]) == (oi["depth_max"], oi["depth_min"], oi["n_nodes"])
    nd, nd2 = ctx.tree_nodes(info["n_nodes"]), o.tree_nodes()
    for k in ("type", "id", "depth", "child"):
        assert np.array_equal(nd[k], nd2[k]), k
    ms, iso = o.iso_constants()
    leaf = nd2["type"] == capi.LEAF
    S, S2 = iso - nd["node"][leaf, 3], iso - nd2["node"][leaf, 3]
    assert np.all(np.abs(S - S2) <= 1e-10 * np.maximum(np.abs(S2), abs(iso)))
    v, t = ctx.mesh()
    v2, t2 = o.mesh()
    assert v.shape == v2.shape and np.array_equal(t, t2)
    if len(v):
        assert np.max(np.linalg.norm(v - v2, axis=1)) < 1e-6 * r
    if mode == "iso":
        # the isotropic poly6 sums stay far below the B-spline iso value (sigma B(0.01)): no surface
        assert len(t) == 0
    else:
        # without normals the curvature is 1 everywhere and nothing refines below D_min
        assert len(t) > (100 if mode == "no_normals" else 1000)


@pytest.mark.parametrize("fif", [1, 4])
def test_batch_24_frames(fif):
    """The 24-frame batch (compensated FP32, FP64 QEF) through one pisurf_run_frames call, one
    frame at a time and four at once: every frame's mesh and tree are bit-identical to a fresh
    context meshing that frame alone, and the meshes arrive in frame order; the smallest frame's
    cell table is bit-exact against the oracle and its field within the compensated tolerance of
    the FP64 oracle at 50 k points."""
    frames = synth.batch_frames()
    got = {}
    order = []

    def on_mesh(k, v, t, tree):
        order.append(k)
        got[k] = (v, t, tree)

    kw = dict(precision=capi.FP32_KAHAN, qef_mode=capi.QEF_SVD64)
    ctx = capi.Context(0)
    ctx.set_params(0.01, frames_in_flight=fif, **kw)
    per = ctx.run_frames([f for f, _ in frames], [r for _, r in frames], on_mesh)
    assert order == list(range(24))
    for k, (xyz, r) in enumerate(frames):
        v, t, tree = got[k]
        ref = gpu_context(xyz, r, **kw)
        info = ref.run_frame()
        v2, t2 = ref.mesh()
        assert tree == info, k
        assert np.array_equal(v, v2) and np.array_equal(t, t2), k
        assert per[k]["field_evals"] == ref.stats()["field_evals"], k
        assert per[k]["qef_solves"] > 0, k
        ref.close()
    k = int(np.argmin([len(f) for f, _ in frames]))
    xyz, r = frames[k]
    o = Oracle(xyz, r, fp64=True)
    go = o.hashgrid_build()
    ctx.set_params(r, **kw)
    ctx.upload(xyz)
    g = ctx.hashgrid_build()
    assert g["cellnum"] == go["cellnum"] and g["n_invalid"] == go["n_invalid"]
    il, s, e = ctx.hashgrid_download(g["cellnum"])
    assert np.array_equal(il, o.index_list())
    o.anisotropy()
    ms, iso = o.iso_constants()
    q = _queries(xyz, r, 50000, seed=6)
    _field_check(ctx.eval_points(q), o.eval_points(q), iso, "fp32k")


def test_pour_12m_depth12():
    """Multi-body pour, 12.1 M particles, r = 0.0005, maximum depth forced to 12 (the derived
    depth is 10): FP64 cell table and index list bit-exact against the FP64 oracle, FP64 field
    values at 100 k points around the surface against it; then the compensated FP32 frame against
    the FP64 frame on the GPU (same octree and triangles, field, signs, vertices within 1e-3 r)."""
    xyz, r = synth.multibody_pour()
    o = Oracle(xyz, r, fp64=True, depth_max_override=12)
    go = o.hashgrid_build()
    ctx = gpu_context(xyz, r, precision=capi.FP64, depth_max_override=12)
    g = ctx.hashgrid_build()
    assert np.array_equal(g["bounding"], go["bounding"]) and g["cellnum"] == go["cellnum"]
    assert g["n_invalid"] == go["n_invalid"]
    il, s, e = ctx.hashgrid_download(g["cellnum"])
    assert np.array_equal(il, o.index_list())
    del il, s, e
    o.anisotropy()
    ms, iso = o.iso_constants()
    q = _queries(xyz, r, 100000, seed=12)
    f64 = ctx.eval_points(q)
    _field_check(f64, o.eval_points(q), iso, "fp64")
    del o
    i64 = ctx.run_frame()
    assert i64["depth_max"] == 12 and i64["depth_min"] == 9
    v64, t64 = ctx.mesh()
    nd = ctx.tree_nodes(i64["n_nodes"])
    internal = nd["type"] == capi.INTERNAL
    assert i64["n_nodes"] == 1 + 8 * internal.sum()
    assert np.all(nd["depth"][nd["type"] == capi.LEAF] >= 9)
    assert int(nd["depth"].max()) == 12
    del nd
    ctx.set_params(r, precision=capi.FP32_KAHAN, depth_max_override=12)
    i32 = ctx.run_frame()
    v32, t32 = ctx.mesh()
    f32 = ctx.eval_points(q)
    assert i32["n_nodes"] == i64["n_nodes"]
    _field_check(f32, f64, iso, "fp32k")
    assert v32.shape == v64.shape and np.array_equal(t32, t64)
    d = np.linalg.norm(v32 - v64, axis=1) / r
    assert np.quantile(d, 0.99) <= 1e-3
    assert np.mean(d <= 1e-3) >= 1 - 1e-4


def test_crown_splash_against_oracle():
    """Crown splash, 4.02 M particles: cell table and index list bit-exact, and field values at
    100 k points around the surface from the FP64 and compensated-FP32 GPU paths against the FP64
    oracle."""
    xyz, r = synth.crown_splash()
    o = Oracle(xyz, r, fp64=True)
    go = o.hashgrid_build()
    ctx = gpu_context(xyz, r, precision=capi.FP64)
    g = ctx.hashgrid_build()
    assert g["cellnum"] == go["cellnum"] and g["n_invalid"] == go["n_invalid"]
    il, s, e = ctx.hashgrid_download(g["cellnum"])
    assert np.array_equal(il, o.index_list())
    o.anisotropy()
    ms, iso = o.iso_constants()
    q = _queries(xyz, r, 100000, seed=13)
    f2 = o.eval_points(q)
    _field_check(ctx.eval_points(q), f2, iso, "fp64")
    ctx.set_params(r, precision=capi.FP32_KAHAN)
    _field_check(ctx.eval_points(q), f2, iso, "fp32k")


def test_neighbors_total_guard():
    """Neighbour lists are offset in 32 bits: a call whose lists total 2^31 ids or more is refused
    before anything is written (400 particles within one cell, 6 M queries there: 2.4e9 ids), and
    the context answers a normal call afterwards."""
    rng = np.random.default_rng(1)
    r = 0.01
    xyz = rng.uniform(0.0, 0.035, (400, 3)).astype(np.float32)
    ctx = gpu_context(xyz, r)
    ctx.hashgrid_build()
    q = np.tile(np.array([[0.017, 0.017, 0.017]], np.float32), (6_000_000, 1))
    with pytest.raises(capi.PisurfError, match="2\\^31"):
        ctx.neighbors(q)
    off, ids = ctx.neighbors(q[:1000])
    assert off[-1] == 400 * 1000 and np.array_equal(np.sort(ids[:400]), np.arange(400))


def _batch_matches_fresh(frames, kw, entry):
    got = {}

    def on_mesh(k, v, t, tree):
        got[k] = (v, t, tree)

    ctx = capi.Context(0)
    ctx.set_params(0.01, **kw)
    per = getattr(ctx, entry)([f for f, _ in frames], [r for _, r in frames], on_mesh)
    assert sorted(got) == list(range(len(frames)))
    for k, (xyz, r) in enumerate(frames):
        v, t, tree = got[k]
        if len(xyz) == 0:
            assert len(v) == 0 and len(t) == 0 and per[k]["field_evals"] == 0
            continue
        ref = gpu_context(xyz, r, **kw)
        info = ref.run_frame()
        v2, t2 = ref.mesh()
        assert tree == info, k
        assert np.array_equal(v, v2) and np.array_equal(t, t2), k
        st = ref.stats()
        for key in ("field_evals", "grad_evals", "qef_solves", "sample_evals", "sample_unique", "candidates"):
            assert per[k][key] == st[key], (k, key)
        ref.close()
    return ctx


BATCH_OFF = ("pisurf_run_batch is switched off: its first device run stopped making progress and "
             "the cause has not been found")


def test_run_batch_disabled():
    """The switched-off entry point fails cleanly with PISURF_E_UNSUPPORTED."""
    ctx = capi.Context(0)
    ctx.set_params(0.01)
    xyz, r = synth.dam_break()
    with pytest.raises(capi.PisurfError, match=r"\(-4\).*disabled"):
        ctx.run_batch([xyz], [r])


@pytest.mark.skip(reason=BATCH_OFF)
@pytest.mark.parametrize("prec", ["fp32", "fp32k", "fp64"])
def test_run_batch_small_frames(prec):
    """Frames of different sizes and radii meshed together (lockstep levels, one run-queue launch
    per level): every frame bit-identical to a fresh context, counters equal; an empty frame; and a
    second call on the same context (frame contexts reused)."""
    frames = [synth.dam_break(), synth.lattice_sphere(), (np.zeros((0, 3), np.float32), 0.004),
              synth.random_cloud(20000, r=0.004, extent=0.2), synth.dam_break()]
    ctx = _batch_matches_fresh(frames, dict(precision=PREC[prec]), "run_batch")
    per = ctx.run_batch([f for f, _ in frames[:2]], [r for _, r in frames[:2]])
    assert per[0]["field_evals"] > 0


@pytest.mark.skip(reason=BATCH_OFF)
def test_run_batch_24_frames():
    """The 24-frame batch (compensated FP32, FP64 QEF) through pisurf_run_batch: every frame's mesh,
    tree and counters equal to a fresh context meshing it alone."""
    _batch_matches_fresh(synth.batch_frames(), dict(precision=capi.FP32_KAHAN, qef_mode=capi.QEF_SVD64),
                         "run_batch")


def test_run_shape_tuning_keeps_results():
    """Run-shape autotuning: a frame meshed while the shapes are being timed (every candidate launch
    on the same data), again with the kept shapes, and with tuning off gives the same mesh, tree and
    counters each time; the tuner keeps one shape per (precision, launch kind, density bucket)."""
    L = capi.load()
    xyz, r = synth.dam_break()
    out = []
    try:
        for enabled, clear in ((1, 1), (1, 0), (0, 0)):
            L.pisurf_run_tuning(enabled, clear)
            ctx = gpu_context(xyz, r, precision=capi.FP32_KAHAN)
            info = ctx.run_frame()
            v, t = ctx.mesh()
            out.append((info, v, t, ctx.stats(), L.pisurf_run_tuning(-1, 0)))
            ctx.close()
    finally:
        L.pisurf_run_tuning(1, 0)
    kept = out[0][4]
    assert kept >= 3  # samples of the sampled levels, leaf-depth centres, normals
    assert out[1][4] == kept  # nothing new to time on the second frame
    for info, v, t, st, _ in out[1:]:
        assert info == out[0][0]
        assert np.array_equal(v, out[0][1]) and np.array_equal(t, out[0][2])
        for key in ("field_evals", "grad_evals", "sample_unique", "sample_evals", "candidates", "run_evals"):
            assert st[key] == out[0][3][key], key
