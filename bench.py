#!/usr/bin/env python3
"""Benchmark: field evals/s and frame mesh time of the MI355X path on the crown-splash frame.

One step = one full frame through the hot path (hash grid, anisotropy, iso constants, particle
normals, octree level sweeps, dual-cell extraction, bisection vertex placement) on particles
already resident in HBM, in the compensated FP32 mode (float terms summed in FP64, FP64 eigensolve
for clamped spectra). The same frame is also meshed by the all-FP64 GPU path, and the two are compared.

Prints one JSON line (see the keys below). The CPU baseline is the oracle (the reference's CPU
algorithm with its std::map cell tables and OpenMP loops) evaluating a bounded sample of the
frame's own octree sample points on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from particleisosurface_amd import capi, synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
# MI355X FP32 vector peak as published (packed FP32) and the rate without packed instructions,
# which this code keeps out: one FP32 FMA wave64 instruction per CU per cycle, 256 CUs at 2.4 GHz
FP32_PEAK_TFLOPS = 157.3
FP32_UNPACKED_TFLOPS = 78.6
N_CU, CLOCK_HZ = 256, 2.4e9
BYTES_PER_CAND = {capi.FP32: 52, capi.FP32_KAHAN: 52, capi.FP64: 104}
BYTES_FIXED = 16 + 27 * 8


def measured_traffic(kernel, workload_tag):
    """HBM bytes per launch of `kernel` on the workload tagged `workload_tag` (e.g. crown_splash_4M)
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    profiles/<kernel>_<workload_tag>_traffic.md (lines `traffic_bytes_per_launch <kernel> = <bytes>`),
    or None when no summary covers that kernel on that workload."""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(HERE, "profiles", "%s_%s_traffic.md" % (kernel, workload_tag)))):
        for line in open(f):
            m = re.match(r"traffic_bytes_per_launch (\w+) = (\d+)", line.strip())
            if m and m.group(1) == kernel:
                return int(m.group(2)), os.path.relpath(f, HERE)
    return None, None


def measured_issue(kernel, workload_tag):
    """VALU wave instructions per launch of `kernel` on the workload from the committed rocprofv3 SQ
    counter summary profiles/field_kernels_issue_<workload_tag>.md (line
    `valu_instr_per_launch <kernel> = <count>`), or (None, None)."""
    import re
    f = os.path.join(HERE, "profiles", "field_kernels_issue_%s.md" % workload_tag)
    if os.path.exists(f):
        for line in open(f):
            m = re.match(r"valu_instr_per_launch (\w+) = (\d+)", line.strip())
            if m and m.group(1) == kernel:
                return int(m.group(2)), os.path.relpath(f, HERE)
    return None, None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def device_memory_in_use():
    """Bytes of device memory in use (hipMemGetInfo), or None."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
            return None
        return total.value - free.value
    except OSError:
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frame(ctx, steps, warmup):
    for _ in range(warmup):
        ctx.run_frame()
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        info = ctx.run_frame()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    return info, st, dt


def sample_points_of_tree(ctx, info, r, limit, seed=3):
    """Octree sample points (27 per sampled node) of the last frame, as the sweep evaluates them."""
    nd = ctx.tree_nodes(info["n_nodes"])
    cw = (2.0 * nd["half"]).astype(np.float32)
    leafd = (cw - np.float32(r)).astype(np.float64) < 1e-8
    sel = (nd["type"] != capi.EMPTY) & (nd["depth"] >= info["depth_min"]) & ~leafd
    c = nd["center"][sel]
    h = nd["half"][sel][:, None]
    mn = (c - h).astype(np.float32)
    mx = (c + h).astype(np.float32)
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(c) * 27, size=min(limit, len(c) * 27), replace=False)
    node, s = pick // 27, pick % 27
    t = np.stack([s % 3, (s // 3) % 3, s // 9], 1).astype(np.float32) / np.float32(2)
    pts = ((np.float32(1) - t) * mn[node] + t * mx[node]).astype(np.float32)
    return pts


def mesh_error(v1, t1, v2, t2, r):
    """Vertex distance / r between two meshes of the same frame."""
    if len(v1) == 0 or len(v2) == 0:
        return None
    if v1.shape == v2.shape and np.array_equal(t1, t2):
        d = np.linalg.norm(v1 - v2, axis=1) / r
        how = "same topology"
    else:
        from scipy.spatial import cKDTree
        d, _ = cKDTree(v2).query(v1)
        d = d / r
        how = "nearest vertex"
    return dict(match=how, p50=float(np.median(d)), p99=float(np.quantile(d, 0.99)), max=float(d.max()),
                frac_within_1e3r=float(np.mean(d <= 1e-3)))


def aniso_roofline(ctx, ms_aniso, n, wtag):
    """The two anisotropy passes (xMean, then covariance and G): their algorithmic bytes per
    particle, [12 + 216 + 12 C + 12] + [12 + 216 + 12 C + 40] with C the candidates of its 27-cell
    window, over the anisotropy phase of a timed frame (host clock around both launches), and their
    VALU issue from the committed SQ summary. The candidates' positions and means are re-read from
    L1 / L2 by neighbouring threads, so the algorithmic rate is not an HBM rate. They issue about one
    VALU instruction per CU and cycle, but cutting their instructions by 15 % did not shorten them:
    the latency of their candidate loads limits them as much as issue does."""
    g = ctx.hashgrid_build()
    _, s, e = ctx.hashgrid_download(g["cellnum"])
    nx, ny, nz = (int(v) for v in g["cellnum3"])
    cnt = (e - s).astype(np.int64).reshape(nz, ny, nx)
    p = np.pad(cnt, 1)
    win = sum(p[z:z + nz, y:y + ny, x:x + nx] for z in range(3) for y in range(3) for x in range(3))
    cand = int((cnt * win).sum())
    alg = n * (12 + 216 + 12 + 12 + 216 + 40) + 24 * cand
    out = dict(kernels="k_xmean + k_gmat", phase_ms=ms_aniso, candidates_per_particle=cand / max(n, 1),
               algorithmic_bytes=alg, algorithmic_GBps=round(alg / (ms_aniso * 1e-3) / 1e9, 1), bound="valu")
    ix, src = measured_issue("k_xmean", wtag)
    ig, _ = measured_issue("k_gmat", wtag)
    if ix and ig:
        achieved = (ix + ig) * 64 * 2 / (ms_aniso * 1e-3) / 1e12
        out.update(valu_wave_instr=ix + ig, achieved=round(achieved, 2), unit="TFLOP/s", peak=FP32_PEAK_TFLOPS,
                   frac=round(achieved / FP32_PEAK_TFLOPS, 4), frac_unpacked=round(achieved / FP32_UNPACKED_TFLOPS, 4),
                   issue_source=src)
    return out


def cpu_baseline(pts_all, xyz, r, ctx, budget_s, box, gpu_subset_ms):
    """The oracle on the host cores: (1) SingleEval throughput on a bounded sample of the frame's
    own octree sample points (the bench's unit), (2) a whole frame (grid, anisotropy, constants,
    normals, octree, extraction) on a stated spatial subset of the workload, with the full-frame
    time extrapolated by particle count."""
    from oracle import pyoracle
    o = pyoracle.Oracle(xyz, r, fp64=False)
    t = time.perf_counter()
    o.hashgrid_build()
    xm, G, d = ctx.aniso()  # FP32 records of the same frame seed the oracle's evaluator
    o.set_aniso(xm, G, d)
    o.iso_constants()
    setup = time.perf_counter() - t
    done, spent = 0, 0.0
    chunk = 250_000
    while done < len(pts_all) and spent < budget_s:
        q = pts_all[done:done + chunk]
        t = time.perf_counter()
        o.eval_points(q)
        spent += time.perf_counter() - t
        done += len(q)
    del o
    sel = ((xyz[:, 0] >= box[0]) & (xyz[:, 0] < box[1]) & (xyz[:, 1] >= box[2]) & (xyz[:, 1] < box[3]))
    sub = xyz[sel]
    os_ = pyoracle.Oracle(sub, r, fp64=False)
    t = time.perf_counter()
    ph = os_.run()
    cpu_frame_s = time.perf_counter() - t
    return dict(value=done / spent, unit="field evals/s", cores=pyoracle.num_threads(), kind="port",
                cpu_model=cpu_model(),
                sample="%d octree sample points of the same frame (SingleEval with std::map cell tables, "
                       "OpenMP over points), %.1f s; setup %.1f s not timed" % (done, spent, setup),
                frame={"subset": "particles with x in [%g, %g) and y in [%g, %g): the crown, the droplets and "
                                 "the pool beneath them" % tuple(box),
                       "particles": int(len(sub)), "cpu_frame_s": round(cpu_frame_s, 3),
                       "cpu_phases_s": {k: round(v, 3) for k, v in ph.items()},
                       "gpu_frame_ms_same_subset": gpu_subset_ms,
                       "cpu_over_gpu_same_subset": round(1e3 * cpu_frame_s / gpu_subset_ms, 1) if gpu_subset_ms else None,
                       "extrapolated_full_frame_s": round(cpu_frame_s * len(xyz) / max(len(sub), 1), 2),
                       "extrapolation": "subset frame time x (frame particles / subset particles)"})


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                      "launches": 5}), flush=True)


def batch_leg(prec, dedupe, qef, lanes=(1, 2, 3, 4)):
    """24 frames of uneven size: one run_frames call per number of frames in flight (1: one
    context, buffers reused, the next frame staged while the current one is meshed; K > 1: K frames
    meshed at once on K contexts and streams), vs. a loop of the plain path (a fresh context per
    frame, as the reference builds a new SurfReconstructor per frame)."""
    t = time.perf_counter()
    frames = synth.batch_frames()
    log("batch: %d frames, %d particles (%.1f s to generate)" % (len(frames), sum(len(f) for f, _ in frames),
                                                                 time.perf_counter() - t))
    xs, rs = [f for f, _ in frames], [r for _, r in frames]
    tris = [0]

    def on_mesh(k, v, t_, tree):
        tris[0] += len(t_)

    ctx = capi.Context(0)
    ctx.set_params(rs[0], precision=prec, sample_dedupe=dedupe, qef_mode=qef)
    per_lanes, tri_counts = {}, set()
    for k in lanes:
        ctx.run_frames(xs, rs, frames_in_flight=k)  # warm-up: buffers, lane contexts, code objects, run shapes
        tris[0] = 0
        t = time.perf_counter()
        ctx.run_frames(xs, rs, on_mesh, frames_in_flight=k)
        per_lanes[k] = len(xs) / (time.perf_counter() - t)
        tri_counts.add(tris[0])
        log("batch: %d frame(s) in flight: %.1f frames/s" % (k, per_lanes[k]))
    ctx.close()
    best = max(per_lanes, key=per_lanes.get)
    dt_one = len(xs) / per_lanes[best]
    tris_one = tris[0] if len(tri_counts) == 1 else -1
    t = time.perf_counter()
    tris_l = 0
    for x, r in zip(xs, rs):
        c = capi.Context(0)
        c.set_params(r, precision=prec, sample_dedupe=dedupe, qef_mode=qef)
        c.upload(x)
        c.run_frame()
        tris_l += len(c.mesh()[1])
        c.close()
    dt_loop = time.perf_counter() - t
    return dict(frames=len(xs), particles=int(sum(len(x) for x in xs)),
                run_frames_frames_per_s=round(len(xs) / dt_one, 3), frames_in_flight=best,
                run_frames_frames_per_s_by_frames_in_flight={str(k): round(v, 3) for k, v in per_lanes.items()},
                loop_frames_per_s=round(len(xs) / dt_loop, 3), speedup=round(dt_loop / dt_one, 3),
                same_triangle_count=bool(tris_one == tris_l), host_input=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pool-layers", type=int, default=42, help="42: ~4.0 M particles, 20: geometry as first stated")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-qef", action="store_true", help="skip the FP64-QEF leg")
    ap.add_argument("--precision", default="fp32k", choices=["fp32", "fp32k", "fp64"])
    ap.add_argument("--no-dedupe", action="store_true", help="evaluate all 27 samples of every node")
    ap.add_argument("--qef", type=int, default=0, choices=[0, 1, 2],
                    help="leaf vertex: 0 centre (the reference as shipped), 1 float QEF, 2 FP64 truncated-SVD QEF")
    ap.add_argument("--workload", default="crown", choices=["crown", "pour"],
                    help="crown: crown splash (~4.0 M, r=0.001); pour: multi-body pour (~12.1 M, r=0.0005, depth 12)")
    ap.add_argument("--depth-max", type=int, default=-1, help="octree depth override (-1: 12 for pour, else derived)")
    ap.add_argument("--batch", action="store_true",
                    help="also mesh the 24-frame batch (50 k - 3 M particles per frame, radii by frame) "
                         "through pisurf_run_frames with 1-4 frames in flight and through a loop of fresh contexts")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) 5 times, to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings, and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    if args.workload == "pour":
        xyz, r = synth.multibody_pour()
        wname = "multi-body pour (config 5), %d particles, r=0.0005, pool 46 layers" % len(xyz)
        wtag = "multibody_pour_12M"
    else:
        xyz, r = synth.crown_splash(pool_layers=args.pool_layers)
        wname = "crown splash (config 3), %d particles, r=0.001, pool %d layers" % (len(xyz), args.pool_layers)
        wtag = "crown_splash_4M" if args.pool_layers == 42 else "crown_splash_%dlayers" % args.pool_layers
    depth = args.depth_max if args.depth_max >= 0 else (12 if args.workload == "pour" else 0)
    if args.precision != "fp32k":  # the committed traffic summaries cover compensated FP32
        wtag += "_" + args.precision
    if args.depth_max >= 0:
        wtag += "_depth%d" % args.depth_max
    log("workload: %s (%.1f s to generate)" % (wname, time.perf_counter() - t))

    prec = {"fp32": capi.FP32, "fp32k": capi.FP32_KAHAN, "fp64": capi.FP64}[args.precision]
    mem0 = device_memory_in_use()
    ctx = capi.Context(0)

    def set_params(precision):
        ctx.set_params(r, precision=precision, sample_dedupe=not args.no_dedupe, qef_mode=args.qef,
                       depth_max_override=depth)

    set_params(prec)
    ctx.upload(xyz)  # resident in HBM before the timed region
    info, st, dt = frame(ctx, args.steps, args.warmup)
    ms = 1e3 * dt / args.steps
    evals_per_frame = st["field_evals"] + st["grad_evals"]
    log("frame %.2f ms  nodes %d  evals/frame %d  phases(ms/frame): grid %.2f aniso %.2f normals %.2f tree %.2f "
        "extract-host %.2f bisect %.2f" % (ms, info["n_nodes"], evals_per_frame / args.steps,
                                            st["ms_grid"] / args.steps, st["ms_aniso"] / args.steps,
                                            st["ms_normals"] / args.steps, st["ms_tree"] / args.steps,
                                            st["ms_extract_host"] / args.steps, st["ms_bisect"] / args.steps))
    performed = st["sample_unique"] if st["sample_unique"] > 0 else st["sample_evals"]
    performed_per_frame = (evals_per_frame - st["sample_evals"] + performed) // args.steps
    # the headline: field evaluations the device performed per second of frame time (node samples
    # shared by neighbouring nodes are evaluated once); the reference's count (27 per sampled node)
    # is reported beside it
    value = performed_per_frame / (dt / args.steps)

    # roofline of the dominant field-evaluation kernel, the node-sampling kernel. Its limiter is VALU
    # instruction issue (SQ counters): achieved = its VALU wave instructions per launch (committed
    # rocprofv3 SQ summary) x 64 lanes x 2 flops, as FMA-equivalent TFLOP/s, over the launch time
    # measured here with HIP events. Measured HBM traffic and the algorithmic bytes (16 + 27*8 + R*C
    # per evaluation, C candidates) are reported beside it.
    def kernel_roof(name, evals, cands, launches, ms):
        k_ms = ms / max(launches, 1)
        k_s = k_ms * 1e-3
        alg = evals * BYTES_FIXED + cands * BYTES_PER_CAND[prec]
        traffic, source = measured_traffic(name, wtag)
        instr, isource = measured_issue(name, wtag)
        roof = dict(bound="valu", unit="TFLOP/s", peak=FP32_PEAK_TFLOPS, kernel=name,
                    launches_per_frame=launches / args.steps, avg_launch_ms=round(k_ms, 4),
                    evals_per_launch=evals / max(launches, 1), candidates_per_eval=cands / max(evals, 1),
                    evals_per_s=evals / (ms * 1e-3) if ms > 0 else None)
        if instr and k_s > 0:
            achieved = instr * 64 * 2 / k_s / 1e12
            roof.update(achieved=round(achieved, 2), frac=round(achieved / FP32_PEAK_TFLOPS, 4),
                        peak_unpacked=FP32_UNPACKED_TFLOPS, frac_unpacked=round(achieved / FP32_UNPACKED_TFLOPS, 4),
                        valu_wave_instr_per_launch=instr,
                        valu_wave_instr_per_cu_cycle=round(instr / (N_CU * CLOCK_HZ * k_s), 3),
                        issue_source=isource)
        else:
            roof.update(achieved=None, frac=None, issue_source=None)
        roof.update(traffic=traffic, traffic_source=source,
                    hbm_achieved_GBps=round(traffic / k_s / 1e9, 1) if traffic and k_s > 0 else None,
                    hbm_frac=round(traffic / k_s / 1e9 / HBM_PEAK_GBS, 4) if traffic and k_s > 0 else None,
                    algorithmic_bytes_per_launch=alg / max(launches, 1),
                    algorithmic_GBps=round(alg / max(launches, 1) / k_s / 1e9, 1) if k_s > 0 else None)
        return roof
    if st["run_launches"] > 0:
        roof = kernel_roof("k_run_samples", st["run_evals"], st["run_candidates"], st["run_launches"],
                           st["ms_run_kernel"])
    else:
        roof = kernel_roof("k_samples", st["sample_evals"], st["sample_candidates"], st["sample_launches"],
                           st["ms_sample_kernel"])
    v32, t32 = ctx.mesh()
    # the same frame from host particle arrays to host mesh arrays (PCIe upload included)
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.upload(xyz)
        ctx.run_frame()
    host_ms = 1e3 * (time.perf_counter() - t) / args.steps
    out = {
        "metric": "field evals/s + frame mesh time @4M particles, FP32-comp vs FP64 at stated err",
        "value": round(value, 1), "unit": "field evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32k": "f32 (sums in f64; f64 eigensolve when clamped)", "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": wname, "precision": args.precision, "qef_mode": args.qef,
                   "depth_max": info["depth_max"], "depth_min": info["depth_min"]},
        "qef_solves_per_frame": st["qef_solves"] // args.steps,
        "device_memory_bytes": device_memory_in_use() - mem0 if mem0 is not None else None,
        "frame_ms": round(ms, 3),
        "frame_ms_from_host_particles": round(host_ms, 3),
        "field_evals_performed_per_frame": performed_per_frame,
        "field_evals_reference_count_per_frame": evals_per_frame // args.steps,
        "field_evals_reference_count_per_s": round(evals_per_frame / dt, 1),
        "phases_ms": {k[3:]: round(st[k] / args.steps, 3) for k in
                      ["ms_grid", "ms_aniso", "ms_normals", "ms_tree", "ms_extract_host", "ms_bisect"]},
        "mesh": {"verts": len(v32), "tris": len(t32), "nodes": info["n_nodes"]},
        "fp64_eigensolves": st["fp64_eigensolves"],
        "roofline": roof,
    }

    # FP64 comparison on the same frame
    if not args.no_fp64 and prec != capi.FP64:
        rng = np.random.default_rng(1)
        qi = rng.choice(len(xyz), 200_000, replace=False)
        q = (xyz[qi] + rng.uniform(-2 * r, 2 * r, (len(qi), 3))).astype(np.float32)
        f32 = ctx.eval_points(q)
        n32 = ctx.particle_normals()
        iso = ctx.iso_constants()[1]
        set_params(capi.FP64)
        info64, st64, dt64 = frame(ctx, 1, 1)
        ms64 = 1e3 * dt64
        v64, t64 = ctx.mesh()
        f64 = ctx.eval_points(q)
        n64 = ctx.particle_normals()
        S32, S64 = iso - f32, iso - f64
        mask = S64 > 1e-3 * iso
        rel = np.abs(S32 - S64)[mask] / S64[mask]
        # the same error against max(S64, iso), the form the parity tests bound: sums far below
        # the iso level are measured against the iso level
        rel_iso = np.abs(S32 - S64) / np.maximum(np.abs(S64), iso)
        nn = np.linalg.norm(n32, axis=1) * np.linalg.norm(n64, axis=1)
        ok = nn > 0
        ang = np.degrees(np.arctan2(np.linalg.norm(np.cross(n32[ok], n64[ok]), axis=1),
                                    np.sum(n32[ok] * n64[ok], axis=1)))
        out["fp64_frame_ms"] = round(ms64, 3)
        out["speedup_vs_fp64"] = round(ms64 / ms, 3)
        out["error_vs_fp64"] = {
            "field_rel_p99": float(np.quantile(rel, 0.99)), "field_rel_max": float(rel.max()),
            "field_rel_iso_floor_max": float(rel_iso.max()),
            "field_points": int(mask.sum()),
            "sign_mismatch_outside_2e-5_iso": int(np.sum(((f32 > 0) != (f64 > 0)) & (np.abs(f64) > 2e-5 * iso))),
            "normal_angle_deg_p99": float(np.quantile(ang, 0.99)), "normal_angle_deg_max": float(ang.max()),
            "vertex_over_r": mesh_error(v32, t32, v64, t64, r),
            "same_tree": bool(info64["n_nodes"] == info["n_nodes"]),
        }
        log("fp64 frame %.2f ms, speedup %.2fx" % (ms64, ms64 / ms))
        set_params(prec)
        info, st, dt = frame(ctx, 1, 0)

    if not args.no_qef and args.qef == 0:
        # the leaf vertex by the FP64 truncated-SVD QEF (TNode::NodeCalcNode, dormant in the
        # reference) in the same precision mode: solves per frame and per second of frame time
        ctx.set_params(r, precision=prec, sample_dedupe=not args.no_dedupe, qef_mode=capi.QEF_SVD64,
                       depth_max_override=depth)
        infoq, stq, dtq = frame(ctx, args.steps, 1)
        msq = 1e3 * dtq / args.steps
        out["qef"] = {"mode": "fp64 truncated-SVD pseudo-inverse (tol 1e-3)", "frame_ms": round(msq, 3),
                      "solves_per_frame": stq["qef_solves"] // args.steps,
                      "solves_per_s": round(stq["qef_solves"] / dtq, 1), "nodes": infoq["n_nodes"]}
        log("qef frame %.2f ms, %d solves" % (msq, out["qef"]["solves_per_frame"]))
        set_params(prec)
        info, st, dt = frame(ctx, 1, 0)

    if args.batch:
        out["batch"] = batch_leg(prec, not args.no_dedupe, args.qef)
        log("batch: %s" % out["batch"])

    if not args.no_cpu:
        pts = sample_points_of_tree(ctx, info, r, limit=20_000_000)
        box = (0.2, 0.4, 0.2, 0.4) if args.workload == "crown" else (0.2, 0.3, 0.2, 0.3)
        sel = (xyz[:, 0] >= box[0]) & (xyz[:, 0] < box[1]) & (xyz[:, 1] >= box[2]) & (xyz[:, 1] < box[3])
        sub_ctx = capi.Context(0)
        sub_ctx.set_params(r, precision=capi.FP32, sample_dedupe=not args.no_dedupe, depth_max_override=depth)
        sub_ctx.upload(xyz[sel])
        _, _, dts = frame(sub_ctx, 3, 1)
        sub_ctx.close()
        out["cpu_baseline"] = cpu_baseline(pts, xyz, r, ctx, args.cpu_seconds, box, round(1e3 * dts / 3, 3))
        log("cpu baseline %.3g evals/s on %d threads; subset frame %.1f s (%d particles)" % (
            out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"], out["cpu_baseline"]["frame"]["cpu_frame_s"],
            out["cpu_baseline"]["frame"]["particles"]))
    else:
        out["cpu_baseline"] = None
    out["aniso_roofline"] = aniso_roofline(ctx, out["phases_ms"]["aniso"], len(xyz), wtag)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
