#!/usr/bin/env python3
"""Benchmark of the batched collision-query hot path on one MI355X.

Primary line: batched TriMesh ray casts (Mrays/s).  The broadphase +
narrowphase frame (contact pairs/s) is measured in the same run and reported
in the "contact" object of the same JSON line.

Workload (one "step" = one batched launch over the whole ray set, inputs
resident in HBM): a noise-displaced icosphere at subdivision 6 (81,920
triangles, 40,962 vertices, radius 1, radial noise U[-0.05, 0.05], seed 1)
and 1,048,576 rays (origins uniform on the radius-3 sphere, aimed at points
uniform in the unit ball, seed 2), max_toi = 10, solid = true, output per ray:
hit flag, triangle/feature index, toi, normal.

Prints ONE JSON line.  `roofline.achieved` counts the algorithmic bytes of
parry's traversal: per ray 24 B (ray) + 20 B (result) + 64 B per Bvh wide
node fetched + 48 B per triangle tested, with the fetch/test counts taken
from the kernel's own instrumented launch (they equal the oracle's), divided
by the kernel's average duration measured with HIP events on its stream.
The CPU baseline is the oracle (C++ restatement of parry3d, f32) over all
host cores on the same inputs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s & contact pairs/s on 1 MI355X; learned vs plain GJK at 1e-5 tol"
HBM_PEAK_GBS = 8000.0
MAX_TOI = 10.0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")


def measured_traffic(kernel: str, field: str = "hbm_bytes_per_launch"):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's readings), or
    another field of its entry; None if absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)[kernel][field]
    except (OSError, KeyError, ValueError):
        return None


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 vector peak


def f64_error(val32, val64, ok=None) -> dict:
    """Error of f32 results (rows of values) against the f64 oracle on the same
    inputs: max and 99.9th-percentile absolute error, 99th-percentile relative
    error (|d| / max(1, |f64|), worst component per row) and the share of rows
    within 1e-5 relative."""
    import numpy as np

    a = np.asarray(val32, np.float64)
    b = np.asarray(val64, np.float64)
    if ok is not None:
        a, b = a[ok], b[ok]
    if a.shape[0] == 0:
        return {"rows": 0}
    d = np.abs(a - b).reshape(a.shape[0], -1)
    rel = (d / np.maximum(1.0, np.abs(b).reshape(a.shape[0], -1))).max(1)
    return {"rows": int(a.shape[0]), "abs_err_max": float(d.max()), "abs_err_p999": float(np.quantile(d.max(1), 0.999)),
            "rel_err_p99": float(np.quantile(rel, 0.99)), "within_1e-5_rel": round(float((rel <= 1e-5).mean()), 6)}


def contact_f64_error(gpu_out, descs, vertices, poses, pairs, prediction, init_dirs=None, threads=None) -> dict:
    """query::contact of the GPU (f32) against the f64 oracle on the same pairs
    and poses (given seeds, if any): status agreement, then dist, normal1 and the
    two points.  The f32 oracle's EPA exit (bit-identical to the GPU's path)
    splits the dist error: a converged EPA returns the best face's plane, which
    the two precisions can reach by different paths."""
    import numpy as np

    from oracle import oracle

    r64 = oracle.contact_pairs(descs, np.asarray(vertices, np.float64), np.asarray(poses, np.float64), pairs,
                               prediction, init_dirs=None if init_dirs is None else np.asarray(init_dirs, np.float64),
                               f64=True, threads=threads)
    r32 = oracle.contact_pairs(descs, vertices, poses, pairs, prediction, init_dirs=init_dirs, threads=threads)
    has = gpu_out.view(np.uint32)[:, 13].astype(np.int64)
    both = (has == 1) & (r64["status"] == 1)
    rec = {"pairs": int(pairs.shape[0]), "status_agreement": round(float((has == r64["status"]).mean()), 7),
           "dist": f64_error(gpu_out[:, 12], r64["out"][:, 12], both),
           "normal1": f64_error(gpu_out[:, 6:9], r64["out"][:, 6:9], both),
           "points": f64_error(gpu_out[:, 0:6], r64["out"][:, 0:6], both), "dist_by_epa_exit": {}}
    for k, name in enumerate(oracle.EPA_EXIT_NAMES):
        sel = both & (r32["epa_exit"] == k)
        if sel.any():
            rec["dist_by_epa_exit"][name] = f64_error(gpu_out[:, 12], r64["out"][:, 12], sel)
    return rec


def ray_f64_error(v, f, rays, hits, threads=None) -> dict:
    """Ray casts of the GPU (f32, (n, 5) pq_ray_hit rows) against the f64 oracle
    on the same mesh and rays: hit/feature agreement, then toi and normal error
    where both hit the same triangle."""
    import numpy as np

    from oracle import oracle

    ref = oracle.OracleTriMesh(np.asarray(v, np.float64), f, f64=True).cast_rays(
        np.asarray(rays, np.float64), MAX_TOI, threads=threads)
    feat = hits[:, 4].view(np.uint32)
    same = feat == ref["feature"]
    hit = same & (feat != np.uint32(0xFFFFFFFF))
    return {"rays": int(rays.shape[0]), "feature_agreement": round(float(same.mean()), 7),
            "toi": f64_error(hits[:, 0], ref["toi"], hit), "normal": f64_error(hits[:, 1:4], ref["normal"], hit)}


def contact_roofline(nbytes: int, flops: int, ms: float, kernel: str) -> dict:
    """HBM fraction on algorithmic bytes and FP32 fraction on the oracle's
    counted flops (add/sub, mul, div, sqrt) of the same pairs."""
    gbs = nbytes / (ms * 1e-3) / 1e9
    tf = flops / (ms * 1e-3) / 1e12
    return {"bound": "latency (GJK/EPA branchy scalar code)", "kernel": kernel, "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
            "algorithmic_bytes_per_launch": int(nbytes), "kernel_ms": round(ms, 4),
            "fp32": {"achieved_tflops": round(tf, 4), "peak_tflops": FP32_PEAK_TFLOPS,
                     "frac": round(tf / FP32_PEAK_TFLOPS, 6), "algorithmic_flops_per_launch": int(flops),
                     "source": "add/sub, mul, div and sqrt counted by the oracle's query::contact on the same "
                               "pairs and seeds"}}


def contact_bytes(descs, pairs) -> int:
    """Algorithmic bytes of query::contact per call: 12 B per polyhedron or
    triangle vertex of both shapes, two 28-B poses and the 56-B result per pair."""
    import numpy as np

    t = np.asarray(descs["type"])
    nv = np.asarray(descs["num_vertices"], np.int64) * ((t == 4) | (t == 5))
    pr = np.asarray(pairs).astype(np.int64)
    return int((12 * (nv[pr[:, 0]] + nv[pr[:, 1]]) + 2 * 28 + 56).sum())


def cpu_baseline(v, f, rays, gpu_feat):
    """Times the oracle on the host cores; returns (record, parity summary)."""
    import numpy as np

    from oracle import oracle

    cores = oracle.default_threads()
    om = oracle.OracleTriMesh(v, f)
    n = rays.shape[0]
    om.cast_rays(rays[: 1 << 14], MAX_TOI, threads=cores)  # warm-up
    times, ref = [], None
    for _ in range(5):
        t0 = time.perf_counter()
        ref = om.cast_rays(rays, MAX_TOI, threads=cores)
        times.append(time.perf_counter() - t0)
    t_all = sorted(times)[len(times) // 2]
    sample = 1 << 17
    t0 = time.perf_counter()
    om.cast_rays(rays[:sample], MAX_TOI, threads=1)
    t_one = time.perf_counter() - t0
    mism = int(np.count_nonzero(ref["feature"] != gpu_feat))
    rec = {
        "value": round(n / t_all / 1e6, 4),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"all {n} rays of the workload, median of 5 runs on {cores} threads "
        f"(C++ restatement of parry3d f32, -O3 -ffp-contract=off); single thread on the first "
        f"{sample} rays: {sample / t_one / 1e6:.4f} Mrays/s",
        "single_thread_value": round(sample / t_one / 1e6, 4),
    }
    return rec, {"feature_mismatches_vs_oracle": mism, "rays_checked": n}


def batch_leg(args, log):
    """64 noisy icospheres at levels 3-7 (uniform), 2^X rays each with X ~ U[10, 20]
    (seed 2): one step = one batched launch over every mesh's rays, replayed
    from a captured hipGraph after the autotuner picked the lane mapping.
    Compared with 64 back-to-back single-mesh launches of the same rays."""
    import numpy as np
    import torch

    from parry_amd import TriMesh, TriMeshBatch, scenes

    spec = scenes.uneven_mesh_batch(64, seed=2)
    meshes, rays, hosts = [], [], []
    for level, nr, mseed, rseed in spec:
        v, f = scenes.noisy_icosphere(level, seed=mseed)
        meshes.append(TriMesh(v, f))
        hosts.append((v, f))
        rays.append(scenes.rays_into_unit_ball(nr, seed=rseed))
    counts = [r.shape[0] for r in rays]
    n = int(sum(counts))
    ntri = int(sum(f.shape[0] for _, f in hosts))
    batch = TriMeshBatch(meshes, counts)
    stream = torch.cuda.current_stream()
    d_rays = torch.from_numpy(np.concatenate(rays)).cuda()
    out = torch.empty((n, 5), dtype=torch.float32, device="cuda")
    _, cnt = batch.cast_local_rays_counted(d_rays, MAX_TOI)
    sum_w = int(cnt[:, 0].sum(dtype=torch.int64).item())
    sum_k = int(cnt[:, 1].sum(dtype=torch.int64).item())
    alg_bytes = n * 44 + 64 * sum_w + 48 * sum_k

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3, e0.elapsed_time(e1) / steps

    # 64 separate single-mesh launches (static mapping) on slices of the same buffers
    offs = batch.offsets.astype(np.int64)
    views = [(m, d_rays[offs[i]:offs[i + 1]], out[offs[i]:offs[i + 1]]) for i, m in enumerate(meshes)]
    for m in meshes:
        m.set_tuning("static")

    def per_mesh():
        for m, r, o in views:
            if r.shape[0]:
                m.cast_local_rays_and_get_normal(r, MAX_TOI, True, out=o, stream=stream)

    sep_wall, sep_gpu = timed(per_mesh, max(1, args.steps // 2))
    ref = out.clone()
    modes = {}
    for mapping, chunk in (("static", 128), ("persistent", 128), ("binned", 128)):
        batch.set_tuning(mapping, chunk)
        modes[f"{mapping}/{chunk}"] = timed(
            lambda: batch.cast_local_rays_and_get_normal(d_rays, MAX_TOI, True, out=out, stream=stream), args.steps)
    if args.batch_tuning:
        m, c, bpc, blk, regs = args.batch_tuning.split(":")
        batch.set_tuning(m, int(c), int(bpc), int(blk), int(regs))
        tuning, tune_ms = batch.tuning(), float("nan")
    else:
        tuning, tune_ms = batch.autotune(d_rays, MAX_TOI, out=out)
    graph = batch.capture(d_rays, MAX_TOI, True, out)
    # counter summary of the kernel the tuned mapping launches (null if the summary is of another mapping)
    batch_traffic_entry = {"static": "cast_rays_batch_kernel", "binned": "cast_rays_batch_step_binned"}.get(
        tuning["mapping"], "cast_rays_batch_persistent_kernel")
    batch_traffic = measured_traffic(batch_traffic_entry)
    batch_traffic_tuning = measured_traffic(batch_traffic_entry, "tuning")
    wall_ms, gpu_ms = timed(lambda: graph.launch(stream), args.steps)
    identical = bool(torch.equal(out.view(torch.int32), ref.view(torch.int32)))
    rec = {
        "metric": "Mrays/s",
        "value": round(n / (wall_ms * 1e-3) / 1e6, 3),
        "unit": "Mrays/s",
        "ms_per_step": round(wall_ms, 4),
        "gpu_ms_per_step": round(gpu_ms, 4),
        "workload": f"64 noisy icospheres, levels 3-7 ({ntri} tris), 2^U(10,20) rays each ({n} rays), seed 2; "
                    "one step = one hipGraph replay of the batched launch (autotuned mapping)",
        "rays": n,
        "triangles": ntri,
        "tuning": tuning,
        "autotune_ms": round(tune_ms, 4),
        "mode_ms": {k: round(v[1], 4) for k, v in modes.items()},
        "separate_launches": {"wall_ms": round(sep_wall, 4), "gpu_ms": round(sep_gpu, 4),
                              "value": round(n / (sep_wall * 1e-3) / 1e6, 3)},
        "roofline": {
            "bound": "hbm",
            "kernel": ("bin_count_batch_kernel + bin_scan_kernel + bin_cells_kernel + bin_scatter_batch_kernel"
                       " + cast_rays_batch_kernel" if tuning["mapping"] == "binned" else "cast_rays_batch_*_kernel"),
            "achieved": round(alg_bytes / (gpu_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(alg_bytes / (gpu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": alg_bytes,
            "traffic": batch_traffic if batch_traffic_tuning and batch_traffic_tuning.get("mapping") == tuning["mapping"]
            else None,
            "traffic_entry": batch_traffic_entry,
            "traffic_tuning": batch_traffic_tuning,
        },
        "parity": {"identical_to_separate_launches": identical},
    }
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        take = 1 << 14
        om = [oracle.OracleTriMesh(v, f) for v, f in hosts]
        t0 = time.perf_counter()
        got = 0
        mism = 0
        host_out = out.cpu().numpy().view(np.uint32)[:, 4]
        for i, (o, r) in enumerate(zip(om, rays)):
            k = min(take, r.shape[0])
            res = o.cast_rays(r[:k], MAX_TOI, threads=cores)
            mism += int(np.count_nonzero(res["feature"] != host_out[offs[i]:offs[i] + k]))
            got += k
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(got / t_cpu / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
                               "sample": f"first min(16384, n_i) rays of each of the 64 meshes ({got} rays), "
                                         f"{cores} threads"}
        rec["parity"]["feature_mismatches_vs_oracle_sample"] = mism
        # f64 error on the same sample (the first 4096 rays of every mesh)
        hits = out.cpu().numpy()
        errs = [ray_f64_error(v, f, r[:4096], hits[offs[i]:offs[i] + min(4096, r.shape[0])], threads=cores)
                for i, ((v, f), r) in enumerate(zip(hosts, rays))]
        nr = sum(e["rays"] for e in errs)
        rec["f64_error"] = {
            "rays": nr, "feature_agreement": round(sum(e["feature_agreement"] * e["rays"] for e in errs) / nr, 7),
            "toi_abs_err_max": max(e["toi"].get("abs_err_max", 0.0) for e in errs),
            "normal_abs_err_max": max(e["normal"].get("abs_err_max", 0.0) for e in errs),
            "sample": "the first min(4096, n_i) rays of each of the 64 meshes"}
    return rec


def mapping_sweep(args, log):
    """Every ray-kernel mapping the tuner considers (lane mappings over workgroup
    sizes and register stack depths, persistent chunk sizes, wave per ray), timed
    on the C0 mesh and rays and on three small-batch workloads, plus the
    autotuner's pick for each; every setting's hits are checked against the
    default mapping's, bit for bit."""
    import numpy as np
    import torch

    from parry_amd import TriMesh, scenes
    from parry_amd import _native

    cands = ([{"mapping": "static", "block_size": b, "stack_regs": r} for b in (64, 128, 256) for r in (0, 1, 2, 4)]
             + [{"mapping": "binned", "block_size": b, "stack_regs": r} for b in (64, 128, 256) for r in (0, 2)]
             + [{"mapping": "persistent", "block_size": b, "stack_regs": r, "chunk": c}
                for b in (64, 256) for r in (0, 2) for c in (64, 128, 256)]
             + [{"mapping": "wave", "chunk": c} for c in (1, 4)])
    out_rec = []
    for level, n, label in ((6, 1 << 20, "C0 mesh, 1048576 rays"), (7, 65536, "icosphere s=7, 65536 rays"),
                            (7, 4096, "icosphere s=7, 4096 rays"), (5, 1024, "icosphere s=5, 1024 rays")):
        v, f = scenes.noisy_icosphere(level, seed=1)
        rays = scenes.rays_into_unit_ball(n, seed=2)
        mesh = TriMesh(v, f)
        d = torch.from_numpy(rays).cuda()
        out = torch.empty((n, 5), dtype=torch.float32, device="cuda")
        mesh.set_tuning("static")
        mesh.cast_local_rays_and_get_normal(d, MAX_TOI, True, out=out)
        torch.cuda.synchronize()
        ref = out.clone()
        rows = []
        for c in cands:
            mesh.set_tuning(**c)
            for _ in range(2):
                mesh.cast_local_rays_and_get_normal(d, MAX_TOI, True, out=out)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            steps = 5 if c["mapping"] == "wave" else 10
            e0.record()
            for _ in range(steps):
                mesh.cast_local_rays_and_get_normal(d, MAX_TOI, True, out=out)
            e1.record()
            torch.cuda.synchronize()
            same = bool(torch.equal(out.view(torch.int32), ref.view(torch.int32)))
            rows.append({**c, "ms": round(e0.elapsed_time(e1) / steps, 4), "identical": same})
        tu, ms = mesh.autotune(d, MAX_TOI, out=out)
        log(f"mapping sweep {label}: autotune chose {tu} at {ms:.4f} ms")
        out_rec.append({"workload": label, "triangles": int(f.shape[0]), "rays": n, "settings": rows,
                        "autotuned": tu, "autotuned_ms": round(ms, 4)})
    return out_rec


def queries_leg(args, log):
    """Two more TriMesh queries on the C0 mesh (noisy icosphere s=6, 81,920 triangles):
    point projection (TriMesh::project_local_point, solid = false) of 1,048,576 points
    uniform in [-1.5, 1.5]^3 (seed 6), and the C0 rays cast with back faces culled
    (TriMesh::cast_local_ray_with_culling); one step = one call over every point
    (ray): for the culled casts with the binned mapping, the sort and one cast launch.  Algorithmic bytes per query: 12 (point) + 20 (record), or 24 (ray) + 20
    (hit), plus 64 per wide node fetched and 48 per triangle tested, the counts
    coming from the kernels' instrumented launch."""
    import numpy as np
    import torch

    from parry_amd import TriMesh, scenes, split_hits, split_projections

    v, f = scenes.noisy_icosphere(6, seed=1)
    mesh = TriMesh(v, f)
    stream = torch.cuda.current_stream()
    n = 1 << 20
    pts = np.random.default_rng(6).uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    rays = scenes.rays_into_unit_ball(n, seed=2)
    d_pts = torch.from_numpy(pts).cuda()
    d_rays = torch.from_numpy(rays).cuda()
    out = torch.empty((n, 5), dtype=torch.float32, device="cuda")
    counts = torch.empty((n, 2), dtype=torch.int32, device="cuda")

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps * 1e3, e0.elapsed_time(e1) / args.steps

    def roofline(kernel, per_query, gpu_ms):
        w = int(counts[:, 0].sum(dtype=torch.int64).item())
        k = int(counts[:, 1].sum(dtype=torch.int64).item())
        alg = n * per_query + 64 * w + 48 * k
        gbs = alg / (gpu_ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": alg,
                "mean_wide_nodes": round(w / n, 3), "mean_triangles": round(k / n, 3)}

    mesh.project_local_points(d_pts, False, out=out, counts=counts)
    p_wall, p_gpu = timed(lambda: mesh.project_local_points(d_pts, False, out=out, stream=stream))
    proj = out.cpu().numpy().copy()
    rec_p = {"metric": "Mpoints/s", "value": round(n / (p_wall * 1e-3) / 1e6, 3), "unit": "Mpoints/s",
             "ms_per_step": round(p_wall, 4), "gpu_ms_per_step": round(p_gpu, 4),
             "workload": "TriMesh::project_local_point, solid=false: noisy icosphere s=6 (81920 tris) x 1048576 "
                         "points uniform in [-1.5,1.5]^3, seed 6",
             "roofline": roofline("project_points_kernel", 32, p_gpu)}
    mesh.cast_local_rays_with_culling(d_rays, MAX_TOI, "ignore_backfaces", out=out, counts=counts)
    mesh.set_tuning("static")
    s_wall, s_gpu = timed(lambda: mesh.cast_local_rays_with_culling(d_rays, MAX_TOI, "ignore_backfaces", out=out,
                                                                    stream=stream))
    static_hits = out.cpu().numpy().copy()
    mesh.set_tuning("binned")  # the C0 rays' mapping: the culled casts sort them too
    c_wall, c_gpu = timed(lambda: mesh.cast_local_rays_with_culling(d_rays, MAX_TOI, "ignore_backfaces", out=out,
                                                                    stream=stream))
    hits = out.cpu().numpy().copy()
    rec_c = {"metric": "Mrays/s", "value": round(n / (c_wall * 1e-3) / 1e6, 3), "unit": "Mrays/s",
             "ms_per_step": round(c_wall, 4), "gpu_ms_per_step": round(c_gpu, 4),
             "workload": "TriMesh::cast_local_ray_with_culling(IgnoreBackfaces): the C0 mesh and its 1048576 rays, "
                         "max_toi=10",
             "mapping": "binned (the sort, then cast_rays_kernel<false, 1, 256, 0> over the sorted order)",
             "static_mapping": {"ms_per_step": round(s_wall, 4), "gpu_ms_per_step": round(s_gpu, 4),
                                "identical": bool(np.array_equal(static_hits.view(np.uint32), hits.view(np.uint32)))},
             "roofline": roofline("bin_*_kernel + cast_rays_kernel<false, 1, 256, 0>", 44, c_gpu)}
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        om = oracle.OracleTriMesh(v, f)

        def cpu_time(fn):
            fn()
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                r = fn()
                ts.append(time.perf_counter() - t0)
            return sorted(ts)[1], r

        t, ref = cpu_time(lambda: om.project_points(pts, False, threads=cores))
        rec_p["cpu_baseline"] = {"value": round(n / t / 1e6, 4), "unit": "Mpoints/s", "cores": cores, "kind": "port",
                                 "sample": f"all {n} points, median of 3 runs on {cores} threads"}
        rec_p["parity"] = {
            "mismatching_records_vs_oracle": int(np.count_nonzero(
                (proj[:, :3] != ref["point"]).any(axis=1) | (proj.view(np.uint32)[:, 3] != ref["is_inside"])
                | (proj.view(np.uint32)[:, 4] != ref["triangle"])))}
        t, ref = cpu_time(lambda: om.cast_rays_with_culling(rays, MAX_TOI, 1, threads=cores))
        rec_c["cpu_baseline"] = {"value": round(n / t / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
                                 "sample": f"all {n} rays, median of 3 runs on {cores} threads"}
        rec_c["parity"] = {"feature_mismatches_vs_oracle": int(np.count_nonzero(
            hits.view(np.uint32)[:, 4] != ref["feature"]))}
        o64 = oracle.OracleTriMesh(np.asarray(v, np.float64), f, f64=True)
        r64 = o64.project_points(pts.astype(np.float64), False, threads=cores)
        same = proj.view(np.uint32)[:, 4] == r64["triangle"]
        rec_p["f64_error"] = {"points": n, "triangle_agreement": round(float(same.mean()), 7),
                              "is_inside_agreement": round(float((proj.view(np.uint32)[:, 3] == r64["is_inside"]).mean()), 7),
                              "projection": f64_error(proj[:, :3], r64["point"], same)}
        r64 = o64.cast_rays_with_culling(rays.astype(np.float64), MAX_TOI, 1, threads=cores)
        feat = hits.view(np.uint32)[:, 4]
        same = (feat == r64["feature"]) & (feat != np.uint32(0xFFFFFFFF))
        rec_c["f64_error"] = {"rays": n, "feature_agreement": round(float((feat == r64["feature"]).mean()), 7),
                              "toi": f64_error(hits[:, 0], r64["toi"], same),
                              "normal": f64_error(hits[:, 1:4], r64["normal"], same)}
    return {"project_points": rec_p, "culling": rec_c}


def mesh_contact_leg(args, log):
    """query::contact between one TriMesh and many convex shapes
    (contact_composite_shape_shape.rs:14-59): 262,144 mixed shapes (Ball, Cuboid,
    Capsule, Cone, ConvexPolyhedron of 8-64 points; scale 0.15, seed 13) resting on
    a noisy icosphere s=6 of radius 8 (81,920 triangles), centres at 0.96-1.04 x
    the radius; prediction 0.01.  One step = one pq_mesh_contact_shapes call: the
    Bvh candidate walk (count, scan, write), query::contact on every (triangle,
    shape) candidate, and the per-shape selection."""
    import numpy as np
    import torch

    from parry_amd import MeshContact, ShapeSet, TriMesh, scenes

    v, f = scenes.noisy_icosphere(6, seed=1, radius=8.0)
    n = 1 << 18
    descs, verts, poses_np = scenes.shapes_around_mesh(n, seed=13, mesh_radius=8.0, scale=0.15, shell=(0.48, 0.52))
    mesh = TriMesh(v, f)
    shapes = ShapeSet(descs, verts)
    mc = MeshContact(mesh, shapes)
    stream = torch.cuda.current_stream()
    mesh_pose = torch.tensor([0, 0, 0, 1, 0, 0, 0], dtype=torch.float32, device="cuda")
    poses = torch.from_numpy(poses_np).cuda()
    needed = torch.zeros(1, dtype=torch.int64, device="cuda")
    out = torch.empty((n, 14), dtype=torch.float32, device="cuda")
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    ws0 = torch.empty(mc.workspace_bytes(0), dtype=torch.uint8, device="cuda")
    mc.contact(mesh_pose, poses, 0.01, 0, out=out, triangles=tri, workspace=ws0, needed=needed)
    cap = int(needed.item())
    del ws0
    ws = torch.empty(mc.workspace_bytes(cap), dtype=torch.uint8, device="cuda")

    def step():
        mc.contact(mesh_pose, poses, 0.01, cap, out=out, triangles=tri, workspace=ws, stream=stream, needed=needed)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    o = out.cpu().numpy()
    has = o.view(np.uint32)[:, 13]
    got_tri = tri.cpu().numpy().view(np.uint32)
    rec = {
        "metric": "shape queries/s",
        "value": round(n / wall, 1),
        "unit": "shapes/s",
        "ms_per_step": round(wall * 1e3, 4),
        "gpu_ms_per_step": round(gpu_ms, 4),
        "workload": "query::contact(TriMesh, shape) for 262144 mixed shapes (scale 0.15, seed 13) on a noisy "
                    "icosphere s=6 of radius 8 (81920 tris), centres at 0.96-1.04 x radius, prediction 0.01",
        "candidates": cap,
        "mean_candidates_per_shape": round(cap / n, 3),
        "contacts": int((has == 1).sum()),
        "candidates_per_s": round(cap / wall, 1),
    }
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        om = oracle.OracleTriMesh(v, f)
        t0 = time.perf_counter()
        ref = om.contact_shapes(np.array([0, 0, 0, 1, 0, 0, 0], np.float32), descs, verts, poses_np, 0.01,
                                threads=cores)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(n / t_cpu, 1), "unit": "shapes/s", "cores": cores, "kind": "port",
                               "sample": f"all {n} shapes, one run on {cores} threads (C++ restatement of parry3d f32)"}
        c = ref["status"] == 1
        rec["parity"] = {
            "has_contact_mismatches": int((has.astype(np.int64) != ref["status"].astype(np.int64)).sum()),
            "triangle_mismatches": int((got_tri[c] != ref["triangle"][c]).sum()),
            "bit_identical_contacts": int(np.all(o[c, :13] == ref["out"][c], axis=1).sum()),
            "candidates_oracle": int(ref["candidates"]),
        }
        r64 = oracle.OracleTriMesh(np.asarray(v, np.float64), f, f64=True).contact_shapes(
            np.array([0, 0, 0, 1, 0, 0, 0], np.float64), descs, np.asarray(verts, np.float64),
            np.asarray(poses_np, np.float64), 0.01, threads=cores)
        both = (has == 1) & (r64["status"] == 1)
        rec["f64_error"] = {"shapes": n, "status_agreement": round(float((has.astype(np.int64) == r64["status"]).mean()), 7),
                            "triangle_agreement": round(float((got_tri[both] == r64["triangle"][both]).mean()), 7),
                            "dist": f64_error(o[:, 12], r64["out"][:, 12], both),
                            "normal1": f64_error(o[:, 6:9], r64["out"][:, 6:9], both)}
    return rec


def cast_leg(args, log):
    """query::cast_shapes (shape_cast.rs:274-286) with parry's default options on
    1,048,576 pairs of mixed shapes in approach (scenes.approach_pairs, seed 7: Ball,
    Cuboid, Capsule, Cone, ConvexPolyhedron of 8-64 points and Triangle; the second
    shape 2-6 units away, closing at 0.5-3 units/s).  One step = one
    pq_cast_shapes_pairs call: the GJK ray cast of every pair, then the impact
    geometry (GJK/EPA contact) of the pairs that start within 1e-5 of touching.
    Algorithmic bytes per pair: 8 (pair) + 2 x 28 (poses) + 2 x 12 (velocities) +
    2 x 32 (shape records) + 12 per vertex of a ConvexPolyhedron or Triangle +
    56 (hit record)."""
    import numpy as np
    import torch

    from parry_amd import ShapeSet, cast_shapes_pairs, cast_shapes_workspace_bytes, scenes

    n = 1 << 20
    descs, verts, poses_np, vels_np, pairs_np = scenes.approach_pairs(n, 7)
    shapes = ShapeSet(descs, verts)
    poses = torch.from_numpy(poses_np).cuda()
    vels = torch.from_numpy(vels_np).cuda()
    pairs = torch.from_numpy(pairs_np.view(np.int32)).cuda()
    out = torch.empty((n, 14), dtype=torch.float32, device="cuda")
    ws = torch.empty(cast_shapes_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        cast_shapes_pairs(shapes, poses, vels, pairs, out=out, workspace=ws, stream=stream, check_indices=False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    o = out.cpu().numpy()
    st = o.view(np.uint32)[:, 13]
    nv = np.where(np.isin(descs["type"], (4, 5)), descs["num_vertices"], 0).astype(np.int64)
    alg = int((8 + 56 + 24 + 64 + 56) * n + 12 * (nv[pairs_np[:, 0]] + nv[pairs_np[:, 1]]).sum())
    gbs = alg / (gpu_ms * 1e-3) / 1e9
    rec = {
        "metric": "shape casts/s",
        "value": round(n / wall, 1),
        "unit": "casts/s",
        "ms_per_step": round(wall * 1e3, 4),
        "gpu_ms_per_step": round(gpu_ms, 4),
        "workload": "query::cast_shapes, default ShapeCastOptions: 1048576 mixed pairs in approach "
                    "(scenes.approach_pairs seed 7, Triangles included)",
        "hits": int((st != 0).sum()),
        "penetrating": int((st == 4).sum()),
        "roofline": {"bound": "latency (GJK ray cast, branchy scalar code)", "kernel": "cast_kernel + contact",
                     "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_launch": alg},
    }
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        t0 = time.perf_counter()
        ref = oracle.cast_shapes(descs, verts, poses_np, vels_np, pairs_np, threads=cores)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(n / t_cpu, 1), "unit": "casts/s", "cores": cores, "kind": "port",
                               "sample": f"all {n} pairs, one run on {cores} threads (C++ restatement of parry3d f32)"}
        h = ref["status"] != 0
        rec["parity"] = {"status_mismatches": int((st != ref["status"]).sum()),
                         "bit_identical_hits": int(np.all(o[h, :13] == ref["out"][h], axis=1).sum()),
                         "hits_oracle": int(h.sum())}
        r64 = oracle.cast_shapes(descs, np.asarray(verts, np.float64), np.asarray(poses_np, np.float64),
                                 np.asarray(vels_np, np.float64), pairs_np, f64=True, threads=cores)
        both = (st != 0) & (r64["status"] != 0)
        rec["f64_error"] = {"pairs": n, "status_agreement": round(float((st == r64["status"]).mean()), 7),
                            "toi": f64_error(o[:, 0], r64["out"][:, 0], both),
                            "normal1": f64_error(o[:, 7:10], r64["out"][:, 7:10], both)}
    return rec


def ploc_leg(args, log):
    """Bvh::from_leaves(BvhBuildStrategy::Ploc, aabbs) (bvh_ploc_build.rs:10-93) built
    on the GPU from the 262,144 shape AABBs of the dynamic scene (scenes.mixed_scene
    seed 3) and from 1,048,576 random boxes (centres U[-200, 200]^3, half extents
    U[0.01, 1]^3, seed 12).  One step = one pq_bvh_from_leaves_ploc call (Morton
    keys, radix sort, every merge level with its host readback, the depth-first
    relayout), inputs and outputs in device memory."""
    import numpy as np
    import torch

    from parry_amd import BVH_NODE_WIDE_DTYPE, ShapeSet, ploc_nodes, scenes
    from parry_amd import _native

    sc = scenes.mixed_scene(262144, 3)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    scene_aabbs = shapes.compute_aabbs(torch.from_numpy(sc["poses"]).cuda())
    rng = np.random.default_rng(12)
    c = rng.uniform(-200, 200, (1 << 20, 3)).astype(np.float32)
    h = rng.uniform(0.01, 1.0, (1 << 20, 3)).astype(np.float32)
    rand_aabbs = torch.from_numpy(np.concatenate([c - h, c + h], 1)).cuda()
    out = {}
    for name, a in (("mixed_scene_262144", scene_aabbs), ("random_1048576", rand_aabbs)):
        n = a.shape[0]
        nodes = torch.empty((n - 1, 64), dtype=torch.uint8, device="cuda")
        ws = torch.empty(int(_native.lib().pq_bvh_ploc_workspace_bytes(n)), dtype=torch.uint8, device="cuda")
        for _ in range(args.warmup):
            ploc_nodes(a, out=nodes, workspace=ws)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, depth = ploc_nodes(a, out=nodes, workspace=ws)
        wall = (time.perf_counter() - t0) / args.steps
        rec = {"metric": "Bvh leaves built/s", "value": round(n / wall, 1), "unit": "leaves/s",
               "ms_per_build": round(wall * 1e3, 4), "leaves": n, "depth": depth}
        if not args.no_cpu:
            from oracle import oracle

            host = a.cpu().numpy()
            t0 = time.perf_counter()
            ob = oracle.OracleBvh(host, ploc=True)
            t_cpu = time.perf_counter() - t0
            ref = ob.nodes()
            got = nodes.cpu().numpy().view(BVH_NODE_WIDE_DTYPE).reshape(-1)
            rec["cpu_baseline"] = {"value": round(n / t_cpu, 1), "unit": "leaves/s", "cores": 1, "kind": "port",
                                   "sample": f"the same {n} leaves, one build on one thread (C++ restatement of "
                                             "parry3d's PLOC, which is single-threaded)"}
            rec["parity"] = {"byte_identical_tree": bool(got.tobytes() == ref.tobytes())}
        out[name] = rec
    return out


def manifold_leg(args, log):
    """contact_manifold_pfm_pfm (contact_manifolds_pfm_pfm.rs:44-162) on fresh manifolds
    for 1,048,576 pairs of Cuboid, Capsule, Cone and Triangle (scenes.feature_map_pairs,
    seed 3, offsets N(0, 0.6^2)): one step = one pq_contact_manifolds_pfm_pairs call
    (GJK/EPA on the feature maps, then the features and their clipping per pair).
    Algorithmic bytes per pair: 8 (pair) + 2 x 28 (poses) + 2 x 32 (records) + 36 per
    Triangle + 32 (manifold header) + 36 per manifold point written."""
    import numpy as np
    import torch

    from parry_amd import CONTACT_MANIFOLD_DTYPE, ShapeSet, contact_manifolds_pfm, contact_manifolds_workspace_bytes
    from parry_amd import scenes

    n = 1 << 20
    d, verts, poses_np, pairs_np = scenes.feature_map_pairs(n, 3)
    shapes = ShapeSet(d, verts)
    poses = torch.from_numpy(poses_np).cuda()
    pairs = torch.from_numpy(pairs_np.view(np.int32)).cuda()
    out = torch.empty((n, CONTACT_MANIFOLD_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    ws = torch.empty(contact_manifolds_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        contact_manifolds_pfm(shapes, poses, pairs, 0.02, out=out, workspace=ws, stream=stream, check_indices=False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    got = out.cpu().numpy().view(CONTACT_MANIFOLD_DTYPE).reshape(-1)
    pts = int(got["num_points"].sum())
    ntri = int((d["type"] == 5).sum())
    alg = (8 + 56 + 64 + 32) * n + 36 * ntri + 36 * pts
    gbs = alg / (gpu_ms * 1e-3) / 1e9
    rec = {
        "metric": "contact manifolds/s",
        "value": round(n / wall, 1),
        "unit": "manifolds/s",
        "ms_per_step": round(wall * 1e3, 4),
        "gpu_ms_per_step": round(gpu_ms, 4),
        "workload": "contact_manifold_pfm_pfm on fresh manifolds, prediction 0.02: 1048576 pairs of Cuboid, "
                    "Capsule, Cone and Triangle (scenes.feature_map_pairs seed 3)",
        "manifolds_in_contact": int((got["status"] == 1).sum()),
        "points": pts,
        "roofline": {"bound": "latency (GJK/EPA and per-pair clipping)", "kernel": "contact kernels + manifold_kernel",
                     "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_launch": alg},
    }
    # The update from the previous frame's manifolds (pq_contact_manifolds_pfm_update):
    # the poses alternate between the scene's and a copy moved by N(0, 1e-4^2) per
    # axis, so most manifolds in contact are kept and the rest are recomputed with
    # GJK seeded by their cached normal.
    from parry_amd import MANIFOLD_UPDATED, contact_manifolds_pfm_update

    dp = np.zeros((2 * n, 7), np.float32)
    dp[:, 4:] = np.random.default_rng(7).normal(0, 1e-4, (2 * n, 3))
    poses_b = poses + torch.from_numpy(dp).cuda()
    mf = out.clone()
    for _ in range(args.warmup):
        contact_manifolds_pfm_update(shapes, poses_b, pairs, 0.02, mf, workspace=ws, stream=stream,
                                     check_indices=False)
        contact_manifolds_pfm_update(shapes, poses, pairs, 0.02, mf, workspace=ws, stream=stream,
                                     check_indices=False)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for t in range(args.steps):
        contact_manifolds_pfm_update(shapes, poses_b if t % 2 == 0 else poses, pairs, 0.02, mf, workspace=ws,
                                     stream=stream, check_indices=False)
    ev[1].record(stream)
    torch.cuda.synchronize()
    upd = mf.cpu().numpy().view(CONTACT_MANIFOLD_DTYPE).reshape(-1)
    rec["update"] = {"gpu_ms_per_step": round(ev[0].elapsed_time(ev[1]) / args.steps, 4),
                     "kept": int((upd["status"] == MANIFOLD_UPDATED).sum()),
                     "recomputed_in_contact": int((upd["status"] == 1).sum()),
                     "workload": "the same pairs, poses alternating with a copy moved by N(0, 1e-4^2) per axis"}
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        t0 = time.perf_counter()
        ref = oracle.manifolds_pfm(d, verts, poses_np, pairs_np, 0.02, threads=cores)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(n / t_cpu, 1), "unit": "manifolds/s", "cores": cores, "kind": "port",
                               "sample": f"all {n} pairs, one run on {cores} threads (C++ restatement of parry3d f32)"}
        ok = ref["status"] == 1
        same = (got["status"] == np.where(ref["status"] < 0, 0xFFFFFFFF, ref["status"]).astype(np.uint32)) & \
               (got["num_points"] == ref["count"])
        rec["parity"] = {"status_or_count_mismatches": int((~same).sum()), "manifolds_oracle": int(ok.sum())}
        r64 = oracle.manifolds_pfm(d, np.asarray(verts, np.float64), np.asarray(poses_np, np.float64), pairs_np, 0.02,
                                   f64=True, threads=cores)
        st64 = np.where(r64["status"] < 0, 0xFFFFFFFF, r64["status"]).astype(np.uint32)
        agree = (got["status"] == st64) & (got["num_points"] == r64["count"])
        both = agree & (got["status"] == 1)
        # GJK's own point (the last of a manifold) and the normal, where both precisions agree on the point count
        last = np.maximum(got["num_points"].astype(np.int64) - 1, 0)
        rows = np.arange(n)
        rec["f64_error"] = {"pairs": n, "status_and_count_agreement": round(float(agree.mean()), 7),
                            "local_n1": f64_error(got["local_n1"], r64["n1"], both),
                            "gjk_point_dist": f64_error(got["points"]["dist"][rows, last], r64["dist"][rows, last],
                                                        both)}
    return rec


def learned_leg(args, log):
    """Learned GJK warm start vs the plain GJK on 1,048,576 mixed convex pairs
    (ConvexPolyhedron 8-64 points, Cuboid, Ball, Capsule, Cone): the first 2^20
    overlapping-AABB pairs of a 320,000-shape scene (seed 3), prediction 0.01.
    One step = query::contact on every pair; the learned step adds the fused
    seed kernel.  A "perfect seed" run (the plain GJK's own final direction as
    the seed) bounds what any seed can gain, and the contact path's phase times
    bound what any GJK speed-up can gain (Amdahl: the call with GJK taking no
    time at all)."""
    import numpy as np
    import torch

    from parry_amd import Bvh, ShapeSet, _native, contact_pairs, contact_workspace_bytes, scenes
    from parry_amd.learned import WEIGHTS_FILE, WarmStartModel, gjk_directions, support_map_pairs

    if not os.path.exists(WEIGHTS_FILE):
        return {"skipped": f"no trained parameters at {os.path.relpath(WEIGHTS_FILE, ROOT)}"}
    sc = scenes.mixed_scene(320000, seed=3)
    shapes = ShapeSet(sc["descs"], sc["vertices"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    pairs = Bvh(shapes.compute_aabbs(poses).cpu().numpy()).self_pairs()
    total_pairs = int(pairs.shape[0])
    if total_pairs < 1 << 20:
        raise RuntimeError(f"the scene yields {total_pairs} pairs, fewer than 2^20")
    pairs = pairs[: 1 << 20].contiguous()
    m = pairs.shape[0]
    net = WarmStartModel.pretrained()
    desc = shapes.descriptors()
    stream = torch.cuda.current_stream()
    out = torch.empty((m, 14), dtype=torch.float32, device="cuda")
    ws = torch.empty(contact_workspace_bytes(m), dtype=torch.uint8, device="cuda")
    seeds = torch.empty((m, 3), dtype=torch.float32, device="cuda")

    def run(seed_fn, steps):
        for _ in range(args.warmup):
            seed_fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        t_seed = t_contact = 0.0
        for _ in range(steps):
            ev[0].record(stream)
            s = seed_fn()
            ev[1].record(stream)
            contact_pairs(shapes, poses, pairs, 0.01, init_dirs=s, out=out, workspace=ws, check_indices=False)
            ev[2].record(stream)
            torch.cuda.synchronize()
            t_seed += ev[0].elapsed_time(ev[1])
            t_contact += ev[1].elapsed_time(ev[2])
        return t_seed / steps, t_contact / steps

    def phases(s):
        """The contact call's phase times (one untimed call with events between phases)."""
        _native.set_contact_phase_timing(True)
        try:
            contact_pairs(shapes, poses, pairs, 0.01, init_dirs=s, out=out, workspace=ws, check_indices=False)
            return {k: round(v, 4) for k, v in _native.contact_phase_ms().items()}
        finally:
            _native.set_contact_phase_timing(False)

    steps = max(3, args.steps // 2)
    plain_ms = run(lambda: None, steps)
    learned_ms = run(lambda: net.seeds(poses, pairs, desc, out=seeds), steps)
    # perfect seeds: the plain GJK's own directions (zero for pairs with a Ball)
    perfect, valid = gjk_directions(shapes, poses, pairs)
    perfect[~valid] = 0.0
    keep = torch.from_numpy(support_map_pairs(sc["descs"], pairs.cpu().numpy().view(np.uint32))).cuda()
    perfect[~keep] = 0.0
    perfect_ms = run(lambda: perfect, steps)
    ph_plain = phases(None)
    ph_learned = phases(net.seeds(poses, pairs, desc, out=seeds))

    def stats(s):
        c, it = contact_pairs(shapes, poses, pairs, 0.01, init_dirs=s, with_iters=True, workspace=ws)
        return c.cpu().numpy(), it.cpu().numpy()

    cp, ip = stats(None)
    learned_seeds = net.seeds(poses, pairs, desc)
    cl, il = stats(learned_seeds)
    _, iq = stats(perfect)
    hp, hl = cp.view(np.uint32)[:, 13], cl.view(np.uint32)[:, 13]
    both = (hp == 1) & (hl == 1)
    derr = np.abs(cp[both, 12] - cl[both, 12]) / np.maximum(1.0, np.abs(cp[both, 12]))
    km = keep.cpu().numpy()
    pr = pairs.cpu().numpy().view(np.uint32)

    def rate(ms):
        return round(m / ((ms[0] + ms[1]) * 1e-3), 1)

    # GJK iterations per shape-type pair (unordered), seeded pairs only
    names = {0: "Ball", 1: "Cuboid", 2: "Capsule", 3: "Cone", 4: "ConvexPolyhedron"}
    t = sc["descs"]["type"]
    ta, tb = np.minimum(t[pr[:, 0]], t[pr[:, 1]]), np.maximum(t[pr[:, 0]], t[pr[:, 1]])
    by_type = {}
    for a in range(5):
        for b in range(a, 5):
            sel = km & (ta == a) & (tb == b)
            if sel.any():
                by_type[f"{names[a]}-{names[b]}"] = {
                    "pairs": int(sel.sum()), "plain": round(float(ip[sel, 0].mean()), 4),
                    "learned": round(float(il[sel, 0].mean()), 4),
                    "perfect_seed": round(float(iq[sel, 0].mean()), 4)}
    contact_ms = plain_ms[1]
    gjk_ms = ph_plain["gjk"]
    rec = {
        "metric": "contact pairs/s, learned warm start vs plain GJK",
        "workload": f"the first {m} of {total_pairs} overlapping-AABB pairs of 320000 mixed convex shapes "
                    "(5 types uniform), seed 3, prediction 0.01",
        "pairs": m,
        "seeded_pairs": int(km.sum()),
        "plain": {"pairs_per_s": rate(plain_ms), "contact_ms": round(plain_ms[1], 4), "phases_ms": ph_plain},
        "learned": {"pairs_per_s": rate(learned_ms), "seed_kernel_ms": round(learned_ms[0], 4),
                    "contact_ms": round(learned_ms[1], 4), "phases_ms": ph_learned},
        "perfect_seed_bound": {"pairs_per_s": round(m / (perfect_ms[1] * 1e-3), 1),
                               "contact_ms": round(perfect_ms[1], 4)},
        "speedup_learned_vs_plain": round(rate(learned_ms) / rate(plain_ms), 4),
        "amdahl_bound_any_gjk_speedup": {
            "gjk_share_of_contact": round(gjk_ms / sum(ph_plain.values()), 4),
            "max_speedup": round(sum(ph_plain.values()) / (sum(ph_plain.values()) - gjk_ms), 4),
            "note": "contact time with the GJK/dispatch kernel taking no time at all; a seed only shortens GJK"},
        "mean_gjk_iterations_seeded_pairs": {"plain": round(float(ip[km, 0].mean()), 4),
                                             "learned": round(float(il[km, 0].mean()), 4),
                                             "perfect_seed": round(float(iq[km, 0].mean()), 4)},
        "gjk_iterations_saved_by_type_pair": by_type,
        "mean_epa_iterations_seeded_pairs": {"plain": round(float(ip[km, 1].mean()), 4),
                                             "learned": round(float(il[km, 1].mean()), 4)},
        "agreement_with_plain": {"has_contact_mismatches": int((hp != hl).sum()),
                                 "dist_rel_diff_p99": float(np.quantile(derr, 0.99)) if derr.size else 0.0,
                                 "dist_rel_diff_max": float(derr.max()) if derr.size else 0.0,
                                 "within_1e-5": float((derr <= 1e-5).mean()) if derr.size else 1.0},
        "mlp_roofline": {"bound": "mfma", "flops_per_pair": 2 * (128 * 64 + 2 * 128 * 128 + 16 * 128),
                         "achieved_tflops": round(m * 2 * (128 * 64 + 2 * 128 * 128 + 16 * 128)
                                                  / (learned_ms[0] * 1e-3) / 1e12, 2),
                         "peak_tflops": 2500.0},
    }
    if not args.no_cpu:
        from oracle import oracle

        nbytes = contact_bytes(sc["descs"], pr)
        log("learned leg: oracle flop count, CPU baseline and f64 error ...")
        fl = oracle.contact_flops(sc["descs"], sc["vertices"], sc["poses"], pr, 0.01)
        rec["roofline"] = contact_roofline(nbytes, fl["flops"], contact_ms, "contact (plain GJK + EPA kernels)")
        rec["roofline"]["fp32"]["flops_per_pair"] = round(fl["flops"] / m, 1)
        cores = oracle.default_threads()
        oracle.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pr[: 1 << 14], 0.01, threads=cores)
        t0 = time.perf_counter()
        ref = oracle.contact_pairs(sc["descs"], sc["vertices"], sc["poses"], pr, 0.01, threads=cores)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(m / t_cpu, 1), "unit": "pairs/s", "cores": cores, "kind": "port",
                               "sample": f"query::contact (parry's default seed) on all {m} pairs, {cores} threads "
                                         "(C++ restatement of parry3d f32)"}
        rec["parity"] = {"plain_has_contact_mismatches":
                         int((ref["status"].astype(np.int64) != hp.astype(np.int64)).sum())}
        rec["f64_error"] = {"plain": contact_f64_error(cp, sc["descs"], sc["vertices"], sc["poses"], pr, 0.01,
                                                       threads=cores),
                            "learned": contact_f64_error(cl, sc["descs"], sc["vertices"], sc["poses"], pr, 0.01,
                                                         init_dirs=learned_seeds.cpu().numpy(), threads=cores)}
    return rec


def dynamic_leg(args, log):
    """262,144 mixed convex shapes in rigid motion (v ~ U[-1,1]^3, w ~ U[-pi,pi]^3 rad/s,
    dt = 1/60), 100 frames, each one hipGraph replay: AABBs, Bvh refit, pairs,
    seeds, query::contact (prediction 0.01), warm-start cache, pose update.
    Run with parry's default seeds, with the cached previous normals, and with
    the learned seeds plus the cache.  The f64 error is given on frames 0, 49
    and 99, the roofline on one frame's algorithmic bytes and oracle-counted
    flops, and the CPU baseline is the oracle's frame 0 on the host cores."""
    import numpy as np
    import torch

    from parry_amd import scenes
    from parry_amd.dynamics import DynamicScene
    from parry_amd.learned import WEIGHTS_FILE

    sc = scenes.mixed_scene(262144, seed=4)
    frames = 100
    rec = {"metric": "contact pairs/s over 100 dynamic frames",
           "workload": "262144 mixed convex shapes (5 types uniform), v ~ U[-1,1]^3, w ~ U[-pi,pi]^3 rad/s, "
                       "dt 1/60, 100 frames, one hipGraph per frame, seed 4", "modes": {}}
    modes = ["plain", "cache"] + (["learned+cache"] if os.path.exists(WEIGHTS_FILE) else [])
    for mode in modes:
        ds = DynamicScene(sc["descs"], sc["vertices"], sc["poses"], seed=104, seeding=mode,
                          pair_capacity=4 * 262144 + (1 << 20))
        ds.capture()
        pair_counts = []
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(ds.stream)
        for f in range(frames):
            ds.step()
        ev1.record(ds.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        m_last = ds.num_pairs()
        # pair totals: replay again from the start is not possible (poses advanced), so
        # count on a second pass with a sync per frame (untimed)
        ds2 = DynamicScene(sc["descs"], sc["vertices"], sc["poses"], seed=104, seeding=mode,
                           pair_capacity=4 * 262144 + (1 << 20))
        ds2.capture()
        for f in range(frames):
            ds2.step()
            pair_counts.append(ds2.num_pairs())
        gpu_ms = ev0.elapsed_time(ev1) / frames
        total = int(sum(pair_counts))
        rec["modes"][mode] = {"ms_per_frame": round(gpu_ms, 4), "wall_ms_per_frame": round(wall * 1e3 / frames, 4),
                              "pairs_per_s": round(total / (gpu_ms * frames * 1e-3), 1),
                              "mean_pairs_per_frame": round(total / frames, 1), "last_frame_pairs": m_last}
        log(f"dynamic {mode}: {gpu_ms:.3f} ms/frame")
        del ds, ds2
    base = rec["modes"]["plain"]["pairs_per_s"]
    for mode in modes:
        rec["modes"][mode]["vs_plain"] = round(rec["modes"][mode]["pairs_per_s"] / base, 4)
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        descs, verts = sc["descs"], sc["vertices"]
        # frames 0, 49 and 99 of a plain run, issued eagerly: the poses each frame
        # queries, its pairs and its contacts
        ds = DynamicScene(descs, verts, sc["poses"], seed=104, seeding="plain", pair_capacity=4 * 262144 + (1 << 20))
        snaps = {}
        for f in range(frames):
            if f in (0, 49, 99):
                torch.cuda.synchronize()
                p_f = ds.poses.cpu().numpy().copy()
            ds.step(use_graph=False)
            if f in (0, 49, 99):
                m_f = ds.num_pairs()
                snaps[f] = (p_f, ds.pairs[:m_f].cpu().numpy().view(np.uint32).copy(), ds.contacts[:m_f].cpu().numpy())
        del ds
        log("dynamic leg: f64 error on frames 0, 49, 99 ...")
        rec["f64_error"] = {f"frame_{f}": contact_f64_error(c, descs, verts, p_f, pr, 0.01, threads=cores)
                            for f, (p_f, pr, c) in snaps.items()}
        # one frame's algorithmic traffic and flops, and the oracle's frame on the host cores
        p0, pr0, c0 = snaps[0]
        t0 = time.perf_counter()
        aabbs = oracle.shape_aabbs(descs, verts, p0)
        obvh = oracle.OracleBvh(aabbs)
        ref_pairs, touched = obvh.self_pairs()
        ref = oracle.contact_pairs(descs, verts, p0, ref_pairs, 0.01, threads=cores)
        t_cpu = time.perf_counter() - t0
        n = len(descs)
        t = descs["type"]
        nv = descs["num_vertices"].astype(np.int64) * (t == 4)
        frame_bytes = {"aabbs": int((28 + 32 + 12 * nv + 24).sum()), "refit": 24 * n + 128 * (n - 1),
                       "pairs": 64 * touched + 8 * int(pr0.shape[0]), "contact": contact_bytes(descs, pr0)}
        fl = oracle.contact_flops(descs, verts, p0, pr0, 0.01)
        plain_ms = rec["modes"]["plain"]["ms_per_frame"]
        rec["roofline"] = contact_roofline(sum(frame_bytes.values()), fl["flops"], plain_ms,
                                           "whole frame (AABBs, refit, pairs, contact), plain seeds")
        rec["roofline"]["bytes_per_frame"] = frame_bytes
        rec["roofline"]["fp32"]["flops_per_pair"] = round(fl["flops"] / max(1, pr0.shape[0]), 1)
        rec["cpu_baseline"] = {"value": round(ref_pairs.shape[0] / t_cpu, 1), "unit": "pairs/s", "cores": cores,
                               "kind": "port",
                               "sample": f"frame 0 on the host: shape AABBs and a new Bvh (1 thread, as parry), "
                                         f"its {ref_pairs.shape[0]} pairs and query::contact on {cores} threads "
                                         "(C++ restatement of parry3d f32)"}
        rec["parity_frame0"] = {
            "pair_sets_equal": bool(oracle.pair_set(ref_pairs) == oracle.pair_set(pr0)),
            "has_contact_mismatches": int((oracle.contact_pairs(descs, verts, p0, pr0, 0.01, threads=cores)["status"]
                                           .astype(np.int64) != c0.view(np.uint32)[:, 13].astype(np.int64)).sum())}
    return rec


def contact_leg(args, log):
    """65,536 ConvexPolyhedra (8-64 hull points), mean ~8 AABB overlaps per shape:
    one step = leaf AABBs + Bvh refit + overlapping pairs + query::contact
    (prediction 0.01) on every pair, all on the device."""
    import numpy as np
    import torch

    from parry_amd import ShapeSet, scenes
    from parry_amd.pipeline import CollisionPipeline

    n = 65536
    sc = scenes.polyhedra_scene(n, seed=3)
    shapes = ShapeSet.convex_polyhedra(sc["vertices"], sc["offsets"])
    poses = torch.from_numpy(sc["poses"]).cuda()
    pipe = CollisionPipeline(shapes, poses, prediction=0.01, pair_capacity=8 * n)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        pipe.step(poses)
    torch.cuda.synchronize()
    m = pipe.num_pairs()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        pipe.step(poses)
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = ev[0].elapsed_time(ev[1]) / args.steps
    # Stage split (one extra untimed frame with events between the stages).
    lib_events = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    from parry_amd import _native
    import ctypes

    lib = _native.lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    lib_events[0].record(stream)
    lib.pq_shapes_compute_aabbs(shapes.handle, poses.data_ptr(), pipe.aabbs.data_ptr(), sp)
    lib_events[1].record(stream)
    lib.pq_bvh_update_and_refit(pipe.bvh.handle, pipe.aabbs.data_ptr(), sp)
    lib_events[2].record(stream)
    lib.pq_bvh_self_pairs(pipe.bvh.handle, pipe.pairs.data_ptr(), pipe.capacity, pipe.count.data_ptr(), sp)
    lib_events[3].record(stream)
    lib.pq_contact_pairs_device_count(shapes.handle, poses.data_ptr(), pipe.pairs.data_ptr(), pipe.capacity,
                                      pipe.count.data_ptr(), pipe.prediction, None, pipe.contacts.data_ptr(), None,
                                      pipe.workspace.data_ptr(), sp)
    lib_events[4].record(stream)
    torch.cuda.synchronize()
    stages = {name: round(lib_events[i].elapsed_time(lib_events[i + 1]), 4)
              for i, name in enumerate(["aabbs", "refit", "pairs", "contact"])}
    pairs = pipe.pairs[:m].cpu().numpy().view(np.uint32)
    has = pipe.contacts[:m].cpu().numpy().view(np.uint32)[:, 13]
    offs = sc["offsets"].astype(np.int64)
    nverts = offs[1:] - offs[:-1]
    cbytes = int((12 * (nverts[pairs[:, 0]] + nverts[pairs[:, 1]]) + 2 * 28 + 56).sum())
    contact_ms = stages["contact"]
    rec = {
        "metric": "contact pairs/s",
        "value": round(m * args.steps / wall, 1),
        "unit": "pairs/s",
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "gpu_ms_per_step": round(step_ms, 4),
        "stages_ms": stages,
        "workload": "65536 ConvexPolyhedra (8-64 hull points, ellipsoid semi-axes U[0.5,1.5], 5% jitter), "
                    f"cube side {sc['side']:.2f} for ~8 AABB overlaps/shape, seed 3; one step = AABBs + refit + "
                    "pairs + query::contact(prediction 0.01) on every pair",
        "pairs": m,
        "contacts": int((has == 1).sum()),
        "roofline": {
            "bound": "latency (GJK/EPA branchy scalar code)",
            "kernel": "contact (GJK + EPA kernels)",
            "achieved": round(cbytes / (contact_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(cbytes / (contact_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "algorithmic_bytes_per_launch": cbytes,
        },
    }
    if not args.no_cpu:
        from oracle import oracle

        cores = oracle.default_threads()
        t0 = time.perf_counter()
        ref = oracle.contact_pairs(shapes.descs, sc["vertices"], sc["poses"], pairs, 0.01, threads=cores)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {
            "value": round(m / t_cpu, 1),
            "unit": "pairs/s",
            "cores": cores,
            "kind": "port",
            "sample": f"query::contact on all {m} pairs of one frame, {cores} threads (C++ restatement of parry3d f32)",
        }
        rec["parity"] = {"has_contact_mismatches": int((ref["status"].astype(np.int64) != has.astype(np.int64)).sum())}
        fl = oracle.contact_flops(shapes.descs, sc["vertices"], sc["poses"], pairs, 0.01)
        rec["roofline"] = contact_roofline(cbytes, fl["flops"], contact_ms, "contact (GJK + EPA kernels)")
        rec["roofline"]["fp32"]["flops_per_pair"] = round(fl["flops"] / m, 1)
        rec["f64_error"] = contact_f64_error(pipe.contacts[:m].cpu().numpy(), shapes.descs, sc["vertices"],
                                             sc["poses"], pairs, 0.01, threads=cores)
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--skip-contact", action="store_true", help="measure only the ray-cast line")
    ap.add_argument("--contact-only", action="store_true", help="profiling aid: run only the contact leg")
    ap.add_argument("--skip-batch", action="store_true", help="skip the batched multi-mesh leg")
    ap.add_argument("--skip-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--skip-dynamic", action="store_true", help="skip the dynamic-scene leg")
    ap.add_argument("--skip-queries", action="store_true", help="skip the point-projection and culling leg")
    ap.add_argument("--queries-only", action="store_true", help="profiling aid: run only the point/culling leg")
    ap.add_argument("--skip-mesh-contact", action="store_true", help="skip the TriMesh-shape contact leg")
    ap.add_argument("--skip-cast", action="store_true", help="skip the shape-cast leg")
    ap.add_argument("--skip-ploc", action="store_true", help="skip the PLOC build leg")
    ap.add_argument("--skip-manifold", action="store_true", help="skip the contact-manifold leg")
    ap.add_argument("--manifold-only", action="store_true", help="profiling aid: run only the contact-manifold leg")
    ap.add_argument("--ploc-only", action="store_true", help="profiling aid: run only the PLOC build leg")
    ap.add_argument("--cast-only", action="store_true", help="profiling aid: run only the shape-cast leg")
    ap.add_argument("--mesh-contact-only", action="store_true", help="profiling aid: run only the mesh contact leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="profiling aid: only copy BYTES of device memory 4 times (counter calibration)")
    ap.add_argument("--dynamic-only", action="store_true", help="profiling aid: run only the dynamic leg")
    ap.add_argument("--learned-only", action="store_true", help="profiling aid: run only the learned leg")
    ap.add_argument("--batch-only", action="store_true", help="profiling aid: run only the batched multi-mesh leg")
    ap.add_argument("--batch-tuning", default="", metavar="MAPPING:CHUNK:BLOCKS_PER_CU:BLOCK:REGS",
                    help="profiling aid: use this ray tuning for the batched leg instead of the autotuner's")
    ap.add_argument("--ray-tuning", default="", metavar="MAPPING:CHUNK:BLOCKS_PER_CU:BLOCK:REGS",
                    help="profiling aid: use this ray tuning for the headline leg instead of the autotuner's")
    ap.add_argument("--mapping-sweep", action="store_true",
                    help="profiling aid: time every ray-kernel mapping on four workloads and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU; --gpus {args.gpus} is not supported here (the path shards "
            "rays/pairs with no exchange; multi-GPU runs are out of scope for this build)",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    from parry_amd import TriMesh, scenes, split_hits

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_probe:
        src = torch.ones(args.copy_probe, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "copies": 4}), flush=True)
        return 0
    if args.contact_only:
        print(json.dumps({"contact": contact_leg(args, log)}), flush=True)
        return 0
    if args.dynamic_only:
        print(json.dumps({"dynamic": dynamic_leg(args, log)}), flush=True)
        return 0
    if args.learned_only:
        print(json.dumps({"learned": learned_leg(args, log)}), flush=True)
        return 0
    if args.batch_only:
        print(json.dumps({"batch": batch_leg(args, log)}), flush=True)
        return 0
    if args.mapping_sweep:
        print(json.dumps({"mapping_sweep": mapping_sweep(args, log)}), flush=True)
        return 0
    if args.queries_only:
        print(json.dumps({"queries": queries_leg(args, log)}), flush=True)
        return 0
    if args.manifold_only:
        print(json.dumps({"manifold": manifold_leg(args, log)}), flush=True)
        return 0
    if args.ploc_only:
        print(json.dumps({"ploc": ploc_leg(args, log)}), flush=True)
        return 0
    if args.cast_only:
        print(json.dumps({"cast": cast_leg(args, log)}), flush=True)
        return 0
    if args.mesh_contact_only:
        print(json.dumps({"mesh_contact": mesh_contact_leg(args, log)}), flush=True)
        return 0
    v, f = scenes.noisy_icosphere(6, seed=1)
    rays = scenes.rays_into_unit_ball(1 << 20, seed=2)
    n = rays.shape[0]
    log(f"workload: icosphere s=6 ({f.shape[0]} tris, {v.shape[0]} verts), {n} rays")
    mesh = TriMesh(v, f)
    rays_d = torch.from_numpy(rays).to(dev)
    out = torch.empty((n, 5), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    # Traffic model from the kernel's instrumented launch (untimed).
    _, counts = mesh.cast_local_rays_counted(rays_d, MAX_TOI, True, stream=stream)
    sum_w = int(counts[:, 0].sum(dtype=torch.int64).item())
    sum_k = int(counts[:, 1].sum(dtype=torch.int64).item())
    alg_bytes = n * (24 + 20) + 64 * sum_w + 48 * sum_k

    # The autotuner picks the mapping (it keeps the fastest in the mesh), untimed.
    if args.ray_tuning:
        m_, c_, bpc_, blk_, regs_ = args.ray_tuning.split(":")
        mesh.set_tuning(m_, int(c_), int(bpc_), int(blk_), int(regs_))
        tuned, tuned_ms = mesh.tuning(), float("nan")
    else:
        tuned, tuned_ms = mesh.autotune(rays_d, MAX_TOI, stream=stream)
    log(f"ray tuning: {tuned} (autotuner's time {tuned_ms:.4f} ms)")

    for _ in range(args.warmup):
        mesh.cast_local_rays_and_get_normal(rays_d, MAX_TOI, True, out=out, stream=stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        mesh.cast_local_rays_and_get_normal(rays_d, MAX_TOI, True, out=out, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    log(f"timed {args.steps} steps: wall {wall * 1e3:.3f} ms, kernel avg {kern_ms:.4f} ms")

    _, _, feat = split_hits(out)
    gpu_feat = feat.cpu().numpy().view(np.uint32)
    host_hits = out.cpu().numpy()
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    binned = tuned["mapping"] == "binned"
    kernels = (["bin_count_kernel", "bin_scan_kernel", "bin_cells_kernel", "bin_scatter_kernel"] if binned else []) + [
        f"cast_rays_kernel<false, 0, {tuned['block_size'] or 256}, {tuned['stack_regs']}>"]
    traffic_entry = "cast_rays_step_binned" if binned else "cast_rays_kernel"
    # the committed counters were taken at one tuning: used for the same mapping
    traffic_tuning = measured_traffic(traffic_entry, "tuning")
    traffic = (measured_traffic(traffic_entry)
               if traffic_tuning is None or traffic_tuning.get("mapping") == tuned["mapping"] else None)
    rec = {
        "metric": METRIC,
        "value": round(n * args.steps / wall / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "trimesh ray cast: noisy icosphere s=6 (81920 tris) x 1048576 rays, max_toi=10, solid",
            "triangles": int(f.shape[0]),
            "rays": n,
            "mesh_seed": 1,
            "ray_seed": 2,
            "parallelism": "single GPU",
            "mapping": tuned,
            "autotuner": {"chosen": tuned, "ms_per_launch": None if tuned_ms != tuned_ms else round(tuned_ms, 5),
                          "forced": bool(args.ray_tuning)},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_entry": traffic_entry,
            "traffic_tuning": traffic_tuning,
            "kernel": " + ".join(kernels),
            "kernel_avg_ms": round(kern_ms, 5),
            "kernel_avg_covers": "every kernel of one cast call (HIP events around the timed calls)",
            "algorithmic_bytes_per_launch": alg_bytes,
            "mean_wide_nodes_per_ray": round(sum_w / n, 4),
            "mean_triangles_per_ray": round(sum_k / n, 4),
            "stream_only_frac": round(n * 44 / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        },
    }
    if not args.no_cpu:
        log("cpu baseline (oracle on host cores) ...")
        cb, parity = cpu_baseline(v, f, rays, gpu_feat)
        rec["cpu_baseline"] = cb
        rec["parity"] = parity
        from oracle import oracle

        rec["f64_error"] = ray_f64_error(v, f, rays, host_hits, threads=oracle.default_threads())
    else:
        rec["cpu_baseline"] = None
    if not args.skip_batch:
        log("batched multi-mesh leg ...")
        rec["batch"] = batch_leg(args, log)
    if not args.skip_queries:
        log("point projection and culling leg ...")
        rec["queries"] = queries_leg(args, log)
    if not args.skip_manifold:
        log("contact-manifold leg ...")
        rec["manifold"] = manifold_leg(args, log)
    if not args.skip_ploc:
        log("PLOC build leg ...")
        rec["ploc"] = ploc_leg(args, log)
    if not args.skip_cast:
        log("shape-cast leg ...")
        rec["cast"] = cast_leg(args, log)
    if not args.skip_mesh_contact:
        log("TriMesh-shape contact leg ...")
        rec["mesh_contact"] = mesh_contact_leg(args, log)
    if not args.skip_learned:
        log("learned warm-start leg ...")
        rec["learned"] = learned_leg(args, log)
    if not args.skip_dynamic:
        log("dynamic-scene leg ...")
        rec["dynamic"] = dynamic_leg(args, log)
    if not args.skip_contact:
        log("contact leg (broadphase + narrowphase frame) ...")
        rec["contact"] = contact_leg(args, log)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
