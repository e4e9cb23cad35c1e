"""Benchmark of the SdfOracle hot path on one MI355X (ring-on-cube at full size).

One step = (1) dense FP64 build of the torus ring mesh (R 50 mm, r 10 mm,
256 x 64 segments = 32 768 triangles) at target resolution 128;
(2) Tricubic value + raw gradient of 10^6 FP32 points uniform in the valid
region of the 128^3 cube SDF (side 0.1 m, top face at z = 55 mm);
(3) contact forward (penalty kn = 1e11 order 2, damping 1e4) of the 7 680
ring-centroid quadrature points (sum w = 1) under 1 000 seeded poses;
(4) batched contact forward + backward over 256 poses with friction (mu 0.3)
against the same work done pose by pose (the plain per-pose GPU path).

``value`` is (2) in Mpts/s; the other rates are reported beside it.  Inputs
are resident in HBM before the timed region.  The cpu_baseline leg times the
FP64 CPU oracle (oracle/) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sdf-contact-module_amd"))

METRIC = "build Mvoxel/s; tricubic value+grad Mpts/s; fwd+bwd contact evals/s vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


TRAFFIC_FILE = "profiles/query_tricubic_cube128_1M_traffic.json"


def _traffic(nq, shape):
    """HBM bytes per query launch from the committed FETCH_SIZE / WRITE_SIZE
    summary (counters scaled by the copy-kernel calibration), when the
    workload is the one it was measured on."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if w.get("query_points") != nq or list(w.get("grid", [])) != list(shape):
        return None
    return float(t["fetch_kib_reported"] * 1024 * t["fetch_factor"] +
                 t["write_kib_reported"] * 1024 * t["write_factor"])


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--query-points", type=int, default=1_000_000)
    ap.add_argument("--query-reps", type=int, default=5,
                    help="query passes per step, timed together (per-pass time reported)")
    ap.add_argument("--poses", type=int, default=1000)
    ap.add_argument("--bwd-poses", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parts", default="build,query,contact,bwd",
                    help="subset of the step to run (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (scales HBM counters)")
    ap.add_argument("--workload", default="ring_on_cube",
                    choices=["ring_on_cube", "sparse_cylinder", "batched_gears"],
                    help="ring_on_cube (default line); sparse_cylinder: 512^3 narrow band of a "
                         "266k-triangle tube + 1e6 ContactAware points; batched_gears: 64 gear "
                         "meshes built in one launch + 64 query streams in one captured graph")
    return ap.parse_args()


def cpu_baseline(args, cube_spec, torus_mesh, cube_mesh, pts_np, X, w, poses):
    """FP64 oracle on bounded samples of the same workload (rank 0, N=1)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if args.cpu_threads > 0:
        O.set_threads(args.cpu_threads)
    threads = O.num_threads()
    res = {}
    # build: a 16-slice z slab of the torus grid (same per-voxel work as the full build)
    tv, tt = torus_mesh
    tg = O.setup_grid(tv, 0.1, 0.0, 128, 128)
    slab = tg.centers()[: tg.shape[0] * tg.shape[1] * 16]
    t0 = time.perf_counter()
    O.signed_nearest(tv, tt, slab)
    res["build_mvoxel_per_s"] = len(slab) / (time.perf_counter() - t0) / 1e6
    # query: the same 1e6 points, FP64 tricubic, on the oracle's cube field
    cv, ct = cube_mesh
    cg = O.Grid(cube_spec.bmin, cube_spec.bmax, cube_spec.voxel, cube_spec.shape)
    cd = O.build_dense(cv, ct, cg)
    p64 = pts_np.astype(np.float64)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        O.query_dense(cg, cd, "Tricubic", p64)
    res["tricubic_mpts_per_s"] = reps * len(p64) / (time.perf_counter() - t0) / 1e6
    # contact: 200 of the poses
    sub = poses[:200]
    t0 = time.perf_counter()
    O.contact_forward(cg, cd["phi"], cd["n"], X, w, sub, kn=1e11, c=1e4, mu=0.0, reg=1e-6, order=2)
    res["contact_fwd_evals_per_s"] = len(sub) * len(X) / (time.perf_counter() - t0)
    return res, threads


def _events(torch, stream):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _busy(torch):
    try:
        torch.cuda._sleep(1_000_000)
    except (AttributeError, RuntimeError):
        pass


def bench_sparse_cylinder(args):
    """Tube R_o 40 mm, R_i 30 mm, H 50 mm, 1024 x 64 segments (266 240 triangles),
    sparse 512^3, 8^3 bricks, band 4 voxels, halo 2; 1e6 FP32 points on the
    surface offset by N(0, (2h)^2) along the normal, ContactAware value + gradient
    (the plain sparse grid has no candidates, so ContactAware runs the sparse
    Tricubic path, SdfOracle.cpp:191-193); then the same with a sparse
    contact-aware build (face normals as corner normals)."""
    import torch

    import sdfo
    from sdfo import meshes, scenes

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 1024, 64)
    prm = dict(target_resolution=512, max_resolution=512, block_dim=8, band_voxels=4,
               halo_voxels=2)
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    cn = np.repeat(fn[:, None, :], 3, axis=1)
    g = sdfo.SdfGrid.build_sparse(v, t, **prm)  # warm-up
    g.close()
    K = max(1, args.steps)
    build_s = 0.0
    for k in range(K):
        t0 = time.perf_counter()
        g = sdfo.SdfGrid.build_sparse(v, t, **prm)
        build_s += (time.perf_counter() - t0) / K
        if k < K - 1:
            g.close()  # outside the timed build
    info = g.info
    h = g.spec.voxel.max()
    pts_np = scenes.near_surface_points(v, t, args.query_points, 2.0 * h, seed=3)
    pts = torch.from_numpy(pts_np).to(dev)
    n = len(pts_np)
    gap = torch.empty(n, device=dev)
    grad = torch.empty(n, 3, device=dev)
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    for _ in range(args.warmup):
        g.query_into(pts, gap, grad, valid, "ContactAware")
    qs = []
    for _ in range(K):
        a, b = _events(torch, stream)
        _busy(torch)
        a.record(stream)
        g.query_into(pts, gap, grad, valid, "ContactAware")
        b.record(stream)
        torch.cuda.synchronize()
        qs.append(a.elapsed_time(b))
    q_ms = float(np.mean(qs))
    frac_valid = float(valid.float().mean().item())
    sdfo.SdfGrid.build_contact_aware(v, t, cn, sparse=True, **prm).close()  # warm-up
    t0 = time.perf_counter()
    gca = sdfo.SdfGrid.build_contact_aware(v, t, cn, sparse=True, **prm)
    ca_build_s = time.perf_counter() - t0
    a, b = _events(torch, stream)
    gca.query_into(pts, gap, grad, valid, "ContactAware")
    _busy(torch)
    a.record(stream)
    gca.query_into(pts, gap, grad, valid, "ContactAware")
    b.record(stream)
    torch.cuda.synchronize()
    ca_q_ms = a.elapsed_time(b)
    vox = info.block_count * 512
    # Algorithmic bytes: 29 B per point plus every distinct voxel the points'
    # cells touch, read once (16 B each): the corner voxels of the valid
    # points' cells, counted on the host with the query's own cell rule.
    spec = g.spec
    f = (pts_np.astype(np.float64) - spec.bmin) / spec.voxel - 0.5
    c0 = np.floor(f).astype(np.int64)
    nn = np.array(spec.shape, np.int64)
    ok = np.all((c0 >= 0) & (c0 + 1 < nn), axis=1)
    c0 = c0[ok]
    corners = [(c0 + np.array([dx, dy, dz])) for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
    ids = np.concatenate([c[:, 0] + nn[0] * (c[:, 1] + nn[1] * c[:, 2]) for c in corners])
    touched = int(np.unique(ids).size)
    alg = 29.0 * n + 16.0 * touched
    line = {
        "metric": "sparse narrow-band build Mvoxel/s; ContactAware value+grad Mpts/s",
        "value": round(n / (q_ms * 1e-3) / 1e6, 3), "unit": "Mpts/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(build_s * 1e3 + q_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 query / f64 build", "data": "synthetic (seeded mesh and points)",
        "config": {"workload": "sparse hollow cylinder 1024x64 (266240 tris) at 512^3, 8^3 bricks, "
                               "band 4, halo 2; 1e6 near-surface ContactAware points",
                   "active_bricks": int(info.block_count), "query_points": n,
                   "touched_voxels": touched},
        "sparse_build_ms": round(build_s * 1e3, 3),
        "sparse_build_mvoxel_per_s": round(vox / build_s / 1e6, 3),
        "query_kernel_ms": round(q_ms, 4), "valid_fraction": round(frac_valid, 4),
        "contact_aware_sparse_build_ms": round(ca_build_s * 1e3, 3),
        "contact_aware_query_kernel_ms": round(ca_q_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(alg / (q_ms * 1e-3) / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (q_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "queryPairKernel<float, float, 3> (ContactAware -> sparse Tricubic "
                               "on sparse cell records)",
                     "algorithmic_bytes": alg},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        if args.cpu_threads > 0:
            O.set_threads(args.cpu_threads)
        bijk = g.download_sparse()["bijk"]
        og = O.setup_grid(v, 0.1, 0.0, 512, 512)
        rng = np.random.default_rng(0)
        pick = bijk[rng.choice(len(bijk), size=min(len(bijk), 400), replace=False)]
        L = np.arange(512)
        ijk = (pick[:, None, :] * 8 + np.stack([L % 8, (L // 8) % 8, L // 64], 1)[None]).reshape(-1, 3)
        c = og.bmin + og.voxel * (ijk + 0.5)
        t0 = time.perf_counter()
        O.signed_nearest(v, t, c)
        cpu_rate = len(c) / (time.perf_counter() - t0) / 1e6
        line["cpu_baseline"] = {
            "value": round(cpu_rate, 4), "unit": "Mvoxel/s (sparse fill)", "cores": O.num_threads(),
            "kind": "port", "sample": "FP64 oracle signed nearest over the voxels of 400 random "
                                      "active bricks (204 800 voxels)"}
    print(json.dumps(line))


def bench_batched_gears(args):
    """64 seeded gear meshes (2k-200k triangles, grids 64^3-256^3 by size)
    built by one batched call (one fill launch), then 64 Tricubic query
    streams of 10k-500k near-surface points against them in one plan launch,
    replayed from a captured graph."""
    import torch

    import sdfo
    from sdfo import _lib, scenes

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    batch = scenes.gear_batch(64, seed=5)
    meshes_ = [(v, t) for v, t, _ in batch]
    prm = [dict(target_resolution=r, max_resolution=r) for _, _, r in batch]
    grids = sdfo.SdfGrid.build_dense_batched(meshes_, prm)  # warm-up
    for g in grids:
        g.close()
    K = max(1, args.steps)
    build_s = 0.0
    for k in range(K):
        t0 = time.perf_counter()
        grids = sdfo.SdfGrid.build_dense_batched(meshes_, prm)
        build_s += (time.perf_counter() - t0) / K
        if k < K - 1:
            for g in grids:  # outside the timed build
                g.close()
    phases = _lib.debug_build_phases()
    vox = sum(int(np.prod(g.spec.shape)) for g in grids)
    rng = np.random.default_rng(9)
    counts = rng.integers(10_000, 500_001, size=len(grids))
    pts_np = np.concatenate([scenes.near_surface_points(v, t, int(c), 2.0 * g.spec.voxel.max(),
                                                        seed=i)
                             for i, ((v, t), c, g) in enumerate(zip(meshes_, counts, grids))])
    plan = sdfo.QueryPlan(grids, counts)
    n = plan.total
    pts = torch.from_numpy(pts_np.astype(np.float32)).to(dev)
    gap = torch.empty(n, device=dev)
    grad = torch.empty(n, 3, device=dev)
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    plan.run(pts, gap, grad, valid, "Tricubic")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(stream)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.run(pts, gap, grad, valid, "Tricubic", stream=s)
    stream.wait_stream(s)
    qs = []
    for _ in range(K):
        a, b = _events(torch, stream)
        _busy(torch)
        a.record(stream)
        graph.replay()
        b.record(stream)
        torch.cuda.synchronize()
        qs.append(a.elapsed_time(b))
    q_ms = float(np.mean(qs))
    line = {
        "metric": "batched build Mvoxel/s; batched tricubic Mpts/s",
        "value": round(vox / build_s / 1e6, 3), "unit": "Mvoxel/s", "n_gpus": 1, "steps": K,
        "warmup": 1, "ms_per_step": round(build_s * 1e3 + q_ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 build / f32 query",
        "data": "synthetic (seeded gear meshes and points)",
        "config": {"workload": "64 gear meshes (12-120 teeth, 2k-200k tris, grids 64-256 per "
                               "axis) in one batched build; 64 query streams 10k-500k points",
                   "meshes": len(grids), "triangles": int(sum(len(t) for _, t in meshes_)),
                   "voxels": vox, "query_points": n},
        "batched_build_ms": round(build_s * 1e3, 3),
        "batched_build_phases_ms": {k: round(float(v), 3) for k, v in phases.items()},
        "batched_query_graph_ms": round(q_ms, 4),
        "batched_query_mpts_per_s": round(n / (q_ms * 1e-3) / 1e6, 3),
        "tuner": sdfo.tuner_report(),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        if args.cpu_threads > 0:
            O.set_threads(args.cpu_threads)
        v, t, r = min(batch, key=lambda b: len(b[1]))
        og = O.setup_grid(v, 0.1, 0.0, r, r)
        t0 = time.perf_counter()
        O.build_dense(v, t, og)
        cpu_rate = og.total / (time.perf_counter() - t0) / 1e6
        line["cpu_baseline"] = {
            "value": round(cpu_rate, 4), "unit": "Mvoxel/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"FP64 oracle dense build of the smallest gear "
                                      f"({len(t)} tris, grid {list(og.shape)})"}
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sparse_cylinder":
        return bench_sparse_cylinder(args)
    if args.workload == "batched_gears":
        return bench_batched_gears(args)
    import torch

    import sdfo
    from sdfo import _lib
    from sdfo import contact as SC
    from sdfo import scenes

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()

    torus = scenes.ring_mesh(256, 64)
    cube = scenes.cube_mesh()
    cube_grid = sdfo.SdfGrid.build_dense(*cube, target_resolution=128, max_resolution=128)
    spec = cube_grid.spec
    pts_np = scenes.valid_region_points(spec, args.query_points, seed=1)
    pts = torch.from_numpy(pts_np).to(dev)
    X_np, w_np = scenes.ring_quadrature(120, 32)
    R, t = scenes.ring_poses(args.poses, seed=0)
    poses_np = scenes.pose_array(R, t)
    Rb, tb = scenes.ring_poses(args.bwd_poses, seed=2)
    v0, Om = scenes.pose_velocities(args.bwd_poses, seed=3)
    poses_b_np = scenes.pose_array(Rb, tb, v0=v0, Om=Om)
    X = torch.from_numpy(X_np).to(dev)
    w = torch.from_numpy(w_np).to(dev)
    poses = torch.from_numpy(poses_np).to(dev)
    poses_b = torch.from_numpy(poses_b_np).to(dev)
    g_out = torch.ones((args.bwd_poses, 6), dtype=torch.float64, device=dev)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.0, regularizer=1e-6, order=2)
    law_f = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    n_vox_torus = int(np.prod(sdfo.setup_grid(torus[0], target_resolution=128,
                                              max_resolution=128).shape))

    lin_cache = SC.LinearizationCache(cube_grid, X, w, law).fill(poses[0])
    q_gap = torch.empty(args.query_points, dtype=torch.float32, device=dev)
    q_grad = torch.empty((args.query_points, 3), dtype=torch.float32, device=dev)
    q_valid = torch.empty(args.query_points, dtype=torch.uint8, device=dev)

    def busy():
        try:
            torch.cuda._sleep(1_000_000)
        except (AttributeError, RuntimeError):
            pass

    def ev():
        return torch.cuda.Event(enable_timing=True)

    parts = set(args.parts.split(","))
    if args.copy_calibration:
        buf = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        buf.clone()
        torch.cuda.synchronize()
        del buf

    def step(rec):
        # (1) build
        if "build" in parts:
            t0 = time.perf_counter()
            g = sdfo.SdfGrid.build_dense(*torus, target_resolution=128, max_resolution=128)
            rec["build_s"] += time.perf_counter() - t0
            rec["phases"].append(_lib.debug_build_phases())
            g.close()
        e = [ev() for _ in range(10)]
        # (2) 1M-point tricubic query; each timed segment starts behind a short
        # device spin so the host's launch latency is not inside the events
        busy()
        e[0].record(stream)
        if "query" in parts:
            # back-to-back passes over the same batch: the per-pass time at
            # steady clocks (a single 35 us launch after the build's host
            # phases would also time the clock ramp)
            for _ in range(args.query_reps):
                cube_grid.query_into(pts, q_gap, q_grad, q_valid, "Tricubic")
        e[1].record(stream)
        # (3) contact forward, 1000 poses
        busy()
        e[2].record(stream)
        if "contact" in parts:
            SC.contact_forward(cube_grid, X, w, poses, law)
        e[3].record(stream)
        # (4) batched fwd+bwd vs per-pose plain path
        busy()
        e[4].record(stream)
        if "bwd" in parts:
            SC.contact_forward(cube_grid, X, w, poses_b, law_f)
            SC.contact_backward(cube_grid, X, w, poses_b, g_out, law_f)
        e[5].record(stream)
        busy()
        e[6].record(stream)
        if "bwd" in parts:
            for b in range(args.bwd_poses):
                pb = poses_b[b:b + 1]
                SC.contact_forward(cube_grid, X, w, pb, law_f)
                SC.contact_backward(cube_grid, X, w, pb, g_out[b:b + 1], law_f)
        e[7].record(stream)
        # (5) the same 1000 poses on the linearisation cache (a Newton iterate)
        busy()
        e[8].record(stream)
        if "contact" in parts:
            lin_cache.forward(poses)
        e[9].record(stream)
        torch.cuda.synchronize()
        rec["query_ms"].append(e[0].elapsed_time(e[1]) / args.query_reps)
        rec["contact_ms"].append(e[2].elapsed_time(e[3]))
        rec["fwdbwd_ms"].append(e[4].elapsed_time(e[5]))
        rec["plain_ms"].append(e[6].elapsed_time(e[7]))
        rec["lin_ms"].append(e[8].elapsed_time(e[9]))

    def new_rec():
        return dict(build_s=0.0, query_ms=[], contact_ms=[], fwdbwd_ms=[], plain_ms=[], lin_ms=[],
                    phases=[])

    warm = new_rec()
    for _ in range(args.warmup):
        step(warm)
    rec = new_rec()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(rec)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    K = args.steps
    q_ms = max(float(np.mean(rec["query_ms"])), 1e-9)
    c_ms = max(float(np.mean(rec["contact_ms"])), 1e-9)
    fb_ms = max(float(np.mean(rec["fwdbwd_ms"])), 1e-9)
    pl_ms = max(float(np.mean(rec["plain_ms"])), 1e-9)
    lin_ms = max(float(np.mean(rec["lin_ms"])), 1e-9)
    build_s = max(rec["build_s"] / K, 1e-12)
    nq = args.query_points
    value = nq / (q_ms * 1e-3) / 1e6
    # Algorithmic HBM bytes of one query launch: 12 B in + 16 B out + 1 B valid
    # per point, plus the FP32 field (16 B per voxel) read once.
    alg_bytes = 29.0 * nq + 16.0 * spec.total
    achieved = alg_bytes / (q_ms * 1e-3) / 1e9
    n_eval = args.poses * len(X_np)
    n_eval_b = args.bwd_poses * len(X_np)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 query / f64 build+contact",
        "data": "synthetic (seeded meshes, points and poses)",
        "config": {
            "workload": "ring-on-cube full size: torus 256x64 (32768 tris) dense build at 128; "
                        "1e6-pt Tricubic value+grad on cube 128^3; 1000 poses x 7680 pts contact; "
                        "256-pose fwd+bwd with friction",
            "query_points": nq, "grid": list(spec.shape), "poses": args.poses,
            "quadrature_points": len(X_np), "bwd_poses": args.bwd_poses,
        },
        "build_mvoxel_per_s": round(n_vox_torus / build_s / 1e6, 3),
        "build_phases_ms": {k: round(float(np.mean([p[k] for p in rec["phases"]])), 4)
                            for k in (rec["phases"][0] if rec["phases"] else {})},
        "build_ms": round(build_s * 1e3, 3),
        "query_kernel_ms": round(q_ms, 4),
        "contact_fwd_evals_per_s": round(n_eval / (c_ms * 1e-3), 1),
        "contact_linearized_evals_per_s": round(n_eval / (lin_ms * 1e-3), 1),
        "linearized_vs_direct": round(c_ms / lin_ms, 3),
        "contact_fwdbwd_evals_per_s": round(n_eval_b / (fb_ms * 1e-3), 1),
        "plain_fwdbwd_evals_per_s": round(n_eval_b / (pl_ms * 1e-3), 1),
        "fwdbwd_vs_plain": round(pl_ms / fb_ms, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "queryPairKernel<float, float, 3> (Tricubic, two points per lane)",
            "algorithmic_bytes": alg_bytes,
        },
    }
    tr = _traffic(nq, spec.shape)
    if tr is not None:
        line["roofline"]["traffic"] = tr
        line["roofline"]["traffic_source"] = TRAFFIC_FILE
    if not args.no_cpu:
        cpu, threads = cpu_baseline(args, spec, torus, cube, pts_np, X_np, w_np, poses_np)
        line["cpu_baseline"] = {
            "value": round(cpu["tricubic_mpts_per_s"], 3), "unit": "Mpts/s", "cores": threads,
            "kind": "port",
            "sample": "FP64 oracle: Tricubic on the same 1e6 points x3; build rate on a 16-slice "
                      "slab of the torus grid; contact on 200 of the 1000 poses",
            "build_mvoxel_per_s": round(cpu["build_mvoxel_per_s"], 4),
            "contact_fwd_evals_per_s": round(cpu["contact_fwd_evals_per_s"], 1),
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
