"""Benchmark of the SdfOracle hot path on one MI355X (ring-on-cube at full size).

One step = (1) dense FP64 build of the torus ring mesh (R 50 mm, r 10 mm,
256 x 64 segments = 32 768 triangles) at target resolution 128;
(2) Tricubic value + raw gradient of 10^6 FP32 points uniform in the valid
region of the 128^3 cube SDF (side 0.1 m, top face at z = 55 mm), over five
distinct pre-generated point sets with their own outputs, so no pass reuses
another's points or outputs in a cache; also the same points five times
(warm) and 10^6 FP64 points (the exact path);
(3) contact forward (penalty kn = 1e11 order 2, damping 1e4) of the 7 680
ring-centroid quadrature points (sum w = 1) under 1 000 seeded poses, in
body B's frame and as the world-frame evaluate with a rotated body B;
(4) batched contact forward + backward over 256 poses with friction (mu 0.3)
against the same work done pose by pose (the plain per-pose GPU path);
(5) the 1 000 poses on the linearisation cache.
Beside the step, the other two benchmark configurations run once each: the
512^3 sparse tube (build + 1e6 ContactAware points, plain and contact-aware)
and the 64-gear batched build + captured 64-stream query.

``value`` is (2) on the distinct point sets, in Mpts/s; the other rates are
reported beside it.  Inputs are resident in HBM before the timed region.  The
cpu_baseline leg times the FP64 CPU oracle (oracle/) on bounded samples of
the same workloads: one warm-up, then the median of five runs.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sdf-contact-module_amd"))

METRIC = "build Mvoxel/s; tricubic value+grad Mpts/s; fwd+bwd contact evals/s vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Algorithmic flops per (pose, point) evaluation of the contact law: 600 for the
# forward (tricubic value + gradient 500, transform, law and torque 100) and
# 1300 for the backward (recomputed forward 600, Hessian 500, chain rule 200).
FWD_FLOPS = 600.0
BWD_FLOPS = 1300.0

TRAFFIC_FILE = "profiles/query_tricubic_cube128_1M_traffic.json"
WORK_FILE = "profiles/algorithmic_work_counts.json"


def _json(path):
    try:
        with open(os.path.join(ROOT, path)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def _traffic(nq, shape):
    """HBM bytes per query launch from the committed FETCH_SIZE / WRITE_SIZE
    summary (counters scaled by the copy-kernel calibration), when the
    workload is the one it was measured on."""
    t = _json(TRAFFIC_FILE)
    if not t:
        return None
    w = t.get("workload", {})
    if w.get("query_points") != nq or list(w.get("grid", [])) != list(shape):
        return None
    return float(t["fetch_kib_reported"] * 1024 * t["fetch_factor"] +
                 t["write_kib_reported"] * 1024 * t["write_factor"])


def _leg_traffic(name, nq):
    """HBM bytes per launch of a leg's query kernel from its committed
    counter summary, when the point count is the one it was measured on."""
    t = _json(os.path.join("profiles", name))
    if not t or t.get("workload", {}).get("query_points") != nq:
        return None
    return float(t["bytes_per_launch"])


def _valu_roofline(flops, seconds, kernel, basis):
    ach = flops / seconds / 1e12
    return {"bound": "valu", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": kernel, "algorithmic_flops": flops, "flops_basis": basis}


def _hbm_roofline(alg_bytes, seconds, kernel, traffic=None):
    ach = alg_bytes / seconds / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "algorithmic_bytes": alg_bytes}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--query-points", type=int, default=1_000_000)
    ap.add_argument("--query-reps", type=int, default=5,
                    help="query passes per step (distinct point sets), timed together")
    ap.add_argument("--poses", type=int, default=1000)
    ap.add_argument("--bwd-poses", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parts", default="build,query,contact,bwd,plain",
                    help="subset of the step to run (profiling)")
    ap.add_argument("--legs", default="sparse_cylinder,batched_gears",
                    help="other configurations measured once beside the step ('' for none)")
    ap.add_argument("--tuner", default="on", choices=["on", "off"],
                    help="autotuner: 'on' times the launch variants on the first call of each "
                         "size class (during the warm-up) and keeps the fastest")
    ap.add_argument("--cell-records", default="on", choices=["on", "off"],
                    help="'off' builds the query cube without its 128-byte cell records, so the "
                         "FP32 queries gather the 8 corners from the mirror (the path of grids "
                         "past the record limit; profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (scales HBM counters)")
    ap.add_argument("--workload", default="ring_on_cube",
                    choices=["ring_on_cube", "sparse_cylinder", "batched_gears"],
                    help="ring_on_cube (default line); sparse_cylinder: 512^3 narrow band of a "
                         "266k-triangle tube + 1e6 ContactAware points; batched_gears: 64 gear "
                         "meshes built in one launch + 64 query streams in one captured graph")
    return ap.parse_args()


def cpu_leg(args, leg):
    """The FP64 oracle on a bounded sample of one workload (rank 0, N=1), in a
    child process with its OpenMP threads bound to cores (OMP_PROC_BIND=close,
    OMP_PLACES=cores) so the binding stays out of this process's host threads:
    one warm-up, then the median of five runs per rate."""
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--leg", leg]
    if args.cpu_threads > 0:
        cmd += ["--threads", str(args.cpu_threads)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def _events(torch, stream):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _busy(torch):
    try:
        torch.cuda._sleep(1_000_000)
    except (AttributeError, RuntimeError):
        pass


def _work_counts():
    return _json(WORK_FILE) or {}


def sparse_cylinder_leg(args, steps):
    """Tube R_o 40 mm, R_i 30 mm, H 50 mm, 1024 x 64 segments (266 240 triangles),
    sparse 512^3, 8^3 bricks, band 4 voxels, halo 2; 1e6 FP32 points on the
    surface offset by N(0, (2h)^2) along the normal, ContactAware value + gradient
    (the plain sparse grid has no candidates, so ContactAware runs the sparse
    Tricubic path, SdfOracle.cpp:191-193); then the same with a sparse
    contact-aware build (face normals as corner normals)."""
    import torch

    import sdfo
    from sdfo import meshes, scenes

    sdfo.tuner_set_mode(1 if args.tuner == "on" else 0)
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
    K = max(1, steps)
    builds = []
    for k in range(K):
        t0 = time.perf_counter()
        g = sdfo.SdfGrid.build_sparse(v, t, **prm)
        builds.append(time.perf_counter() - t0)
        if k < K - 1:
            g.close()  # outside the timed build
    build_s = float(np.median(builds))
    info = g.info
    h = g.spec.voxel.max()
    pts_np = scenes.near_surface_points(v, t, args.query_points, 2.0 * h, seed=3)
    pts = torch.from_numpy(pts_np).to(dev)
    n = len(pts_np)
    gap = torch.empty(n, device=dev)
    grad = torch.empty(n, 3, device=dev)
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    for _ in range(2):
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
    q_ms = float(np.median(qs))
    frac_valid = float(valid.float().mean().item())
    sdfo.SdfGrid.build_contact_aware(v, t, cn, sparse=True, **prm).close()  # warm-up
    t0 = time.perf_counter()
    gca = sdfo.SdfGrid.build_contact_aware(v, t, cn, sparse=True, **prm)
    ca_build_s = time.perf_counter() - t0
    gca.query_into(pts, gap, grad, valid, "ContactAware")
    cqs = []
    for _ in range(K):
        a, b = _events(torch, stream)
        _busy(torch)
        a.record(stream)
        gca.query_into(pts, gap, grad, valid, "ContactAware")
        b.record(stream)
        torch.cuda.synchronize()
        cqs.append(a.elapsed_time(b))
    ca_q_ms = float(np.median(cqs))
    gca.close()
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
    work = _work_counts().get("tube_1024x64_sparse_512")
    out = {
        "workload": "sparse hollow cylinder 1024x64 (266240 tris) at 512^3, 8^3 bricks, band 4, "
                    "halo 2; 1e6 near-surface ContactAware points",
        "active_bricks": int(info.block_count), "query_points": n, "touched_voxels": touched,
        "sparse_build_ms": round(build_s * 1e3, 3),
        "sparse_build_mvoxel_per_s": round(vox / build_s / 1e6, 3),
        "query_kernel_ms": round(q_ms, 4), "query_mpts_per_s": round(n / (q_ms * 1e-3) / 1e6, 3),
        "valid_fraction": round(frac_valid, 4),
        "contact_aware_sparse_build_ms": round(ca_build_s * 1e3, 3),
        "contact_aware_query_kernel_ms": round(ca_q_ms, 4),
        "contact_aware_query_mpts_per_s": round(n / (ca_q_ms * 1e-3) / 1e6, 3),
        "roofline": _hbm_roofline(alg, q_ms * 1e-3, "queryPairKernel<float, float, 3> "
                                  "(ContactAware -> sparse Tricubic on sparse cell records)",
                                  _leg_traffic("query_sparse_tube512_1M_traffic.json", n)),
    }
    if work:
        out["build_roofline"] = _valu_roofline(
            work["flops"], build_s, "sparse build (activation + fill + Hessian), whole call",
            "reference-order nearest-search tests: fill voxels x " +
            str(work["fill"]["flops_per_voxel"]) + " + brick centres x " +
            str(work["activation"]["flops_per_voxel"]) + " flops (" + WORK_FILE + ")")
    if not args.no_cpu:
        c = cpu_leg(args, "sparse_cylinder")
        out["cpu_baseline"] = {
            "value": round(c["sparse_fill_mvoxel_per_s"], 4), "unit": "Mvoxel/s (sparse fill)",
            "cores": c["threads"], "kind": "port", "cpu_model": c["cpu_model"],
            "sample": "FP64 oracle signed nearest over the voxels of 400 random active bricks "
                      "(204 800 voxels), warm-up + median of 5"}
    g.close()
    return out


def bench_sparse_cylinder(args):
    leg = sparse_cylinder_leg(args, args.steps)
    line = {
        "metric": "sparse narrow-band build Mvoxel/s; ContactAware value+grad Mpts/s",
        "value": leg["query_mpts_per_s"], "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 2, "ms_per_step": round(leg["sparse_build_ms"] + leg["query_kernel_ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 query / f64 build", "data": "synthetic (seeded mesh and points)",
        "config": {"workload": leg["workload"], "active_bricks": leg["active_bricks"],
                   "query_points": leg["query_points"]},
    }
    line.update({k: v for k, v in leg.items() if k not in ("workload",)})
    print(json.dumps(line))


def batched_gears_leg(args, steps):
    """64 seeded gear meshes (2k-200k triangles, grids 64^3-256^3 by size)
    built by one batched call (one fill launch), then 64 Tricubic query
    streams of 10k-500k near-surface points against them in one plan launch,
    replayed from a captured graph."""
    import torch

    import sdfo
    from sdfo import _lib, scenes

    sdfo.tuner_set_mode(1 if args.tuner == "on" else 0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    batch = scenes.gear_batch(64, seed=5)
    meshes_ = [(v, t) for v, t, _ in batch]
    prm = [dict(target_resolution=r, max_resolution=r) for _, _, r in batch]
    grids = sdfo.SdfGrid.build_dense_batched(meshes_, prm)  # warm-up
    for g in grids:
        g.close()
    K = max(1, steps)
    builds, phases = [], []
    for k in range(K):
        t0 = time.perf_counter()
        grids = sdfo.SdfGrid.build_dense_batched(meshes_, prm)
        builds.append(time.perf_counter() - t0)
        phases.append(_lib.debug_build_phases())
        if k < K - 1:
            for g in grids:  # outside the timed build
                g.close()
    build_s = float(np.median(builds))
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
    q_ms = float(np.median(qs))
    fill_ms = float(np.median([p["fill_kernel"] for p in phases]))
    # 29 B per point plus each grid's FP32 field (16 B per voxel) read once
    alg = 29.0 * n + 16.0 * vox
    work = _work_counts().get("gears_64_batched")
    out = {
        "workload": "64 gear meshes (12-120 teeth, 2k-200k tris, grids 64-256 per axis) in one "
                    "batched build; 64 query streams 10k-500k near-surface points, Tricubic, one "
                    "captured graph",
        "meshes": len(grids), "triangles": int(sum(len(t) for _, t in meshes_)), "voxels": vox,
        "query_points": n,
        "batched_build_ms": round(build_s * 1e3, 3),
        "batched_build_mvoxel_per_s": round(vox / build_s / 1e6, 3),
        "batched_build_phases_ms": {k: round(float(np.median([p[k] for p in phases])), 3)
                                    for k in phases[0]},
        "batched_query_graph_ms": round(q_ms, 4),
        "batched_query_mpts_per_s": round(n / (q_ms * 1e-3) / 1e6, 3),
        "query_roofline": _hbm_roofline(alg, q_ms * 1e-3, "queryBatchedKernel (Tricubic plan, "
                                        "one graph replay)",
                                        _leg_traffic("query_batched_gears_traffic.json", n)),
        "tuner": sdfo.tuner_report(),
    }
    if work:
        out["build_roofline"] = _valu_roofline(
            work["flops"], fill_ms * 1e-3, "denseFillBatchedKernel + frontier passes (fill phase)",
            "reference-order nearest-search tests, " + str(work["flops_per_voxel"]) +
            " flops per voxel averaged over the 64 grids (" + WORK_FILE + ")")
    if not args.no_cpu:
        c = cpu_leg(args, "batched_gears")
        out["cpu_baseline"] = {
            "value": round(c["build_mvoxel_per_s"], 4), "unit": "Mvoxel/s", "cores": c["threads"],
            "kind": "port", "cpu_model": c["cpu_model"],
            "sample": f"FP64 oracle dense build of the smallest gear ({c['triangles']} tris, grid "
                      f"{c['grid']}), warm-up + median of 5"}
    for g in grids:
        g.close()
    plan.close()
    return out


def bench_batched_gears(args):
    leg = batched_gears_leg(args, args.steps)
    line = {
        "metric": "batched build Mvoxel/s; batched tricubic Mpts/s",
        "value": leg["batched_build_mvoxel_per_s"], "unit": "Mvoxel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1,
        "ms_per_step": round(leg["batched_build_ms"] + leg["batched_query_graph_ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 build / f32 query", "data": "synthetic (seeded gear meshes and points)",
        "config": {"workload": leg["workload"], "meshes": leg["meshes"],
                   "triangles": leg["triangles"], "voxels": leg["voxels"],
                   "query_points": leg["query_points"]},
    }
    line.update({k: v for k, v in leg.items() if k != "workload"})
    print(json.dumps(line))


def _tuned_variant(report, prefix):
    """The variant the autotuner chose (marked *) for the first key that
    starts with prefix, or "default" when it timed none."""
    for line in report.splitlines():
        key, _, rest = line.partition(": ")
        if key.startswith(prefix):
            for item in rest.split(", "):
                if item.endswith("*"):
                    return item.split("=")[0]
    return "default"


def _morton_order(pts, spec):
    """Order of points by the Morton (z-order) code of their grid cells."""
    cell = np.floor((pts - np.asarray(spec.bmin)) / np.asarray(spec.voxel)).astype(np.int64)
    cell = np.clip(cell, 0, (1 << 20) - 1)
    code = np.zeros(len(pts), np.uint64)
    for b in range(20):
        for a in range(3):
            code |= ((cell[:, a] >> b) & 1).astype(np.uint64) << np.uint64(3 * b + a)
    return np.argsort(code, kind="stable")


def _random_read_floor(nq, shape, q_ms):
    """The memory system's floor for the record-path query: nq uniformly
    random 128-byte reads from an array the size of the grid's cell records
    (and the same count of 64-byte reads), best of 10 launches with fresh
    random records each."""
    from sdfo import _lib

    cells = int(np.prod(np.asarray(shape, np.int64) - 1))
    t128 = _lib.debug_random_read_floor(128, nq, cells * 128, 10)
    t64 = _lib.debug_random_read_floor(64, nq, cells * 64, 10)
    return {"reads": nq, "array_mb": round(cells * 128 / 1e6, 1), "read_128B_us": round(t128, 2),
            "read_64B_us": round(t64, 2), "query_us": round(q_ms * 1e3, 2),
            "query_over_128B_floor": round(q_ms * 1e3 / t128, 3)}


def _host_api_latency(grid, pts):
    """Round trip of the per-point C++ API path (SdfOracle::query and
    queryBatch call sdfo_query_host_f64: copy in, FP64 kernel, copy out,
    synchronise): median microseconds per call for one point and for a batch
    of 7680 points, Tricubic."""
    import ctypes as C

    from sdfo import _lib

    out = {}
    for n in (1, 7680):
        p = np.ascontiguousarray(pts[:n], np.float64)
        gap, nrm, wit = np.empty(n), np.empty((n, 3)), np.empty((n, 3))
        val = np.empty(n, np.uint8)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        ts = []
        for k in range(220):
            t0 = time.perf_counter()
            _lib.check(_lib.lib.sdfo_query_host_f64(grid.handle, 3, ptr(p), n, ptr(gap), ptr(nrm),
                                                   ptr(wit), ptr(val), None, None, None, None,
                                                   None, None, None))
            if k >= 20:
                ts.append(time.perf_counter() - t0)
        out["query_1_point" if n == 1 else "query_batch_7680_points"] = round(
            float(np.median(ts)) * 1e6, 1)
    return out


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
    sdfo.tuner_set_mode(1 if args.tuner == "on" else 0)

    torus = scenes.ring_mesh(256, 64)
    cube = scenes.cube_mesh()
    _lib.debug_skip_cell_records(args.cell_records == "off")
    cube_grid = sdfo.SdfGrid.build_dense(*cube, target_resolution=128, max_resolution=128)
    # the same grid without cell records: the per-corner gather from the FP32
    # mirror, for the spatially coherent stream below
    _lib.debug_skip_cell_records(True)
    cube_gather = sdfo.SdfGrid.build_dense(*cube, target_resolution=128, max_resolution=128)
    _lib.debug_skip_cell_records(False)
    spec = cube_grid.spec
    nq = args.query_points
    reps = max(1, args.query_reps)
    # distinct point sets (and outputs) for the passes of one step
    sets_np = [scenes.valid_region_points(spec, nq, seed=1 + r) for r in range(reps)]
    sets = [torch.from_numpy(p).to(dev) for p in sets_np]
    # the same point sets in Morton order of their cells (a spatially coherent
    # stream, as a caller that bins its points would hand over)
    sets_z = [torch.from_numpy(p[_morton_order(p, spec)]).to(dev) for p in sets_np]
    outs = [(torch.empty(nq, dtype=torch.float32, device=dev),
             torch.empty((nq, 3), dtype=torch.float32, device=dev),
             torch.empty(nq, dtype=torch.uint8, device=dev)) for _ in range(reps)]
    pts64 = torch.from_numpy(sets_np[0].astype(np.float64)).to(dev)
    out64 = (torch.empty(nq, dtype=torch.float64, device=dev),
             torch.empty((nq, 3), dtype=torch.float64, device=dev),
             torch.empty(nq, dtype=torch.uint8, device=dev))
    X_np, w_np = scenes.ring_quadrature(120, 32)
    R, t = scenes.ring_poses(args.poses, seed=0)
    poses_np = scenes.pose_array(R, t)
    Rb, tb = scenes.ring_poses(args.bwd_poses, seed=2)
    v0, Om = scenes.pose_velocities(args.bwd_poses, seed=3)
    poses_b_np = scenes.pose_array(Rb, tb, v0=v0, Om=Om)
    # the same 1000 poses as world-frame body states with a rotated, moving body B
    rng = np.random.default_rng(4)
    st = np.zeros((args.poses, 36))
    for b in range(args.poses):
        ax = rng.normal(size=3)
        RB = scenes.rot_axis(ax, rng.uniform(0.0, np.pi))
        rB = rng.normal(0, 0.3, 3)
        vA, wA, vB, wB = rng.normal(0, 0.05, (4, 3))
        st[b] = np.concatenate([rB + RB @ t[b], (RB @ R[b]).ravel(), vA, wA, rB, RB.ravel(), vB, wB])
    X = torch.from_numpy(X_np).to(dev)
    w = torch.from_numpy(w_np).to(dev)
    poses = torch.from_numpy(poses_np).to(dev)
    poses_b = torch.from_numpy(poses_b_np).to(dev)
    states = torch.from_numpy(st).to(dev)
    g_out = torch.ones((args.bwd_poses, 6), dtype=torch.float64, device=dev)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.0, regularizer=1e-6, order=2)
    law_f = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    law_e = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.0, regularizer=1e-6, narrow_margin=0.07,
                          order=2)
    n_vox_torus = int(np.prod(sdfo.setup_grid(torus[0], target_resolution=128,
                                              max_resolution=128).shape))

    lin_cache = SC.LinearizationCache(cube_grid, X, w, law).fill(poses[0])

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
            rec["build_s"].append(time.perf_counter() - t0)
            rec["phases"].append(_lib.debug_build_phases())
            g.close()
        e = [ev() for _ in range(22)]
        # (2) 1e6-point tricubic queries: each timed segment starts behind a
        # short device spin so the host's launch latency is not in the events.
        # (2a) distinct point sets and outputs, back to back
        busy()
        e[0].record(stream)
        if "query" in parts:
            for r in range(reps):
                cube_grid.query_into(sets[r], *outs[r], "Tricubic")
        e[1].record(stream)
        # (2b) the same points and outputs again and again (cache-warm)
        busy()
        e[2].record(stream)
        if "query" in parts:
            for _ in range(reps):
                cube_grid.query_into(sets[0], *outs[0], "Tricubic")
        e[3].record(stream)
        # (2b') the point sets in Morton order, back to back
        busy()
        e[18].record(stream)
        if "query" in parts:
            for r in range(reps):
                cube_grid.query_into(sets_z[r], *outs[r], "Tricubic")
        e[19].record(stream)
        # (2b'') the same on the per-corner gather path
        busy()
        e[20].record(stream)
        if "query" in parts:
            for r in range(reps):
                cube_gather.query_into(sets_z[r], *outs[r], "Tricubic")
        e[21].record(stream)
        # (2c) FP64 points, the exact path
        busy()
        e[4].record(stream)
        if "query" in parts:
            cube_grid.query_into(pts64, *out64, "Tricubic")
        e[5].record(stream)
        # (3) contact forward, 1000 poses, B frame and world-frame evaluate
        busy()
        e[6].record(stream)
        if "contact" in parts:
            SC.contact_forward(cube_grid, X, w, poses, law)
        e[7].record(stream)
        busy()
        e[8].record(stream)
        if "contact" in parts:
            SC.contact_evaluate(cube_grid, X, w, states, law_e)
        e[9].record(stream)
        # (4) batched fwd+bwd vs per-pose plain path
        busy()
        e[10].record(stream)
        if "bwd" in parts:
            # a training step: the forward keeps its pairs in contact for the
            # backward (what ContactFunction does)
            _, saved = SC.contact_forward(cube_grid, X, w, poses_b, law_f, save=True)
            SC.contact_backward(cube_grid, X, w, poses_b, g_out, law_f, saved=saved)
        e[11].record(stream)
        # (4b) the same with a backward that screens every pair itself
        busy()
        e[16].record(stream)
        if "bwd" in parts:
            SC.contact_forward(cube_grid, X, w, poses_b, law_f)
            SC.contact_backward(cube_grid, X, w, poses_b, g_out, law_f)
        e[17].record(stream)
        busy()
        e[12].record(stream)
        if "plain" in parts:
            for b in range(args.bwd_poses):
                pb = poses_b[b:b + 1]
                SC.contact_forward(cube_grid, X, w, pb, law_f)
                SC.contact_backward(cube_grid, X, w, pb, g_out[b:b + 1], law_f)
        e[13].record(stream)
        # (5) the same 1000 poses on the linearisation cache (a Newton iterate)
        busy()
        e[14].record(stream)
        if "contact" in parts:
            lin_cache.forward(poses)
        e[15].record(stream)
        torch.cuda.synchronize()
        rec["query_ms"].append(e[0].elapsed_time(e[1]) / reps)
        rec["query_warm_ms"].append(e[2].elapsed_time(e[3]) / reps)
        rec["query_z_ms"].append(e[18].elapsed_time(e[19]) / reps)
        rec["query_zg_ms"].append(e[20].elapsed_time(e[21]) / reps)
        rec["query64_ms"].append(e[4].elapsed_time(e[5]))
        rec["contact_ms"].append(e[6].elapsed_time(e[7]))
        rec["eval_ms"].append(e[8].elapsed_time(e[9]))
        rec["fwdbwd_ms"].append(e[10].elapsed_time(e[11]))
        rec["fwdbwd_screen_ms"].append(e[16].elapsed_time(e[17]))
        rec["plain_ms"].append(e[12].elapsed_time(e[13]))
        rec["lin_ms"].append(e[14].elapsed_time(e[15]))

    def new_rec():
        return dict(build_s=[], query_ms=[], query_warm_ms=[], query_z_ms=[], query_zg_ms=[],
                    query64_ms=[],
                    contact_ms=[],
                    eval_ms=[], fwdbwd_ms=[], fwdbwd_screen_ms=[], plain_ms=[], lin_ms=[],
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
    med = lambda k: max(float(np.median(rec[k])), 1e-9) if rec[k] else 1e-9  # noqa: E731
    q_ms, qw_ms, q64_ms = med("query_ms"), med("query_warm_ms"), med("query64_ms")
    qz_ms, qzg_ms = med("query_z_ms"), med("query_zg_ms")
    c_ms, e_ms, fb_ms = med("contact_ms"), med("eval_ms"), med("fwdbwd_ms")
    pl_ms, lin_ms, fbs_ms = med("plain_ms"), med("lin_ms"), med("fwdbwd_screen_ms")
    build_s = med("build_s")
    phase = {k: float(np.median([p[k] for p in rec["phases"]])) for k in
             (rec["phases"][0] if rec["phases"] else {})}
    value = nq / (q_ms * 1e-3) / 1e6
    # Algorithmic HBM bytes of one query launch: 12 B in + 16 B out + 1 B valid
    # per point, plus the FP32 field (16 B per voxel) read once.
    alg_bytes = 29.0 * nq + 16.0 * spec.total
    n_eval = args.poses * len(X_np)
    n_eval_b = args.bwd_poses * len(X_np)
    work = _work_counts().get("torus_256x64_dense_128")
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
                        "1e6-pt Tricubic value+grad on cube 128^3 (five distinct point sets per "
                        "step); 1000 poses x 7680 pts contact; 256-pose fwd+bwd with friction",
            "query_points": nq, "grid": list(spec.shape), "poses": args.poses,
            "quadrature_points": len(X_np), "bwd_poses": args.bwd_poses,
            "cell_records": args.cell_records == "on",
        },
        "build_mvoxel_per_s": round(n_vox_torus / build_s / 1e6, 3),
        "build_phases_ms": {k: round(v, 4) for k, v in phase.items()},
        "build_ms": round(build_s * 1e3, 3),
        "query_kernel_ms": round(q_ms, 4),
        "query_same_points_kernel_ms": round(qw_ms, 4),
        "query_same_points_mpts_per_s": round(nq / (qw_ms * 1e-3) / 1e6, 3),
        "query_morton_order_kernel_ms": round(qz_ms, 4),
        "query_morton_order_mpts_per_s": round(nq / (qz_ms * 1e-3) / 1e6, 3),
        "query_morton_order_roofline": _hbm_roofline(
            alg_bytes, qz_ms * 1e-3, "queryPairKernel<float, float, 3>, the same five point "
            "sets in Morton order of their cells (sorted on the host before timing)"),
        "query_morton_order_gather_kernel_ms": round(qzg_ms, 4),
        "query_morton_order_gather_mpts_per_s": round(nq / (qzg_ms * 1e-3) / 1e6, 3),
        "query_morton_order_gather_roofline": _hbm_roofline(
            alg_bytes, qzg_ms * 1e-3, "queryPairKernel<float, float, 3> gathering corners from "
            "the FP32 mirror (grid built without cell records), Morton-ordered point sets"),
        "tricubic_f64_points_mpts_per_s": round(nq / (q64_ms * 1e-3) / 1e6, 3),
        "contact_fwd_evals_per_s": round(n_eval / (c_ms * 1e-3), 1),
        "contact_evaluate_world_evals_per_s": round(n_eval / (e_ms * 1e-3), 1),
        "contact_linearized_evals_per_s": round(n_eval / (lin_ms * 1e-3), 1),
        "linearized_vs_direct": round(c_ms / lin_ms, 3),
        "contact_fwdbwd_evals_per_s": round(n_eval_b / (fb_ms * 1e-3), 1),
        "contact_fwdbwd_screening_backward_evals_per_s": round(n_eval_b / (fbs_ms * 1e-3), 1),
        "plain_fwdbwd_evals_per_s": round(n_eval_b / (pl_ms * 1e-3), 1),
        "fwdbwd_vs_plain": round(pl_ms / fb_ms, 3),
        "roofline": _hbm_roofline(alg_bytes, q_ms * 1e-3,
                                  "queryPairKernel<float, float, 3> (Tricubic, tuned launch "
                                  "shape " + _tuned_variant(sdfo.tuner_report(),
                                                            "query dense tricubic f32") +
                                  "), distinct point sets"),
        "contact_roofline": _valu_roofline(FWD_FLOPS * n_eval, c_ms * 1e-3,
                                           "contactForwardKernel + finalize (1000 poses)",
                                           "600 flops per (pose, point) evaluation"),
        "fwdbwd_roofline": _valu_roofline((FWD_FLOPS + BWD_FLOPS) * n_eval_b, fb_ms * 1e-3,
                                          "contact forward + backward launches (256 poses)",
                                          "1900 flops per (pose, point) evaluation"),
        "tuner": sdfo.tuner_report(),
        "host_api_latency_us": _host_api_latency(cube_grid, sets_np[0]),
        "query_random_read_floor": _random_read_floor(nq, spec.shape, q_ms),
    }
    tr = _traffic(nq, spec.shape)
    if tr is not None:
        line["roofline"]["traffic"] = tr
        line["roofline"]["traffic_source"] = TRAFFIC_FILE
    if work and "fill_kernel" in phase:
        line["build_roofline"] = _valu_roofline(
            work["flops"], phase["fill_kernel"] * 1e-3,
            "denseFillKernel + frontier and replay passes (fill phase, device events)",
            "reference-order nearest-search tests, " + str(work["flops_per_voxel"]) +
            " flops per voxel (" + WORK_FILE + ")")
    if not args.no_cpu:
        cpu = cpu_leg(args, "ring_on_cube")
        line["cpu_baseline"] = {
            "value": round(cpu["tricubic_mpts_per_s"], 3), "unit": "Mpts/s",
            "cores": cpu["threads"], "kind": "port", "cpu_model": cpu["cpu_model"],
            "sample": "FP64 oracle, one warm-up then the median of 5 runs each: Tricubic on the "
                      "first 1e6-point set; build rate on 262144 uniformly drawn voxels of the "
                      "torus grid; contact on all 1000 poses; fwd+bwd on 64 of the 256 poses; "
                      "the 0.1 m cube at 64^3 with the demo's 7680 Trilinear points",
            "build_mvoxel_per_s": round(cpu["build_mvoxel_per_s"], 4),
            "contact_fwd_evals_per_s": round(cpu["contact_fwd_evals_per_s"], 1),
            "fwdbwd_evals_per_s": round(cpu["fwdbwd_evals_per_s"], 1),
            "cube64_demo": cpu["cube64_demo"],
        }
    # a part left out of --parts was not timed: its figures are dropped, not
    # reported from empty timings
    skipped = {
        "build": ("build_mvoxel_per_s", "build_phases_ms", "build_ms", "build_roofline"),
        "query": ("query_kernel_ms", "query_same_points_kernel_ms", "query_same_points_mpts_per_s",
                  "query_morton_order_kernel_ms", "query_morton_order_mpts_per_s",
                  "query_morton_order_roofline", "query_morton_order_gather_kernel_ms",
                  "query_morton_order_gather_mpts_per_s", "query_morton_order_gather_roofline",
                  "tricubic_f64_points_mpts_per_s", "roofline",
                  "query_random_read_floor"),
        "contact": ("contact_fwd_evals_per_s", "contact_evaluate_world_evals_per_s",
                    "contact_linearized_evals_per_s", "linearized_vs_direct", "contact_roofline"),
        "bwd": ("contact_fwdbwd_evals_per_s", "contact_fwdbwd_screening_backward_evals_per_s",
                "fwdbwd_vs_plain", "fwdbwd_roofline"),
        "plain": ("plain_fwdbwd_evals_per_s", "fwdbwd_vs_plain"),
    }
    for part, keys in skipped.items():
        if part not in parts:
            for k in keys:
                line.pop(k, None)
    if "query" not in parts:
        line["value"] = None
    legs = [x for x in args.legs.split(",") if x]
    if "sparse_cylinder" in legs:
        line["sparse_cylinder"] = sparse_cylinder_leg(args, min(3, max(1, K)))
    if "batched_gears" in legs:
        line["batched_gears"] = batched_gears_leg(args, min(3, max(1, K)))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
