"""Batched builds (many meshes, one persistent fill launch), batched query
plans (many grids and point streams, one launch, HIP-graph capturable) and
the autotuner: every result equals the single-mesh / single-grid path, and
the dense fields equal the FP64 oracle bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import ograd  # noqa: F401
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu


def _small_batch():
    # uneven meshes and grids: gears of different sizes, a torus, a cube
    b = [(v, t, r) for v, t, r in scenes.gear_batch(4, seed=11, min_tris=500, max_tris=6000)]
    b[1] = (b[1][0], b[1][1], 40)
    b.append(meshes.torus(0.05, 0.01, 48, 12) + (32,))
    b.append(meshes.cube(0.1, (0.01, 0.0, -0.02)) + (24,))
    return [(v, t, min(r, 48)) for v, t, r in b]


def _params(r):
    return dict(target_resolution=r, max_resolution=r)


def test_batched_build_matches_single_and_oracle(O, sdfo_gpu):
    batch = _small_batch()
    print(f"[gpu] batched build of {len(batch)} meshes", flush=True)
    grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t) for v, t, _ in batch],
                                                 [_params(r) for _, _, r in batch])
    assert len(grids) == len(batch)
    for m, ((v, t, r), g) in enumerate(zip(batch, grids)):
        single = sdfo_gpu.SdfGrid.build_dense(v, t, **_params(r))
        a, b = g.download_dense(), single.download_dense()
        for key in ("phi", "n", "witness", "H", "tri"):
            assert np.array_equal(a[key], b[key]), (m, key)
        if m in (0, 4, 5):
            og = O.setup_grid(v, 0.1, 0.0, r, r)
            ref = O.build_dense(v, t, og)
            for key in ("phi", "n", "witness", "H", "tri"):
                assert np.array_equal(a[key], ref[key]), (m, key)


def test_batched_build_replay_pass(O, sdfo_gpu):
    from sdfo import _lib

    batch = _small_batch()[3:]
    print("[gpu] batched build with forced replay", flush=True)
    _lib.debug_force_replay(True)
    try:
        grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t) for v, t, _ in batch],
                                                     [_params(r) for _, _, r in batch])
    finally:
        _lib.debug_force_replay(False)
    for (v, t, r), g in zip(batch, grids):
        ref = O.build_dense(v, t, O.setup_grid(v, 0.1, 0.0, r, r))
        got = g.download_dense()
        for key in ("phi", "n", "tri"):
            assert np.array_equal(got[key], ref[key]), key


@pytest.mark.parametrize("budget,cap", [(16, 0), (16, 3)])
def test_batched_build_frontier_pass(O, sdfo_gpu, budget, cap):
    """A tiny step budget hands most bricks to the frontier pass (and, past a
    heavy-list capacity of 3 bricks, to the replay pass); results stay exact."""
    from sdfo import _lib

    batch = _small_batch()[2:]
    print(f"[gpu] batched build, budget {budget}, heavy cap {cap}", flush=True)
    _lib.debug_fill_budget(budget, cap)
    try:
        grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t) for v, t, _ in batch],
                                                     [_params(r) for _, _, r in batch])
        replayed = _lib.debug_build_phases()["replayed_voxels"]
    finally:
        _lib.debug_fill_budget(0, 0)
    if cap:
        assert replayed > 0
    for (v, t, r), g in zip(batch, grids):
        ref = O.build_dense(v, t, O.setup_grid(v, 0.1, 0.0, r, r))
        got = g.download_dense()
        for key in ("phi", "n", "witness", "tri"):
            assert np.array_equal(got[key], ref[key]), key


def test_batched_build_errors(sdfo_gpu):
    v, t = meshes.cube(0.1)
    with pytest.raises(sdfo_gpu.SdfoError, match="Empty mesh"):
        sdfo_gpu.SdfGrid.build_dense_batched([(v, t), (np.zeros((0, 3)), np.zeros((0, 3), np.int32))])
    assert sdfo_gpu.SdfGrid.build_dense_batched([]) == []


def _plan_setup(sdfo):
    cube = sdfo.SdfGrid.build_dense(*meshes.cube(0.1), target_resolution=32, max_resolution=32)
    tor = sdfo.SdfGrid.build_dense(*meshes.torus(0.05, 0.01, 32, 12), target_resolution=24,
                                   max_resolution=24)
    sp = sdfo.SdfGrid.build_sparse(*meshes.uv_sphere(0.05, 12, 24), target_resolution=32,
                                   max_resolution=32, band_voxels=3)
    grids = [cube, tor, sp, cube, tor]
    counts = [1000, 0, 777, 64, 4097]
    pts = [scenes.box_points(g.spec, n, seed=k, dtype=np.float32) for k, (g, n) in
           enumerate(zip(grids, counts))]
    P = torch.from_numpy(np.concatenate(pts).astype(np.float32)).cuda()
    return grids, counts, P


@pytest.mark.parametrize("mode", ["Trilinear", "Tricubic", "FirstOrder", "SecondOrder",
                                  "ContactAware"])
def test_query_plan_equals_single_queries(sdfo_gpu, mode):
    grids, counts, P = _plan_setup(sdfo_gpu)
    print(f"[gpu] query plan {mode} over {len(grids)} streams", flush=True)
    plan = sdfo_gpu.QueryPlan(grids, counts)
    n = plan.total
    gap = torch.empty(n, device="cuda")
    grad = torch.empty(n, 3, device="cuda")
    valid = torch.empty(n, dtype=torch.uint8, device="cuda")
    plan.run(P, gap, grad, valid, mode)
    torch.cuda.synchronize()
    off = plan.offsets
    for s, g in enumerate(grids):
        a, b = int(off[s]), int(off[s + 1])
        if a == b:
            continue
        g0, gr0, v0 = g.query(P[a:b].contiguous(), mode)
        assert torch.equal(gap[a:b], g0) and torch.equal(grad[a:b], gr0)
        assert torch.equal(valid[a:b].bool(), v0)


def test_query_batched_convenience(sdfo_gpu):
    from sdfo import _lib

    grids, counts, P = _plan_setup(sdfo_gpu)
    plan = sdfo_gpu.QueryPlan(grids, counts)
    n = plan.total
    outs = [torch.empty(n, device="cuda"), torch.empty(n, 3, device="cuda"),
            torch.empty(n, dtype=torch.uint8, device="cuda")]
    plan.run(P, *outs, "Tricubic")
    o2 = [torch.zeros_like(t) for t in outs]
    handles = (C.c_void_p * len(grids))(*[g._h for g in grids])
    _lib.check(_lib.lib.sdfo_query_batched(len(grids), handles, 3,
                                           plan.offsets.ctypes.data_as(C.c_void_p),
                                           C.c_void_p(P.data_ptr()), C.c_void_p(o2[0].data_ptr()),
                                           C.c_void_p(o2[1].data_ptr()), C.c_void_p(o2[2].data_ptr()),
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    for a, b in zip(outs, o2):
        assert torch.equal(a, b)


def test_query_plan_graph_replays(sdfo_gpu):
    """The plan launch captured in a HIP graph: every replay (with eager work
    and point updates in between) equals an eager run."""
    grids, counts, P = _plan_setup(sdfo_gpu)
    plan = sdfo_gpu.QueryPlan(grids, counts)
    n = plan.total
    gap = torch.empty(n, device="cuda")
    grad = torch.empty(n, 3, device="cuda")
    valid = torch.empty(n, dtype=torch.uint8, device="cuda")
    pts = P.clone()
    plan.run(pts, gap, grad, valid, "Tricubic")  # warm-up outside the capture
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.run(pts, gap, grad, valid, "Tricubic", stream=s)
    torch.cuda.current_stream().wait_stream(s)
    eg, er, ev = (torch.empty_like(gap), torch.empty_like(grad), torch.empty_like(valid))
    for rep in range(4):
        pts.copy_(P + 1e-3 * rep)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        plan.run(pts, eg, er, ev, "Tricubic")
        torch.cuda.synchronize()
        assert torch.equal(gap, eg) and torch.equal(grad, er) and torch.equal(valid, ev), rep


def test_tuner_chooses_and_keeps_results(sdfo_gpu):
    sdfo = sdfo_gpu
    v, t = meshes.torus(0.05, 0.01, 48, 12)
    ref = sdfo.SdfGrid.build_dense(v, t, target_resolution=32, max_resolution=32).download_dense()
    sdfo.tuner_reset()
    sdfo.tuner_set_mode(1)
    try:
        print("[gpu] tuner: dense fill shapes", flush=True)
        g = sdfo.SdfGrid.build_dense(v, t, target_resolution=32, max_resolution=32)
        got = g.download_dense()
        grids, counts, P = _plan_setup(sdfo)
        plan = sdfo.QueryPlan(grids, counts)
        n = plan.total
        outs = [torch.empty(n, device="cuda"), torch.empty(n, 3, device="cuda"),
                torch.empty(n, dtype=torch.uint8, device="cuda")]
        print("[gpu] tuner: query plan block sizes", flush=True)
        plan.run(P, *outs, "Tricubic")
        print("[gpu] tuner: record-path query block sizes", flush=True)
        q_pts = P[: counts[0]].contiguous()
        q_tuned = grids[0].query(q_pts, "Tricubic")
        rep = sdfo.tuner_report()
    finally:
        sdfo.tuner_set_mode(0)
    print(rep, flush=True)
    assert "dense_fill" in rep and "query_plan" in rep and "query dense tricubic" in rep
    assert "*" in rep
    # the tuned query gives the bits of every other block size
    q_ref = grids[0].query(q_pts, "Tricubic")
    for a, b in zip(q_tuned, q_ref):
        assert torch.equal(a, b)
    for key in ("phi", "n", "witness", "tri"):
        assert np.array_equal(got[key], ref[key]), key
    # a tuned choice is reused (and still exact) with the tuner off
    again = sdfo.SdfGrid.build_dense(v, t, target_resolution=32, max_resolution=32).download_dense()
    assert np.array_equal(again["phi"], ref["phi"])
    sdfo.tuner_reset()
    assert sdfo.tuner_report() == ""
