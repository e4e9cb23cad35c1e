"""Parity of the remaining benchmark-size legs against the oracle.

* The sparse tube (1024 x 64 segments, 266 240 triangles) at 512^3: 1e6
  near-surface FP32 points in ContactAware mode on the plain sparse grid
  (no candidates: the sparse Tricubic path) and on the sparse contact-aware
  grid (face normals as corner normals): valid mask equal, gap within 1e-6 x
  bbox diagonal, raw gradient within 2e-4; 100 000 of the points as FP64
  bit for bit.
* The sparse contact-aware build of that tube: its brick mask against the
  oracle's activation rule, and every candidate of the voxels of 300 seeded
  bricks bit for bit (FP32 storage of the FP64 values).
* The 64-gear batched plan at full size (20.1 M points over 64 streams),
  captured in a HIP graph and replayed: 2000 seeded points per stream
  against the oracle's Tricubic on each grid's FP32 field.
"""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu

PRM = dict(target_resolution=512, max_resolution=512, block_dim=8, band_voxels=4, halo_voxels=2)


@pytest.fixture(scope="module")
def tube():
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 1024, 64)
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    return v, t, np.repeat(fn[:, None, :], 3, axis=1)


def _check_queries(O, g, sp, pts32, label):
    print(f"[gpu] 1e6 ContactAware points, {label}", flush=True)
    og = ograd(O, g.spec)
    gap, grad, valid = g.query(torch.from_numpy(pts32).cuda(), "ContactAware")
    rq = O.query_sparse(og, sp, "ContactAware", pts32.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"])
    assert vv.mean() > 0.9
    assert np.abs(gap.cpu().numpy()[vv] - rq["gap"][vv]).max() <= 1e-6 * g.spec.diagonal
    assert np.abs(grad.cpu().numpy()[vv] - rq["grad"][vv]).max() <= 2e-4
    p64 = pts32[:100_000].astype(np.float64)
    p64 += 1e-9 * np.random.default_rng(4).standard_normal(p64.shape)   # off the FP32 lattice
    out = g.query(torch.from_numpy(p64).cuda(), "ContactAware", full=True)
    r64 = O.query_sparse(og, sp, "ContactAware", p64)
    for k, a in zip(("gap", "grad", "valid", "normal", "witness", "branch"), out):
        assert np.array_equal(a.cpu().numpy(), r64[k]), k


def test_tube_plain_sparse_contact_aware_queries_1m(O, sdfo_gpu, tube):
    v, t, _ = tube
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, **PRM)
    try:
        sp = g.download_sparse()
        pts = scenes.near_surface_points(v, t, 1_000_000, 2.0 * g.spec.voxel.max(), seed=3)
        _check_queries(O, g, sp, pts, "plain sparse tube (tricubic fallback)")
    finally:
        g.close()


def test_tube_contact_aware_build_and_queries_full_size(O, sdfo_gpu, tube):
    v, t, cn = tube
    print("[gpu] sparse contact-aware tube build at 512^3", flush=True)
    g = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, sparse=True, **PRM)
    try:
        sp = g.download_sparse()
        info = g.info
        S = info.candidate_stride
        assert S == 4 and len(sp["bijk"]) > 40000
        og = ograd(O, g.spec)
        rng = np.random.default_rng(5)
        pick = rng.choice(len(sp["bijk"]), 300, replace=False)
        L = np.arange(512)
        off = np.stack([L % 8, (L // 8) % 8, L // 64], 1)
        ijk = (sp["bijk"][pick][:, None, :] * 8 + off[None]).reshape(-1, 3)
        inside = np.all(ijk < np.array(g.spec.shape)[None], axis=1)
        ref = O.contact_aware_probe(v, t, cn, og, ijk[inside], max_candidates=4,
                                    sparse_block_dim=8, halo=2, band_voxels=4)
        assert np.array_equal(sp["active"], ref["mask"])
        cnt = sp["ccount"][pick].reshape(-1)[inside]
        assert np.array_equal(cnt, ref["ccount"])
        assert cnt.mean() > 1.0
        got_phi = sp["cphi"][pick].reshape(-1, S)[inside]
        got_n = sp["cn"][pick].reshape(-1, S, 3)[inside]
        got_b = sp["cbranch"][pick].reshape(-1, S)[inside]
        for q in range(S):
            m = cnt > q
            assert np.array_equal(got_phi[m, q], ref["cphi"][m, q].astype(np.float32))
            assert np.array_equal(got_n[m, q], ref["cn"][m, q].astype(np.float32))
            assert np.array_equal(got_b[m, q], ref["cbranch"][m, q])
        phi0 = sp["phi"][pick].reshape(-1)[inside]
        has = cnt > 0
        assert np.array_equal(phi0[has], ref["cphi"][has, 0].astype(np.float32))
        pts = scenes.near_surface_points(v, t, 1_000_000, 2.0 * g.spec.voxel.max(), seed=3)
        _check_queries(O, g, sp, pts, "sparse contact-aware tube")
    finally:
        g.close()


def test_batched_gears_plan_full_size_captured(O, sdfo_gpu):
    batch = scenes.gear_batch(64, seed=5)
    print("[gpu] 64-stream captured plan at full size", flush=True)
    grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t) for v, t, _ in batch],
                                                 [dict(target_resolution=r, max_resolution=r)
                                                  for _, _, r in batch])
    try:
        rng = np.random.default_rng(9)
        counts = rng.integers(10_000, 500_001, size=len(grids))
        pts_np = np.concatenate([scenes.near_surface_points(v, t, int(c), 2.0 * g.spec.voxel.max(),
                                                            seed=i)
                                 for i, ((v, t, _), c, g) in enumerate(zip(batch, counts, grids))])
        plan = sdfo_gpu.QueryPlan(grids, counts)
        n = plan.total
        assert n > 15_000_000
        pts = torch.from_numpy(pts_np.astype(np.float32)).cuda()
        gap = torch.empty(n, device="cuda")
        grad = torch.empty((n, 3), device="cuda")
        valid = torch.empty(n, dtype=torch.uint8, device="cuda")
        plan.run(pts, gap, grad, valid, "Tricubic")   # outside the capture (tuner, warm-up)
        torch.cuda.synchronize()
        main = torch.cuda.current_stream()
        s = torch.cuda.Stream()
        s.wait_stream(main)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            plan.run(pts, gap, grad, valid, "Tricubic", stream=s)
        main.wait_stream(s)
        gap.fill_(np.nan)
        valid.fill_(7)
        graph.replay()
        torch.cuda.synchronize()
        gap_h, grad_h, valid_h = gap.cpu().numpy(), grad.cpu().numpy(), valid.cpu().numpy()
        off = plan.offsets
        for sidx, g in enumerate(grids):
            ref = g.download_dense()
            pick = off[sidx] + rng.choice(int(counts[sidx]), 2000, replace=False)
            rq = O.query_dense(ograd(O, g.spec), dict(phi=f32(ref["phi"]), n=f32(ref["n"])),
                               "Tricubic", pts_np[pick].astype(np.float32).astype(np.float64))
            vv = valid_h[pick].astype(bool)
            assert np.array_equal(vv, rq["valid"]), sidx
            assert np.abs(gap_h[pick][vv] - rq["gap"][vv]).max() <= 1e-6 * g.spec.diagonal, sidx
            assert np.abs(grad_h[pick][vv] - rq["grad"][vv]).max() <= 2e-4, sidx
        plan.close()
    finally:
        for g in grids:
            g.close()


def test_dense_512_grid_keeps_cell_records(O, sdfo_gpu):
    """The record limit scales with device memory (an eighth of it): a dense
    512^3 grid (17.2 GB of records) keeps its records, and 1e6 Tricubic points
    on them match the oracle on the FP32-rounded field."""
    import ctypes as C

    from sdfo import _lib
    v, t = scenes.cube_mesh()
    print("[gpu] dense 512^3 cube grid with cell records", flush=True)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=512, max_resolution=512)
    try:
        n = np.array(g.spec.shape, np.int64)
        assert tuple(n) == (512, 512, 512)
        cells = int(np.prod(n - 1))
        assert g.info.device_bytes > cells * 128 + int(np.prod(n)) * 100
        pts = scenes.valid_region_points(g.spec, 1_000_000, seed=6)
        gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), "Tricubic")
        N = int(np.prod(n))
        phi, n3 = np.empty(N), np.empty((N, 3))
        _lib.check(_lib.lib.sdfo_grid_download_dense(g.handle, phi.ctypes.data_as(C.c_void_p),
                                                     n3.ctypes.data_as(C.c_void_p), None, None,
                                                     None))
        rq = O.query_dense(ograd(O, g.spec), dict(phi=f32(phi), n=f32(n3)), "Tricubic",
                           pts.astype(np.float64))
        vv = valid.cpu().numpy()
        assert np.array_equal(vv, rq["valid"]) and vv.mean() > 0.99
        assert np.abs(gap.cpu().numpy()[vv] - rq["gap"][vv]).max() <= 1e-6 * g.spec.diagonal
        assert np.abs(grad.cpu().numpy()[vv] - rq["grad"][vv]).max() <= 2e-4
    finally:
        g.close()
