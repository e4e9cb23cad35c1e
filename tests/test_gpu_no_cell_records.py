"""The per-corner gather path of grids that keep no cell records.

Grids whose packed 8-corner cell records would exceed the library's limit
(an eighth of device memory, at least 4 GiB: dense grids past about 650^3 on
a 288 GB device) query by gathering each corner from the
FP32 mirror instead.  The debug switch builds small grids that way, and the
FP32 queries (all five modes, dense and sparse), the batched plan and the
contact kernels are checked against the oracle at the same bars as the
record path.
"""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import _lib
from sdfo import contact as SC
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu

MODES = ["FirstOrder", "SecondOrder", "Trilinear", "Tricubic", "ContactAware"]


@pytest.fixture(scope="module")
def grids(O, sdfo_gpu):
    vt, tt = meshes.torus(0.05, 0.01, 64, 16)
    vc, tc = meshes.hollow_cylinder(0.04, 0.03, 0.05, 64, 8)
    vq, tq = scenes.cube_mesh()
    with_rec = sdfo_gpu.SdfGrid.build_dense(vt, tt, target_resolution=48, max_resolution=48)
    _lib.debug_skip_cell_records(True)
    try:
        dense = sdfo_gpu.SdfGrid.build_dense(vt, tt, target_resolution=48, max_resolution=48)
        sparse = sdfo_gpu.SdfGrid.build_sparse(vc, tc, target_resolution=64, max_resolution=64,
                                               block_dim=8, band_voxels=4, halo_voxels=2)
        cube = sdfo_gpu.SdfGrid.build_dense(vq, tq, target_resolution=64, max_resolution=64)
    finally:
        _lib.debug_skip_cell_records(False)
    # the records are 8 float4 per cell: skipping them shrinks the grid by ~128 B per voxel
    assert with_rec.info.device_bytes - dense.info.device_bytes > 100 * dense.spec.total
    with_rec.close()
    yield dict(dense=(dense, dense.download_dense(), vt, tt),
               sparse=(sparse, sparse.download_sparse(), vc, tc), cube=(cube, cube.download_dense()))
    for g in (dense, sparse, cube):
        g.close()


def _points(spec, n, seed):
    p = scenes.box_points(spec, n, seed=seed)
    return np.concatenate([p, [spec.bmin - 1.0, spec.bmax + 1.0, spec.bmin]]).astype(np.float32)


@pytest.mark.parametrize("mode", MODES)
def test_dense_modes_without_records(O, grids, mode):
    g, ref, _, _ = grids["dense"]
    print(f"[gpu] no-record dense {mode}", flush=True)
    pts = _points(g.spec, 20000, seed=21)
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), mode)
    rq = O.query_dense(ograd(O, g.spec), dict(phi=f32(ref["phi"]), n=f32(ref["n"]), H=f32(ref["H"])),
                       mode, pts.astype(np.float64))
    v = valid.cpu().numpy()
    assert np.array_equal(v, rq["valid"]) and 0.5 < v.mean() < 1.0
    assert np.abs(gap.cpu().numpy() - rq["gap"]).max() <= 1e-6 * g.spec.diagonal
    assert np.abs(grad.cpu().numpy() - rq["grad"]).max() <= 2e-4


@pytest.mark.parametrize("mode", MODES)
def test_sparse_modes_without_records(O, grids, mode):
    g, sp, v, t = grids["sparse"]
    print(f"[gpu] no-record sparse {mode}", flush=True)
    pts = np.concatenate([scenes.near_surface_points(v, t, 15000, 2 * g.spec.voxel[0], seed=5),
                          _points(g.spec, 3000, seed=6)]).astype(np.float32)
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), mode)
    rq = O.query_sparse(ograd(O, g.spec), sp, mode, pts.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"]) and 0.3 < vv.mean() < 1.0
    assert np.abs(gap.cpu().numpy() - rq["gap"]).max() <= 1e-6 * g.spec.diagonal
    assert np.abs(grad.cpu().numpy() - rq["grad"]).max() <= 2e-4


def test_plan_without_records(O, grids, sdfo_gpu):
    gd, ref, _, _ = grids["dense"]
    gs, sp, v, t = grids["sparse"]
    print("[gpu] no-record batched plan", flush=True)
    pd = _points(gd.spec, 5000, seed=31)
    ps = scenes.near_surface_points(v, t, 7000, 2 * gs.spec.voxel[0], seed=32).astype(np.float32)
    plan = sdfo_gpu.QueryPlan([gd, gs], [len(pd), len(ps)])
    P = torch.from_numpy(np.concatenate([pd, ps])).cuda()
    n = len(P)
    gap = torch.empty(n, device="cuda")
    grad = torch.empty((n, 3), device="cuda")
    valid = torch.empty(n, dtype=torch.uint8, device="cuda")
    plan.run(P, gap, grad, valid, "Tricubic")
    rd = O.query_dense(ograd(O, gd.spec), dict(phi=f32(ref["phi"]), n=f32(ref["n"])), "Tricubic",
                       pd.astype(np.float64))
    rs = O.query_sparse(ograd(O, gs.spec), sp, "Tricubic", ps.astype(np.float64))
    gap, valid = gap.cpu().numpy(), valid.cpu().numpy().astype(bool)
    assert np.array_equal(valid[:len(pd)], rd["valid"]) and np.array_equal(valid[len(pd):], rs["valid"])
    assert np.abs(gap[:len(pd)] - rd["gap"]).max() <= 1e-6 * gd.spec.diagonal
    assert np.abs(gap[len(pd):] - rs["gap"]).max() <= 1e-6 * gs.spec.diagonal
    plan.close()


def test_contact_without_records(O, grids):
    g, ref = grids["cube"]
    print("[gpu] no-record contact forward/backward/evaluate", flush=True)
    X, w = scenes.ring_quadrature(60, 16)
    R, t = scenes.ring_poses(32, seed=3)
    v0, Om = scenes.pose_velocities(32, seed=4)
    poses = scenes.pose_array(R, t, v0=v0, Om=Om)
    Xt, wt, Pt = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (X, w, poses))
    kw = dict(kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2, mode="Tricubic")
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    og, phi, n3 = ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"])
    out, act = SC.contact_forward(g, Xt, wt, Pt, law, active=True)
    o_ref, a_ref = O.contact_forward(og, phi, n3, X, w, poses, **kw)
    assert np.array_equal(act.cpu().numpy(), a_ref) and (a_ref > 0).sum() > 8
    scale = np.abs(o_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out.cpu().numpy() - o_ref) <= 1e-9 * scale)
    gout = np.random.default_rng(0).normal(size=(len(poses), 6))
    got = SC.contact_backward(g, Xt, wt, Pt, torch.from_numpy(gout).cuda(), law)
    exp = O.contact_backward(og, phi, n3, X, w, poses, gout, **kw)
    for a, e in zip(got, exp):
        assert np.abs(a.cpu().numpy() - e).max() <= 1e-7 * (np.abs(e).max() + 1e-300)
    st = np.zeros((len(poses), 36))
    for b in range(len(poses)):
        st[b] = O.pair_states(t[b][None], R[b][None], v0[b][None], Om[b][None], np.zeros((1, 3)),
                              np.eye(3)[None], np.zeros((1, 3)), np.zeros((1, 3)))[0]
    law_e = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, narrow_margin=0.07,
                          order=2)
    out_e = SC.contact_evaluate(g, Xt, wt, torch.from_numpy(st).cuda(), law_e)
    e_ref, _ = O.contact_evaluate(og, phi, n3, X, w, st, margin=0.07, **kw)
    scale = np.abs(e_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out_e.cpu().numpy() - e_ref) <= 1e-9 * scale)
