"""Parity at the full sizes of the benchmark workloads.

* Ring build: the torus of 256 x 64 segments (32 768 triangles) at target
  128 (128 x 128 x 64 voxels).  The oracle's signed nearest search on 40 000
  seeded voxels plus every voxel of the columns around the torus axis (points
  nearly equidistant from the inner equator, the hardest for the search):
  distance, witness and triangle bit for bit.
* 1e6 FP32 points, Tricubic, on the cube grid at 128^3: every point against
  the oracle's FP64 Tricubic on the FP32-rounded field (valid mask equal, gap
  within 1e-6 x bbox diagonal).
* Contact: 1000 seeded poses x 7680 ring quadrature points on that grid,
  every pose's force and torque against the oracle to 1e-9 of each
  component's largest magnitude.
* Sparse tube (1024 x 64 segments, 266 240 triangles) at 512^3 in 8^3 bricks:
  every brick's activation against the oracle's |signed distance| of the
  brick centre, and the voxels of 300 seeded active bricks bit for bit.
* Differentiation: 256 poses with twists x 7680 points, friction on: forward
  and every gradient output of the backward against the oracle, with the
  backward's own screen and with the forward's record of the pairs in
  contact; 1200 poses (the backward in pose groups) the same; and poses out of
  contact give zeros.
* Batched gears: 64 seeded meshes (2.88 M triangles) in one batched build,
  1500 seeded voxels of every grid bit for bit.
* The reference's own CPU-runnable case: the 0.1 m cube at 64^3, every
  voxel, and the demo's 7680 LCG points in Trilinear mode.
Plus size-independent properties: the dense field is finite with unit
normals, and a batch of queries equals the same points queried in pieces.
"""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import contact as SC
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ring(sdfo_gpu):
    v, t = scenes.ring_mesh(256, 64)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
    yield v, t, g
    g.close()


def test_ring_build_full_size_sampled(O, ring):
    v, t, g = ring
    nx, ny, nz = g.spec.shape
    assert (nx, ny, nz) == (128, 128, 64)
    print("[gpu] full-size ring build, sampled voxels", flush=True)
    got = g.download_dense()
    rng = np.random.default_rng(11)
    sample = rng.choice(nx * ny * nz, 40000, replace=False)
    ii, jj = np.meshgrid(np.arange(nx // 2 - 3, nx // 2 + 3), np.arange(ny // 2 - 3, ny // 2 + 3),
                         indexing="ij")
    axis = (ii.ravel()[:, None] + nx * (jj.ravel()[:, None] + ny * np.arange(nz)[None, :])).ravel()
    idx = np.unique(np.concatenate([sample, axis]))
    og = ograd(O, g.spec)
    c = og.centers()[idx]
    ref = O.signed_nearest(v, t, c)
    assert np.array_equal(got["phi"][idx], ref["sd"])
    assert np.array_equal(got["witness"][idx], ref["cp"])
    assert np.array_equal(got["tri"][idx], ref["tri"])
    # size-independent properties of the whole field
    assert np.isfinite(got["phi"]).all()
    nn = np.linalg.norm(got["n"], axis=1)
    assert np.abs(nn - 1.0).max() < 1e-12
    assert (got["phi"] < 0).sum() > 0 and (got["phi"] > 0).sum() > 0


@pytest.fixture(scope="module")
def cube_field(O, sdfo_gpu):
    v, t = scenes.cube_mesh()
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
    ref = g.download_dense()
    yield g, ref
    g.close()


def test_tricubic_1m_points_full_size(O, cube_field):
    g, ref = cube_field
    print("[gpu] full-size Tricubic, 1e6 points", flush=True)
    pts = scenes.valid_region_points(g.spec, 1_000_000, seed=1)
    P = torch.from_numpy(pts).cuda()
    gap, grad, valid = g.query(P, "Tricubic")
    d32 = dict(phi=f32(ref["phi"]), n=f32(ref["n"]), H=f32(ref["H"]))
    rq = O.query_dense(ograd(O, g.spec), d32, "Tricubic", pts.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"])
    assert vv.mean() > 0.99
    assert np.abs(gap.cpu().numpy()[vv] - rq["gap"][vv]).max() <= 1e-6 * g.spec.diagonal
    assert np.abs(grad.cpu().numpy()[vv] - rq["grad"][vv]).max() <= 2e-4
    # the same points in pieces give the same bits (no cross-point state)
    parts = [g.query(P[a:a + 250_000], "Tricubic")[0] for a in range(0, len(pts), 250_000)]
    assert torch.equal(torch.cat(parts), gap)


def test_tricubic_1m_fp64_points_full_size(O, cube_field):
    """The benchmark's 1e6 points as FP64 points: the exact path (FP64 cell
    records, made by this first FP64 query) returns the oracle's FP64
    Tricubic on the FP64 field bit for bit, gap, gradient and valid mask."""
    g, ref = cube_field
    print("[gpu] full-size FP64-point Tricubic, 1e6 points", flush=True)
    pts = scenes.valid_region_points(g.spec, 1_000_000, seed=1).astype(np.float64)
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), "Tricubic")
    rq = O.query_dense(ograd(O, g.spec), dict(phi=ref["phi"], n=ref["n"], H=ref["H"]),
                       "Tricubic", pts)
    assert np.array_equal(valid.cpu().numpy(), rq["valid"])
    assert np.array_equal(gap.cpu().numpy(), rq["gap"])
    assert np.array_equal(grad.cpu().numpy(), rq["grad"])


def test_contact_1000_poses_full_size(O, cube_field):
    g, ref = cube_field
    print("[gpu] full-size contact, 1000 poses x 7680 points", flush=True)
    X, w = scenes.ring_quadrature(120, 32)
    assert len(X) == 7680
    R, t = scenes.ring_poses(1000, seed=0)
    poses = scenes.pose_array(R, t)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.0, regularizer=1e-6, order=2)
    out, act = SC.contact_forward(g, torch.from_numpy(X).cuda(), torch.from_numpy(w).cuda(),
                                  torch.from_numpy(poses).cuda(), law, active=True)
    o_ref, a_ref = O.contact_forward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses,
                                     kn=1e11, c=1e4, mu=0.0, reg=1e-6, order=2)
    assert np.array_equal(act.cpu().numpy(), a_ref)
    assert (a_ref > 0).mean() > 0.5
    out = out.cpu().numpy()
    scale = np.abs(o_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out - o_ref) <= 1e-9 * scale)


def test_sparse_tube_full_size(O, sdfo_gpu):
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 1024, 64)
    print("[gpu] full-size sparse tube at 512^3", flush=True)
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=512, max_resolution=512,
                                      block_dim=8, band_voxels=4, halo_voxels=2)
    try:
        got = g.download_sparse()
        info = g.info
        bd = info.block_dim
        nbx, nby, nbz = info.block_grid[:]
        n = np.array(g.spec.shape)
        # every brick: |signed distance of the clamped brick centre| <= activation
        # (SdfBuilder.cpp:870-889), the distance from the oracle
        cells = np.arange(nbx * nby * nbz)
        b3 = np.stack([cells % nbx, (cells // nbx) % nby, cells // (nbx * nby)], axis=1)
        lo = b3 * bd
        hi = np.minimum(lo + bd, n[None, :])
        centre = g.spec.bmin[None, :] + g.spec.voxel[None, :] * (0.5 * (lo + hi).astype(np.float64))
        ref = O.signed_nearest(v, t, centre)
        want = (np.abs(ref["sd"]) <= info.activation_distance).astype(np.uint8)
        assert np.array_equal(got["active"], want)
        assert len(got["bijk"]) == int(want.sum()) > 30000
        # the voxels of 300 seeded active bricks, bit for bit
        rng = np.random.default_rng(5)
        pick = rng.choice(len(got["bijk"]), 300, replace=False)
        loc = np.arange(bd ** 3)
        off = np.stack([loc % bd, (loc // bd) % bd, loc // (bd * bd)], axis=1)
        ijk = (got["bijk"][pick][:, None, :] * bd + off[None, :, :]).reshape(-1, 3)
        inside = np.all(ijk < n[None, :], axis=1)
        c = g.spec.bmin[None, :] + g.spec.voxel[None, :] * (ijk[inside] + 0.5)
        sd = O.signed_nearest(v, t, c)["sd"]
        phi = got["phi"][pick].reshape(-1)
        assert np.array_equal(phi[inside], sd.astype(np.float32))
        assert (phi[~inside] == 0).all()
    finally:
        g.close()


def test_batched_gears_full_size_sampled(O, sdfo_gpu):
    """64 seeded gears (2k-200k triangles, grids 64-256 per axis) in one batched
    build: 1500 seeded voxels of every grid against the oracle, bit for bit."""
    batch = scenes.gear_batch(64, seed=5)
    print("[gpu] full-size batched gears", flush=True)
    grids = sdfo_gpu.SdfGrid.build_dense_batched([(v, t) for v, t, _ in batch],
                                                 [dict(target_resolution=r, max_resolution=r)
                                                  for _, _, r in batch])
    try:
        assert sum(len(t) for _, t, _ in batch) > 2_500_000
        rng = np.random.default_rng(13)
        for (v, t, r), g in zip(batch, grids):
            got = g.download_dense()
            idx = rng.choice(len(got["phi"]), 1500, replace=False)
            ref = O.signed_nearest(v, t, ograd(O, g.spec).centers()[idx])
            assert np.array_equal(got["phi"][idx], ref["sd"])
            assert np.array_equal(got["tri"][idx], ref["tri"])
    finally:
        for g in grids:
            g.close()


def test_fwd_bwd_256_poses_full_size(O, cube_field):
    """The differentiation workload: 256 seeded poses with twists x 7680 points,
    friction on; every gradient output against the oracle's FP64 backward."""
    g, ref = cube_field
    print("[gpu] full-size fwd+bwd, 256 poses", flush=True)
    X, w = scenes.ring_quadrature(120, 32)
    R, t = scenes.ring_poses(256, seed=2)
    v0, Om = scenes.pose_velocities(256, seed=3)
    poses = scenes.pose_array(R, t, v0=v0, Om=Om)
    gout = np.random.default_rng(17).normal(size=(256, 6))
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    Xd, wd, Pd = (torch.from_numpy(a).cuda() for a in (X, w, poses))
    out, _ = SC.contact_forward(g, Xd, wd, Pd, law, active=True)
    o_ref, _ = O.contact_forward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses,
                                 kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2)
    scale = np.abs(o_ref).max(axis=0) + 1e-300
    assert np.all(np.abs(out.cpu().numpy() - o_ref) <= 1e-9 * scale)
    grads = SC.contact_backward(g, Xd, wd, Pd, torch.from_numpy(gout).cuda(), law)
    refs = O.contact_backward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses, gout,
                              kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2)
    for got, exp in zip(grads, refs):
        got = got.cpu().numpy()
        assert np.abs(got - exp).max() <= 1e-7 * (np.abs(exp).max() + 1e-300)
    # the training-step form: the forward records its pairs for the backward
    out2, rec = SC.contact_forward(g, Xd, wd, Pd, law, save=True)
    assert torch.equal(out2, out)
    grads2 = SC.contact_backward(g, Xd, wd, Pd, torch.from_numpy(gout).cuda(), law, saved=rec)
    for got, exp in zip(grads2, refs):
        got = got.cpu().numpy()
        assert np.abs(got - exp).max() <= 1e-7 * (np.abs(exp).max() + 1e-300)


def test_fwd_bwd_1200_poses_in_pose_groups(O, cube_field):
    """Past 944 poses x 7680 points the backward's per-(pose, point) buffers
    are split into pose groups (separate finalize, parameter sum over the
    groups, point gradients accumulated group by group): every output of the
    screening and of the recorded backward against the oracle."""
    g, ref = cube_field
    print("[gpu] fwd+bwd, 1200 poses in pose groups", flush=True)
    X, w = scenes.ring_quadrature(120, 32)
    R, t = scenes.ring_poses(1200, seed=12)
    v0, Om = scenes.pose_velocities(1200, seed=13)
    poses = scenes.pose_array(R, t, v0=v0, Om=Om)
    gout = np.random.default_rng(18).normal(size=(1200, 6))
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    Xd, wd, Pd = (torch.from_numpy(a).cuda() for a in (X, w, poses))
    G = torch.from_numpy(gout).cuda()
    refs = O.contact_backward(ograd(O, g.spec), f32(ref["phi"]), f32(ref["n"]), X, w, poses, gout,
                              kn=1e11, c=1e4, mu=0.3, reg=1e-6, order=2)
    _, rec = SC.contact_forward(g, Xd, wd, Pd, law, save=True)
    for grads in (SC.contact_backward(g, Xd, wd, Pd, G, law),
                  SC.contact_backward(g, Xd, wd, Pd, G, law, saved=rec)):
        for got, exp in zip(grads, refs):
            got = got.cpu().numpy()
            assert np.abs(got - exp).max() <= 1e-7 * (np.abs(exp).max() + 1e-300)


def test_fwd_bwd_without_contact_is_zero(cube_field):
    """Poses 50 mm above the cube: no pair is in contact, every output is 0."""
    g, _ = cube_field
    X, w = scenes.ring_quadrature(120, 32)
    R, t = scenes.ring_poses(64, seed=4, z_range=(0.11, 0.12))
    poses = scenes.pose_array(R, t)
    law = SC.ContactLaw(kn=1e11, damping=1e4, mu=0.3, regularizer=1e-6, order=2)
    Xd, wd, Pd = (torch.from_numpy(a).cuda() for a in (X, w, poses))
    out, act, rec = SC.contact_forward(g, Xd, wd, Pd, law, active=True, save=True)
    assert int(act.sum()) == 0 and float(out.abs().max()) == 0.0
    G = torch.ones((64, 6), dtype=torch.float64, device="cuda")
    for grads in (SC.contact_backward(g, Xd, wd, Pd, G, law),
                  SC.contact_backward(g, Xd, wd, Pd, G, law, saved=rec)):
        for t_ in grads:
            assert float(t_.abs().max()) == 0.0


def test_cube64_demo_trilinear(O, sdfo_gpu):
    """The reference's CPU-runnable case: the 0.1 m cube (12 triangles) at 64^3
    with 10% padding, every voxel bit for bit, and 7680 points from the demo's
    LCG (SdfOracleDemo.cpp:202-274), Trilinear value + gradient: FP64 points
    read the FP64 field, so they equal the oracle's (the reference's own
    numbers) bit for bit."""
    v, t = scenes.cube_mesh()
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=64, max_resolution=64)
    try:
        assert tuple(g.spec.shape) == (64, 64, 64)
        og = ograd(O, g.spec)
        ref = O.build_dense(v, t, og)
        got = g.download_dense()
        for k in ("phi", "n", "witness", "H"):
            assert np.array_equal(got[k], ref[k]), k
        pts = scenes.demo_benchmark_points(g.spec, 7680)
        gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), "Trilinear")
        vv = valid.cpu().numpy()
        assert vv.all()
        r64 = O.query_dense(og, ref, "Trilinear", pts)
        assert np.array_equal(gap.cpu().numpy(), r64["gap"])
        assert np.array_equal(grad.cpu().numpy(), r64["grad"])
        # the configuration's own sampling: 7680 seeded points uniform over the
        # padded box [bmin, bmax], about (63/64)^3 = 95.4% of them interpolable
        ub = scenes.box_points(g.spec, 7680, seed=11)
        gap, grad, valid = g.query(torch.from_numpy(ub).cuda(), "Trilinear")
        rb = O.query_dense(og, ref, "Trilinear", ub)
        vv = valid.cpu().numpy()
        assert np.array_equal(vv, rb["valid"]) and 0.93 < vv.mean() < 0.98
        assert np.array_equal(gap.cpu().numpy(), rb["gap"])
        assert np.array_equal(grad.cpu().numpy(), rb["grad"])
    finally:
        g.close()
