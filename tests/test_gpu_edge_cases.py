"""Edge cases of the query and contact paths on the GPU: empty inputs, points
that are NaN or infinite, points on the grid's faces and corners, and a
single point; each compared with the oracle (SdfOracle.cpp:543-549: queries
never throw, invalid points return valid = false)."""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import contact as SC
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu

MODES = ["FirstOrder", "SecondOrder", "Trilinear", "Tricubic", "ContactAware"]


@pytest.fixture(scope="module")
def grids(O, sdfo_gpu):
    v, t = meshes.torus(0.05, 0.01, 48, 16)
    d = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=32, max_resolution=32)
    s = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=32, max_resolution=32, band_voxels=3)
    yield d, s
    d.close()
    s.close()


def _odd_points(spec):
    lo, hi, h = spec.bmin, spec.bmax, spec.voxel
    nan, inf = np.nan, np.inf
    pts = [
        [nan, 0.0, 0.0], [0.0, nan, 0.0], [0.0, 0.0, nan],
        [inf, 0.0, 0.0], [-inf, 0.0, 0.0], [0.0, inf, -inf],
        lo, hi, lo + 0.5 * h, hi - 0.5 * h,  # faces, corners, first / last voxel centres
        [lo[0], 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])],
        [1e30, 1e30, 1e30], [-1e30, 0.0, 0.0],
    ]
    return np.array(pts, np.float64).astype(np.float32)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_empty_and_non_finite_queries(O, grids, kind, mode):
    d, s = grids
    g = d if kind == "dense" else s
    print(f"[gpu] edge-case query {kind} {mode}", flush=True)
    gap, grad, valid = g.query(torch.zeros((0, 3), dtype=torch.float32, device="cuda"), mode)
    assert gap.shape == (0,) and grad.shape == (0, 3) and valid.shape == (0,)
    pts = _odd_points(g.spec)
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), mode)
    og = ograd(O, g.spec)
    if kind == "dense":
        ref = g.download_dense()
        rq = O.query_dense(og, dict(phi=f32(ref["phi"]), n=f32(ref["n"]), H=f32(ref["H"])), mode,
                           pts.astype(np.float64))
    else:
        rq = O.query_sparse(og, g.download_sparse(), mode, pts.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"])
    assert not vv[:6].any()  # NaN / infinite points are never valid
    gg = gap.cpu().numpy()
    assert np.isfinite(gg).all()
    assert np.abs(gg[vv] - rq["gap"][vv]).max(initial=0.0) <= 1e-6 * g.spec.diagonal
    assert (gg[~vv] == 0).all()
    # one point alone gives the bits it gives in a batch
    one = g.query(torch.from_numpy(pts[7:8]).cuda(), mode)[0]
    assert torch.equal(one, gap[7:8])


def test_contact_empty_inputs(sdfo_gpu, grids):
    d, _ = grids
    X, w = scenes.ring_quadrature(8, 4)
    R, t = scenes.ring_poses(3, seed=1)
    poses = torch.from_numpy(scenes.pose_array(R, t)).cuda()
    law = SC.ContactLaw()
    Xd, wd = torch.from_numpy(X).cuda(), torch.from_numpy(w).cuda()
    out = SC.contact_forward(d, Xd, wd, poses[:0], law)
    assert out.shape == (0, 6)
    out = SC.contact_forward(d, Xd[:0].contiguous(), wd[:0].contiguous(), poses, law)
    assert out.shape == (3, 6) and (out == 0).all()
