"""Mesh distance queries (sdf::KdTree, KdTree.h:86-170) on the GPU against the
oracle's restatement of the same search, bit for bit."""
import numpy as np
import pytest
import torch

from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu

MESHES = [
    ("cube", lambda: meshes.cube(0.1, (0.01, 0.0, -0.02))),
    ("torus", lambda: meshes.torus(0.05, 0.01, 48, 16)),
    ("gear", lambda: meshes.gear(12, 6, 0.02, 0.002, 3)),
    ("tube", lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 64, 8)),
]


def _points(v, t, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = v.min(0), v.max(0)
    span = hi - lo
    box = rng.uniform(lo - 0.3 * span, hi + 0.3 * span, (n, 3))
    near = scenes.near_surface_points(v, t, n, 0.02 * span.max(), seed=seed, dtype=np.float64)
    on = v[rng.integers(0, len(v), 64)]  # exactly on vertices: zero distance
    return np.concatenate([box, near, on])


@pytest.mark.parametrize("name,mk", MESHES, ids=[m[0] for m in MESHES])
def test_mesh_queries_match_oracle(O, sdfo_gpu, name, mk):
    v, t = mk()
    print(f"[gpu] KdTree queries {name}", flush=True)
    tree = sdfo_gpu.KdTree(v, t)
    assert tree.triangle_count() == len(t)
    p = _points(v, t, 3000, seed=3)
    ref = O.signed_nearest(v, t, p)
    pd = torch.from_numpy(p).cuda()
    sd, cp, nrm, tri = (x.cpu().numpy() for x in tree.find_signed_nearest(pd))
    assert np.array_equal(tri, ref["tri"])
    assert np.array_equal(sd, ref["sd"])
    assert np.array_equal(cp, ref["cp"])
    assert np.array_equal(nrm, ref["face"])
    d, cp2, _, tri2 = (x.cpu().numpy() for x in tree.find_nearest(pd))
    assert np.array_equal(d, np.abs(ref["sd"]))
    assert np.array_equal(cp2, ref["cp"]) and np.array_equal(tri2, ref["tri"])
    assert (sd < 0).any() and (sd > 0).any()
    tree.close()


def test_empty_mesh_and_errors(sdfo_gpu):
    tree = sdfo_gpu.KdTree(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert tree.triangle_count() == 0
    d, cp, nrm, tri = (x.cpu().numpy() for x in tree.find_nearest(torch.ones(5, 3, dtype=torch.float64).cuda()))
    assert (tri == -1).all() and (d == np.finfo(np.float64).max).all()
    assert (nrm == [0, 0, 1]).all() and (cp == 0).all()
    with pytest.raises(sdfo_gpu.SdfoError, match="invalid vertex"):
        sdfo_gpu.KdTree(np.zeros((3, 3)), np.array([[0, 1, 5]], np.int32))
    with pytest.raises(ValueError):
        tree.find_nearest(torch.ones(5, 3))
