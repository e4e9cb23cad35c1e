"""SDFO 0.1.0 files written and read by the library (SdfIO.cpp:93-523)."""
import numpy as np
import pytest
import torch

import sdfo_format as F
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu


def test_dense_save_is_byte_identical(sdfo_gpu, tmp_path):
    v, t = meshes.torus(0.05, 0.01, 32, 12)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=24, max_resolution=24)
    d = g.download_dense()
    path = str(tmp_path / "t.sdf")
    g.save(path)
    got = open(path, "rb").read()
    s = g.spec
    want = F.dense_bytes(s.shape, s.bmin, s.bmax, s.voxel, d["phi"], d["n"], d["H"], d["witness"])
    assert got == want
    # the prototype reads phi as (nz, ny, nx) then transposes to [i, j, k] (sdf_grid.py:58-64)
    phi = F.python_reader_phi(got)
    nx, ny, nz = s.shape
    i, j, k = 3, 5, 7
    assert phi[i, j, k] == d["phi"][i + nx * (j + ny * k)]


def test_dense_round_trip(sdfo_gpu, tmp_path):
    v, t = meshes.gear(12, 6, 0.02, 0.002, 3)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=32, max_resolution=32)
    path = str(tmp_path / "g.sdf")
    g.save(path)
    h = sdfo_gpu.SdfGrid.load(path)
    a, b = g.download_dense(), h.download_dense()
    for k in ("phi", "n", "H", "witness"):
        assert np.array_equal(a[k], b[k]), k
    assert np.all(b["tri"] == -1)  # the file holds no triangle index
    pts = torch.from_numpy(scenes.box_points(g.spec, 5000, seed=1, dtype=np.float32)).cuda()
    for mode in ("Tricubic", "SecondOrder"):
        ra, rb = g.query(pts, mode), h.query(pts, mode)
        assert all(torch.equal(x, y) for x, y in zip(ra, rb))
    path2 = str(tmp_path / "g2.sdf")
    h.save(path2)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_sparse_save_and_round_trip(sdfo_gpu, tmp_path):
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 48, 6)
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=48, max_resolution=48, band_voxels=3)
    sp = g.download_sparse()
    path = str(tmp_path / "s.sdf")
    g.save(path)
    got = open(path, "rb").read()
    s = g.spec
    want = F.sparse_bytes(s.shape, s.bmin, s.bmax, s.voxel, 8, 2, g.info.band_distance,
                          sp["bijk"], sp["phi"], sp["n"], sp["H"])
    assert got == want
    assert F.parse_header(got)["flags"] == F.FLAG_SPARSE | F.FLAG_HESSIAN
    h = sdfo_gpu.SdfGrid.load(path)
    assert h.is_sparse and h.info.block_count == g.info.block_count
    sb = h.download_sparse()
    for k in ("bijk", "phi", "n", "H"):
        assert np.array_equal(sp[k], sb[k]), k
    pts = torch.from_numpy(scenes.near_surface_points(v, t, 4000, 0.001, seed=2)).cuda()
    for mode in ("Tricubic", "Trilinear", "SecondOrder"):
        ra, rb = g.query(pts, mode), h.query(pts, mode)
        assert all(torch.equal(x, y) for x, y in zip(ra, rb))


def test_sparse_without_hessian(sdfo_gpu, tmp_path):
    v, t = meshes.uv_sphere(0.05, 10, 20)
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=32, max_resolution=32,
                                      store_hessian=0, band_voxels=2)
    path = str(tmp_path / "nh.sdf")
    g.save(path)
    b = open(path, "rb").read()
    assert F.parse_header(b)["flags"] == F.FLAG_SPARSE
    h = sdfo_gpu.SdfGrid.load(path)
    assert not h.info.has_hessian
    pts = torch.from_numpy(scenes.near_surface_points(v, t, 2000, 0.0005, seed=3)).cuda()
    # second order degrades to first order without a Hessian (SdfOracle.cpp:626-629)
    a = h.query(pts, "SecondOrder")
    b1 = h.query(pts, "FirstOrder")
    assert all(torch.equal(x, y) for x, y in zip(a, b1))
