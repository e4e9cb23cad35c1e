"""GPU query parity for the five modes (SdfOracle.cpp:176-1092), dense and sparse.

The HIP query reads the FP32 field mirror and computes in FP32 (FP64 for the
cell location); the oracle is run on the same FP32-rounded fields in FP64.
Bar: valid mask identical; gap within 1e-6 * bbox diagonal (the accuracy
rule for field values); raw gradient within 2e-4 absolute (an FP32
interpolant of unit normals).
"""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu

MODES = ["FirstOrder", "SecondOrder", "Trilinear", "Tricubic", "ContactAware"]


def _points(spec, n, seed):
    p = scenes.box_points(spec, n, seed=seed)
    # a few far outside, one on the lower corner, one exactly on a voxel centre
    extra = np.array([spec.bmin - 1.0, spec.bmax + 1.0, spec.bmin,
                      spec.bmin + spec.voxel * np.array([3.5, 4.5, 5.5])])
    return np.concatenate([p, extra]).astype(np.float32)


@pytest.fixture(scope="module")
def torus_grid(O, sdfo_gpu):
    v, t = meshes.torus(0.05, 0.01, 64, 16)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=48, max_resolution=48)
    ref = g.download_dense()
    yield g, ref
    g.close()


@pytest.mark.parametrize("mode", MODES)
def test_dense_query_modes(O, torus_grid, mode):
    g, ref = torus_grid
    print(f"[gpu] dense query {mode}", flush=True)
    og = ograd(O, g.spec)
    pts = _points(g.spec, 20000, seed=11)
    gap, grad, valid, normal, wit, _ = g.query(torch.from_numpy(pts).cuda(), mode, full=True)
    rq = O.query_dense(og, dict(phi=f32(ref["phi"]), n=f32(ref["n"]), H=f32(ref["H"])), mode,
                       pts.astype(np.float64))
    v = valid.cpu().numpy()
    assert np.array_equal(v, rq["valid"])
    assert 0.5 < v.mean() < 1.0
    diag = g.spec.diagonal
    gg = gap.cpu().numpy().astype(np.float64)
    assert np.abs(gg - rq["gap"]).max() <= 1e-6 * diag
    assert np.abs(grad.cpu().numpy() - rq["grad"]).max() <= 2e-4
    nm = normal.cpu().numpy()
    assert np.abs(nm[v] - rq["normal"][v]).max() <= 2e-4
    assert np.abs(wit.cpu().numpy() - rq["witness"]).max() <= 2e-6 * diag
    assert np.all(gg[~v] == 0.0)


@pytest.mark.parametrize("mode", MODES)
def test_dense_query_f64_points(O, torus_grid, mode):
    """FP64 points on a dense grid read its FP64 fields in FP64 arithmetic
    without FMA contraction: every output is the oracle's on the same FP64
    SdfData, bit for bit."""
    g, ref = torus_grid
    print(f"[gpu] dense FP64-point query {mode}", flush=True)
    og = ograd(O, g.spec)
    pts = _points(g.spec, 5000, seed=12).astype(np.float64)
    pts[:4000] = scenes.box_points(g.spec, 4000, seed=13)  # off the FP32 lattice
    gap, grad, valid, normal, wit, br = g.query(torch.from_numpy(pts).cuda(), mode, full=True)
    rq = O.query_dense(og, ref, mode, pts)
    assert np.array_equal(valid.cpu().numpy(), rq["valid"])
    assert np.array_equal(gap.cpu().numpy(), rq["gap"])
    assert np.array_equal(grad.cpu().numpy(), rq["grad"])
    assert np.array_equal(normal.cpu().numpy(), rq["normal"])
    assert np.array_equal(wit.cpu().numpy(), rq["witness"])
    assert np.array_equal(br.cpu().numpy(), rq["branch"])


def test_query_exact_on_separable_cubic(O, sdfo_gpu):
    """Tricubic reproduces phi = f(x)+g(y)+h(z) (cubics) with exact gradient
    corner data: closed form, no oracle involved."""
    spec = sdfo_gpu.setup_grid(meshes.cube(1.0)[0], target_resolution=16, max_resolution=16)
    nx, ny, nz = spec.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = spec.bmin + spec.voxel * (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5)
    f = lambda x, y, z: 0.3 * x ** 3 - 0.2 * x + 0.5 * y ** 2 - 0.1 * z ** 3 + 0.05  # noqa: E731
    df = lambda x, y, z: np.stack([0.9 * x ** 2 - 0.2, y, -0.3 * z ** 2], -1)  # noqa: E731
    phi = f(c[:, 0], c[:, 1], c[:, 2])
    n = df(c[:, 0], c[:, 1], c[:, 2])
    g = sdfo_gpu.SdfGrid.from_dense(spec, phi, n)
    p = scenes.valid_region_points(spec, 4000, seed=2, dtype=np.float64)
    gap, grad, valid = g.query(torch.from_numpy(p).cuda(), "Tricubic")
    assert valid.all()
    assert np.abs(gap.cpu().numpy() - f(p[:, 0], p[:, 1], p[:, 2])).max() < 1e-6
    assert np.abs(grad.cpu().numpy() - df(p[:, 0], p[:, 1], p[:, 2])).max() < 1e-5


# brick sizes 8 (cell-record validity bits one 64-bit word per wave) and 3
# (27 voxels per brick: bits or-ed into shared words)
@pytest.fixture(scope="module", params=[8, 3], ids=["b8", "b3"])
def sparse_grid(O, sdfo_gpu, request):
    v, t = meshes.hollow_cylinder(0.04, 0.03, 0.05, 64, 8)
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=64, max_resolution=64,
                                      block_dim=request.param, band_voxels=4, halo_voxels=2)
    sp = g.download_sparse()
    yield g, sp, v, t
    g.close()


@pytest.mark.parametrize("mode", MODES)
def test_sparse_query_modes(O, sparse_grid, mode):
    g, sp, v, t = sparse_grid
    print(f"[gpu] sparse query {mode}", flush=True)
    og = ograd(O, g.spec)
    pts = np.concatenate([
        scenes.near_surface_points(v, t, 15000, 2 * g.spec.voxel[0], seed=5),
        _points(g.spec, 3000, seed=6)]).astype(np.float32)
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), mode)
    rq = O.query_sparse(og, sp, mode, pts.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"])
    assert 0.3 < vv.mean() < 1.0
    diag = g.spec.diagonal
    assert np.abs(gap.cpu().numpy() - rq["gap"]).max() <= 1e-6 * diag
    assert np.abs(grad.cpu().numpy() - rq["grad"]).max() <= 2e-4


@pytest.mark.parametrize("mode", MODES)
def test_sparse_query_f64_points(O, sparse_grid, mode):
    """FP64 points on a sparse grid: FP64 arithmetic on the FP32 brick fields
    widened, as the CPU path does, so every output is the oracle's bit for bit."""
    g, sp, v, t = sparse_grid
    print(f"[gpu] sparse FP64-point query {mode}", flush=True)
    pts = np.concatenate([scenes.near_surface_points(v, t, 6000, 2 * g.spec.voxel[0], seed=8),
                          _points(g.spec, 2000, seed=9)]).astype(np.float64)
    pts[:3000] += 1e-9 * np.random.default_rng(10).standard_normal((3000, 3))  # off the FP32 lattice
    out = g.query(torch.from_numpy(pts).cuda(), mode, full=True)
    rq = O.query_sparse(ograd(O, g.spec), sp, mode, pts)
    assert 0.3 < rq["valid"].mean() < 1.0
    for k, a in zip(("gap", "grad", "valid", "normal", "witness", "branch"), out):
        assert np.array_equal(a.cpu().numpy(), rq[k]), k


def test_query_batch_contract(torus_grid):
    g, _ = torus_grid
    pts = _points(g.spec, 1000, seed=3)
    gap, grad, gn, unit, valid = g.query_batch(torch.from_numpy(pts).cuda(), "Trilinear")
    v = valid.cpu().numpy()
    assert np.all(gap.cpu().numpy()[~v] == 0)
    assert np.all(unit.cpu().numpy()[~v] == [0, 0, 1])
    assert np.allclose(gn.cpu().numpy(), np.linalg.norm(grad.cpu().numpy(), axis=1))
