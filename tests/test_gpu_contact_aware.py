"""GPU contact-aware build (SdfBuilder.cpp:416-821) against the oracle, bit for bit,
and ContactAware queries (SdfOracle.cpp:226-432) on the result."""
import numpy as np
import pytest
import torch

from conftest import f32, ograd
from sdfo import meshes, scenes

pytestmark = pytest.mark.gpu


def face_corner_normals(v, t):
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n[:, None, :], 3, axis=1)


def vertex_corner_normals(v, t):
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    vn = np.zeros_like(v)
    for c in range(3):
        np.add.at(vn, t[:, c], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    return vn[t]


CASES = [
    ("cube_faces_24", lambda: meshes.cube(0.1), face_corner_normals, 24, {}),
    ("torus_smooth_32", lambda: meshes.torus(0.05, 0.01, 48, 16), vertex_corner_normals, 32, {}),
    ("gear_faces_32", lambda: meshes.gear(12, 6, 0.02, 0.002, 3), face_corner_normals, 32,
     dict(max_candidates=6)),
    ("cube_no_tiles_16", lambda: meshes.cube(0.1), face_corner_normals, 16,
     dict(use_tile_acceleration=0, max_candidates=3)),
    ("sphere_tol_16", lambda: meshes.uv_sphere(0.05, 8, 16), vertex_corner_normals, 16,
     dict(feature_distance_tolerance=0.004, normal_merge_angle_degrees=20.0, tile_dim=4)),
    # a wide tolerance and a narrow merge angle leave more merged candidates
    # than the first pass keeps in registers: those voxels take the second pass
    ("torus_wide_tol_16", lambda: meshes.torus(0.05, 0.01, 48, 16), vertex_corner_normals, 16,
     dict(feature_distance_tolerance=0.02, normal_merge_angle_degrees=2.0, max_candidates=8)),
]


def _oracle_kw(p):
    return dict(max_candidates=p.get("max_candidates", 4),
                feature_tol=p.get("feature_distance_tolerance", 0.0),
                merge_deg=p.get("normal_merge_angle_degrees", 5.0), tile_dim=p.get("tile_dim", 8),
                use_tiles=bool(p.get("use_tile_acceleration", 1)))


@pytest.mark.parametrize("name,mk,nf,res,params", CASES, ids=[c[0] for c in CASES])
def test_dense_contact_aware_bit_exact(O, sdfo_gpu, name, mk, nf, res, params):
    v, t = mk()
    cn = nf(v, t)
    print(f"[gpu] contact-aware dense {name}", flush=True)
    g = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, target_resolution=res, max_resolution=res,
                                             **params)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    ref = O.build_contact_aware_dense(v, t, cn, og, **_oracle_kw(params))
    got = g.download_dense()
    c = got["candidates"]
    S = ref["stride"]
    assert c["stride"] == S
    assert np.array_equal(c["count"], ref["ccount"])
    assert np.array_equal(c["phi"], ref["cphi"][:, :S])
    assert np.array_equal(c["n"], ref["cn"][:, :S])
    assert np.array_equal(c["branch"], ref["cbranch"][:, :S])
    for k in ("phi", "n", "witness", "H"):
        assert np.array_equal(got[k], ref[k]), k
    assert (ref["ccount"] > 1).any()

    # ContactAware queries on the built grid
    pts = scenes.box_points(g.spec, 8000, seed=4, dtype=np.float32)
    gap, grad, valid, normal, wit, branch = g.query(torch.from_numpy(pts).cuda(), "ContactAware",
                                                    full=True)
    d32 = dict(phi=f32(ref["phi"]), n=f32(ref["n"]), H=f32(ref["H"]), stride=S,
               ccount=ref["ccount"], cphi=f32(ref["cphi"]), cn=f32(ref["cn"]),
               cbranch=ref["cbranch"])
    rq = O.query_dense(ograd(O, g.spec), d32, "ContactAware", pts.astype(np.float64))
    vv = valid.cpu().numpy()
    assert np.array_equal(vv, rq["valid"])
    assert np.abs(gap.cpu().numpy() - rq["gap"]).max() <= 1e-6 * g.spec.diagonal
    bb = branch.cpu().numpy()
    # branch ids are exact unless two candidates tie on |gap| within FP32 rounding
    assert (bb[vv] == rq["branch"][vv]).mean() > 0.999

    # FP64 points read the FP64 candidates: the oracle's answer bit for bit
    p64 = scenes.box_points(g.spec, 8000, seed=5)
    out = g.query(torch.from_numpy(p64).cuda(), "ContactAware", full=True)
    r64 = O.query_dense(ograd(O, g.spec), ref, "ContactAware", p64)
    for k, a in zip(("gap", "grad", "valid", "normal", "witness", "branch"), out):
        assert np.array_equal(a.cpu().numpy(), r64[k]), k


SPARSE = [
    ("cube_faces_32", lambda: meshes.cube(0.1), face_corner_normals, 32, dict(band_voxels=2)),
    ("torus_smooth_48", lambda: meshes.torus(0.05, 0.01, 48, 16), vertex_corner_normals, 48,
     dict(band_voxels=2, halo_voxels=1)),
    ("cube_no_tiles_32", lambda: meshes.cube(0.1), face_corner_normals, 32,
     dict(band_voxels=2, use_tile_acceleration=0)),
    ("torus_wide_tol_32", lambda: meshes.torus(0.05, 0.01, 48, 16), vertex_corner_normals, 32,
     dict(band_voxels=2, feature_distance_tolerance=0.02, normal_merge_angle_degrees=2.0,
          max_candidates=8)),
]


@pytest.mark.parametrize("name,mk,nf,res,params", SPARSE, ids=[c[0] for c in SPARSE])
def test_sparse_contact_aware_bit_exact(O, sdfo_gpu, name, mk, nf, res, params):
    v, t = mk()
    cn = nf(v, t)
    print(f"[gpu] contact-aware sparse {name}", flush=True)
    g = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, sparse=True, target_resolution=res,
                                             max_resolution=res, **params)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    kw = _oracle_kw(params)
    ref = O.build_contact_aware_sparse(v, t, cn, og, block_dim=8, halo=params.get("halo_voxels", 2),
                                       band_voxels=params["band_voxels"], **kw)
    got = g.download_sparse()
    assert np.array_equal(got["active"], ref["active"])
    assert np.array_equal(got["bijk"], ref["bijk"])
    for k in ("phi", "n", "H", "ccount", "cphi", "cn", "cbranch"):
        assert np.array_equal(got[k], ref[k]), k
    pts = np.concatenate([scenes.near_surface_points(v, t, 6000, g.spec.voxel[0], seed=6),
                          scenes.box_points(g.spec, 2000, seed=7, dtype=np.float32)])
    gap, grad, valid = g.query(torch.from_numpy(pts).cuda(), "ContactAware")
    rq = O.query_sparse(ograd(O, g.spec), ref, "ContactAware", pts.astype(np.float64))
    assert np.array_equal(valid.cpu().numpy(), rq["valid"])
    assert np.abs(gap.cpu().numpy() - rq["gap"]).max() <= 1e-6 * g.spec.diagonal
    # FP64 points: the FP32 candidates widened, as the CPU path does, bit for bit
    p64 = pts.astype(np.float64) + 1e-9
    out = g.query(torch.from_numpy(p64).cuda(), "ContactAware", full=True)
    r64 = O.query_sparse(ograd(O, g.spec), ref, "ContactAware", p64)
    for k, a in zip(("gap", "grad", "valid", "normal", "witness", "branch"), out):
        assert np.array_equal(a.cpu().numpy(), r64[k]), k


@pytest.mark.parametrize("cap", [(1, 0), (8, 0), (1, 2)], ids=["1-64", "8-64", "1-2"])
@pytest.mark.parametrize("idx", [1, 2, 5])
def test_contact_aware_merge_capacity(O, sdfo_gpu, cap, idx):
    """Other capacities of the passes that hold merged candidates (first pass in
    registers, second pass in device memory, then room for the longest list):
    voxels move between the passes, results stay exact."""
    from sdfo import _lib

    name, mk, nf, res, params = CASES[idx]
    v, t = mk()
    cn = nf(v, t)
    print(f"[gpu] contact-aware merge capacity {cap} {name}", flush=True)
    _lib.debug_ca_merge_cap(*cap)
    try:
        g = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, target_resolution=res,
                                                 max_resolution=res, **params)
        sp = dict(band_voxels=2, **params)
        gs = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, sparse=True, target_resolution=res,
                                                  max_resolution=res, **sp)
    finally:
        _lib.debug_ca_merge_cap(0, 0)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    ref = O.build_contact_aware_dense(v, t, cn, og, **_oracle_kw(params))
    got = g.download_dense()
    S = ref["stride"]
    assert np.array_equal(got["candidates"]["count"], ref["ccount"])
    assert np.array_equal(got["candidates"]["phi"], ref["cphi"][:, :S])
    assert np.array_equal(got["candidates"]["n"], ref["cn"][:, :S])
    assert np.array_equal(got["candidates"]["branch"], ref["cbranch"][:, :S])
    for k in ("phi", "n", "witness"):
        assert np.array_equal(got[k], ref[k]), k
    refs = O.build_contact_aware_sparse(v, t, cn, og, block_dim=8, halo=2, band_voxels=2,
                                        **_oracle_kw(params))
    gots = gs.download_sparse()
    for k in ("bijk", "phi", "n", "ccount", "cphi", "cn", "cbranch"):
        assert np.array_equal(gots[k], refs[k]), k


def test_contact_aware_errors(sdfo_gpu):
    v, t = meshes.cube(0.1)
    cn = face_corner_normals(v, t)
    with pytest.raises(sdfo_gpu.SdfoError, match="maxCandidates must be positive"):
        sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, max_candidates=0)
    with pytest.raises(ValueError, match="one corner-normal triplet per triangle"):
        sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn[:5])
    with pytest.raises(sdfo_gpu.SdfoError, match="Empty mesh"):
        sdfo_gpu.SdfGrid.build_contact_aware(np.zeros((0, 3)), np.zeros((0, 3), np.int32),
                                             np.zeros((0, 3, 3)))


def test_contact_aware_round_trip(sdfo_gpu, tmp_path):
    import sdfo_format as F

    v, t = meshes.cube(0.1)
    cn = face_corner_normals(v, t)
    g = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, target_resolution=16, max_resolution=16)
    d = g.download_dense()
    p = str(tmp_path / "ca.sdf")
    g.save(p)
    s = g.spec
    want = F.dense_bytes(s.shape, s.bmin, s.bmax, s.voxel, d["phi"], d["n"], d["H"], d["witness"],
                         d["candidates"])
    assert open(p, "rb").read() == want
    h = sdfo_gpu.SdfGrid.load(p)
    e = h.download_dense()
    for k in ("count", "phi", "n", "branch"):
        assert np.array_equal(d["candidates"][k], e["candidates"][k])
    gs = sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, sparse=True, target_resolution=16,
                                              max_resolution=16, band_voxels=2)
    ps = str(tmp_path / "cas.sdf")
    gs.save(ps)
    hs = sdfo_gpu.SdfGrid.load(ps)
    a, b = gs.download_sparse(), hs.download_sparse()
    for k in ("bijk", "phi", "n", "H", "ccount", "cphi", "cn", "cbranch"):
        assert np.array_equal(a[k], b[k]), k
