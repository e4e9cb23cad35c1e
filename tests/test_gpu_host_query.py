"""The host-pointer FP64 query (the C++ SdfOracle::query / queryBatch path).

Calls up to 8 MiB of arrays run the kernels on pinned, device-mapped host
memory; larger calls copy through a device block.  Both must return what the
device-pointer FP64 query returns, bit for bit (that query equals the
oracle bit for bit in the other GPU tests), on dense and sparse
contact-aware grids in all five modes, and the candidate lists of the mapped
path must equal the device-block path's for the same points.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from sdfo import _lib, scenes

pytestmark = pytest.mark.gpu

MODES = ("FirstOrder", "SecondOrder", "Trilinear", "Tricubic", "ContactAware")
MODE_ID = {m: i for i, m in enumerate(MODES)}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _host_query(g, mode, pts, cands=False):
    n = len(pts)
    out = dict(gap=np.empty(n), normal=np.empty((n, 3)), witness=np.empty((n, 3)),
               valid=np.empty(n, np.uint8), branch=np.empty(n, np.int32))
    cc = [None] * 5
    if cands:
        out.update(ccount=np.empty(n, np.uint8), cgap=np.empty((n, 8)), cn=np.empty((n, 8, 3)),
                   cw=np.empty((n, 8, 3)), cb=np.empty((n, 8), np.int32))
        cc = [_ptr(out[k]) for k in ("ccount", "cgap", "cn", "cw", "cb")]
    p = np.ascontiguousarray(pts, np.float64)
    _lib.check(_lib.lib.sdfo_query_host_f64(g.handle, MODE_ID[mode], _ptr(p), n, _ptr(out["gap"]),
                                           _ptr(out["normal"]), _ptr(out["witness"]),
                                           _ptr(out["valid"]), _ptr(out["branch"]), *cc, None))
    return out


@pytest.fixture(scope="module")
def grids(sdfo_gpu):
    v, t = scenes.ring_mesh(64, 16)
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    cn = np.repeat(fn[:, None, :], 3, axis=1)
    prm = dict(target_resolution=64, max_resolution=64)
    out = {"dense": sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, **prm),
           "sparse": sdfo_gpu.SdfGrid.build_contact_aware(v, t, cn, sparse=True, block_dim=8,
                                                           band_voxels=4, halo_voxels=2, **prm)}
    out["points"] = scenes.near_surface_points(v, t, 120_000, 2.0 * out["dense"].spec.voxel.max(),
                                               seed=11, dtype=np.float64)
    yield out
    out["dense"].close()
    out["sparse"].close()


@pytest.mark.parametrize("kind", ["dense", "sparse"])
@pytest.mark.parametrize("mode", MODES)
def test_host_query_equals_device_query(grids, kind, mode):
    g, pts = grids[kind], grids["points"]
    print(f"[gpu] host FP64 query {kind} {mode}", flush=True)
    gap, grad, valid, normal, witness, branch = g.query(torch.from_numpy(pts).cuda(), mode,
                                                        full=True)
    ref = dict(gap=gap.cpu().numpy(), normal=normal.cpu().numpy(), witness=witness.cpu().numpy(),
               valid=valid.cpu().numpy().astype(np.uint8), branch=branch.cpu().numpy())
    for n in (1, 1000, len(pts)):   # the last one is past the mapped-memory size
        got = _host_query(g, mode, pts[:n])
        for k, a in ref.items():
            assert np.array_equal(got[k], a[:n]), (n, k)


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_host_candidates_mapped_equal_device_block(grids, kind):
    g, pts = grids[kind], grids["points"]
    print(f"[gpu] host FP64 candidates {kind}", flush=True)
    big = _host_query(g, "ContactAware", pts[:20_000], cands=True)     # 11.8 MB: device block
    small = _host_query(g, "ContactAware", pts[:3000], cands=True)     # mapped
    again = _host_query(g, "ContactAware", pts[:3000], cands=True)
    assert big["ccount"].mean() > 1.0
    for k in ("gap", "normal", "witness", "valid", "branch", "ccount"):
        assert np.array_equal(small[k], big[k][:3000]), k
        assert np.array_equal(again[k], small[k]), k
    cnt = small["ccount"]
    for k in ("cgap", "cn", "cw", "cb"):
        for q in range(8):
            m = cnt > q
            assert np.array_equal(small[k][m, q], big[k][:3000][m, q]), (k, q)
