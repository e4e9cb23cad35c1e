"""GPU build parity: dense and sparse SDF builds against the FP64 oracle.

The build kernels run FP64 without FMA contraction and traverse the same BVH
in the same near-first order as SdfBuilder / KdTree, so every field is
compared bit for bit (the accuracy rule asks for nearest-triangle indices and
the sparse brick mask bit-exact; here the values are bit-exact too).
"""
import numpy as np
import pytest

from conftest import ograd
from sdfo import meshes

pytestmark = pytest.mark.gpu

DENSE_CASES = [
    ("cube_64", lambda: meshes.cube(0.1), 64),
    ("cube_offcentre_32", lambda: meshes.cube(0.07, (0.01, -0.02, 0.03)), 32),
    ("torus_48", lambda: meshes.torus(0.05, 0.01, 64, 16), 48),
    ("sphere_32", lambda: meshes.uv_sphere(0.05, 12, 24), 32),
    ("gear_64", lambda: meshes.gear(16, 8, 0.02, 0.002, 4), 64),
    ("hollow_cylinder_32", lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 48, 6), 32),
]


@pytest.mark.parametrize("name,mk,res", DENSE_CASES, ids=[c[0] for c in DENSE_CASES])
def test_dense_build_bit_exact(O, sdfo_gpu, name, mk, res):
    v, t = mk()
    print(f"[gpu] dense build {name} res={res} tris={len(t)}", flush=True)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=res, max_resolution=res)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    assert tuple(g.spec.shape) == og.shape
    assert np.array_equal(g.spec.bmin, og.bmin)
    assert np.array_equal(g.spec.bmax, og.bmax)
    assert np.array_equal(g.spec.voxel, og.voxel)
    ref = O.build_dense(v, t, og)
    got = g.download_dense()
    assert np.array_equal(got["tri"], ref["tri"]), "nearest-triangle index"
    for key in ("phi", "n", "witness", "H"):
        assert np.array_equal(got[key], ref[key]), key
    g.close()


def test_dense_build_non_cubic_grid(O, sdfo_gpu):
    # a flat box: the ladder snaps each axis separately (GridSpec.h:65-74)
    v, t = meshes.cube(0.1)
    v = v * np.array([1.0, 0.3, 0.1])
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=40, max_resolution=64)
    og = O.setup_grid(v, 0.1, 0.0, 40, 64)
    assert tuple(g.spec.shape) == og.shape and len(set(og.shape)) > 1
    ref = O.build_dense(v, t, og)
    got = g.download_dense()
    for key in ("phi", "n", "witness", "H", "tri"):
        assert np.array_equal(got[key], ref[key]), key


SPARSE_CASES = [
    ("hollow_cylinder_64_b8", lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 64, 8), 64, 8, 4, 2),
    ("torus_64_b8", lambda: meshes.torus(0.05, 0.01, 64, 16), 64, 8, 2, 1),
    ("sphere_48_b6", lambda: meshes.uv_sphere(0.05, 12, 24), 48, 6, 3, 2),
]


@pytest.mark.parametrize("name,mk,res,bd,band,halo", SPARSE_CASES, ids=[c[0] for c in SPARSE_CASES])
def test_sparse_build_bit_exact(O, sdfo_gpu, name, mk, res, bd, band, halo):
    v, t = mk()
    print(f"[gpu] sparse build {name}", flush=True)
    g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=res, max_resolution=res,
                                      block_dim=bd, band_voxels=band, halo_voxels=halo)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    ref = O.build_sparse(v, t, og, block_dim=bd, halo=halo, band_voxels=band)
    got = g.download_sparse()
    assert g.info.activation_distance == ref["activation"]
    assert g.info.band_distance == ref["band"]
    assert np.array_equal(got["active"], ref["active"]), "brick activation mask"
    assert np.array_equal(got["bijk"], ref["bijk"])
    assert 0 < len(got["bijk"]) < len(got["active"])
    for key in ("phi", "n", "H"):
        assert np.array_equal(got[key], ref[key]), key
    g.close()


def test_empty_mesh_rejected(sdfo_gpu):
    with pytest.raises(sdfo_gpu.SdfoError, match="Empty mesh"):
        sdfo_gpu.SdfGrid.build_dense(np.zeros((0, 3)), np.zeros((0, 3), np.int32))


def test_invalid_vertex_rejected(sdfo_gpu):
    v, t = meshes.cube(0.1)
    t = t.copy()
    t[3, 1] = 99
    with pytest.raises(sdfo_gpu.SdfoError, match="invalid vertex"):
        sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=16, max_resolution=16)


@pytest.mark.parametrize("name,mk,res", [DENSE_CASES[0], DENSE_CASES[2], DENSE_CASES[4]],
                         ids=["cube_64", "torus_48", "gear_64"])
def test_dense_replay_pass_bit_exact(O, sdfo_gpu, name, mk, res):
    """Every voxel forced through the per-lane replay pass still matches."""
    from sdfo import _lib

    v, t = mk()
    print(f"[gpu] dense build with forced replay {name}", flush=True)
    _lib.debug_force_replay(True)
    try:
        g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=res, max_resolution=res)
        ph = _lib.debug_build_phases()
    finally:
        _lib.debug_force_replay(False)
    assert ph["replayed_voxels"] == np.prod(g.spec.shape)
    og = O.setup_grid(v, 0.1, 0.0, res, res)
    ref = O.build_dense(v, t, og)
    got = g.download_dense()
    for key in ("tri", "phi", "n", "witness", "H"):
        assert np.array_equal(got[key], ref[key]), key


@pytest.mark.parametrize("name,mk,res", [DENSE_CASES[2], DENSE_CASES[4]], ids=["torus_48", "gear_64"])
def test_dense_frontier_pass_bit_exact(O, sdfo_gpu, name, mk, res):
    """A step budget of 16 hands nearly every brick to the frontier pass."""
    from sdfo import _lib

    v, t = mk()
    print(f"[gpu] dense build, budget 16, {name}", flush=True)
    _lib.debug_fill_budget(16, 0)
    try:
        g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=res, max_resolution=res)
    finally:
        _lib.debug_fill_budget(0, 0)
    ref = O.build_dense(v, t, O.setup_grid(v, 0.1, 0.0, res, res))
    got = g.download_dense()
    for key in ("tri", "phi", "n", "witness", "H"):
        assert np.array_equal(got[key], ref[key]), key


def test_sparse_frontier_pass_bit_exact(O, sdfo_gpu):
    from sdfo import _lib

    v, t = meshes.torus(0.05, 0.01, 64, 16)
    print("[gpu] sparse build, budget 16, torus_64_b8", flush=True)
    _lib.debug_fill_budget(16, 0)
    try:
        g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=64, max_resolution=64,
                                          block_dim=8, band_voxels=2, halo_voxels=1)
    finally:
        _lib.debug_fill_budget(0, 0)
    ref = O.build_sparse(v, t, O.setup_grid(v, 0.1, 0.0, 64, 64), block_dim=8, halo=1,
                         band_voxels=2)
    got = g.download_sparse()
    assert np.array_equal(got["active"], ref["active"])
    for key in ("phi", "n", "H"):
        assert np.array_equal(got[key], ref[key]), key


def test_sparse_replay_pass_bit_exact(O, sdfo_gpu):
    from sdfo import _lib

    v, t = meshes.torus(0.05, 0.01, 64, 16)
    print("[gpu] sparse build with forced replay torus_64_b8", flush=True)
    _lib.debug_force_replay(True)
    try:
        g = sdfo_gpu.SdfGrid.build_sparse(v, t, target_resolution=64, max_resolution=64,
                                          block_dim=8, band_voxels=2, halo_voxels=1)
    finally:
        _lib.debug_force_replay(False)
    og = O.setup_grid(v, 0.1, 0.0, 64, 64)
    ref = O.build_sparse(v, t, og, block_dim=8, halo=1, band_voxels=2)
    got = g.download_sparse()
    assert np.array_equal(got["active"], ref["active"])
    for key in ("phi", "n", "H"):
        assert np.array_equal(got[key], ref[key]), key


def test_wave_search_replays_few_voxels(sdfo_gpu):
    """The wave-cooperative search certifies almost every voxel by itself."""
    from sdfo import _lib, scenes

    v, t = scenes.ring_mesh(256, 64)
    g = sdfo_gpu.SdfGrid.build_dense(v, t, target_resolution=128, max_resolution=128)
    ph = _lib.debug_build_phases()
    n = np.prod(g.spec.shape)
    print(f"[gpu] ring 128: replayed {ph['replayed_voxels']} of {n}; phases {ph}", flush=True)
    assert 0 <= ph["replayed_voxels"] < 0.01 * n


def _soup_with_signed_zeros(n_tris, seed):
    """A triangle soup whose coordinates include +0.0 and -0.0 (the host tree's
    min / max keep the first of equal arguments, which fixes a zero bound's sign)."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1.0, 1.0, (3 * n_tris, 3))
    z = rng.random(v.shape) < 0.2
    v[z] = np.where(rng.random(int(z.sum())) < 0.5, -0.0, 0.0)
    return v, np.arange(3 * n_tris, dtype=np.int32).reshape(-1, 3)


def test_device_assembled_bvh_equals_host_tree(sdfo_gpu):
    """The builds assemble the BVH on the device from the host's slot order
    (shape from the triangle count, boxes folded as on the host): nodes, links,
    slot triangles and slot boxes equal the host tree bit for bit, for one mesh
    and for several uneven meshes concatenated as the batched build does."""
    from sdfo import _lib

    cases = [meshes.torus(0.05, 0.01, 64, 32), meshes.cube(0.1), _soup_with_signed_zeros(5, 1),
             _soup_with_signed_zeros(9, 2), _soup_with_signed_zeros(1000, 3),
             meshes.gear(40, 12, 0.02, 0.002, 12), _soup_with_signed_zeros(17, 4),
             meshes.hollow_cylinder(0.04, 0.03, 0.05, 256, 32)]
    groups = [[c] for c in cases] + [cases]
    for gi, group in enumerate(groups):
        print(f"[gpu] device BVH group {gi}: {[len(t) for _, t in group]} tris", flush=True)
        n6, i4, perm, tri9, box6, depth = _lib.debug_device_bvh_export(group)
        node_base = slot_base = 0
        want_depth = 0
        for v, t in group:
            h6, hi4, hperm, hdepth = _lib.debug_bvh_export(v, t)
            nn, ns = len(h6), len(t)
            want_depth = max(want_depth, hdepth)
            hi4 = hi4.copy()
            leaf = hi4[:, 3] > 0
            hi4[leaf, 2] += slot_base
            hi4[~leaf, 0] += node_base
            hi4[~leaf, 1] += node_base
            assert np.array_equal(n6[node_base:node_base + nn].view(np.uint64), h6.view(np.uint64))
            assert np.array_equal(i4[node_base:node_base + nn], hi4)
            assert np.array_equal(perm[slot_base:slot_base + ns], hperm)
            vv = np.asarray(v, np.float64).reshape(-1, 3)
            tt = np.asarray(t, np.int64).reshape(-1, 3)
            corners = vv[tt[hperm]]  # (ns, 3, 3)
            assert np.array_equal(tri9[slot_base:slot_base + ns].view(np.uint64),
                                  corners.reshape(ns, 9).view(np.uint64))
            a, b, c = corners[:, 0], corners[:, 1], corners[:, 2]
            lo = np.where(c < np.where(b < a, b, a), c, np.where(b < a, b, a))
            hm = np.where(a < b, b, a)
            hi = np.where(hm < c, c, hm)
            assert np.array_equal(box6[slot_base:slot_base + ns].view(np.uint64),
                                  np.concatenate([lo, hi], 1).view(np.uint64))
            node_base += nn
            slot_base += ns
        assert len(n6) == node_base
        assert depth == want_depth
