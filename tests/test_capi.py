"""C ABI checks that need no GPU: exports, host-side grid and BVH set-up
against the oracle, file validation and its error messages."""
import os
import re
import struct

import numpy as np
import pytest

import sdfo
from sdfo import _lib, meshes

import sdfo_format as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(sdfo_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    declared = _declared_symbols()
    assert len(declared) >= 24
    missing = [n for n in declared if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_default_params_are_the_reference_defaults():
    # SdfBuilder.h:24-55
    p = sdfo.default_build_params()
    assert (p.padding, p.voxel_size, p.target_resolution, p.max_resolution) == (0.1, 0.0, 128, 512)
    assert (p.block_dim, p.halo_voxels, p.band_voxels, p.band_distance, p.store_hessian) == \
        (8, 2, 8, 0.0, 1)
    assert (p.max_candidates, p.feature_distance_tolerance, p.normal_merge_angle_degrees,
            p.tile_dim, p.use_tile_acceleration) == (4, 0.0, 5.0, 8, 1)


@pytest.mark.parametrize("scale,target,maxr,voxel", [((1, 1, 1), 64, 64, 0.0),
                                                     ((1, 0.3, 0.07), 100, 512, 0.0),
                                                     ((2, 1, 1), 128, 96, 0.0),
                                                     ((1, 1, 1), 64, 512, 0.003)])
def test_setup_grid_matches_oracle(O, scale, target, maxr, voxel):
    v, _ = meshes.cube(0.1, (0.01, -0.02, 0.003))
    v = v * np.array(scale)
    s = sdfo.setup_grid(v, target_resolution=target, max_resolution=maxr, voxel_size=voxel)
    o = O.setup_grid(v, 0.1, voxel, target, maxr)
    assert tuple(s.shape) == o.shape
    for a in ("bmin", "bmax", "voxel"):
        assert np.array_equal(getattr(s, a), getattr(o, a))


@pytest.mark.parametrize("mk", [lambda: meshes.cube(0.1), lambda: meshes.torus(0.05, 0.01, 64, 32),
                                lambda: meshes.gear(40, 12, 0.02, 0.002, 12),
                                lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 256, 32),
                                lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 1024, 40)])
def test_host_bvh_matches_oracle_tree(O, mk):
    """The library's (multi-threaded) host BVH is node-for-node the KdTree of
    the reference as restated by the oracle, including the nth_element leaf order."""
    v, t = mk()
    n6, i4, perm, depth = _lib.debug_bvh_export(v, t)
    o6, oi4, operm = O.bvh_export(v, t)
    assert np.array_equal(n6, o6)
    assert np.array_equal(i4, oi4)
    assert np.array_equal(perm, operm)
    assert depth + 2 <= 32


def test_host_bvh_concurrent_builds_match():
    """Builds on several threads at once (each thread reuses its own host
    scratch between builds, and the worker pool interleaves their subtree
    tasks) give the single-threaded trees, build after build."""
    import threading

    ms = [meshes.torus(0.05, 0.01, 64, 32), meshes.gear(40, 12, 0.02, 0.002, 12),
          meshes.hollow_cylinder(0.04, 0.03, 0.05, 256, 32), meshes.cube(0.1)]
    ref = [_lib.debug_bvh_export(v, t) for v, t in ms]
    bad = []

    def run(k):
        for r in range(6):
            m = (k + r) % len(ms)
            got = _lib.debug_bvh_export(*ms[m])
            if not all(np.array_equal(a, b) for a, b in zip(got[:3], ref[m][:3])):
                bad.append((k, r, m))

    th = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not bad


def test_empty_mesh_is_rejected_without_device():
    with pytest.raises(sdfo.SdfoError, match="Empty mesh"):
        sdfo.setup_grid(np.zeros((0, 3)))


# ------------------------------------------------------------- file checks
def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def _dense_file(tmp_path, **over):
    shape = (2, 2, 2)
    N = 8
    b = F.dense_bytes(shape, (0, 0, 0), (2, 2, 2), (1, 1, 1), np.zeros(N), np.zeros((N, 3)),
                      np.zeros((N, 6)), np.zeros((N, 3)))
    b = bytearray(b)
    for off, val in over.items():
        b[int(off[1:]):int(off[1:]) + len(val)] = val
    return bytes(b)


def test_file_predicates(tmp_path):
    good = _write(tmp_path, "d.sdf", _dense_file(tmp_path))
    assert _lib.lib.sdfo_is_valid_file(good.encode()) == 1
    assert _lib.lib.sdfo_is_sparse_file(good.encode()) == 0
    sp = F.sparse_bytes((8, 8, 8), (0, 0, 0), (8, 8, 8), (1, 1, 1), 8, 2, 1.0, np.zeros((0, 3)),
                        [], [])
    sp = _write(tmp_path, "s.sdf", sp)
    assert _lib.lib.sdfo_is_sparse_file(sp.encode()) == 1
    bad = _write(tmp_path, "b.sdf", b"SDFX" + _dense_file(tmp_path)[4:])
    assert _lib.lib.sdfo_is_valid_file(bad.encode()) == 0
    short = _write(tmp_path, "short.sdf", b"SDFO\x64\0\0\0")
    assert _lib.lib.sdfo_is_valid_file(short.encode()) == 0
    assert _lib.lib.sdfo_is_valid_file(str(tmp_path / "missing.sdf").encode()) == 0


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda b: b"SDFX" + b[4:], -4, "Invalid SdfOracle data magic number"),
    (lambda b: b[:4] + struct.pack("<I", 101) + b[8:], -4,
     "Unsupported SdfOracle data version: 101 \\(only 0.1.0 is supported\\)"),
    (lambda b: b[:96] + struct.pack("<I", 7) + b[100:], -4, "not dense 0.1.0 layout"),
    (lambda b: b[:92] + struct.pack("<I", 1 | 4) + b[96:], -4,
     "Dense SDF 0.1.0 contact-aware flag/layout mismatch"),
    (lambda b: b[:200], -4, "Failed to read dense SDF payload"),
    (lambda b: b[:60], -4, "Failed to read SdfOracle data header"),
])
def test_load_validation_messages(tmp_path, mutate, code, msg):
    path = _write(tmp_path, "x.sdf", mutate(_dense_file(tmp_path)))
    with pytest.raises(sdfo.SdfoError, match=msg) as ei:
        sdfo.SdfGrid.load(path)
    assert ei.value.code == code


def test_load_missing_file(tmp_path):
    with pytest.raises(sdfo.SdfoError, match="Cannot open SdfOracle data file") as ei:
        sdfo.SdfGrid.load(str(tmp_path / "nope.sdf"))
    assert ei.value.code == -3


def test_sparse_load_validation(tmp_path):
    base = F.sparse_bytes((8, 8, 8), (0, 0, 0), (8, 8, 8), (1, 1, 1), 8, 2, 1.0,
                          np.zeros((1, 3), np.int32), np.zeros((1, 512)), np.zeros((1, 512, 3)),
                          np.zeros((1, 512, 6)))
    # sparse flag set but contact flag without the contact layout
    bad = base[:92] + struct.pack("<I", 2 | 1 | 4) + base[96:]
    with pytest.raises(sdfo.SdfoError, match="contact-aware flag/layout mismatch"):
        sdfo.SdfGrid.load(_write(tmp_path, "a.sdf", bad))
    # blockDim 0
    bad = base[:100] + struct.pack("<I", 0) + base[104:]
    with pytest.raises(sdfo.SdfoError, match="blockDim must be positive"):
        sdfo.SdfGrid.load(_write(tmp_path, "b.sdf", bad))
    # truncated block payload
    with pytest.raises(sdfo.SdfoError, match="Failed to read sparse SDF block"):
        sdfo.SdfGrid.load(_write(tmp_path, "c.sdf", base[:128 + 16 + 100]))
    # contact layout with stride 9
    res = struct.pack("<IIII", 4, 8, 2, 0) + struct.pack("<d", 1.0) + struct.pack("<I", 9)
    hdr = F.header((8, 8, 8), (0, 0, 0), (8, 8, 8), (1, 1, 1), 2 | 4, res)
    with pytest.raises(sdfo.SdfoError, match="Invalid sparse contact-aware candidate stride"):
        sdfo.SdfGrid.load(_write(tmp_path, "d.sdf", hdr))


def test_dense_contact_stride_and_count_checks(tmp_path):
    N = 8
    cand = dict(stride=9, count=np.zeros(N), phi=np.zeros((N, 9)), n=np.zeros((N, 9, 3)),
                branch=np.zeros((N, 9)))
    b = F.dense_bytes((2, 2, 2), (0, 0, 0), (2, 2, 2), (1, 1, 1), np.zeros(N), np.zeros((N, 3)),
                      np.zeros((N, 6)), np.zeros((N, 3)), cand)
    with pytest.raises(sdfo.SdfoError, match="Invalid contact-aware candidate stride"):
        sdfo.SdfGrid.load(_write(tmp_path, "s.sdf", b))
    cand = dict(stride=2, count=np.array([0, 1, 3, 0, 0, 0, 0, 0]), phi=np.zeros((N, 2)),
                n=np.zeros((N, 2, 3)), branch=np.zeros((N, 2)))
    b = F.dense_bytes((2, 2, 2), (0, 0, 0), (2, 2, 2), (1, 1, 1), np.zeros(N), np.zeros((N, 3)),
                      np.zeros((N, 6)), np.zeros((N, 3)), cand)
    with pytest.raises(sdfo.SdfoError, match="Invalid contact-aware candidate count"):
        sdfo.SdfGrid.load(_write(tmp_path, "c.sdf", b))


def test_candidate_order_matches_std_sort(O):
    """The builder's |phi| ordering equals std::sort's on arrays full of ties,
    across the insertion-sort (<= 16), introsort and heap-sort regimes."""
    rng = np.random.default_rng(0)
    for n in list(range(0, 65)) * 8:
        kind = rng.integers(4)
        if kind == 0:
            vals = rng.integers(-3, 4, n).astype(np.float64)            # many |x| ties
        elif kind == 1:
            vals = rng.normal(size=n)
        elif kind == 2:
            vals = np.sort(rng.integers(0, 5, n)).astype(np.float64)[::-1].copy()  # adversarial
        else:
            vals = np.where(rng.random(n) < 0.5, 1.0, -1.0) * rng.integers(1, 3, n)
        assert np.array_equal(_lib.debug_sort_by_abs(vals), O.sort_by_abs(vals)), (n, vals)


@pytest.mark.parametrize("seed,leaf", [(0, 16), (1, 3), (2, 1), (3, 64)])
def test_point_bvh_matches_oracle(O, seed, leaf):
    """The library's broad-phase tree (host only) returns the oracle's
    indices in the oracle's order (bvh.py:56-75)."""
    from sdfo import contact as SC
    rng = np.random.default_rng(seed)
    P = np.concatenate([rng.normal(size=(700, 3)) * [0.05, 0.02, 0.01],
                        rng.uniform(-0.05, 0.05, (300, 3))])
    P[::37] = P[5]
    P[1::41, 0] = P[2, 0]
    bvh = SC.PointBvh(P, leaf_size=leaf)
    nodes = O.point_bvh_build(P, leaf_size=leaf)
    assert bvh.n_nodes == len(nodes)
    assert bvh.n_leaves == sum(nd["left"] < 0 for nd in nodes)
    for k in range(40):
        c = rng.uniform(-0.06, 0.06, 3)
        h = rng.uniform(0.0, 0.05, 3)
        lo, hi = (c - h, c + h) if k % 5 else (c + h, c - h)     # inverted boxes too
        m = 0.0 if k % 3 else 0.004
        assert np.array_equal(bvh.query(lo, hi, m), O.point_bvh_query(nodes, lo, hi, m))
    assert np.array_equal(bvh.query([-1] * 3, [1] * 3), np.arange(len(P)))
    bvh.close()


def test_point_bvh_empty_and_errors():
    from sdfo import contact as SC
    b = SC.PointBvh(np.zeros((0, 3)))
    assert (b.n_nodes, b.n_leaves) == (0, 0)
    assert b.query([0, 0, 0], [1, 1, 1]).size == 0
    with pytest.raises(sdfo.SdfoError, match="leaf_size"):
        SC.PointBvh(np.zeros((4, 3)), leaf_size=0)
