"""Pins the oracle's contact-aware build (SdfBuilder.cpp:28-359, 416-549) and
its sparse Hessian stencil (SdfBuilder.cpp:1154-1225) on the CPU.

Each expectation comes from geometry worked out in the test (closest points
on axis-aligned squares, barycentric blends on a right triangle, one-sided
differences of a linear field), not from the oracle's own code path, so a
wrong blend weight, a dropped tolerance, a reversed merge rule or a missing
fallback fails.
"""
import numpy as np
import pytest

from sdfo import meshes

H_CUBE = 0.05
# meshes.cube lists its faces as triangle pairs: -z, +z, -y, +y, -x, +x
FACE_NORMALS = np.array([[0, 0, -1], [0, 0, 1], [0, -1, 0], [0, 1, 0], [-1, 0, 0], [1, 0, 0]],
                        np.float64)


def fnv1a_branch(normal):
    """Branch id of a unit normal: FNV-1a (offset basis 2166136261, prime
    16777619) over the three components quantised to round(4095 x)."""
    h = 2166136261
    for x in normal:
        q = int(np.floor(abs(np.clip(x, -1.0, 1.0)) * 4095.0 + 0.5)) * (1 if x >= 0 else -1)
        h ^= q & 0xFFFFFFFF
        h = (h * 16777619) & 0xFFFFFFFF
    return 1 + (h & 0x3FFFFFFF)


def _square_distance(p, axis, sign):
    """Distance from p to the cube face {x_axis = sign h, |other| <= h}."""
    cp = np.clip(p, -H_CUBE, H_CUBE)
    cp[axis] = sign * H_CUBE
    return np.linalg.norm(p - cp)


def _cube_expected(p, tol, max_cand):
    faces = []
    for f, n in enumerate(FACE_NORMALS):
        axis = int(np.flatnonzero(n)[0])
        sign = n[axis]
        faces.append((_square_distance(p, axis, sign), sign * p[axis] - H_CUBE, f))
    m = min(d for d, _, _ in faces)
    allowed = m + tol
    if any(abs(d - allowed) < 1e-12 for d, _, _ in faces):
        return None                       # too close to the reach boundary to call
    kept = [(phi, f) for d, phi, f in faces if d <= allowed]
    kept.sort(key=lambda e: abs(e[0]))    # stable: ties keep triangle order
    return kept[:max_cand]


@pytest.mark.parametrize("tol_h,max_cand,tiles", [(1.5, 4, True), (1.5, 4, False), (3.7, 8, True),
                                                  (3.7, 2, False)])
def test_cube_face_candidates_closed_form(O, tol_h, max_cand, tiles):
    v, t = meshes.cube(0.1)
    cn = np.repeat(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])[:, None, :], 3, 1)
    cn /= np.linalg.norm(cn, axis=2, keepdims=True)
    g = O.setup_grid(v, 0.1, 0.0, 16, 16)
    h = g.voxel.max()
    r = O.build_contact_aware_dense(v, t, cn, g, max_candidates=max_cand, feature_tol=tol_h * h,
                                    use_tiles=tiles)
    branch = {f: fnv1a_branch(n) for f, n in enumerate(FACE_NORMALS)}
    assert branch[5] == fnv1a_branch([1.0, 0.0, 0.0]) and len(set(branch.values())) == 6
    c = g.centers()
    checked = 0
    for idx in range(len(c)):
        want = _cube_expected(c[idx], tol_h * h, max_cand)
        if want is None:
            continue
        checked += 1
        cnt = int(r["ccount"][idx])
        assert cnt == len(want), idx
        for q, (phi, f) in enumerate(want):
            assert r["cphi"][idx, q] == phi
            assert np.array_equal(r["cn"][idx, q], FACE_NORMALS[f])
            assert r["cbranch"][idx, q] == branch[f]
        assert r["phi"][idx] == want[0][0] and np.array_equal(r["n"][idx], FACE_NORMALS[want[0][1]])
    assert checked > 0.9 * len(c)
    # corner region: three faces in reach, edge region: two
    assert (r["ccount"] == min(3, max_cand)).any() and (r["ccount"] == 2).any()


def test_blended_corner_normal_closed_form(O):
    """One triangle a=(0,0,0), b=(1,0,0), c=(0,1,0) with distinct corner
    normals: a voxel centre above the interior has closest point (x, y, 0),
    barycentric weights (1-x-y, x, y), normal = normalised blend and
    phi = (p - cp) . normal."""
    v = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0]])
    t = np.array([[0, 1, 2]], np.int32)
    na = np.array([0.0, 0.0, 1.0])
    nb = np.array([0.6, 0.0, 0.8])
    nc = np.array([0.0, -0.28, 0.96])
    cn = np.stack([na, nb, nc])[None]
    g = O.Grid(np.array([0.0, 0.0, 0.0]), np.array([1.0, 1.0, 0.5]), np.array([0.1, 0.1, 0.1]),
               (10, 10, 5))
    r = O.build_contact_aware_dense(v, t, cn, g, max_candidates=2, use_tiles=False)
    c = g.centers()
    inside = (c[:, 0] + c[:, 1] < 1.0 - 1e-9)
    assert inside.sum() > 100
    for idx in np.flatnonzero(inside):
        p = c[idx]
        x, y = p[0], p[1]
        blend = (1 - x - y) * na + x * nb + y * nc
        n = blend / np.linalg.norm(blend)
        phi = p[2] * n[2]                       # (p - cp) = (0, 0, z)
        assert r["ccount"][idx] == 1
        assert np.allclose(r["cn"][idx, 0], n, rtol=0, atol=1e-14)
        assert np.isclose(r["cphi"][idx, 0], phi, rtol=0, atol=1e-14)
        assert np.allclose(r["witness"][idx], [x, y, 0.0], rtol=0, atol=1e-15)
        q = np.abs(np.clip(n, -1, 1)) * 4095.0
        if np.all(np.abs(q - np.floor(q) - 0.5) > 1e-6):   # away from a rounding tie
            assert r["cbranch"][idx, 0] == fnv1a_branch(n)
    # a blend that swaps the b and c weights moves the normal by ~0.2 at (0.65, 0.15)
    k = np.flatnonzero(np.all(np.isclose(c, [0.65, 0.15, 0.25]), axis=1))[0]
    swapped = (1 - 0.8) * na + 0.15 * nb + 0.65 * nc
    assert np.linalg.norm(r["cn"][k, 0] - swapped / np.linalg.norm(swapped)) > 0.1


@pytest.mark.parametrize("theta_deg,merged", [(4.0, True), (6.0, False)])
def test_merge_angle_keeps_smaller_abs_phi(O, theta_deg, merged):
    """Two stacked triangles with a voxel centre between them, the lower
    one's normals tilted by theta: within the 5 degree merge angle the pair
    collapses to the entry with the smaller |phi| (the later, tilted one);
    beyond it both stay, sorted by |phi|."""
    th = np.radians(theta_deg)
    dz = 0.01
    v = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0],
                  [0.0, 0, -dz], [1.0, 0, -dz], [0.0, 1, -dz]])
    t = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    n1 = np.array([0.0, 0.0, 1.0])
    n2 = np.array([np.sin(th), 0.0, np.cos(th)])
    cn = np.stack([np.tile(n1, (3, 1)), np.tile(n2, (3, 1))])
    p = np.array([0.25, 0.25, -0.006])
    g = O.Grid(p - 0.05, p + 0.05, np.array([0.1, 0.1, 0.1]), (1, 1, 1))
    r = O.build_contact_aware_dense(v, t, cn, g, max_candidates=4, feature_tol=0.05,
                                    use_tiles=False)
    phi1 = p[2]                                  # (p - cp1) = (0, 0, z), normal +z
    phi2 = (p[2] + dz) * np.cos(th)              # (p - cp2) = (0, 0, z + dz)
    assert abs(phi2) < abs(phi1)
    if merged:
        assert r["ccount"][0] == 1
        assert np.isclose(r["cphi"][0, 0], phi2, rtol=0, atol=1e-15)
        assert np.allclose(r["cn"][0, 0], n2, rtol=0, atol=1e-15)
    else:
        assert r["ccount"][0] == 2
        assert np.allclose(r["cphi"][0, :2], [phi2, phi1], rtol=0, atol=1e-15)
        assert np.allclose(r["cn"][0, 1], n1, rtol=0, atol=0)


def test_tiles_equal_full_scan_where_the_expansion_covers_the_reach(O):
    """The tile index grows each triangle box by tol + 2h (influence) + the
    tile radius + h, so a voxel whose nearest distance m satisfies m + tol <=
    tol + 3h sees every triangle in reach through its tile: tiles on and off
    agree there, voxel for voxel."""
    v, t = meshes.torus(0.05, 0.012, 24, 12)
    cen = np.zeros_like(v)
    ang = np.arctan2(v[:, 1], v[:, 0])
    cen[:, 0], cen[:, 1] = 0.05 * np.cos(ang), 0.05 * np.sin(ang)
    vn = v - cen
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    cn = vn[t]
    g = O.setup_grid(v, 0.1, 0.0, 24, 24)
    h = g.voxel.max()
    kw = dict(max_candidates=4, feature_tol=1.5 * h)
    on = O.build_contact_aware_dense(v, t, cn, g, use_tiles=True, **kw)
    off = O.build_contact_aware_dense(v, t, cn, g, use_tiles=False, **kw)
    dense = O.build_dense(v, t, g)
    near = np.abs(dense["phi"]) <= 2.5 * h
    assert near.sum() > 500
    for key in ("ccount", "cphi", "cn", "cbranch", "phi", "n"):
        assert np.array_equal(on[key][near], off[key][near]), key
    # the list invariants on every voxel
    for idx in range(g.total):
        cnt = int(on["ccount"][idx])
        assert 1 <= cnt <= 4
        ph = np.abs(on["cphi"][idx, :cnt])
        assert np.all(np.diff(ph) >= 0)
        nn = on["cn"][idx, :cnt]
        assert np.allclose(np.linalg.norm(nn, axis=1), 1.0, atol=1e-15)
        assert on["phi"][idx] == on["cphi"][idx, 0]


# ------------------------------------------------------------ sparse Hessian
def _sparse_linear_normals(A, b, active):
    g_shape = (24, 16, 16)
    from oracle import Grid
    g = Grid(np.zeros(3), np.array(g_shape) * 0.125, np.full(3, 0.125), g_shape)
    bijk = np.array(active, np.int32)
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    loc = np.stack([lx.ravel(), ly.ravel(), lz.ravel()], 1)
    c = (bijk[:, None, :] * 8 + loc[None] + 0.5) * 0.125          # (nb, 512, 3)
    n = np.einsum("ij,bkj->bki", A, c) + b
    assert np.array_equal(n.astype(np.float32).astype(np.float64), n)   # exact in FP32
    return g, bijk, loc, n.astype(np.float32)


def test_sparse_hessian_one_sided_at_missing_bricks(O):
    A = np.array([[1.0, 2.0, -1.0], [0.5, -3.0, 4.0], [2.0, 0.0, 1.5]])
    b = np.array([0.25, -0.5, 1.0])
    active = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 1)]   # (2,*,*) and others missing
    g, bijk, loc, n = _sparse_linear_normals(A, b, active)
    H = O.sparse_hessian(g, 8, bijk, n)
    have = {tuple(x) for x in active}

    def avail(i, j, k):
        if not (0 <= i < 24 and 0 <= j < 16 and 0 <= k < 16):
            return False
        return (i // 8, j // 8, k // 8) in have

    for bi, blk in enumerate(bijk):
        for li, (lx, ly, lz) in enumerate(loc):
            ijk = blk * 8 + np.array([lx, ly, lz])
            cols = np.zeros((3, 3))
            for a in range(3):
                up, dn = ijk.copy(), ijk.copy()
                up[a] += 1
                dn[a] -= 1
                scale = (avail(*up) + avail(*dn)) / 2.0     # 1 central, 1/2 one-sided, 0 none
                cols[:, a] = A[:, a] * scale
            S = 0.5 * (cols + cols.T)
            want = np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]], np.float32)
            assert np.array_equal(H[bi, li], want), (bi, li)
    # the inner voxels of brick (1,0,0) next to the missing brick (2,0,0): x column halved
    one_sided = np.array([0.5 * A[0, 0], 0.5 * (A[0, 1] + 0.5 * A[1, 0])], np.float32)
    li = 7 + 8 * (3 + 8 * 3)
    assert np.array_equal(H[1, li, :2], one_sided)


def test_sparse_second_order_exact_on_quadratic(O):
    """Sparse SecondOrder (SdfOracle.cpp:624-654) on bricks holding an exact
    quadratic field: phi0 + n0 . d + 1/2 d^T H0 d reproduces it; without
    stored Hessians the mode degrades to first order (SdfOracle.cpp:636-640)."""
    from oracle import Grid
    Q = np.array([[0.5, 0.25, 0.0], [0.25, -0.25, 0.5], [0.0, 0.5, 1.0]])
    bvec = np.array([0.25, -0.75, 0.5])
    g = Grid(np.zeros(3), np.full(3, 2.0), np.full(3, 0.125), (16, 16, 16))
    bijk = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)], np.int32)
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    loc = np.stack([lx.ravel(), ly.ravel(), lz.ravel()], 1)
    c = (bijk[:, None, :] * 8 + loc[None] + 0.5) * 0.125
    f = lambda x: 0.125 + x @ bvec + 0.5 * np.einsum("...i,ij,...j->...", x, Q, x)  # noqa: E731
    phi, nrm = f(c), bvec + c @ Q
    Hs = np.tile([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]], (4, 512, 1))
    for a in (phi, nrm):
        assert np.array_equal(a.astype(np.float32).astype(np.float64), a)
    sp = dict(block_dim=8, bijk=bijk, phi=phi.astype(np.float32), n=nrm.astype(np.float32),
              H=Hs.astype(np.float32), stride=0)
    rng = np.random.default_rng(8)
    blk = bijk[rng.integers(0, 4, 600)]
    p = (blk * 8 + rng.uniform(0.5, 7.5, (600, 3))) * 0.125       # inside active bricks
    r = O.query_sparse(g, sp, "SecondOrder", p)
    assert r["valid"].all()
    assert np.allclose(r["gap"], f(p), rtol=0, atol=1e-13)
    assert np.allclose(r["grad"], bvec + p @ Q, rtol=0, atol=1e-13)
    sp_noh = dict(sp, H=None)
    r2 = O.query_sparse(g, sp_noh, "SecondOrder", p)
    r1 = O.query_sparse(g, sp, "FirstOrder", p)
    assert np.array_equal(r2["gap"], r1["gap"]) and np.array_equal(r2["grad"], r1["grad"])
    # first order is off by the dropped quadratic term
    assert np.abs(r1["gap"] - f(p)).max() > 1e-4


def test_probe_equals_the_full_builds(O):
    """The sampled contact-aware probe (used to check large GPU builds) gives
    the full dense and sparse builds' candidates voxel for voxel, and the
    sparse build's brick mask."""
    v, t = meshes.torus(0.05, 0.012, 24, 12)
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    cn = np.repeat(fn[:, None], 3, 1)
    g = O.setup_grid(v, 0.1, 0.0, 24, 24)
    nx, ny, nz = g.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    d = O.build_contact_aware_dense(v, t, cn, g)
    p = O.contact_aware_probe(v, t, cn, g, ijk)
    for key in ("ccount", "cphi", "cn", "cbranch"):
        assert np.array_equal(d[key], p[key]), key
    s = O.build_contact_aware_sparse(v, t, cn, g, block_dim=8, halo=2, band_voxels=4)
    L = np.arange(512)
    loc = np.stack([L % 8, (L // 8) % 8, L // 64], 1)
    for b in range(len(s["bijk"])):
        vox = s["bijk"][b] * 8 + loc
        inb = np.all(vox < np.array(g.shape), 1)
        pp = O.contact_aware_probe(v, t, cn, g, vox[inb], sparse_block_dim=8, halo=2, band_voxels=4)
        if b == 0:
            assert np.array_equal(s["active"], pp["mask"])
        assert np.array_equal(s["ccount"][b][inb], pp["ccount"])
        assert np.array_equal(s["cphi"][b][inb], pp["cphi"].astype(np.float32))
        assert np.array_equal(s["cbranch"][b][inb], pp["cbranch"])
