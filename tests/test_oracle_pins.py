"""Pins the CPU oracle (oracle/sdfo_oracle.cpp) on the CPU.

The reference cannot be compiled here (its CMake requires Eigen3, absent from
this image) and holds no golden vectors for this path, so every oracle
function is pinned by something independent of its own code: closed forms,
brute force, textbook special cases or invariants, each chosen so that a
plausible mistake (wrong sign, dropped term, swapped index) fails.
"""
import numpy as np
import pytest

from sdfo import meshes


# ---------------------------------------------------------------- grid setup
def test_snap_ladder(O):
    # GridSpec.h:65-74: smallest ladder value >= n, capped at maxRes
    cases = {(1, 512): 8, (8, 512): 8, (9, 512): 16, (100, 512): 128, (129, 512): 256,
             (300, 256): 256, (3000, 4096): 4096, (70, 64): 64, (5, 4): 4}
    # past the end of the ladder the cap itself is returned (GridSpec.h:73)
    for (n, m), want in cases.items():
        assert O.snap_resolution(n, m) == want


def test_setup_grid_closed_form(O):
    v, _ = meshes.cube(0.1)
    g = O.setup_grid(v, 0.1, 0.0, 64, 512)
    diag = np.sqrt(3) * 0.1
    pad = 0.1 * diag
    assert np.allclose(g.bmin, -0.05 - pad, rtol=0, atol=1e-15)
    voxel = (0.1 + 2 * pad) / 64
    assert np.allclose(g.voxel, voxel, rtol=1e-15)
    assert g.shape == (64, 64, 64)
    assert np.array_equal(g.bmax, g.bmin + 64 * g.voxel)
    # flat box: each axis snapped separately, bmax recomputed from n * voxel
    g2 = O.setup_grid(v * [1, 0.3, 0.05], 0.1, 0.0, 100, 512)
    assert g2.shape[0] == 128 and g2.shape[1] < 128 and g2.shape[2] <= g2.shape[1]
    # explicit voxel size overrides the target resolution
    g3 = O.setup_grid(v, 0.0, 0.01, 64, 512)
    assert g3.shape == (16, 16, 16) and np.all(g3.voxel == 0.01)


# ------------------------------------------------- closest point / ray test
def _sample_triangle(a, b, c, n=400):
    s, t = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = s + t <= 1
    s, t = s[m], t[m]
    return a + s[:, None] * (b - a) + t[:, None] * (c - a)


def test_closest_point_vs_dense_sampling(O):
    rng = np.random.default_rng(0)
    for _ in range(60):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        cp = O.closest_point(p, a, b, c)
        d = np.linalg.norm(p - cp)
        samples = _sample_triangle(a, b, c)
        dmin = np.linalg.norm(samples - p, axis=1).min()
        assert d <= dmin + 1e-12          # nothing on the triangle is closer
        assert dmin - d < 2e-2            # and the sampling gets close to it
        # cp lies on the triangle: barycentric coordinates in [0, 1]
        M = np.stack([b - a, c - a], 1)
        st = np.linalg.lstsq(M, cp - a, rcond=None)[0]
        assert st.min() > -1e-9 and st.sum() < 1 + 1e-9


def test_closest_point_voronoi_regions(O):
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    assert np.array_equal(O.closest_point([-1, -1, 0.5], a, b, c), a)      # vertex a
    assert np.array_equal(O.closest_point([2, -0.5, 0], a, b, c), b)       # vertex b
    assert np.array_equal(O.closest_point([-0.1, 3, 0], a, b, c), c)       # vertex c
    assert np.allclose(O.closest_point([0.3, -2, 1], a, b, c), [0.3, 0, 0])  # edge ab
    assert np.allclose(O.closest_point([1, 1, -1], a, b, c), [0.5, 0.5, 0])  # edge bc
    assert np.allclose(O.closest_point([0.2, 0.3, 7], a, b, c), [0.2, 0.3, 0])  # face


def test_ray_triangle_vs_plane_intersection(O):
    rng = np.random.default_rng(1)
    d = np.array([1.0, 0.3713906763541037, 0.52999894000318])
    d /= np.linalg.norm(d)
    hits = 0
    for _ in range(3000):
        a, b, c = rng.normal(size=(3, 3)) * 2.0
        o = rng.normal(size=3) * 0.5
        n = np.cross(b - a, c - a)
        if abs(n @ d) < 1e-3:
            continue
        t = n @ (a - o) / (n @ d)
        x = o + t * d
        M = np.stack([b - a, c - a], 1)
        s, u = np.linalg.lstsq(M, x - a, rcond=None)[0]
        margin = min(s, u, 1 - s - u)
        if abs(margin) < 1e-6 or abs(t) < 1e-6:
            continue
        want = margin > 0 and t > 0
        assert O.ray_triangle(o, a, b, c) == want
        hits += want
    assert hits > 100


# ------------------------------------------------------------ BVH / nearest
MESHES = {
    "cube": lambda: meshes.cube(0.1),
    "torus": lambda: meshes.torus(0.05, 0.01, 32, 12),
    "sphere": lambda: meshes.uv_sphere(0.05, 10, 20),
    "gear": lambda: meshes.gear(12, 6, 0.02, 0.002, 3),
    "tube": lambda: meshes.hollow_cylinder(0.04, 0.03, 0.05, 24, 3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_nearest_equals_brute_force(O, name):
    v, t = MESHES[name]()
    rng = np.random.default_rng(2)
    lo, hi = v.min(0), v.max(0)
    pts = rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), (1500, 3))
    sn = O.signed_nearest(v, t, pts)
    sq, tri = O.brute_nearest(v, t, pts)
    # the BVH search finds the exact minimum of the O(N_t) scan ...
    assert np.array_equal(np.abs(sn["sd"]), np.sqrt(sq))
    # ... and its triangle attains it (ties are resolved by traversal order)
    for i in np.flatnonzero(sn["tri"] != tri)[:50]:
        a, b, c = v[t[sn["tri"][i]]]
        cp = O.closest_point(pts[i], a, b, c)
        assert np.sum((pts[i] - cp) ** 2) == sq[i]


@pytest.mark.parametrize("name", list(MESHES))
def test_ray_parity_sign_vs_winding_number(O, name):
    v, t = MESHES[name]()
    rng = np.random.default_rng(3)
    lo, hi = v.min(0), v.max(0)
    pts = rng.uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), (3000, 3))
    wn = O.winding_number(v, t, pts)
    sn = O.signed_nearest(v, t, pts)
    far = np.abs(sn["sd"]) > 1e-6
    assert np.all((sn["sd"][far] < 0) == (wn[far] > 0.5))
    assert 0.05 < (wn > 0.5).mean() < 0.95


def test_dense_build_cube_is_box_sdf(O):
    v, t = meshes.cube(0.1)
    g = O.setup_grid(v, 0.1, 0.0, 32, 32)
    d = O.build_dense(v, t, g)
    c = g.centers()
    q = np.abs(c) - 0.05
    box = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)
    assert np.array_equal(d["phi"], box)
    assert np.allclose(np.linalg.norm(d["n"], axis=1), 1.0, atol=1e-14)
    # witness lies on the surface and phi = +-|c - w|
    assert np.allclose(np.abs(d["phi"]), np.linalg.norm(c - d["witness"], axis=1), atol=1e-15)


def test_dense_build_sphere_distance(O):
    v, t = meshes.uv_sphere(0.05, 24, 48)
    g = O.setup_grid(v, 0.1, 0.0, 24, 24)
    d = O.build_dense(v, t, g)
    r = np.linalg.norm(g.centers(), axis=1)
    # polygonal sphere: inscribed faces lie within R (1 - cos(pi/24)) of it
    assert np.abs(d["phi"] - (r - 0.05)).max() < 0.05 * (1 - np.cos(np.pi / 24)) * 1.5 + 1e-12


def test_hessian_of_linear_normal_field(O):
    # central differences are exact for a linear field; H = sym(A) inside
    g = O.Grid(np.zeros(3), np.ones(3) * 0.8, np.ones(3) * 0.1, (8, 8, 8))
    A = np.array([[1.0, 2.0, -1.0], [0.5, -3.0, 4.0], [2.0, 0.0, 1.5]])
    c = g.centers()
    n = c @ A.T
    H = O.dense_hessian(g, n)
    S = 0.5 * (A + A.T)
    want = np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]])
    idx = g.centers()
    inner = np.all((idx > 0.1) & (idx < 0.7), axis=1)
    assert np.allclose(H[inner], want, rtol=1e-12, atol=1e-12)
    # clamped border: the one-sided stencil halves the x-derivative at i = 0
    i0 = np.flatnonzero((np.abs(c[:, 0] - 0.05) < 1e-12) & inner.__invert__() &
                        np.all((c[:, 1:] > 0.1) & (c[:, 1:] < 0.7), axis=1))
    assert np.allclose(H[i0, 0], 0.5 * A[0, 0])


# --------------------------------------------------------------- queries
def _field_grid(O, f, df, n=12, ddf=None):
    g = O.Grid(np.array([-0.3, -0.2, -0.25]), None, np.array([0.05, 0.04, 0.06]), (n, n + 1, n + 2))
    g.bmax = g.bmin + g.voxel * np.array(g.shape)
    c = g.centers()
    data = dict(phi=f(c), n=df(c), H=np.zeros((len(c), 6)) if ddf is None else ddf(c))
    return g, data


def _pts(g, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(g.bmin + g.voxel * 0.5, g.bmax - g.voxel * 0.5, (n, 3))


def test_trilinear_exact_on_multilinear_field(O):
    f = lambda c: 0.3 + c[:, 0] - 2 * c[:, 1] + 0.5 * c[:, 2] + 4 * c[:, 0] * c[:, 1] * c[:, 2]  # noqa
    df = lambda c: np.zeros((len(c), 3))  # noqa: E731 (trilinear ignores normals)
    g, d = _field_grid(O, f, df)
    p = _pts(g, 500, 4)
    r = O.query_dense(g, d, "Trilinear", p, hessian=True)
    assert r["valid"].all()
    assert np.allclose(r["gap"], f(p), atol=1e-13)
    x, y, z = p.T
    grad = np.stack([1 + 4 * y * z, -2 + 4 * x * z, 0.5 + 4 * x * y], 1)
    assert np.allclose(r["grad"], grad, atol=1e-12)
    assert np.allclose(r["hess"][:, 1], 4 * z, atol=1e-11)   # d2/dxdy
    assert np.allclose(r["hess"][:, 0], 0.0)


def test_tricubic_exact_on_separable_cubic(O):
    f = lambda c: 0.3 * c[:, 0] ** 3 - 0.2 * c[:, 0] + 0.5 * c[:, 1] ** 2 - 0.1 * c[:, 2] ** 3  # noqa
    df = lambda c: np.stack([0.9 * c[:, 0] ** 2 - 0.2, c[:, 1], -0.3 * c[:, 2] ** 2], 1)  # noqa
    g, d = _field_grid(O, f, df)
    p = _pts(g, 500, 5)
    r = O.query_dense(g, d, "Tricubic", p, hessian=True)
    assert r["valid"].all()
    assert np.allclose(r["gap"], f(p), atol=1e-13)
    assert np.allclose(r["grad"], df(p), atol=1e-12)
    x, y, z = p.T
    assert np.allclose(r["hess"][:, 0], 1.8 * x, atol=1e-10)
    assert np.allclose(r["hess"][:, 3], 1.0, atol=1e-10)
    assert np.allclose(r["hess"][:, 5], -0.6 * z, atol=1e-10)
    assert np.allclose(r["hess"][:, [1, 2, 4]], 0.0, atol=1e-10)
    # unit normal and witness contract (SdfOracle.cpp:1083-1089)
    nrm = df(p) / np.linalg.norm(df(p), axis=1, keepdims=True)
    assert np.allclose(r["normal"], nrm, atol=1e-12)
    assert np.allclose(r["witness"], p - f(p)[:, None] * nrm, atol=1e-12)


def test_taylor_modes_exact_on_quadratic(O):
    Q = np.array([[0.4, 0.1, 0.0], [0.1, -0.2, 0.3], [0.0, 0.3, 0.5]])
    b = np.array([0.2, -0.7, 0.4])
    f = lambda c: 0.1 + c @ b + 0.5 * np.einsum("ni,ij,nj->n", c, Q, c)  # noqa: E731
    df = lambda c: b + c @ Q  # noqa: E731
    hh = lambda c: np.tile([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]], (len(c), 1))  # noqa
    g, d = _field_grid(O, f, df, ddf=hh)
    p = _pts(g, 400, 6)
    r2 = O.query_dense(g, d, "SecondOrder", p)
    assert np.allclose(r2["gap"], f(p), atol=1e-13)
    assert np.allclose(r2["grad"], df(p), atol=1e-12)
    r1 = O.query_dense(g, d, "FirstOrder", p)
    # first order = value + gradient at the nearest voxel centre, gradient not updated
    ii = np.round((p - g.bmin) / g.voxel - 0.5)
    cen = g.bmin + g.voxel * (ii + 0.5)
    assert np.allclose(r1["gap"], f(cen) + np.sum(df(cen) * (p - cen), 1), atol=1e-13)
    assert np.allclose(r1["grad"], df(cen))


def test_validity_rules(O):
    g, d = _field_grid(O, lambda c: c[:, 0], lambda c: np.tile([1.0, 0, 0], (len(c), 1)))
    nx, ny, nz = g.shape
    # continuous index exactly 0 -> valid; just below -> invalid; last cell's upper face -> invalid
    f = np.array([[0.0, 1.0, 1.0], [-1e-9, 1.0, 1.0], [nx - 1.0, 1.0, 1.0], [nx - 1.0 - 1e-9, 1, 1]])
    p = g.bmin + g.voxel * (f + 0.5)
    r = O.query_dense(g, d, "Tricubic", p)
    assert list(r["valid"]) == [True, False, False, True]
    # nearest-voxel modes: round half away from zero (GridSpec.h:42-44)
    f2 = np.array([[-0.5, 1, 1], [-0.51, 1, 1], [nx - 0.5001, 1, 1], [nx - 0.5, 1, 1]])
    r1 = O.query_dense(g, d, "FirstOrder", g.bmin + g.voxel * (f2 + 0.5))
    assert list(r1["valid"]) == [False, False, True, False]


def test_contact_aware_candidate_selection(O):
    g = O.Grid(np.zeros(3), np.ones(3) * 0.4, np.ones(3) * 0.1, (4, 4, 4))
    N = 64
    S = 3
    d = dict(phi=np.zeros(N), n=np.tile([1.0, 0, 0], (N, 1)), H=np.zeros((N, 6)), stride=S,
             ccount=np.zeros(N, np.uint8), cphi=np.zeros((N, S)), cn=np.zeros((N, S, 3)),
             cbranch=-np.ones((N, S), np.int32))
    v = 1 + 4 * (1 + 4 * 1)  # voxel (1,1,1), centre (0.15, 0.15, 0.15)
    d["ccount"][v] = 2
    d["cphi"][v] = [0.02, -0.01, 9.0]
    d["cn"][v] = [[0, 0, 2.0], [1.0, 0, 0], [0, 1.0, 0]]   # first normal not unit
    d["cbranch"][v] = [11, 22, 33]
    p = np.array([[0.17, 0.14, 0.13]])
    r = O.query_dense(g, d, "ContactAware", p)
    gaps = [0.02 + (0.13 - 0.15), -0.01 + (0.17 - 0.15)]   # normalised normals
    k = int(np.argmin(np.abs(gaps)))
    assert r["valid"][0] and np.isclose(r["gap"][0], gaps[k]) and r["branch"][0] == [11, 22][k]
    # zero candidates at the voxel -> first order (SdfOracle.cpp:382-385)
    d["ccount"][v] = 0
    r = O.query_dense(g, d, "ContactAware", p)
    assert np.isclose(r["gap"][0], 0.0 + (0.17 - 0.15)) and r["branch"][0] == -1


# ---------------------------------------------------------------- sparse
def test_sparse_matches_dense_in_band(O):
    v, t = meshes.torus(0.05, 0.01, 48, 12)
    g = O.setup_grid(v, 0.1, 0.0, 48, 48)
    dense = O.build_dense(v, t, g)
    sp = O.build_sparse(v, t, g, block_dim=8, halo=1, band_voxels=3)
    nx, ny, nz = g.shape
    assert 0 < len(sp["bijk"]) < len(sp["active"])
    for b, (bi, bj, bk) in enumerate(sp["bijk"][:20]):
        lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
        i, j, k = bi * 8 + lx.ravel(), bj * 8 + ly.ravel(), bk * 8 + lz.ravel()
        inb = (i < nx) & (j < ny) & (k < nz)
        lin = i[inb] + nx * (j[inb] + ny * k[inb])
        assert np.array_equal(sp["phi"][b][inb], dense["phi"][lin].astype(np.float32))
        assert np.array_equal(sp["n"][b][inb], dense["n"][lin].astype(np.float32))
        assert np.all(sp["phi"][b][~inb] == 0) and np.all(sp["n"][b][~inb, 0] == 1)


def test_sparse_activation_rule_brute_force(O):
    v, t = meshes.uv_sphere(0.05, 10, 20)
    g = O.setup_grid(v, 0.1, 0.0, 32, 32)
    sp = O.build_sparse(v, t, g, block_dim=8, halo=2, band_voxels=2)
    h = g.voxel.max()
    assert sp["band"] == 2 * h
    assert np.isclose(sp["activation"], 2 * h + 2 * h + 0.5 * np.sqrt(3) * 8 * h)
    nbx, nby, nbz = sp["nb_xyz"]
    bl = np.arange(nbx * nby * nbz)
    bi, bj, bk = bl % nbx, (bl // nbx) % nby, bl // (nbx * nby)
    lo = np.stack([bi, bj, bk], 1) * 8
    hi = np.minimum(lo + 8, g.shape)
    centre = g.bmin + g.voxel * 0.5 * (lo + hi)
    sq, _ = O.brute_nearest(v, t, centre)
    wn = O.winding_number(v, t, centre)
    sd = np.sqrt(sq) * np.where(wn > 0.5, -1, 1)
    want = np.abs(sd) <= sp["activation"]
    assert np.array_equal(sp["active"].astype(bool), want)
    # bricks listed in (k, j, i) order
    key = sp["bijk"][:, 0] + nbx * (sp["bijk"][:, 1] + nby * sp["bijk"][:, 2])
    assert np.all(np.diff(key) > 0)


def test_sparse_queries_equal_dense_inside_band(O):
    v, t = meshes.torus(0.05, 0.01, 48, 12)
    g = O.setup_grid(v, 0.1, 0.0, 48, 48)
    dense = O.build_dense(v, t, g)
    sp = O.build_sparse(v, t, g, block_dim=8, halo=2, band_voxels=4)
    from sdfo import scenes
    p = scenes.near_surface_points(v, t, 3000, 0.5 * g.voxel[0], seed=7, dtype=np.float64)
    d32 = dict(phi=dense["phi"].astype(np.float32).astype(np.float64),
               n=dense["n"].astype(np.float32).astype(np.float64), H=dense["H"])
    for mode in ("Trilinear", "Tricubic", "FirstOrder"):
        rs = O.query_sparse(g, sp, mode, p)
        rd = O.query_dense(g, d32, mode, p)
        m = rs["valid"]
        assert m.mean() > 0.95
        assert np.allclose(rs["gap"][m], rd["gap"][m], atol=1e-15, rtol=0)


# --------------------------------------------------------------- contact
def _plane_pose(depth, omega=1.0):
    # body A at z = -depth relative to the plane frame, spinning about z
    R = np.eye(3)[None]
    t = np.array([[0.0, 0.0, -depth]])
    from sdfo import scenes
    return scenes.pose_array(R, t, Om=np.array([[0.0, 0.0, omega]]))


@pytest.mark.parametrize("n_radial", [20, 40])
def test_torsion_closed_form(O, n_radial):
    """Level-0 torsion (ring_on_cube_torsion.py:55-60): uniform pressure on an
    annulus under a spinning body; midpoint-ring quadrature has the exact
    relative error -(Ro-Ri)^3 / (4 n^2 (Ro^3 - Ri^3))."""
    Ri, Ro, mu, kn, pen = 0.4, 0.6, 0.5, 1e6, 0.002
    X, w = meshes.annular_quadrature(Ri, Ro, n_radial, 128)
    out, act = O.contact_forward_plane(0.0, X, w, _plane_pose(pen), kn=kn, c=0.0, mu=mu, reg=0.0,
                                       order=1)
    N = kn * pen * np.pi * (Ro ** 2 - Ri ** 2)
    Tz = -mu * N * (2.0 / 3.0) * (Ro ** 3 - Ri ** 3) / (Ro ** 2 - Ri ** 2)
    quad = -(Ro - Ri) ** 3 / (4 * n_radial ** 2 * (Ro ** 3 - Ri ** 3))
    assert act[0] == len(X)
    assert np.isclose(out[0, 2], N, rtol=1e-13)                 # normal force = p * area
    assert np.isclose(out[0, 5], Tz * (1 + quad), rtol=1e-12)   # torque incl. quadrature error
    assert np.hypot(out[0, 0], out[0, 1]) < 1e-15 * mu * N * 10  # symmetric tangential residual
    if n_radial == 20:
        assert np.isclose(abs(quad), 3.29e-5, rtol=2e-3)         # README.md:84 figure


def test_penalty_damping_closed_form(O):
    """Flat plane, constant penetration, pure normal velocity: F_z = sum w
    (kn delta^2 + c (-v_n)) (contacts_sdf.py:339-362)."""
    X, w = meshes.annular_quadrature(0.1, 0.2, 5, 16)
    w = w / w.sum()
    pen, vz, kn, c = 1e-3, -0.02, 1e9, 1e4
    from sdfo import scenes
    P = scenes.pose_array(np.eye(3)[None], np.array([[0, 0, -pen]]), v0=np.array([[0, 0, vz]]))
    out, _ = O.contact_forward_plane(0.0, X, w, P, kn=kn, c=c, mu=0.0, reg=1e-6, order=2)
    assert np.isclose(out[0, 2], kn * pen ** 2 + c * (-vz), rtol=1e-13)
    assert np.allclose(out[0, [0, 1]], 0, atol=1e-12)


def test_contact_backward_matches_central_differences(O):
    """The oracle's analytic backward against FP64 central differences of its forward."""
    from sdfo import scenes
    v, t = scenes.cube_mesh()
    g = O.setup_grid(v, 0.1, 0.0, 32, 32)
    d = O.build_dense(v, t, g)
    X, w = scenes.ring_quadrature(30, 8)
    R, tt = scenes.ring_poses(3, seed=11, z_range=(0.058, 0.060))
    v0, Om = scenes.pose_velocities(3, seed=12)
    P = scenes.pose_array(R, tt, v0=v0, Om=Om)
    law = dict(kn=1e11, c=1e4, mu=0.3, reg=1e-3, order=2, mode="Tricubic")
    rng = np.random.default_rng(0)
    G = rng.normal(size=(3, 6))
    gt, gp, gx, gw = O.contact_backward(g, d["phi"], d["n"], X, w, P, G, **law)
    L = lambda P_, X_=X, w_=w, **kw: np.sum(  # noqa: E731
        G * O.contact_forward(g, d["phi"], d["n"], X_, w_, P_, **{**law, **kw})[0])

    def twist(P_, xi):
        from sdfo.scenes import rot_axis
        Q = P_.copy()
        for b in range(len(Q)):
            th = np.linalg.norm(xi[b, :3])
            Rr = rot_axis(xi[b, :3], th) if th > 0 else np.eye(3)
            R0 = Q[b, :9].reshape(3, 3)
            Q[b, :9] = (Rr @ R0).ravel()
            Q[b, 12:21] = (Rr @ Q[b, 12:21].reshape(3, 3)).ravel()
            Q[b, 9:12] += xi[b, 3:]
        return Q
    h = 1e-7
    for b in range(3):
        for k in range(6):
            xi = np.zeros((3, 6))
            xi[b, k] = h
            fd = (L(twist(P, xi)) - L(twist(P, -xi))) / (2 * h)
            assert np.isclose(gt[b, k], fd, rtol=1e-4, atol=1e-6 * np.abs(gt).max())
    for name, idx in (("kn", 0), ("c", 1), ("mu", 2)):
        val = law[name]
        fd = (L(P, **{name: val * (1 + 1e-6)}) - L(P, **{name: val * (1 - 1e-6)})) / (2e-6 * val)
        assert np.isclose(gp[idx], fd, rtol=1e-5)
    for i in np.argsort(-np.abs(gx).sum(1))[:5]:
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, k] += h
            Xm[i, k] -= h
            fd = (L(P, Xp) - L(P, Xm)) / (2 * h)
            assert np.isclose(gx[i, k], fd, rtol=1e-4, atol=1e-6 * np.abs(gx).max())
    i = int(np.argmax(np.abs(gw)))
    wp, wm = w.copy(), w.copy()
    wp[i] += 1e-6
    wm[i] -= 1e-6
    assert np.isclose(gw[i], (L(P, X, wp) - L(P, X, wm)) / 2e-6, rtol=1e-6)


# ------------------------------------------------------ linearisation cache
def _cube_contact_scene(O, n_poses=3, seed=11):
    from sdfo import scenes
    v, t = scenes.cube_mesh()
    g = O.setup_grid(v, 0.1, 0.0, 32, 32)
    d = O.build_dense(v, t, g)
    X, w = scenes.ring_quadrature(30, 8)
    R, tt = scenes.ring_poses(n_poses, seed=seed, z_range=(0.058, 0.060))
    v0, Om = scenes.pose_velocities(n_poses, seed=seed + 1)
    return g, d, X, w, scenes.pose_array(R, tt, v0=v0, Om=Om)


@pytest.mark.parametrize("margin", [0.0, 0.059])
def test_linearization_cache_at_reference_pose_is_the_forward(O, margin):
    """contacts_sdf.py:311-327: at the pose the cache was filled at, the
    prediction g0 + grad0 . 0 is the query itself, so the linearised forward
    equals the direct one bit for bit; the cache holds the dense query at
    Y = R X + t behind the narrow filter."""
    g, d, X, w, P = _cube_contact_scene(O)
    law = dict(kn=1e11, c=1e4, mu=0.3, reg=1e-3, order=2, margin=margin)
    for b in range(len(P)):
        g0, grad0, valid0 = O.contact_linearize(g, d["phi"], d["n"], X, P[b], mode="Tricubic",
                                                **law)
        lin, la = O.contact_forward_linearized(g0, grad0, valid0, P[b, 9:12], X, w, P[b:b + 1],
                                               **law)
        ref, ra = O.contact_forward(g, d["phi"], d["n"], X, w, P[b:b + 1], mode="Tricubic", **law)
        assert ra[0] > 0 and np.array_equal(la, ra)
        assert np.array_equal(lin, ref)
        Y = X @ P[b, :9].reshape(3, 3).T + P[b, 9:12]
        q = O.query_dense(g, d, "Tricubic", Y)
        keep = q["valid"] & ((np.abs(Y[:, 2]) < margin) if margin > 0 else True)
        assert np.array_equal(valid0.astype(bool), keep)
        assert np.allclose(g0[keep], q["gap"][keep], rtol=0, atol=1e-13)
        assert np.allclose(grad0[keep], q["grad"][keep], rtol=0, atol=1e-11)
        assert not g0[~keep].any() and not grad0[~keep].any()


def test_linearization_exact_on_linear_field(O):
    """On a linear field the first-order prediction is exact: translated poses
    give the direct forward up to rounding (a wrong sign or a dropped
    component of grad0 . (t - t_ref) moves the gaps by 2e-3)."""
    a = np.array([0.1, -0.2, 1.0])
    f = lambda c: c @ a - 0.1  # noqa: E731
    df = lambda c: np.tile(a, (len(c), 1))  # noqa: E731
    g, d = _field_grid(O, f, df, n=16)
    from sdfo import scenes
    X, w = meshes.annular_quadrature(0.05, 0.15, 6, 24)
    R0 = scenes.rot_axis(np.array([1.0, 0.5, 0.0]), 0.05)[None]
    t0 = np.array([[0.1, 0.15, 0.095]])
    P0 = scenes.pose_array(R0, t0, Om=np.array([[0.1, -0.2, 1.5]]))
    law = dict(kn=1e8, c=1e3, mu=0.3, reg=1e-4, order=2)
    g0, grad0, valid0 = O.contact_linearize(g, d["phi"], d["n"], X, P0[0], mode="Tricubic", **law)
    assert valid0.all()
    rng = np.random.default_rng(3)
    P = np.repeat(P0, 6, axis=0)
    P[:, 9:12] += rng.uniform(-2e-3, 2e-3, (6, 3))
    lin, la = O.contact_forward_linearized(g0, grad0, valid0, P0[0, 9:12], X, w, P, **law)
    ref, ra = O.contact_forward(g, d["phi"], d["n"], X, w, P, mode="Tricubic", **law)
    assert np.array_equal(la, ra) and (ra > 0).all() and (ra < len(X)).any()
    assert np.allclose(lin, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_linearization_ignores_rotation_change(O):
    """The cache predicts from the translation only (contacts_sdf.py:312):
    a pose that differs from the reference by a rotation keeps every cached
    gap, so the contact set is the reference pose's."""
    g, d, X, w, P = _cube_contact_scene(O, n_poses=1)
    law = dict(kn=1e11, c=0.0, mu=0.0, reg=1e-3, order=1)
    g0, grad0, valid0 = O.contact_linearize(g, d["phi"], d["n"], X, P[0], mode="Tricubic", **law)
    from sdfo.scenes import rot_axis
    Q = P.copy()
    Q[0, :9] = (rot_axis(np.array([1.0, 0.0, 0.0]), 0.05) @ P[0, :9].reshape(3, 3)).ravel()
    lin, la = O.contact_forward_linearized(g0, grad0, valid0, P[0, 9:12], X, w, Q, **law)
    _, ra = O.contact_forward(g, d["phi"], d["n"], X, w, P, mode="Tricubic", **law)
    _, rq = O.contact_forward(g, d["phi"], d["n"], X, w, Q, mode="Tricubic", **law)
    assert la[0] == ra[0] and rq[0] != ra[0]
    # order 1, no damping or friction: F = sum w kn max(-g0, 0) grad0
    pen = np.maximum(-g0, 0.0) * valid0
    assert np.allclose(lin[0, :3], (1e11 * w * pen) @ grad0, rtol=1e-12)


# --------------------------------------------------------------- broad phase
def test_point_bvh_known_answer(O):
    """Four points on a line, leaf size 1 (bvh.py:27-75 by hand): splits at
    x = 1.5, 0.5, 2.5; a box from x = 0.5 rejects the leaf {0}; leaves come
    out right subtree first; an enclosing box returns arange."""
    P = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]])
    nodes = O.point_bvh_build(P, leaf_size=1)
    assert len(nodes) == 7
    assert O.point_bvh_query(nodes, [0.5, -1, -1], [3.5, 1, 1]).tolist() == [3, 2, 1]
    assert O.point_bvh_query(nodes, [-0.5, -1, -1], [2.2, 1, 1]).tolist() == [2, 1, 0]
    assert O.point_bvh_query(nodes, [-1, -1, -1], [4, 1, 1]).tolist() == [0, 1, 2, 3]
    assert O.point_bvh_query(nodes, [0.5, -1, -1], [3.5, 1, 1], margin=0.6).tolist() == [3, 2, 1, 0]
    assert O.point_bvh_query(nodes, [5, 5, 5], [6, 6, 6]).tolist() == []
    # identical points: every split is the median split of the stable order
    same = O.point_bvh_build(np.zeros((5, 3)), leaf_size=2)
    leaves = [nd["idx"].tolist() for nd in same if nd["left"] < 0]
    assert leaves == [[0, 1], [2], [3, 4]]


@pytest.mark.parametrize("seed,leaf", [(0, 16), (1, 4), (2, 1)])
def test_point_bvh_invariants(O, seed, leaf):
    """Brute force over the tree: leaves partition the points, node boxes are
    the min/max of their points, midpoint splits separate the children, and a
    query returns every point inside the box and only points of leaves whose
    box overlaps it."""
    rng = np.random.default_rng(seed)
    P = np.concatenate([rng.normal(size=(300, 3)) * [1, 0.3, 0.1],
                        rng.uniform(-1, 1, (200, 3))])
    P[::50] = P[1]                                     # duplicates
    nodes = O.point_bvh_build(P, leaf_size=leaf)

    def pts(i):
        nd = nodes[i]
        return nd["idx"] if nd["left"] < 0 else np.concatenate([pts(nd["left"]), pts(nd["right"])])
    allp = np.sort(np.concatenate([nd["idx"] for nd in nodes if nd["left"] < 0]))
    assert np.array_equal(allp, np.arange(len(P)))
    for i, nd in enumerate(nodes):
        idx = pts(i)
        assert np.array_equal(nd["lo"], P[idx].min(0)) and np.array_equal(nd["hi"], P[idx].max(0))
        if nd["left"] < 0:
            assert len(idx) <= leaf
            continue
        a = int(np.argmax(nd["hi"] - nd["lo"]))
        mid = (nd["lo"][a] + nd["hi"][a]) / 2
        L, R = pts(nd["left"]), pts(nd["right"])
        assert len(L) + len(R) == len(idx)
        if (P[idx, a] <= mid).all() or (P[idx, a] > mid).all():
            assert abs(len(L) - len(R)) <= 1           # median split
        else:
            assert (P[L, a] <= mid).all() and (P[R, a] > mid).all()
    leafs = [nd for nd in nodes if nd["left"] < 0]
    for _ in range(20):
        c = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.05, 0.6, 3)
        q = O.point_bvh_query(nodes, c - h, c + h)
        assert len(set(q.tolist())) == len(q)
        inside = np.nonzero(((P >= c - h) & (P <= c + h)).all(1))[0]
        assert set(inside.tolist()) <= set(q.tolist())
        hit = [nd for nd in leafs if (nd["hi"] >= c - h).all() and (nd["lo"] <= c + h).all()]
        want = np.concatenate([nd["idx"] for nd in hit]) if hit else np.zeros(0)
        assert sorted(q.tolist()) == sorted(want.tolist()) or len(q) == len(P)


def test_broad_phase_boxes_closed_form(O):
    """bmin_local = R^T (bmin - t) - margin per corner (contacts_sdf.py:252-258),
    a quarter turn about z swapping x and y with a sign."""
    c, s = 0.0, 1.0
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    from sdfo import scenes
    P = scenes.pose_array(np.stack([np.eye(3), R]), np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3]]))
    b = O.broad_phase_boxes(P, [-1, -2, -3], [1, 2, 3], 0.05)
    assert np.allclose(b[0], [-1.15, -2.25, -3.35, 0.95, 1.85, 2.75], atol=1e-15)
    # R^T (x, y, z) = (y, -x, z): corner-wise, so qmin.y > qmax.y here
    assert np.allclose(b[1], [-2.25, 1.05, -3.35, 1.85, -0.85, 2.75], atol=1e-15)
    nodes = O.point_bvh_build(np.zeros((3, 3)) + 10.0)
    assert O.culled_points(nodes, b[0], 3).tolist() == [0, 1, 2]     # empty cull -> all
