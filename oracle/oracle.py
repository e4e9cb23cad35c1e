"""ctypes front end of the CPU FP64 oracle (oracle/sdfo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, as the checker.  The product library
(sdfo) never imports this module.

Arrays follow the reference's layouts: dense fields are x-fastest
(``i + nx*(j + ny*k)``, GridSpec.h:56-59), normals are (N,3), Hessians are
(N,6) = Hxx, Hxy, Hxz, Hyy, Hyz, Hzz (SdfIO.cpp:140-146).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsdfo_oracle.so")
_lib = None

MODES = {"FirstOrder": 0, "SecondOrder": 1, "Trilinear": 2, "Tricubic": 3, "ContactAware": 4}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP, -ffp-contract=off)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.oracle_build_sparse.restype = C.c_void_p
        _lib.oracle_build_contact_aware_sparse.restype = C.c_void_p
        _lib.oracle_sparse_info.argtypes = [C.c_void_p, _ip, _ip, _dp]
        _lib.oracle_sparse_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 9
        _lib.oracle_sparse_free.argtypes = [C.c_void_p]
    return _lib


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class Grid:
    bmin: np.ndarray
    bmax: np.ndarray
    voxel: np.ndarray
    shape: tuple

    @property
    def grid9(self):
        return np.concatenate([self.bmin, self.bmax, self.voxel]).astype(np.float64)

    @property
    def shape3(self):
        return np.asarray(self.shape, dtype=np.int32)

    @property
    def total(self):
        return int(np.prod(self.shape))

    def centers(self):
        nx, ny, nz = self.shape
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
        return self.bmin[None, :] + self.voxel[None, :] * (idx + 0.5)


def num_threads() -> int:
    return lib().oracle_num_threads()


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def setup_grid(verts, padding=0.1, voxel_size=0.0, target=128, max_res=512) -> Grid:
    v = _f64(verts)
    g9 = np.zeros(9)
    sh = np.zeros(3, np.int32)
    lib().oracle_setup_grid(_p(v), C.c_int64(len(v)), C.c_double(padding), C.c_double(voxel_size),
                            C.c_int(target), C.c_int(max_res), _p(g9), _p(sh, _ip))
    return Grid(g9[0:3].copy(), g9[3:6].copy(), g9[6:9].copy(), tuple(int(x) for x in sh))


def snap_resolution(n, max_res):
    return lib().oracle_snap_resolution(int(n), int(max_res))


def closest_point(p, a, b, c):
    out = np.zeros(3)
    lib().oracle_closest_point(_p(_f64(p)), _p(_f64(a)), _p(_f64(b)), _p(_f64(c)), _p(out))
    return out


def ray_triangle(o, a, b, c) -> bool:
    return bool(lib().oracle_ray_triangle(_p(_f64(o)), _p(_f64(a)), _p(_f64(b)), _p(_f64(c))))


def bvh_export(verts, tris):
    v, t = _f64(verts), _i32(tris)
    nt = len(t)
    cap = 2 * nt + 1
    n6 = np.zeros((cap, 6))
    i4 = np.zeros((cap, 4), np.int32)
    perm = np.zeros(nt, np.int32)
    nn = lib().oracle_bvh_export(_p(v), _p(t, _ip), C.c_int(nt), _p(n6), _p(i4, _ip),
                                 _p(perm, _ip), C.c_int(cap))
    return n6[:nn].copy(), i4[:nn].copy(), perm


def signed_nearest(verts, tris, pts):
    v, t, p = _f64(verts), _i32(tris), _f64(pts, (-1, 3))
    n = len(p)
    out = dict(sd=np.zeros(n), cp=np.zeros((n, 3)), tri=np.zeros(n, np.int32),
               face=np.zeros((n, 3)), hits=np.zeros(n, np.int32),
               tri_tests=np.zeros(n, np.int64), box_tests=np.zeros(n, np.int64))
    lib().oracle_signed_nearest(_p(v), _p(t, _ip), C.c_int(len(t)), _p(p), C.c_int64(n),
                                _p(out["sd"]), _p(out["cp"]), _p(out["tri"], _ip),
                                _p(out["face"]), _p(out["hits"], _ip),
                                _p(out["tri_tests"], _i64p), _p(out["box_tests"], _i64p))
    return out


def brute_nearest(verts, tris, pts):
    v, t, p = _f64(verts), _i32(tris), _f64(pts, (-1, 3))
    sq = np.zeros(len(p))
    tri = np.zeros(len(p), np.int32)
    lib().oracle_brute_nearest(_p(v), _p(t, _ip), C.c_int(len(t)), _p(p), C.c_int64(len(p)),
                               _p(sq), _p(tri, _ip))
    return sq, tri


def winding_number(verts, tris, pts):
    v, t, p = _f64(verts), _i32(tris), _f64(pts, (-1, 3))
    wn = np.zeros(len(p))
    lib().oracle_winding_number(_p(v), _p(t, _ip), C.c_int(len(t)), _p(p), C.c_int64(len(p)),
                                _p(wn))
    return wn


def build_dense(verts, tris, grid: Grid, counts: bool = False):
    v, t = _f64(verts), _i32(tris)
    N = grid.total
    out = dict(phi=np.zeros(N), n=np.zeros((N, 3)), H=np.zeros((N, 6)),
               witness=np.zeros((N, 3)), tri=np.zeros(N, np.int32))
    tt = np.zeros(N, np.int64) if counts else None
    bt = np.zeros(N, np.int64) if counts else None
    rc = lib().oracle_build_dense(_p(v), C.c_int64(len(v)), _p(t, _ip), C.c_int(len(t)),
                                  _p(grid.grid9), _p(grid.shape3, _ip), _p(out["phi"]),
                                  _p(out["n"]), _p(out["H"]), _p(out["witness"]),
                                  _p(out["tri"], _ip), _p(tt, _i64p), _p(bt, _i64p))
    if rc != 0:
        raise RuntimeError("Empty mesh provided to SdfBuilder")
    if counts:
        out["tri_tests"], out["box_tests"] = tt, bt
    return out


def dense_hessian(grid: Grid, n3):
    n3 = _f64(n3, (-1, 3))
    H = np.zeros((len(n3), 6))
    lib().oracle_dense_hessian(_p(grid.grid9), _p(grid.shape3, _ip), _p(n3), _p(H))
    return H


def _sparse_collect(h, block_dim, with_hessian):
    L = lib()
    nb = C.c_int()
    nbxyz = np.zeros(3, np.int32)
    ba = np.zeros(2)
    stride = L.oracle_sparse_info(h, C.byref(nb), _p(nbxyz, _ip), _p(ba))
    nb = nb.value
    bvc = block_dim ** 3
    out = dict(active=np.zeros(int(np.prod(nbxyz)), np.uint8), bijk=np.zeros((nb, 3), np.int32),
               phi=np.zeros((nb, bvc), np.float32), n=np.zeros((nb, bvc, 3), np.float32),
               H=np.zeros((nb, bvc, 6), np.float32) if with_hessian else None,
               nb_xyz=tuple(int(x) for x in nbxyz), band=float(ba[0]), activation=float(ba[1]),
               stride=int(stride), block_dim=block_dim)
    if stride > 0:
        out.update(ccount=np.zeros((nb, bvc), np.uint8),
                   cphi=np.zeros((nb, bvc, stride), np.float32),
                   cn=np.zeros((nb, bvc, stride, 3), np.float32),
                   cbranch=np.zeros((nb, bvc, stride), np.int32))
    else:
        out.update(ccount=None, cphi=None, cn=None, cbranch=None)
    vp = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)  # noqa: E731
    L.oracle_sparse_copy(h, vp(out["active"]), vp(out["bijk"]), vp(out["phi"]), vp(out["n"]),
                         vp(out["H"]), vp(out["ccount"]), vp(out["cphi"]), vp(out["cn"]),
                         vp(out["cbranch"]))
    L.oracle_sparse_free(h)
    return out


def build_sparse(verts, tris, grid: Grid, block_dim=8, halo=2, band_voxels=8, band_distance=0.0,
                 store_hessian=True):
    v, t = _f64(verts), _i32(tris)
    h = lib().oracle_build_sparse(_p(v), C.c_int64(len(v)), _p(t, _ip), C.c_int(len(t)),
                                  _p(grid.grid9), _p(grid.shape3, _ip), C.c_int(block_dim),
                                  C.c_int(halo), C.c_int(band_voxels), C.c_double(band_distance),
                                  C.c_int(1 if store_hessian else 0))
    if not h:
        raise RuntimeError("Sparse SDF build rejected its inputs")
    return _sparse_collect(h, block_dim, store_hessian)


def build_contact_aware_dense(verts, tris, corner_normals, grid: Grid, max_candidates=4,
                              feature_tol=0.0, merge_deg=5.0, tile_dim=8, use_tiles=True,
                              store_hessian=True):
    v, t = _f64(verts), _i32(tris)
    cn = _f64(corner_normals, (-1, 3, 3))
    N = grid.total
    S = max(0, min(int(max_candidates), 8))
    out = dict(phi=np.zeros(N), n=np.zeros((N, 3)), H=np.zeros((N, 6)), witness=np.zeros((N, 3)),
               ccount=np.zeros(N, np.uint8), cphi=np.zeros((N, max(S, 1))),
               cn=np.zeros((N, max(S, 1), 3)), cbranch=np.zeros((N, max(S, 1)), np.int32))
    rc = lib().oracle_build_contact_aware_dense(
        _p(v), C.c_int64(len(v)), _p(t, _ip), C.c_int(len(t)), _p(cn), _p(grid.grid9),
        _p(grid.shape3, _ip), C.c_int(max_candidates), C.c_double(feature_tol),
        C.c_double(merge_deg), C.c_int(tile_dim), C.c_int(1 if use_tiles else 0),
        C.c_int(1 if store_hessian else 0), _p(out["phi"]), _p(out["n"]), _p(out["H"]),
        _p(out["witness"]), _p(out["ccount"], _u8p), _p(out["cphi"]), _p(out["cn"]),
        _p(out["cbranch"], _ip))
    if rc < 0:
        raise RuntimeError("contact-aware build rejected its inputs (%d)" % rc)
    out["stride"] = rc
    return out


def build_contact_aware_sparse(verts, tris, corner_normals, grid: Grid, max_candidates=4,
                               feature_tol=0.0, merge_deg=5.0, tile_dim=8, use_tiles=True,
                               store_hessian=True, block_dim=8, halo=2, band_voxels=8,
                               band_distance=0.0):
    v, t = _f64(verts), _i32(tris)
    cn = _f64(corner_normals, (-1, 3, 3))
    h = lib().oracle_build_contact_aware_sparse(
        _p(v), C.c_int64(len(v)), _p(t, _ip), C.c_int(len(t)), _p(cn), _p(grid.grid9),
        _p(grid.shape3, _ip), C.c_int(max_candidates), C.c_double(feature_tol),
        C.c_double(merge_deg), C.c_int(tile_dim), C.c_int(1 if use_tiles else 0),
        C.c_int(1 if store_hessian else 0), C.c_int(block_dim), C.c_int(halo),
        C.c_int(band_voxels), C.c_double(band_distance))
    if not h:
        raise RuntimeError("sparse contact-aware build rejected its inputs")
    return _sparse_collect(h, block_dim, store_hessian)


def query_dense(grid: Grid, data: dict, mode, pts, hessian=False):
    mode = MODES.get(mode, mode)
    p = _f64(pts, (-1, 3))
    n = len(p)
    phi = _f64(data["phi"])
    n3 = _f64(data["n"], (-1, 3))
    H = _f64(data["H"], (-1, 6)) if data.get("H") is not None else None
    stride = int(data.get("stride", 0) or 0)
    cc = cp = cn = cb = None
    if stride > 0:
        cc = np.ascontiguousarray(data["ccount"], np.uint8)
        cp = _f64(data["cphi"])
        cn = _f64(data["cn"])
        cb = _i32(data["cbranch"])
    out = dict(gap=np.zeros(n), grad=np.zeros((n, 3)), normal=np.zeros((n, 3)),
               witness=np.zeros((n, 3)), valid=np.zeros(n, np.uint8),
               branch=np.zeros(n, np.int32), hess=np.zeros((n, 6)) if hessian else None)
    lib().oracle_query_dense(_p(grid.grid9), _p(grid.shape3, _ip), _p(phi), _p(n3), _p(H),
                             C.c_int(stride), _p(cc, _u8p), _p(cp), _p(cn), _p(cb, _ip),
                             C.c_int(mode), _p(p), C.c_int64(n), _p(out["gap"]),
                             _p(out["grad"]), _p(out["normal"]), _p(out["witness"]),
                             _p(out["valid"], _u8p), _p(out["branch"], _ip), _p(out["hess"]))
    out["valid"] = out["valid"].astype(bool)
    return out


def query_sparse(grid: Grid, sp: dict, mode, pts):
    mode = MODES.get(mode, mode)
    p = _f64(pts, (-1, 3))
    n = len(p)
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)  # noqa: E731
    stride = int(sp.get("stride", 0) or 0)
    cc = np.ascontiguousarray(sp["ccount"], np.uint8) if stride > 0 else None
    cb = _i32(sp["cbranch"]) if stride > 0 else None
    out = dict(gap=np.zeros(n), grad=np.zeros((n, 3)), normal=np.zeros((n, 3)),
               witness=np.zeros((n, 3)), valid=np.zeros(n, np.uint8), branch=np.zeros(n, np.int32))
    lib().oracle_query_sparse(
        _p(grid.grid9), _p(grid.shape3, _ip), C.c_int(sp["block_dim"]), C.c_int(len(sp["bijk"])),
        _p(_i32(sp["bijk"]), _ip), _p(f32(sp["phi"]), _fp), _p(f32(sp["n"]), _fp),
        _p(f32(sp.get("H")), _fp), C.c_int(stride), _p(cc, _u8p),
        _p(f32(sp.get("cphi")) if stride > 0 else None, _fp),
        _p(f32(sp.get("cn")) if stride > 0 else None, _fp), _p(cb, _ip), C.c_int(mode), _p(p),
        C.c_int64(n), _p(out["gap"]), _p(out["grad"]), _p(out["normal"]), _p(out["witness"]),
        _p(out["valid"], _u8p), _p(out["branch"], _ip))
    out["valid"] = out["valid"].astype(bool)
    return out


def _contact_params(kn, c, mu, reg, margin, order, mode):
    return (np.array([kn, c, mu, reg, margin, 0.0], np.float64),
            np.array([int(order), MODES.get(mode, mode)], np.int32))


def contact_forward(grid: Grid, phi, n3, X, w, poses, kn=1e11, c=1e4, mu=0.0, reg=1e-6,
                    margin=0.0, order=2, mode="Tricubic"):
    """Per-pose (F_A, tau_A) in the SDF frame; poses are (B,30) records."""
    X = _f64(X, (-1, 3))
    w = _f64(w)
    P = _f64(poses, (-1, 30))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, mode)
    out = np.zeros((len(P), 6))
    act = np.zeros(len(P), np.int32)
    lib().oracle_contact_forward(_p(grid.grid9), _p(grid.shape3, _ip), _p(_f64(phi)),
                                 _p(_f64(n3, (-1, 3))), _p(X), _p(w), C.c_int64(len(X)), _p(P),
                                 C.c_int(len(P)), _p(prm), _p(iprm, _ip), _p(out), _p(act, _ip))
    return out, act


def contact_backward(grid: Grid, phi, n3, X, w, poses, g_out, kn=1e11, c=1e4, mu=0.0, reg=1e-6,
                     margin=0.0, order=2, mode="Tricubic"):
    """Returns (g_twist (B,6) [omega, dt], g_params (kn, c, mu), g_X (N,3), g_w (N))."""
    X = _f64(X, (-1, 3))
    w = _f64(w)
    P = _f64(poses, (-1, 30))
    G = _f64(g_out, (-1, 6))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, mode)
    gt = np.zeros((len(P), 6))
    gp = np.zeros(4)
    gx = np.zeros((len(X), 3))
    gw = np.zeros(len(X))
    lib().oracle_contact_backward(_p(grid.grid9), _p(grid.shape3, _ip), _p(_f64(phi)),
                                  _p(_f64(n3, (-1, 3))), _p(X), _p(w), C.c_int64(len(X)), _p(P),
                                  C.c_int(len(P)), _p(prm), _p(iprm, _ip), _p(G), _p(gt), _p(gp),
                                  _p(gx), _p(gw))
    return gt, gp[:3], gx, gw


def contact_linearize(grid: Grid, phi, n3, X, pose, kn=1e11, c=1e4, mu=0.0, reg=1e-6, margin=0.0,
                      order=2, mode="Tricubic"):
    """Linearisation cache at one reference pose (contacts_sdf.py:319-327):
    (g0 (N,), grad0 (N,3), valid0 (N,) uint8)."""
    X = _f64(X, (-1, 3))
    P = _f64(pose, (30,))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, mode)
    n = len(X)
    g0, grad0, valid0 = np.zeros(n), np.zeros((n, 3)), np.zeros(n, np.uint8)
    lib().oracle_contact_linearize(_p(grid.grid9), _p(grid.shape3, _ip), _p(_f64(phi)),
                                   _p(_f64(n3, (-1, 3))), _p(X), C.c_int64(n), _p(P), _p(prm),
                                   _p(iprm, _ip), _p(g0), _p(grad0), _p(valid0, _u8p))
    return g0, grad0, valid0


def contact_forward_linearized(g0, grad0, valid0, t_ref, X, w, poses, kn=1e11, c=1e4, mu=0.0,
                               reg=1e-6, margin=0.0, order=2):
    """Forward on the cached prediction g0 + grad0 . (t - t_ref)
    (contacts_sdf.py:311-318)."""
    X = _f64(X, (-1, 3))
    w = _f64(w)
    P = _f64(poses, (-1, 30))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, 3)
    out = np.zeros((len(P), 6))
    act = np.zeros(len(P), np.int32)
    lib().oracle_contact_forward_linearized(
        _p(_f64(g0)), _p(_f64(grad0, (-1, 3))), _p(np.ascontiguousarray(valid0, np.uint8), _u8p),
        _p(_f64(t_ref, (3,))), _p(X), _p(w), C.c_int64(len(X)), _p(P), C.c_int(len(P)), _p(prm),
        _p(iprm, _ip), _p(out), _p(act, _ip))
    return out, act


# --------------------------------------------------------------- broad phase
def point_bvh_build(points, leaf_size=16):
    """AABB_BVH._build (prototype/src/geometry/bvh.py:27-54) as a node list:
    dict(lo, hi, left, right, idx); node 0 is the root, children after their
    parent, left subtree first.  The degenerate-split sort is stable."""
    P = _f64(points, (-1, 3))
    nodes = []

    def build(idx):
        me = len(nodes)
        nodes.append(None)
        pts = P[idx]
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        if len(idx) <= leaf_size:
            nodes[me] = dict(lo=lo, hi=hi, left=-1, right=-1, idx=idx.copy())
            return me
        axis = int(np.argmax(hi - lo))
        mid = (lo[axis] + hi[axis]) * 0.5
        left, right = idx[pts[:, axis] <= mid], idx[pts[:, axis] > mid]
        if len(left) == 0 or len(right) == 0:
            s = np.argsort(pts[:, axis], kind="stable")
            half = len(idx) // 2
            left, right = idx[s[:half]], idx[s[half:]]
        li = build(left)
        ri = build(right)
        nodes[me] = dict(lo=lo, hi=hi, left=li, right=ri, idx=None)
        return me

    if len(P):
        build(np.arange(len(P)))
    return nodes


def point_bvh_query(nodes, qmin, qmax, margin=0.0):
    """AABB_BVH.query (bvh.py:56-75): point indices in the reference's order."""
    qmin, qmax = _f64(qmin, (3,)), _f64(qmax, (3,))
    if not nodes:
        return np.zeros(0, np.int32)
    r = nodes[0]
    n = sum(len(nd["idx"]) for nd in nodes if nd["left"] < 0)
    if (qmin <= r["lo"] - margin).all() and (qmax >= r["hi"] + margin).all():
        return np.arange(n, dtype=np.int32)
    stack, parts = [0], []
    while stack:
        nd = nodes[stack.pop()]
        if np.any(nd["hi"] + margin < qmin) or np.any(nd["lo"] - margin > qmax):
            continue
        if nd["left"] < 0:
            parts.append(nd["idx"])
        else:
            stack.append(nd["left"])
            stack.append(nd["right"])
    return np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)


def broad_phase_boxes(poses, bmin, bmax, margin):
    """The cull box of SDFContactEngine.evaluate with markers
    (contacts_sdf.py:247-258): the SDF box corners mapped into A's frame by
    the inverse of Y = R X + t, grown by margin; (B,6) = (qmin, qmax)."""
    P = _f64(poses, (-1, 30))
    out = np.zeros((len(P), 6))
    for b in range(len(P)):
        R_ab, t_ab = P[b, :9].reshape(3, 3), P[b, 9:12]
        R_ba = R_ab.T
        t_ba = -R_ab.T @ t_ab
        out[b, :3] = R_ba @ _f64(bmin, (3,)) + t_ba - margin
        out[b, 3:] = R_ba @ _f64(bmax, (3,)) + t_ba + margin
    return out


def culled_points(nodes, box6, n):
    """bvh_indices of evaluate (contacts_sdf.py:266-271): the query's points,
    or every point when the query returns none."""
    idx = point_bvh_query(nodes, box6[:3], box6[3:])
    return idx if len(idx) else np.arange(n, dtype=np.int32)


def contact_forward_plane(z_top, X, w, poses, kn=1e6, c=0.0, mu=0.5, reg=0.0, margin=0.0,
                          order=1):
    X = _f64(X, (-1, 3))
    w = _f64(w)
    P = _f64(poses, (-1, 30))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, 3)
    out = np.zeros((len(P), 6))
    act = np.zeros(len(P), np.int32)
    lib().oracle_contact_forward_plane(C.c_double(z_top), _p(X), _p(w), C.c_int64(len(X)), _p(P),
                                       C.c_int(len(P)), _p(prm), _p(iprm, _ip), _p(out),
                                       _p(act, _ip))
    return out, act


def sort_by_abs(values):
    """Index order of std::sort with |a| < |b| (the contact-aware candidate sort)."""
    v = _f64(values)
    order = np.zeros(len(v), np.int32)
    lib().oracle_sort_by_abs(_p(v), C.c_int(len(v)), _p(order, _ip))
    return order


# --------------------------------------------------------- evaluate (world)
def pair_states(rA, RA, vA, wA, rB, RB, vB, wB):
    """(B,36) records rA RA vA wA rB RB vB wB of SDFContactEngine.evaluate's
    body states (contacts_sdf.py:231-240)."""
    parts = [rA, RA, vA, wA, rB, RB, vB, wB]
    B = len(np.asarray(rA).reshape(-1, 3))
    return np.concatenate([np.asarray(a, np.float64).reshape(B, -1) for a in parts], axis=1)


def _state_fields(s):
    s = _f64(s, (36,))
    return (s[0:3], s[3:12].reshape(3, 3), s[12:15], s[15:18], s[18:21], s[21:30].reshape(3, 3),
            s[30:33], s[33:36])


def evaluate_keep_mask(nodes, states, frames, bmin, bmax, narrow_margin, n):
    """Broad-phase cull of evaluate (contacts_sdf.py:243-271): the SDF box
    (bmin, bmax) mapped into the quadrature frame, grown by max(narrow, 0.05),
    queried in the point tree; every point when nothing overlaps.  (B,N) uint8."""
    S = _f64(states, (-1, 36))
    margin = max(narrow_margin, 0.05)
    bmin, bmax = _f64(bmin, (3,)), _f64(bmax, (3,))
    keep = np.zeros((len(S), n), np.uint8)
    for b in range(len(S)):
        rA, RA, _, _, _, _, _, _ = _state_fields(S[b])
        if frames is not None:
            f = _f64(frames, (-1, 24))[b]
            R_ab, t_ab = f[:9].reshape(3, 3), f[9:12]
            R_ba, t_ba = R_ab.T, -R_ab.T @ t_ab
            lo, hi = R_ba @ bmin + t_ba - margin, R_ba @ bmax + t_ba + margin
        else:
            lo, hi = RA.T @ (bmin - rA) - margin, RA.T @ (bmax - rA) + margin
        idx = point_bvh_query(nodes, lo, hi)
        if len(idx) == 0:
            idx = np.arange(n)
        keep[b, idx] = 1
    return keep


def evaluate_body_check(states, center_a, half_a, half_b):
    """_check_broad_phase (contacts_sdf.py:68-89): 1 where A's rotated local
    AABB around rA overlaps B's half extents around rB.  (B,) uint8."""
    S = _f64(states, (-1, 36))
    out = np.zeros(len(S), np.uint8)
    for b in range(len(S)):
        rA, RA, _, _, rB, _, _, _ = _state_fields(S[b])
        c = rA + RA @ _f64(center_a, (3,))
        h = np.sum(np.abs(RA) * _f64(half_a, (3,))[None, :], axis=1)
        diff = c - rB
        out[b] = 1 if all(abs(diff[i]) <= h[i] + half_b[i] for i in range(3)) else 0
    return out


def contact_evaluate(grid: Grid, phi, n3, X, w, states, frames=None, keep=None, pair_on=None,
                     kn=1e11, c=1e4, mu=0.0, reg=1e-6, margin=0.01, order=2, mode="Tricubic"):
    """SDFContactEngine.evaluate in the world frame: (B,12) = F_A, tau_A, F_B,
    tau_B per pair, and the active-point counts."""
    X = _f64(X, (-1, 3))
    w = _f64(w)
    S = _f64(states, (-1, 36))
    Fr = None if frames is None else _f64(frames, (-1, 24))
    K = None if keep is None else np.ascontiguousarray(keep, np.uint8).reshape(len(S), len(X))
    On = None if pair_on is None else np.ascontiguousarray(pair_on, np.uint8).reshape(len(S))
    prm, iprm = _contact_params(kn, c, mu, reg, margin, order, mode)
    out = np.zeros((len(S), 12))
    act = np.zeros(len(S), np.int32)
    lib().oracle_contact_evaluate(_p(grid.grid9), _p(grid.shape3, _ip), _p(_f64(phi)),
                                  _p(_f64(n3, (-1, 3))), _p(X), _p(w), C.c_int64(len(X)), _p(S),
                                  _p(Fr), C.c_int(len(S)), _p(prm), _p(iprm, _ip), _p(K, _u8p),
                                  _p(On, _u8p), _p(out), _p(act, _ip))
    return out, act


def sparse_hessian(grid: Grid, block_dim, bijk, n3):
    """Sparse Hessian stencil (SdfBuilder.cpp:1154-1225) of FP32 brick normals
    n3 (nb, B^3, 3) -> (nb, B^3, 6) FP32."""
    bijk = _i32(bijk).reshape(-1, 3)
    n3 = np.ascontiguousarray(n3, np.float32).reshape(len(bijk), block_dim ** 3, 3)
    H = np.zeros((len(bijk), block_dim ** 3, 6), np.float32)
    lib().oracle_sparse_hessian(_p(grid.grid9), _p(grid.shape3, _ip), C.c_int(block_dim),
                                C.c_int(len(bijk)), _p(bijk, _ip), _p(n3, _fp), _p(H, _fp))
    return H


def contact_aware_probe(verts, tris, corner_normals, grid: Grid, ijk, max_candidates=4,
                        feature_tol=0.0, merge_deg=5.0, tile_dim=8, use_tiles=True,
                        sparse_block_dim=0, halo=2, band_voxels=8, band_distance=0.0):
    """Contact-aware candidates at the voxels ijk (n,3) as the dense
    (sparse_block_dim = 0) or sparse build computes them, and for the sparse
    build its brick activation mask: dict(ccount, cphi, cn, cbranch, mask)."""
    v, t = _f64(verts), _i32(tris)
    cn = _f64(corner_normals, (-1, 3, 3))
    ijk = _i32(ijk).reshape(-1, 3)
    n = len(ijk)
    S = max(0, min(int(max_candidates), 8))
    out = dict(ccount=np.zeros(n, np.uint8), cphi=np.zeros((n, S)), cn=np.zeros((n, S, 3)),
               cbranch=np.zeros((n, S), np.int32), mask=None)
    if sparse_block_dim > 0:
        nb = [(s + sparse_block_dim - 1) // sparse_block_dim for s in grid.shape]
        out["mask"] = np.zeros(int(np.prod(nb)), np.uint8)
    rc = lib().oracle_contact_aware_probe(
        _p(v), C.c_int64(len(v)), _p(t, _ip), C.c_int(len(t)), _p(cn), _p(grid.grid9),
        _p(grid.shape3, _ip), C.c_int(max_candidates), C.c_double(feature_tol),
        C.c_double(merge_deg), C.c_int(tile_dim), C.c_int(1 if use_tiles else 0),
        C.c_int(sparse_block_dim), C.c_int(halo), C.c_int(band_voxels), C.c_double(band_distance),
        _p(ijk, _ip), C.c_int64(n), _p(out["ccount"], _u8p), _p(out["cphi"]), _p(out["cn"]),
        _p(out["cbranch"], _ip), _p(out["mask"], _u8p))
    if rc < 0:
        raise RuntimeError("contact-aware probe rejected its inputs")
    return out
