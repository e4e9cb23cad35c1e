"""Seeded synthetic workloads for the SDF build / query / contact path.

Ring-on-cube contact (torus quadrature against a cube SDF), benchmark point
streams and the demo's LCG points.  All generators are numpy + a seed, so the
same inputs can be fed to the GPU library and to a CPU checker.
"""
from __future__ import annotations

import numpy as np

from . import meshes

# Ring-on-cube geometry: cube of side 0.1 m with its top face at z = 55 mm,
# ring (torus) R = 50 mm, r = 10 mm.
CUBE_SIDE = 0.1
CUBE_CENTER = (0.0, 0.0, 0.005)
RING_R, RING_r = 0.05, 0.01


def cube_mesh():
    return meshes.cube(CUBE_SIDE, CUBE_CENTER)


def ring_mesh(n_major=256, n_minor=64):
    return meshes.torus(RING_R, RING_r, n_major, n_minor)


def ring_quadrature(n_major=120, n_minor=32):
    """7 680 triangle centroids of a 120 x 32 torus, area weights with sum 1
    (quadrature.py:32-42, contacts_sdf.py:55-57)."""
    v, t = meshes.torus(RING_R, RING_r, n_major, n_minor)
    return meshes.triangle_centroid_quadrature(v, t, normalize=True)


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def rot_axis(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def ring_poses(n_poses=1000, seed=0, z_range=(0.060, 0.070), max_tilt_deg=5.0):
    """Ring poses over the cube: z ~ U[60, 70] mm, tilt axis uniform in the
    xy-plane with angle ~ U[0, 5 deg], yaw ~ U[0, 2 pi).  Returns R (B,3,3), t (B,3)."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(z_range[0], z_range[1], n_poses)
    phi = rng.uniform(0.0, 2.0 * np.pi, n_poses)
    tilt = np.deg2rad(rng.uniform(0.0, max_tilt_deg, n_poses))
    yaw = rng.uniform(0.0, 2.0 * np.pi, n_poses)
    R = np.stack([rot_axis([np.cos(p), np.sin(p), 0.0], a) @ rot_z(y)
                  for p, a, y in zip(phi, tilt, yaw)])
    t = np.stack([np.zeros(n_poses), np.zeros(n_poses), z], axis=1)
    return R, t


def pose_velocities(n_poses, seed=1, v_scale=0.05, w_scale=2.0):
    """Seeded relative velocities (v0, Om) per pose (m/s, rad/s)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, v_scale, (n_poses, 3)), rng.normal(0.0, w_scale, (n_poses, 3))


def pose_array(R, t, M=None, m=None, v0=None, Om=None):
    """(B,30) float64 pose records (see include/sdfo.h)."""
    B = len(R)
    M = R if M is None else M
    z = np.zeros((B, 3))
    m = z if m is None else m
    v0 = z if v0 is None else v0
    Om = z if Om is None else Om
    return np.concatenate([np.reshape(R, (B, 9)), t, np.reshape(M, (B, 9)), m, v0, Om], axis=1)


def valid_region_points(spec, n, seed=0, dtype=np.float32):
    """Uniform points whose interpolation cell is inside the grid."""
    rng = np.random.default_rng(seed)
    lo = spec.bmin + 0.5 * spec.voxel + 1e-3 * spec.voxel
    hi = spec.bmax - 0.5 * spec.voxel - 1e-3 * spec.voxel
    return rng.uniform(lo, hi, (n, 3)).astype(dtype)


def box_points(spec, n, seed=0, dtype=np.float64):
    """Uniform points in [bmin, bmax] (about (n-1)^3/n^3 of them interpolable)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(spec.bmin, spec.bmax, (n, 3)).astype(dtype)


def demo_benchmark_points(spec, count):
    """Dense point set of SdfOracleDemo's --benchmark (SdfOracleDemo.cpp:202-274):
    64-bit LCG (Knuth MMIX constants), seed 0x9e3779b97f4a7c15, continuous
    index in [0.05, n - 1.05] per axis."""
    mask64 = (1 << 64) - 1
    state = 0x9E3779B97F4A7C15

    def next_unit():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & mask64
        return float((state >> 11) & ((1 << 53) - 1)) / float(1 << 53)

    nx, ny, nz = spec.shape
    lo = [0.05 if n > 3 else 0.0 for n in (nx, ny, nz)]
    hi = [float(n) - 1.05 if n > 3 else 0.0 for n in (nx, ny, nz)]
    pts = np.empty((count, 3))
    for i in range(count):
        f = [lo[a] + (hi[a] - lo[a]) * next_unit() for a in range(3)]
        pts[i] = spec.bmin + spec.voxel * (np.array(f) + 0.5)
    return pts


def near_surface_points(verts, tris, n, sigma, seed=0, dtype=np.float32):
    """Random points on the mesh surface offset along the face normal by N(0, sigma^2)."""
    rng = np.random.default_rng(seed)
    v = np.asarray(verts, np.float64)[np.asarray(tris)]
    a = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    idx = rng.choice(len(v), size=n, p=a / a.sum())
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    bary = np.stack([1 - s, s * (1 - r2), s * r2], axis=1)
    p = np.einsum("nk,nkd->nd", bary, v[idx])
    nrm = np.cross(v[idx, 1] - v[idx, 0], v[idx, 2] - v[idx, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-300
    return (p + nrm * rng.normal(0.0, sigma, (n, 1))).astype(dtype)


def gear_batch(n_meshes=64, seed=5, min_tris=2_000, max_tris=200_000):
    """Seeded gear-like meshes of uneven size for batched builds.

    Teeth 12..120, face width 10..40 mm, module 1..3 mm, triangle counts
    log-uniform in [min_tris, max_tris] (set by the axial and per-tooth
    resolution), each centred at the origin.  Returns a list of
    ``(verts, tris, target_resolution)`` with the resolution 64, 128 or 256
    chosen by the triangle count."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_meshes):
        teeth = int(rng.integers(12, 121))
        face = float(rng.uniform(0.010, 0.040))
        module = float(rng.uniform(0.001, 0.003))
        want = float(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        spt = int(np.clip(round(np.sqrt(want / (2.0 * teeth * 4.0))), 4, 16))
        P = teeth * spt
        n_axial = int(np.clip(round((want - 2 * P) / (2 * P)), 1, 400))
        v, t = meshes.gear(teeth, spt, face, module, n_axial)
        res = 64 if len(t) < 8_000 else (128 if len(t) < 60_000 else 256)
        out.append((v, t, res))
    return out
