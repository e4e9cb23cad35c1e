"""Seeded synthetic triangle meshes and quadrature sets for the SDF workloads.

All meshes are closed and outward oriented (counter-clockwise seen from
outside), returned as ``(verts float64 (V,3), tris int32 (T,3))`` -- the
inputs of ``SdfBuilder::build`` (SdfBuilder.h:58-62).
"""
from __future__ import annotations

import numpy as np


def cube(side: float = 0.1, center=(0.0, 0.0, 0.0)):
    """Axis-aligned cube: 8 vertices, 12 outward triangles."""
    h = 0.5 * side
    c = np.asarray(center, np.float64)
    v = np.array([[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h],
                  [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]], np.float64) + c
    t = np.array([[0, 2, 1], [0, 3, 2],   # -z
                  [4, 5, 6], [4, 6, 7],   # +z
                  [0, 1, 5], [0, 5, 4],   # -y
                  [2, 3, 7], [2, 7, 6],   # +y
                  [0, 4, 7], [0, 7, 3],   # -x
                  [1, 2, 6], [1, 6, 5]],  # +x
                 np.int32)
    return v, t


def torus(R: float = 0.05, r: float = 0.01, n_major: int = 256, n_minor: int = 64,
          center=(0.0, 0.0, 0.0)):
    """Parametric torus around +z; ``2 * n_major * n_minor`` triangles."""
    th = 2.0 * np.pi * np.arange(n_major) / n_major
    ph = 2.0 * np.pi * np.arange(n_minor) / n_minor
    T, P = np.meshgrid(th, ph, indexing="ij")
    rad = R + r * np.cos(P)
    v = np.stack([rad * np.cos(T), rad * np.sin(T), r * np.sin(P)], axis=-1).reshape(-1, 3)
    v = v + np.asarray(center, np.float64)
    i = np.arange(n_major)[:, None]
    j = np.arange(n_minor)[None, :]
    i1 = (i + 1) % n_major
    j1 = (j + 1) % n_minor
    v00 = (i * n_minor + j).ravel()
    v10 = (i1 * n_minor + j).ravel()
    v11 = (i1 * n_minor + j1).ravel()
    v01 = (i * n_minor + j1).ravel()
    t = np.empty((2 * n_major * n_minor, 3), np.int32)
    t[0::2] = np.stack([v00, v10, v11], axis=1)
    t[1::2] = np.stack([v00, v11, v01], axis=1)
    return v.astype(np.float64), t


def uv_sphere(radius: float = 0.05, n_lat: int = 16, n_lon: int = 32, center=(0.0, 0.0, 0.0)):
    """Latitude/longitude sphere with single pole vertices (no degenerate faces)."""
    verts = [[0.0, 0.0, -radius]]
    for a in range(1, n_lat):
        lat = -0.5 * np.pi + np.pi * a / n_lat
        for b in range(n_lon):
            lon = 2.0 * np.pi * b / n_lon
            verts.append([radius * np.cos(lat) * np.cos(lon), radius * np.cos(lat) * np.sin(lon),
                          radius * np.sin(lat)])
    verts.append([0.0, 0.0, radius])
    v = np.asarray(verts, np.float64) + np.asarray(center, np.float64)
    top = len(verts) - 1
    ring = lambda a, b: 1 + (a - 1) * n_lon + (b % n_lon)  # noqa: E731
    tris = []
    for b in range(n_lon):
        tris.append([0, ring(1, b + 1), ring(1, b)])
    for a in range(1, n_lat - 1):
        for b in range(n_lon):
            tris.append([ring(a, b), ring(a, b + 1), ring(a + 1, b + 1)])
            tris.append([ring(a, b), ring(a + 1, b + 1), ring(a + 1, b)])
    for b in range(n_lon):
        tris.append([top, ring(n_lat - 1, b), ring(n_lat - 1, b + 1)])
    return v, np.asarray(tris, np.int32)


def hollow_cylinder(r_out: float = 0.04, r_in: float = 0.03, height: float = 0.05,
                    n_az: int = 1024, n_ax: int = 64, center=(0.0, 0.0, 0.0)):
    """Tube along z with annular end caps.

    Triangles: outer and inner walls ``2 * 2 * n_az * n_ax`` plus two annuli
    ``2 * 2 * n_az`` (1024 x 64 gives 266 240).
    """
    th = 2.0 * np.pi * np.arange(n_az) / n_az
    zs = -0.5 * height + height * np.arange(n_ax + 1) / n_ax
    Tz, Zz = np.meshgrid(th, zs, indexing="ij")  # (n_az, n_ax+1)
    outer = np.stack([r_out * np.cos(Tz), r_out * np.sin(Tz), Zz], -1).reshape(-1, 3)
    inner = np.stack([r_in * np.cos(Tz), r_in * np.sin(Tz), Zz], -1).reshape(-1, 3)
    v = np.concatenate([outer, inner]) + np.asarray(center, np.float64)
    no = n_az * (n_ax + 1)
    idx = lambda a, z: (a % n_az) * (n_ax + 1) + z  # noqa: E731
    a = np.arange(n_az)[:, None]
    z = np.arange(n_ax)[None, :]
    o00, o10 = idx(a, z).ravel(), idx(a + 1, z).ravel()
    o11, o01 = idx(a + 1, z + 1).ravel(), idx(a, z + 1).ravel()
    walls_out = np.concatenate([np.stack([o00, o10, o11], 1), np.stack([o00, o11, o01], 1)])
    i00, i10, i11, i01 = o00 + no, o10 + no, o11 + no, o01 + no
    walls_in = np.concatenate([np.stack([i00, i11, i10], 1), np.stack([i00, i01, i11], 1)])
    aa = np.arange(n_az)
    ob0, ob1 = idx(aa, 0), idx(aa + 1, 0)
    ib0, ib1 = ob0 + no, ob1 + no
    bottom = np.concatenate([np.stack([ob0, ib0, ib1], 1), np.stack([ob0, ib1, ob1], 1)])
    ot0, ot1 = idx(aa, n_ax), idx(aa + 1, n_ax)
    it0, it1 = ot0 + no, ot1 + no
    top = np.concatenate([np.stack([ot0, ot1, it1], 1), np.stack([ot0, it1, it0], 1)])
    t = np.concatenate([walls_out, walls_in, bottom, top]).astype(np.int32)
    return v.astype(np.float64), t


def gear(n_teeth: int = 24, samples_per_tooth: int = 8, face_width: float = 0.02,
         module: float = 0.002, n_axial: int = 8, center=(0.0, 0.0, 0.0)):
    """Extruded spur-gear-like profile (trapezoidal teeth on an involute-sized blank).

    Pitch radius = module * n_teeth / 2, addendum = module, dedendum = 1.25 module.
    Side walls ``2 * P * n_axial`` triangles plus two fan caps ``2 * P``,
    with ``P = n_teeth * samples_per_tooth`` profile points.
    """
    P = n_teeth * samples_per_tooth
    rp = 0.5 * module * n_teeth
    r_tip, r_root = rp + module, rp - 1.25 * module
    th = 2.0 * np.pi * np.arange(P) / P
    phase = (np.arange(P) % samples_per_tooth + 0.5) / samples_per_tooth  # in (0,1)
    # trapezoid: flank 0.2, top 0.3, flank 0.2, root 0.3 of the pitch
    prof = np.clip(np.minimum((phase - 0.0) / 0.2, (0.7 - phase) / 0.2), 0.0, 1.0)
    rad = r_root + (r_tip - r_root) * prof
    zs = -0.5 * face_width + face_width * np.arange(n_axial + 1) / n_axial
    ring = np.stack([rad * np.cos(th), rad * np.sin(th)], -1)
    side = np.concatenate([np.concatenate([ring, np.full((P, 1), z)], 1) for z in zs])
    c_bot = np.array([[0.0, 0.0, zs[0]]])
    c_top = np.array([[0.0, 0.0, zs[-1]]])
    v = np.concatenate([side, c_bot, c_top]) + np.asarray(center, np.float64)
    vid = lambda a, z: z * P + (a % P)  # noqa: E731
    a = np.arange(P)[:, None]
    z = np.arange(n_axial)[None, :]
    v00, v10 = vid(a, z).ravel(), vid(a + 1, z).ravel()
    v11, v01 = vid(a + 1, z + 1).ravel(), vid(a, z + 1).ravel()
    walls = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)])
    cb, ct = len(side), len(side) + 1
    aa = np.arange(P)
    bot = np.stack([np.full(P, cb), vid(aa + 1, 0), vid(aa, 0)], 1)
    top = np.stack([np.full(P, ct), vid(aa, n_axial), vid(aa + 1, n_axial)], 1)
    t = np.concatenate([walls, bot, top]).astype(np.int32)
    return v.astype(np.float64), t


def triangle_centroid_quadrature(verts, tris, normalize: bool = True):
    """Centroids and area weights (quadrature.py:32-42); ``sum(w) = 1`` when
    ``normalize`` (contacts_sdf.py:55-57)."""
    v = np.asarray(verts, np.float64)[np.asarray(tris)]
    X = v.mean(axis=1)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    w = 0.5 * np.linalg.norm(n, axis=1)
    if normalize and w.sum() > 1e-30:
        w = w / w.sum()
    return X, w


def annular_quadrature(r_in: float, r_out: float, n_radial: int = 20, n_azimuthal: int = 64):
    """Midpoint rings (mesh_quadrature.py:34-58): points at z = 0, weight = ring-sector area."""
    dr = (r_out - r_in) / n_radial
    dth = 2.0 * np.pi / n_azimuthal
    i = np.arange(n_radial)
    ri, ro = r_in + i * dr, r_in + (i + 1) * dr
    rc = 0.5 * (ri + ro)
    area = 0.5 * (ro ** 2 - ri ** 2) * dth
    th = (np.arange(n_azimuthal) + 0.5) * dth
    RC, TH = np.meshgrid(rc, th, indexing="ij")
    X = np.stack([RC * np.cos(TH), RC * np.sin(TH), np.zeros_like(RC)], -1).reshape(-1, 3)
    w = np.repeat(area, n_azimuthal)
    return X, w


def load_obj_triangles(path: str):
    """load_obj_triangles (quadrature.py:17-29): ``v`` lines give vertices,
    ``f`` lines give zero-based vertex indices (the part before the first
    ``/``), and only faces with exactly three indices are kept; quads and
    larger polygons are skipped, not split."""
    verts, tris = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                verts.append([float(x) for x in line.strip().split()[1:]])
            elif line.startswith("f "):
                parts = line.strip().split()
                idx = [int(p.split("/")[0]) - 1 for p in parts[1:]]
                if len(idx) == 3:
                    tris.append(idx)
    return np.array(verts), np.array(tris)


def quadrature_from_obj(path: str):
    """quadrature_mesh_from_obj (quadrature.py:45-48): triangle centroids,
    raw triangle areas as weights (the engine normalises them when a pair is
    registered), unit face normals and triangle ids."""
    verts, tris = load_obj_triangles(path)
    v = verts[tris]
    X = v.mean(axis=1)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    nn = np.linalg.norm(n, axis=1)
    w = 0.5 * nn
    normals = n / (nn[:, None] + 1e-30)
    return X, w, normals, np.arange(len(tris))
