"""Mesh inputs of the builders: OBJ and the contact-aware surface format.

Restates ``loadObjFile`` and ``loadContactAwareSurfaceFile``
(SdfSurfaceMesh.cpp:111-344) with the same parsing rules and error texts.
Host-side text parsing; returns numpy arrays ready for ``SdfGrid.build_*``.
"""
from __future__ import annotations

import math
import re

import numpy as np

_INT = re.compile(r"^\s*[+-]?\d+")
_FLT = re.compile(r"^\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?|inf|nan)", re.I)


class SurfaceParseError(RuntimeError):
    pass


def _stoi(tok: str) -> int:
    """std::stoi: leading integer prefix, else an error."""
    m = _INT.match(tok)
    if not m:
        raise SurfaceParseError("stoi")
    return int(m.group(0))


class _Stream:
    """Whitespace token stream with istream >> semantics for int / double."""

    def __init__(self, text: str):
        self.toks = text.split()
        self.i = 0
        self.ok = True

    def word(self):
        if not self.ok or self.i >= len(self.toks):
            self.ok = False
            return None
        self.i += 1
        return self.toks[self.i - 1]

    def _num(self, rx, conv):
        tok = self.word()
        if tok is None:
            return None
        m = rx.match(tok)
        if not m:
            self.ok = False
            return None
        rest = tok[m.end():]
        if rest:  # the unparsed suffix stays in the stream
            self.toks[self.i - 1] = rest
            self.i -= 1
        return conv(m.group(0))

    def int(self):
        return self._num(_INT, int)

    def double(self):
        return self._num(_FLT, float)


def load_obj(path: str):
    """loadObjFile (SdfSurfaceMesh.cpp:111-154): ``v x y z`` and the first
    three indices of ``f`` lines (``v``, ``v/vt``, ``v/vt/vn``, ``v//vn``);
    extra face indices are ignored, as in the reference."""
    try:
        f = open(path, "r")
    except OSError:
        raise SurfaceParseError("Cannot open OBJ file: " + path)
    verts, tris = [], []
    with f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "#":
                continue
            s = _Stream(line)
            token = s.word()
            if token == "v":
                xyz = []
                for _ in range(3):
                    x = s.double()
                    xyz.append(0.0 if x is None else x)  # failed extraction yields 0
                verts.append(xyz)
            elif token == "f":
                tri = []
                for _ in range(3):
                    tok = s.word() or ""
                    tri.append(_stoi(tok.split("/")[0]) - 1)
                tris.append(tri)
    return (np.asarray(verts, np.float64).reshape(-1, 3), np.asarray(tris, np.int32).reshape(-1, 3))


def _strip(line: str) -> str:
    c = line.find("#")
    if c >= 0:
        line = line[:c]
    return line.strip(" \t\r\n")


def load_contact_aware_surface(path: str):
    """loadContactAwareSurfaceFile (SdfSurfaceMesh.cpp:156-344): the
    ``NEXDYN_CONTACT_AWARE_SURFACE 1`` text format.  Returns
    ``(verts (V,3), tris (T,3), corner_normals (T,3,3))``."""
    try:
        f = open(path, "r")
    except OSError:
        raise SurfaceParseError("Cannot open contact-aware surface file: " + path)

    def err(ln, msg):
        return SurfaceParseError(f"{path}:{ln}: {msg}")

    verts, tris, cnorm = [], [], []

    def fresh():
        return dict(active=False, name="", ev=-1, et=-1, sv=0, st=0,
                    T=[[1.0 if r == c else 0.0 for c in range(4)] for r in range(4)],
                    N=[[1.0 if r == c else 0.0 for c in range(3)] for r in range(3)], ids={})

    st = fresh()

    def close(ln):
        nonlocal st
        if not st["active"]:
            return
        if st["ev"] < 0:
            raise err(ln, f"surface '{st['name']}' is missing vertices block")
        if st["et"] < 0:
            raise err(ln, f"surface '{st['name']}' is missing triangles block")
        if st["sv"] != st["ev"]:
            raise err(ln, f"surface '{st['name']}' vertex count mismatch")
        if st["st"] != st["et"]:
            raise err(ln, f"surface '{st['name']}' triangle count mismatch")
        st = fresh()

    header = False
    ln = 0
    with f:
        for raw in f:
            ln += 1
            line = _strip(raw)
            if not line:
                continue
            s = _Stream(line)
            token = s.word()
            if not header:
                ver = s.int()
                if token != "NEXDYN_CONTACT_AWARE_SURFACE" or ver is None or ver != 1:
                    raise err(ln, "expected header 'NEXDYN_CONTACT_AWARE_SURFACE 1'")
                header = True
                continue
            if token == "surface":
                close(ln)
                st["active"] = True
                name = s.word()
                if not name:
                    raise err(ln, "surface requires a name")
                st["name"] = name
                continue
            if token == "end":
                close(ln)
                continue
            if not st["active"]:
                raise err(ln, "expected surface block")
            if token == "transform":
                vals = [s.double() for _ in range(16)]
                if any(v is None for v in vals):
                    raise err(ln, "transform requires 16 row-major values")
                T = [[float(vals[4 * r + c]) for c in range(4)] for r in range(4)]
                det = (T[0][0] * (T[1][1] * T[2][2] - T[1][2] * T[2][1]) -
                       T[0][1] * (T[1][0] * T[2][2] - T[1][2] * T[2][0]) +
                       T[0][2] * (T[1][0] * T[2][1] - T[1][1] * T[2][0]))
                if abs(det) <= 1.0e-15:
                    raise err(ln, "transform linear part must be invertible")
                st["T"] = T
                # inverse transpose of the linear part = cofactors / det
                st["N"] = [[(T[(r + 1) % 3][(c + 1) % 3] * T[(r + 2) % 3][(c + 2) % 3] -
                             T[(r + 1) % 3][(c + 2) % 3] * T[(r + 2) % 3][(c + 1) % 3]) / det
                            for c in range(3)] for r in range(3)]
                continue
            if token == "vertices":
                n = s.int()
                if n is None or n < 0:
                    raise err(ln, "vertices requires a non-negative count")
                st["ev"] = n
                continue
            if token == "triangles":
                n = s.int()
                if n is None or n < 0:
                    raise err(ln, "triangles requires a non-negative count")
                st["et"] = n
                continue
            if token == "v":
                if st["ev"] < 0:
                    raise err(ln, "vertex appears before vertices block")
                vid = s.int()
                xyz = [s.double() for _ in range(3)]
                if vid is None or any(x is None for x in xyz):
                    raise err(ln, "vertex requires: v <id> <x> <y> <z>")
                if vid in st["ids"]:
                    raise err(ln, "duplicate vertex id in surface")
                st["ids"][vid] = len(verts)
                T = st["T"]
                h = [((T[r][0] * xyz[0] + T[r][1] * xyz[1]) + T[r][2] * xyz[2]) + T[r][3]
                     for r in range(4)]
                p = [h[0] / h[3], h[1] / h[3], h[2] / h[3]] if abs(h[3]) > 1.0e-15 else h[:3]
                verts.append(np.asarray(p, np.float64))
                st["sv"] += 1
                continue
            if token == "tri":
                if st["et"] < 0:
                    raise err(ln, "triangle appears before triangles block")
                tid = s.int()
                vid = [s.int() for _ in range(3)]
                nums = [s.double() for _ in range(9)]
                if tid is None or any(v is None for v in vid) or any(x is None for x in nums):
                    raise err(ln, "triangle requires id, 3 vertex ids, and 3 corner normals")
                tri, nrm = [], []
                for c in range(3):
                    if vid[c] not in st["ids"]:
                        raise err(ln, "triangle references unknown vertex id")
                    tri.append(st["ids"][vid[c]])
                    N = st["N"]
                    x = nums[3 * c:3 * c + 3]
                    n = [(N[r][0] * x[0] + N[r][1] * x[1]) + N[r][2] * x[2] for r in range(3)]
                    ln_ = math.sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])
                    if not all(math.isfinite(q) for q in n) or not math.isfinite(ln_) or ln_ <= 1.0e-15:
                        raise err(ln, "corner normal must be a finite non-zero vector")
                    nrm.append([n[0] / ln_, n[1] / ln_, n[2] / ln_])
                e0 = verts[tri[1]] - verts[tri[0]]
                e1 = verts[tri[2]] - verts[tri[0]]
                if np.linalg.norm(np.cross(e0, e1)) <= 1.0e-15:
                    raise err(ln, "triangle is degenerate")
                tris.append(tri)
                cnorm.append(nrm)
                st["st"] += 1
                continue
            raise err(ln, f"unknown token '{token}'")
    if not header:
        raise SurfaceParseError(path + ": missing contact-aware surface header")
    close(ln)
    if not verts or not tris:
        raise SurfaceParseError(path + ": contact-aware surface contains no triangles")
    return (np.asarray(verts, np.float64), np.asarray(tris, np.int32),
            np.asarray(cnorm, np.float64))


def write_contact_aware_surface(path: str, verts, tris, corner_normals, name="surface"):
    """Writes a mesh in the NEXDYN_CONTACT_AWARE_SURFACE 1 format (identity transform)."""
    v = np.asarray(verts, np.float64)
    t = np.asarray(tris)
    cn = np.asarray(corner_normals, np.float64).reshape(-1, 3, 3)
    with open(path, "w") as f:
        f.write("NEXDYN_CONTACT_AWARE_SURFACE 1\n")
        f.write(f"surface {name}\nvertices {len(v)}\ntriangles {len(t)}\n")
        for i, p in enumerate(v):
            f.write(f"v {i} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        for i, (tri, n) in enumerate(zip(t, cn)):
            f.write(f"tri {i} {tri[0]} {tri[1]} {tri[2]} " +
                    " ".join(repr(float(x)) for x in n.ravel()) + "\n")
        f.write("end\n")
