"""Mesh loaders (SdfSurfaceMesh.cpp:111-344) on crafted files (CPU)."""
import numpy as np
import pytest

from sdfo import meshes
from sdfo.surface_io import (SurfaceParseError, load_contact_aware_surface, load_obj,
                             write_contact_aware_surface)


def test_obj_faces_and_index_forms(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# comment\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvn 0 0 1\n"
                 "f 1 2 3\nf 1/1 2/2 4/4\nf 1//1 3//1 4//1\nf 2/1/1 3/1/1 4/1/1 1/1/1\n")
    v, t = load_obj(str(p))
    assert v.shape == (4, 3) and np.array_equal(v[1], [1, 0, 0])
    # quads keep their first three indices (SdfSurfaceMesh.cpp:139-151)
    assert t.tolist() == [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]


def test_obj_missing(tmp_path):
    with pytest.raises(SurfaceParseError, match="Cannot open OBJ file"):
        load_obj(str(tmp_path / "none.obj"))


def test_contact_aware_round_trip_and_transform(tmp_path):
    v, t = meshes.cube(0.1)
    cn = np.repeat(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])[:, None], 3, 1)
    cn = cn / np.linalg.norm(cn, axis=-1, keepdims=True)
    p = str(tmp_path / "c.surf")
    write_contact_aware_surface(p, v, t, cn)
    v2, t2, n2 = load_contact_aware_surface(p)
    assert np.array_equal(v2, v) and np.array_equal(t2, t) and np.allclose(n2, cn, atol=1e-15)
    # transform: scale x by 2 and translate; normals transform by the inverse transpose
    text = open(p).read().replace("vertices", "transform 2 0 0 1  0 1 0 0  0 0 1 0  0 0 0 1\nvertices")
    open(p, "w").write(text)
    v3, _, n3 = load_contact_aware_surface(p)
    assert np.allclose(v3, v * [2, 1, 1] + [1, 0, 0])
    want = cn * [0.5, 1, 1]
    assert np.allclose(n3, want / np.linalg.norm(want, axis=-1, keepdims=True))


@pytest.mark.parametrize("body,msg", [
    ("NEXDYN_CONTACT_AWARE_SURFACE 2\n", ":1: expected header 'NEXDYN_CONTACT_AWARE_SURFACE 1'"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nv 0 0 0 0\n", ":2: expected surface block"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nv 0 0 0 0\n", ":3: vertex appears before"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 1\nv 0 0 0\n", ":4: vertex requires"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 2\nv 0 0 0 0\nv 0 1 0 0\n",
     ":5: duplicate vertex id"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 3\ntriangles 1\nv 0 0 0 0\nv 1 1 0 0\n"
     "v 2 0 1 0\ntri 0 0 1 2 0 0 1 0 0 1 0 0 0\n", ":8: corner normal must be a finite non-zero"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 3\ntriangles 1\nv 0 0 0 0\nv 1 1 0 0\n"
     "v 2 2 0 0\ntri 0 0 1 2 0 0 1 0 0 1 0 0 1\n", ":8: triangle is degenerate"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 3\ntriangles 1\nv 0 0 0 0\nv 1 1 0 0\n"
     "v 2 0 1 0\ntri 0 0 1 9 0 0 1 0 0 1 0 0 1\n", ":8: triangle references unknown vertex id"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 2\ntriangles 0\nv 0 0 0 0\nend\n",
     ":6: surface 's' vertex count mismatch"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\ntransform 1 0 0 0 0 1 0 0 0 0 0 0 0 0 0 1\n",
     ":3: transform linear part must be invertible"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nbogus\n", ":3: unknown token 'bogus'"),
    ("NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 0\ntriangles 0\nend\n",
     "contact-aware surface contains no triangles"),
    ("# only a comment\n", "missing contact-aware surface header"),
])
def test_contact_aware_errors(tmp_path, body, msg):
    p = tmp_path / "e.surf"
    p.write_text(body)
    with pytest.raises(SurfaceParseError, match=msg.replace("(", r"\(").replace("'", "'")):
        load_contact_aware_surface(str(p))


# ---- the C++ API's loaders (include/SdfOracle/SdfSurfaceMesh.h), host only --
import os  # noqa: E402
import subprocess  # noqa: E402

_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "sdf-contact-module_amd", "_build", "sdf_api_check")


def _cpp_load(path, tmp_path):
    if not os.path.exists(_BIN):
        pytest.skip("sdf_api_check not built")
    out = tmp_path / "dump.bin"
    subprocess.run([_BIN, "loaders", str(path), str(out)], check=True, timeout=60)
    b = out.read_bytes()
    if b.startswith(b"ERROR: "):
        return b[7:].decode()
    nv, nt, nn = np.frombuffer(b[:24], np.int64)
    o = 24
    v = np.frombuffer(b[o:o + 24 * nv], np.float64).reshape(-1, 3)
    o += 24 * nv
    t = np.frombuffer(b[o:o + 12 * nt], np.int32).reshape(-1, 3)
    o += 12 * nt
    n = np.frombuffer(b[o:o + 72 * nn], np.float64).reshape(-1, 3, 3)
    return v, t, n


def test_cpp_loaders_equal_python(tmp_path):
    v, t = meshes.torus(0.05, 0.01, 16, 8)
    cn = np.repeat(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])[:, None], 3, 1)
    p = str(tmp_path / "s.surf")
    write_contact_aware_surface(p, v, t, cn)
    text = open(p).read().replace("vertices", "transform 0.5 0.1 0 0.01  0 2 0.3 0  0.2 0 1 0  0 0 0 1\nvertices")
    open(p, "w").write(text)
    pv, pt, pn = load_contact_aware_surface(p)
    cv, ct, cnn = _cpp_load(p, tmp_path)
    assert np.array_equal(pv, cv) and np.array_equal(pt, ct) and np.array_equal(pn, cnn)
    o = tmp_path / "m.obj"
    o.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1/1 2/2 4/4\nf 2//1 3//1 4//1 1//1\n")
    pv, pt = load_obj(str(o))
    cv, ct, _ = _cpp_load(o, tmp_path)
    assert np.array_equal(pv, cv) and np.array_equal(pt, ct)


@pytest.mark.parametrize("body", [
    "NEXDYN_CONTACT_AWARE_SURFACE 2\n",
    "NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 2\nv 0 0 0 0\nv 0 1 0 0\n",
    "NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nvertices 3\ntriangles 1\nv 0 0 0 0\nv 1 1 0 0\n"
    "v 2 2 0 0\ntri 0 0 1 2 0 0 1 0 0 1 0 0 1\n",
    "NEXDYN_CONTACT_AWARE_SURFACE 1\nsurface s\nbogus\n",
    "# only a comment\n",
])
def test_cpp_loader_errors_equal_python(tmp_path, body):
    p = tmp_path / "e.surf"
    p.write_text(body)
    with pytest.raises(SurfaceParseError) as ei:
        load_contact_aware_surface(str(p))
    assert _cpp_load(p, tmp_path) == str(ei.value)
