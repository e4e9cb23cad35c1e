"""Host-side argument checks and host helpers that need no GPU."""
import ctypes as C

import numpy as np
import pytest
import torch

from sdfo import meshes


def test_query_into_rejects_wrong_point_dtypes_before_launch():
    from sdfo.grid import SdfGrid
    g = object.__new__(SdfGrid)            # no device handle: nothing may launch
    n = 8
    gap, grad = torch.empty(n, dtype=torch.float16), torch.empty((n, 3), dtype=torch.float16)
    valid = torch.empty(n, dtype=torch.uint8)
    for dt in (torch.float16, torch.bfloat16, torch.int32):
        with pytest.raises(ValueError, match="float32/float64"):
            g.query_into(torch.zeros((n, 3), dtype=dt), gap, grad, valid)
    with pytest.raises(ValueError, match="float32/float64"):
        g.query_into(torch.zeros((n, 4), dtype=torch.float32), gap, grad, valid)
    # float32 points with float16 outputs: the outputs are checked too
    with pytest.raises(ValueError, match="dtypes"):
        g.query_into(torch.zeros((n, 3), dtype=torch.float32), gap, grad, valid)


def test_query_plan_run_rejects_wrong_dtypes_before_launch():
    from sdfo.grid import QueryPlan
    plan = object.__new__(QueryPlan)
    plan.offsets = np.array([0, 4, 8], np.int64)
    plan._h = None
    f32 = lambda *s: torch.zeros(s, dtype=torch.float32)  # noqa: E731
    ok = (f32(8, 3), f32(8), f32(8, 3), torch.zeros(8, dtype=torch.uint8))
    for i, bad in enumerate((torch.zeros((8, 3), dtype=torch.float64), torch.zeros(8, dtype=torch.float16),
                             torch.zeros((8, 3), dtype=torch.bfloat16), torch.zeros(8, dtype=torch.bool))):
        args = list(ok)
        args[i] = bad
        with pytest.raises(ValueError, match="expected"):
            plan.run(*args)


def test_contact_evaluate_rejects_null_grid():
    from sdfo import _lib
    p = _lib.ContactParams(1e11, 1e4, 0.3, 1e-6, 0.01, 2, 3)
    rc = _lib.lib.sdfo_contact_evaluate(None, None, None, None, 0, None, None, 0, C.byref(p), None,
                                        None, None, None)
    assert rc == -1 and b"grid" in _lib.lib.sdfo_last_error()


def test_obj_quadrature_skips_non_triangles(tmp_path):
    """quadrature.py:17-48: quads are skipped, not split; weights are raw
    areas, points are centroids, normals are unit face normals."""
    obj = tmp_path / "m.obj"
    obj.write_text("# two triangles and a quad\n"
                   "v 0 0 0\nv 2 0 0\nv 0 1 0\nv 0 0 3\nv 1 1 1\n"
                   "f 1 2 3\n"
                   "f 1/1/1 4/2/2 2/3/3\n"
                   "f 1 2 5 3\n"
                   "vn 0 0 1\n")
    X, w, nrm, ids = meshes.quadrature_from_obj(str(obj))
    assert len(X) == 2 and list(ids) == [0, 1]
    assert np.allclose(X[0], [2 / 3, 1 / 3, 0]) and np.allclose(X[1], [2 / 3, 0, 1])
    assert np.allclose(w, [1.0, 3.0])                 # 1/2 |e0 x e1|: 1/2 * 2 * 1, 1/2 * 3 * 2
    assert np.allclose(nrm[0], [0, 0, 1]) and np.allclose(nrm[1], [0, 1, 0])
    # the surface loader keeps the quad's first three indices instead
    from sdfo import surface_io
    _, t = surface_io.load_obj(str(obj))
    assert len(t) == 3
