"""Device-resident SDF grids: build, load/save, download and query.

Python mirror of the reference's builder and query API
(``sdf::SdfBuilder``, SdfBuilder.h:58-113; ``sdf::SdfOracle``,
SdfOracle.h:20-67; ``SdfIO.h:29-45``) on top of the C ABI.  Builds take host
numpy meshes; queries take CUDA/HIP torch tensors and run on the current
torch stream.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MODES, check, lib


def _stream_handle(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _mesh(verts, tris):
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    return v, t


@dataclass
class GridSpec:
    """Cell-centred grid (GridSpec.h:17-60)."""

    bmin: np.ndarray
    bmax: np.ndarray
    voxel: np.ndarray
    shape: tuple

    @property
    def total(self) -> int:
        return int(np.prod(self.shape))

    @property
    def diagonal(self) -> float:
        return float(np.linalg.norm(self.bmax - self.bmin))


def _spec_of(info: _lib.GridInfo) -> GridSpec:
    return GridSpec(np.array(info.bmin[:]), np.array(info.bmax[:]), np.array(info.voxel[:]),
                    tuple(int(x) for x in info.shape))


def setup_grid(verts, **params) -> GridSpec:
    """SdfBuilder::setupGrid (SdfBuilder.cpp:994-1035)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
    p = _lib.default_build_params(**params)
    info = _lib.GridInfo()
    check(lib.sdfo_setup_grid(_np_ptr(v), len(v), C.byref(p), C.byref(info)))
    return _spec_of(info)


class SdfGrid:
    """An SDF owned by libsdfo on the current HIP device."""

    def __init__(self, handle: int):
        if not handle:
            raise ValueError("null grid handle")
        self._h = C.c_void_p(handle)
        info = _lib.GridInfo()
        check(lib.sdfo_grid_get_info(self._h, C.byref(info)))
        self.info = info
        self.spec = _spec_of(info)

    # ---- construction ----------------------------------------------------
    @classmethod
    def build_dense(cls, verts, tris, stream=None, **params) -> "SdfGrid":
        """SdfBuilder::build (SdfBuilder.cpp:364-407) on the GPU."""
        v, t = _mesh(verts, tris)
        p = _lib.default_build_params(**params)
        out = C.c_void_p()
        check(lib.sdfo_build_dense(_np_ptr(v), len(v), _np_ptr(t), len(t), C.byref(p),
                                   C.byref(out), _stream_handle(stream)))
        return cls(out.value)

    @classmethod
    def build_dense_batched(cls, meshes, params=None, stream=None) -> list:
        """Many dense builds in one persistent fill launch.

        ``meshes``: list of (verts, tris); ``params``: one dict of builder
        options for all meshes, or a list of dicts (one per mesh).  Returns
        one grid per mesh, each equal to ``build_dense`` of that mesh."""
        n = len(meshes)
        if n == 0:
            return []
        plist = params if isinstance(params, (list, tuple)) else [params or {}] * n
        if len(plist) != n:
            raise ValueError("one params dict per mesh")
        vs, ts = zip(*[_mesh(v, t) for v, t in meshes])
        vo = np.zeros(n + 1, np.int64)
        to = np.zeros(n + 1, np.int64)
        vo[1:] = np.cumsum([len(v) for v in vs])
        to[1:] = np.cumsum([len(t) for t in ts])
        V = np.ascontiguousarray(np.concatenate(vs), np.float64)
        T = np.ascontiguousarray(np.concatenate(ts), np.int32)
        P = (_lib.BuildParams * n)(*[_lib.default_build_params(**q) for q in plist])
        outs = (C.c_void_p * n)()
        check(lib.sdfo_build_dense_batched(n, _np_ptr(V), _np_ptr(vo), _np_ptr(T), _np_ptr(to), P,
                                           outs, _stream_handle(stream)))
        return [cls(outs[i]) for i in range(n)]

    @classmethod
    def build_sparse(cls, verts, tris, stream=None, **params) -> "SdfGrid":
        """SdfBuilder::buildSparse (SdfBuilder.cpp:823-983) on the GPU."""
        v, t = _mesh(verts, tris)
        p = _lib.default_build_params(**params)
        out = C.c_void_p()
        check(lib.sdfo_build_sparse(_np_ptr(v), len(v), _np_ptr(t), len(t), C.byref(p),
                                    C.byref(out), _stream_handle(stream)))
        return cls(out.value)

    @classmethod
    def build_contact_aware(cls, verts, tris, corner_normals, sparse=False, stream=None,
                            **params) -> "SdfGrid":
        """SdfBuilder::buildContactAware / buildSparseContactAware (SdfBuilder.cpp:416-821)."""
        v, t = _mesh(verts, tris)
        cn = np.ascontiguousarray(corner_normals, dtype=np.float64).reshape(-1, 3, 3)
        if len(cn) != len(t):
            raise ValueError("Contact-aware SDF requires one corner-normal triplet per triangle")
        p = _lib.default_build_params(**params)
        out = C.c_void_p()
        check(lib.sdfo_build_contact_aware(_np_ptr(v), len(v), _np_ptr(t), len(t), _np_ptr(cn),
                                           C.byref(p), 1 if sparse else 0, C.byref(out),
                                           _stream_handle(stream)))
        return cls(out.value)

    @classmethod
    def from_dense(cls, spec: GridSpec, phi, n, H=None, witness=None, candidates=None) -> "SdfGrid":
        """SdfOracle(SdfData) (SdfOracle.cpp:154-157) from host FP64 fields."""
        bmin = np.ascontiguousarray(spec.bmin, np.float64)
        bmax = np.ascontiguousarray(spec.bmax, np.float64)
        vox = np.ascontiguousarray(spec.voxel, np.float64)
        shape = np.ascontiguousarray(spec.shape, np.int32)
        phi = np.ascontiguousarray(phi, np.float64).ravel()
        n = np.ascontiguousarray(n, np.float64).reshape(-1, 3)
        if len(phi) != int(np.prod(shape)) or len(n) != len(phi):
            raise ValueError("field size does not match the grid shape")
        Hp = None if H is None else np.ascontiguousarray(H, np.float64).reshape(-1, 6)
        Wp = None if witness is None else np.ascontiguousarray(witness, np.float64).reshape(-1, 3)
        stride = 0
        cc = cp = cn = cb = None
        if candidates is not None:
            cc = np.ascontiguousarray(candidates["count"], np.uint8)
            stride = int(candidates["stride"])
            cp = np.ascontiguousarray(candidates["phi"], np.float64)
            cn = np.ascontiguousarray(candidates["n"], np.float64)
            cb = np.ascontiguousarray(candidates["branch"], np.int32)
        out = C.c_void_p()
        nil = lambda a: None if a is None else _np_ptr(a)  # noqa: E731
        check(lib.sdfo_grid_from_dense(_np_ptr(bmin), _np_ptr(bmax), _np_ptr(vox), _np_ptr(shape),
                                       _np_ptr(phi), _np_ptr(n), nil(Hp), nil(Wp), stride, nil(cc),
                                       nil(cp), nil(cn), nil(cb), C.byref(out)))
        return cls(out.value)

    @classmethod
    def load(cls, path: str) -> "SdfGrid":
        """SdfOracle::loadFromFile (SdfOracle.cpp:165-172)."""
        out = C.c_void_p()
        check(lib.sdfo_grid_load(os.fsencode(path), C.byref(out)))
        return cls(out.value)

    def save(self, path: str) -> None:
        """saveSdf / saveSparseSdf (SdfIO.cpp:93-275)."""
        check(lib.sdfo_grid_save(self._h, os.fsencode(path)))

    # ---- lifetime ----------------------------------------------------------
    def close(self, stream=None) -> None:
        """Releases the grid.  With ``stream`` (a torch stream) the release is
        ordered after the work already on that stream, with no device-wide
        synchronisation; the caller guarantees no other stream still uses the
        grid.  Without it the device is synchronised first."""
        if getattr(self, "_h", None) is not None and self._h.value:
            if stream is None:
                lib.sdfo_grid_free(self._h)
            else:
                lib.sdfo_grid_free_async(self._h, _stream_handle(stream))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def is_sparse(self) -> bool:
        return bool(self.info.sparse)

    @property
    def bmin(self):
        return self.spec.bmin

    @property
    def bmax(self):
        return self.spec.bmax

    # ---- host copies -------------------------------------------------------
    def download_dense(self) -> dict:
        N = self.spec.total
        out = dict(phi=np.zeros(N), n=np.zeros((N, 3)), H=np.zeros((N, 6)),
                   witness=np.zeros((N, 3)), tri=np.zeros(N, np.int32))
        check(lib.sdfo_grid_download_dense(self._h, _np_ptr(out["phi"]), _np_ptr(out["n"]),
                                           _np_ptr(out["H"]), _np_ptr(out["witness"]),
                                           _np_ptr(out["tri"])))
        if self.info.candidate_stride > 0:
            S = self.info.candidate_stride
            cand = dict(count=np.zeros(N, np.uint8), phi=np.zeros((N, S)), n=np.zeros((N, S, 3)),
                        branch=np.zeros((N, S), np.int32), stride=S)
            check(lib.sdfo_grid_download_candidates(self._h, _np_ptr(cand["count"]),
                                                    _np_ptr(cand["phi"]), _np_ptr(cand["n"]),
                                                    _np_ptr(cand["branch"])))
            out["candidates"] = cand
        return out

    def download_sparse(self) -> dict:
        info = self.info
        nb, bd = info.block_count, info.block_dim
        bvc = bd ** 3
        cells = int(np.prod(info.block_grid[:]))
        S = info.candidate_stride
        out = dict(bijk=np.zeros((nb, 3), np.int32), phi=np.zeros((nb, bvc), np.float32),
                   n=np.zeros((nb, bvc, 3), np.float32), H=np.zeros((nb, bvc, 6), np.float32),
                   active=np.zeros(cells, np.uint8), block_dim=bd, stride=S)
        if S > 0:
            out.update(ccount=np.zeros((nb, bvc), np.uint8), cphi=np.zeros((nb, bvc, S), np.float32),
                       cn=np.zeros((nb, bvc, S, 3), np.float32),
                       cbranch=np.zeros((nb, bvc, S), np.int32))
        nil = lambda k: _np_ptr(out[k]) if k in out else None  # noqa: E731
        check(lib.sdfo_grid_download_sparse(self._h, nil("bijk"), nil("phi"), nil("n"), nil("H"),
                                            nil("active"), nil("ccount"), nil("cphi"), nil("cn"),
                                            nil("cbranch")))
        if not info.has_hessian:
            out["H"] = None
        return out

    # ---- queries -------------------------------------------------------------
    def query(self, points, mode="Tricubic", full=False, stream=None):
        """FP32 point stream -> (gap, raw gradient, valid[, unit normal, witness, branch]).

        ``points`` is an (N,3) float32 torch tensor on the GPU (contiguous).
        """
        import torch

        mode = MODES.get(mode, mode)
        pts = points.contiguous()
        if pts.dtype not in (torch.float32, torch.float64) or pts.dim() != 2 or pts.shape[1] != 3:
            raise ValueError("points must be an (N,3) float32/float64 tensor")
        if not pts.is_cuda:
            raise ValueError("points must be a GPU tensor")
        n = pts.shape[0]
        dt = pts.dtype
        gap = torch.empty(n, dtype=dt, device=pts.device)
        grad = torch.empty((n, 3), dtype=dt, device=pts.device)
        valid = torch.empty(n, dtype=torch.uint8, device=pts.device)
        normal = witness = branch = None
        if full:
            normal = torch.empty((n, 3), dtype=dt, device=pts.device)
            witness = torch.empty((n, 3), dtype=dt, device=pts.device)
            branch = torch.empty(n, dtype=torch.int32, device=pts.device)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
        fn = lib.sdfo_query if dt == torch.float32 else lib.sdfo_query_f64
        check(fn(self._h, int(mode), ptr(pts), n, ptr(gap), ptr(grad), ptr(valid), ptr(normal),
                 ptr(witness), ptr(branch), _stream_handle(stream)))
        if full:
            return gap, grad, valid.bool(), normal, witness, branch
        return gap, grad, valid.bool()

    def query_into(self, points, gap, grad, valid, mode="Tricubic", stream=None):
        """Launches one query kernel into caller-owned outputs (no allocation,
        no conversion): ``gap`` (N), ``grad`` (N,3) of the points' dtype and
        ``valid`` (N) uint8, all contiguous GPU tensors."""
        import torch

        mode = MODES.get(mode, mode)
        if points.dtype not in (torch.float32, torch.float64) or points.dim() != 2 or \
                points.shape[1] != 3:
            raise ValueError("points must be an (N,3) float32/float64 tensor")
        n = points.shape[0]
        if gap.shape != (n,) or grad.shape != (n, 3) or valid.shape != (n,) or \
                valid.dtype != torch.uint8 or gap.dtype != points.dtype or grad.dtype != points.dtype:
            raise ValueError("query_into output shapes/dtypes do not match the points")
        for t in (points, gap, grad, valid):
            if not (t.is_cuda and t.is_contiguous()):
                raise ValueError("query_into needs contiguous GPU tensors")
        fn = lib.sdfo_query if points.dtype == torch.float32 else lib.sdfo_query_f64
        ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(fn(self._h, int(mode), ptr(points), n, ptr(gap), ptr(grad), ptr(valid), None, None,
                 None, _stream_handle(stream)))

    def query_batch(self, Y, mode="Tricubic"):
        """The prototype's SDF contract (sdf_grid.py:138-204):
        ``(g, raw_grad, |grad|, unit_normal, valid)``; invalid points give
        g = 0, raw_grad = 0, unit_normal = (0, 0, 1)."""
        import torch

        gap, grad, valid = self.query(Y, mode)
        gn = torch.linalg.vector_norm(grad, dim=1)
        unit = torch.zeros_like(grad)
        unit[:, 2] = 1.0
        m = gn > 1e-30
        unit[m] = grad[m] / gn[m, None]
        gap = torch.where(valid, gap, torch.zeros_like(gap))
        return gap, grad, gn, unit, valid


class QueryPlan:
    """Batched FP32 query of many point streams against many grids in one
    launch.  ``run`` allocates, copies and synchronises nothing, so it can be
    captured in a ``torch.cuda.CUDAGraph``.  The grids must stay alive."""

    def __init__(self, grids, counts):
        import ctypes as _C

        self.grids = list(grids)
        counts = [int(c) for c in counts]
        if len(counts) != len(self.grids):
            raise ValueError("one point count per grid")
        self.offsets = np.zeros(len(counts) + 1, np.int64)
        self.offsets[1:] = np.cumsum(counts)
        handles = (_C.c_void_p * max(1, len(self.grids)))(*[g._h for g in self.grids])
        out = _C.c_void_p()
        check(lib.sdfo_query_plan_create(len(self.grids), handles, _np_ptr(self.offsets),
                                         _C.byref(out)))
        self._h = out.value

    @property
    def total(self) -> int:
        return int(self.offsets[-1])

    def run(self, points, gap, grad, valid, mode="Tricubic", stream=None):
        """points (N,3) f32, gap (N) f32, grad (N,3) f32, valid (N) u8: GPU
        tensors packed stream after stream (stream s owns rows
        offsets[s]:offsets[s+1])."""
        import torch

        mode = MODES.get(mode, mode)
        n = self.total
        specs = ((points, (n, 3), torch.float32), (gap, (n,), torch.float32),
                 (grad, (n, 3), torch.float32), (valid, (n,), torch.uint8))
        for t, _, dt in specs:
            if t.dtype != dt:
                raise ValueError(f"plan tensors: expected {dt}, got {t.dtype}")
        for t, shape, _ in specs:
            if tuple(t.shape) != shape or not t.is_cuda or not t.is_contiguous():
                raise ValueError("plan tensors must be contiguous GPU tensors of the plan's size")
        ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(lib.sdfo_query_plan_run(self._h, int(mode), ptr(points), ptr(gap), ptr(grad),
                                      ptr(valid), _stream_handle(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib.sdfo_query_plan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def library_version() -> str:
    return _lib.version()
