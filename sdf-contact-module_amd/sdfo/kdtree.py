"""Exact nearest-triangle and signed-distance queries on a triangle mesh.

Mirror of the reference's ``sdf::KdTree`` (plugins/SdfOracle/include/SdfOracle/
KdTree.h:86-170): the tree is built on the host exactly as the SDF builders
build it and kept on the GPU; ``find_nearest`` / ``find_signed_nearest``
answer a whole point stream in one launch, each point bit-identical to the
reference's recursive search (and ray-parity sign for the signed form).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .grid import _stream_handle


class KdTree:
    """``KdTree::build`` over a host mesh (vertices (V,3), triangles (T,3))."""

    def __init__(self, vertices, triangles):
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        h = C.c_void_p()
        check(lib.sdfo_mesh_build(v.ctypes.data_as(C.c_void_p), len(v), t.ctypes.data_as(C.c_void_p),
                                  len(t), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib.sdfo_mesh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass

    def triangle_count(self) -> int:
        n = C.c_int64()
        check(lib.sdfo_mesh_triangle_count(self.handle, C.byref(n)))
        return n.value

    def _query(self, points, signed, stream):
        if not points.is_cuda or points.dtype != torch.float64 or points.dim() != 2 \
                or points.shape[1] != 3:
            raise ValueError("points must be an (N,3) float64 GPU tensor")
        p = points.contiguous()
        n = p.shape[0]
        dev = p.device
        dist = torch.empty(n, dtype=torch.float64, device=dev)
        cp = torch.empty((n, 3), dtype=torch.float64, device=dev)
        nrm = torch.empty((n, 3), dtype=torch.float64, device=dev)
        tri = torch.empty(n, dtype=torch.int32, device=dev)
        ptr = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
        check(lib.sdfo_mesh_query(self.handle, ptr(p), n, 1 if signed else 0, ptr(dist), ptr(cp),
                                  ptr(nrm), ptr(tri), _stream_handle(stream)))
        return dist, cp, nrm, tri

    def find_nearest(self, points, stream=None):
        """``findNearest`` per point: (distance, closest point, face normal,
        triangle index); distance DBL_MAX and index -1 for an empty mesh."""
        return self._query(points, False, stream)

    def find_signed_nearest(self, points, stream=None):
        """``findSignedNearest`` per point: (signed distance, closest point,
        face normal, triangle index); the sign from the ray-parity count, or
        from the face normal when the ray crosses nothing."""
        return self._query(points, True, stream)
