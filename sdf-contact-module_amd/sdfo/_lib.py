"""ctypes binding of libsdfo (include/sdfo.h).

The library is built in-tree (``sdf-contact-module_amd/sdfo/_lib/libsdfo.so``).
There is no fallback: importing this module raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsdfo.so")

SDFO_OK = 0
ERRORS = {
    -1: "SDFO_ERR_INVALID_ARGUMENT",
    -2: "SDFO_ERR_EMPTY_MESH",
    -3: "SDFO_ERR_IO",
    -4: "SDFO_ERR_FORMAT",
    -5: "SDFO_ERR_DEVICE",
    -6: "SDFO_ERR_OUT_OF_MEMORY",
    -7: "SDFO_ERR_UNSUPPORTED",
}

MODES = {"FirstOrder": 0, "SecondOrder": 1, "Trilinear": 2, "Tricubic": 3, "ContactAware": 4}


class BuildParams(C.Structure):
    _fields_ = [
        ("padding", C.c_double),
        ("voxel_size", C.c_double),
        ("target_resolution", C.c_int32),
        ("max_resolution", C.c_int32),
        ("block_dim", C.c_int32),
        ("halo_voxels", C.c_int32),
        ("band_voxels", C.c_int32),
        ("store_hessian", C.c_int32),
        ("band_distance", C.c_double),
        ("max_candidates", C.c_int32),
        ("tile_dim", C.c_int32),
        ("feature_distance_tolerance", C.c_double),
        ("normal_merge_angle_degrees", C.c_double),
        ("use_tile_acceleration", C.c_int32),
        ("reserved", C.c_int32),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("bmin", C.c_double * 3),
        ("bmax", C.c_double * 3),
        ("voxel", C.c_double * 3),
        ("shape", C.c_int32 * 3),
        ("sparse", C.c_int32),
        ("has_hessian", C.c_int32),
        ("candidate_stride", C.c_int32),
        ("block_dim", C.c_int32),
        ("halo_voxels", C.c_int32),
        ("block_count", C.c_int32),
        ("block_grid", C.c_int32 * 3),
        ("band_distance", C.c_double),
        ("activation_distance", C.c_double),
        ("device_bytes", C.c_int64),
    ]


class ContactParams(C.Structure):
    _fields_ = [
        ("kn", C.c_double),
        ("damping", C.c_double),
        ("mu", C.c_double),
        ("regularizer", C.c_double),
        ("narrow_margin", C.c_double),
        ("order", C.c_int32),
        ("mode", C.c_int32),
    ]


class SdfoError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code
        self.message = message


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libsdfo is not built ({LIB_PATH} missing); run __graft_entry__.build() "
        "or `make -C sdf-contact-module_amd`")

# torch is loaded first: its bundled HIP runtime has the same soname
# (libamdhip64.so.7) as the one libsdfo links against, so the library binds to
# the runtime that owns torch's device memory and streams instead of loading a
# second HIP runtime into the process.
import torch  # noqa: E402,F401

lib = C.CDLL(LIB_PATH)

_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_P = C.POINTER


def _sig(name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("sdfo_version", [], C.c_char_p)
_sig("sdfo_last_error", [], C.c_char_p)
_sig("sdfo_default_build_params", [_P(BuildParams)], None)
_sig("sdfo_setup_grid", [_vp, _i64, _P(BuildParams), _P(GridInfo)])
_sig("sdfo_build_dense", [_vp, _i64, _vp, _i64, _P(BuildParams), _P(_vp), _vp])
_sig("sdfo_build_sparse", [_vp, _i64, _vp, _i64, _P(BuildParams), _P(_vp), _vp])
_sig("sdfo_build_contact_aware", [_vp, _i64, _vp, _i64, _vp, _P(BuildParams), _i32, _P(_vp), _vp])
_sig("sdfo_build_dense_batched", [_i32, _vp, _vp, _vp, _vp, _P(BuildParams), _P(_vp), _vp])
_sig("sdfo_grid_free", [_vp], None)
_sig("sdfo_grid_free_async", [_vp, _vp], None)
_sig("sdfo_grid_get_info", [_vp, _P(GridInfo)])
_sig("sdfo_grid_download_dense", [_vp, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_grid_download_candidates", [_vp, _vp, _vp, _vp, _vp])
_sig("sdfo_grid_download_sparse", [_vp] + [_vp] * 9)
_sig("sdfo_grid_save", [_vp, C.c_char_p])
_sig("sdfo_grid_load", [C.c_char_p, _P(_vp)])
_sig("sdfo_is_valid_file", [C.c_char_p])
_sig("sdfo_is_sparse_file", [C.c_char_p])
_sig("sdfo_grid_from_dense", [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp,
                              _P(_vp)])
_sig("sdfo_query", [_vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_query_f64", [_vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_query_host_f64", [_vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp, _vp])
_sig("sdfo_query_batched", [_i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_query_plan_create", [_i32, _vp, _vp, _P(_vp)])
_sig("sdfo_query_plan_run", [_vp, _i32, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_query_plan_free", [_vp], None)
_sig("sdfo_tuner_set_mode", [_i32])
_sig("sdfo_tuner_reset", [])
_sig("sdfo_tuner_report", [_vp, _i64], C.c_int64)
_sig("sdfo_contact_forward", [_vp, _vp, _vp, _i64, _vp, _i64, _P(ContactParams), _vp, _vp, _vp])
_sig("sdfo_contact_evaluate", [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _P(ContactParams), _vp,
                               _vp, _vp, _vp])
_sig("sdfo_contact_evaluate_noalloc", [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _P(ContactParams),
                                       _vp, _vp, _vp, _vp])
_sig("sdfo_contact_linearize", [_vp, _vp, _i64, _vp, _P(ContactParams), _vp, _vp, _vp, _vp])
_sig("sdfo_contact_forward_linearized", [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64,
                                         _P(ContactParams), _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_mesh_build", [_vp, _i64, _vp, _i64, _P(_vp)])
_sig("sdfo_mesh_free", [_vp], None)
_sig("sdfo_mesh_triangle_count", [_vp, _P(_i64)])
_sig("sdfo_mesh_query", [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp])
_sig("sdfo_point_bvh_build", [_vp, _i64, _i32, _P(_vp)])
_sig("sdfo_point_bvh_free", [_vp], None)
_sig("sdfo_point_bvh_get_info", [_vp, _P(_i64), _P(_i64), _P(_i64)])
_sig("sdfo_point_bvh_query", [_vp, _vp, _vp, C.c_double, _vp, _P(_i64)])
_sig("sdfo_point_bvh_cull", [_vp, _vp, _i64, C.c_double, _vp, _vp])
_sig("sdfo_contact_forward_culled", [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _P(ContactParams),
                                     _vp, _vp, _vp])
_sig("sdfo_contact_backward", [_vp, _vp, _vp, _i64, _vp, _i64, _P(ContactParams), _vp, _vp, _vp,
                               _vp, _vp, _vp])
_sig("sdfo_contact_saved_bytes", [_i64, _i64], C.c_int64)
_sig("sdfo_contact_forward_save", [_vp, _vp, _vp, _i64, _vp, _i64, _P(ContactParams), _vp, _vp,
                                   _vp, _i64, _vp])
_sig("sdfo_contact_backward_saved", [_vp, _vp, _vp, _i64, _vp, _i64, _P(ContactParams), _vp, _vp,
                                     _i64, _vp, _vp, _vp, _vp, _vp])

_sig("sdfo_debug_bvh_export", [_vp, _i64, _vp, _i64, _vp, _vp, _vp, _i32, _vp])
_sig("sdfo_debug_device_bvh_export", [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp])
_sig("sdfo_debug_random_read_floor", [_i32, _i64, _i64, _i32, _P(C.c_double)])
_sig("sdfo_debug_sort_by_abs", [_vp, _i32, _vp])
_sig("sdfo_debug_build_phases", [_vp, _i32])
_sig("sdfo_debug_force_replay", [_i32])
_sig("sdfo_debug_fill_budget", [_i32, _i32])
_sig("sdfo_debug_ca_merge_cap", [_i32, _i32])
_sig("sdfo_debug_skip_cell_records", [_i32])
_sig("sdfo_debug_tuner_pin", [C.c_char_p, _i32])

BUILD_PHASES = ("setup_grid", "host_bvh", "upload_alloc", "fill_kernel", "hessian_kernel",
                "mirror_kernels", "total", "replayed_voxels")


def debug_fill_budget(budget: int = 0, heavy_cap: int = 0):
    """Test switch: fill step budget and batched heavy-list capacity (0 = default)."""
    check(lib.sdfo_debug_fill_budget(int(budget), int(heavy_cap)))


def debug_ca_merge_cap(first: int = 0, second: int = 0):
    """Test switch: merged entries per voxel in the contact-aware first pass (1, 4, 8)
    and in the second pass's first launch (0 = defaults)."""
    check(lib.sdfo_debug_ca_merge_cap(int(first), int(second)))


def debug_random_read_floor(record_bytes: int, n_reads: int, array_bytes: int, reps: int = 10):
    """Best microseconds for n_reads uniformly random record reads (the
    record-path query's memory floor; measurement support)."""
    out = C.c_double()
    check(lib.sdfo_debug_random_read_floor(int(record_bytes), int(n_reads), int(array_bytes),
                                           int(reps), C.byref(out)))
    return out.value


def debug_skip_cell_records(on: bool):
    """Grids built or loaded on this thread while on keep no cell records
    (test switch: the per-corner gather path of very large grids)."""
    check(lib.sdfo_debug_skip_cell_records(1 if on else 0))


def debug_force_replay(on: bool):
    """Test switch: send every voxel of this thread's builds to the replay pass."""
    check(lib.sdfo_debug_force_replay(1 if on else 0))


def debug_build_phases():
    """Phase timings (ms) of this thread's last dense build."""
    import numpy as np

    ms = np.zeros(len(BUILD_PHASES))
    lib.sdfo_debug_build_phases(ms.ctypes.data_as(C.c_void_p), len(ms))
    return dict(zip(BUILD_PHASES, ms.tolist()))


def debug_sort_by_abs(values):
    """Order the contact-aware builder gives to merged candidates (test support)."""
    import numpy as np

    v = np.ascontiguousarray(values, np.float64)
    order = np.zeros(len(v), np.int32)
    check(lib.sdfo_debug_sort_by_abs(v.ctypes.data_as(C.c_void_p), len(v),
                                     order.ctypes.data_as(C.c_void_p)))
    return order


def debug_bvh_export(verts, tris):
    """Host BVH (nodes6, nodeI4, perm, depth) -- test support, no device use."""
    import numpy as np

    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    cap = 2 * len(t) + 1
    n6 = np.zeros((cap, 6))
    i4 = np.zeros((cap, 4), np.int32)
    perm = np.zeros(len(t), np.int32)
    depth = C.c_int32()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    nn = lib.sdfo_debug_bvh_export(p(v), len(v), p(t), len(t), p(n6), p(i4), p(perm), cap,
                                   C.byref(depth))
    check(nn if nn < 0 else 0)
    return n6[:nn].copy(), i4[:nn].copy(), perm, depth.value


def debug_device_bvh_export(meshes):
    """The device-assembled BVH of [(verts, tris), ...] concatenated:
    (nodes6, nodeI4, perm, tri9, box6, depth) -- test support, current device."""
    import numpy as np

    vs = [np.ascontiguousarray(v, np.float64).reshape(-1, 3) for v, _ in meshes]
    ts = [np.ascontiguousarray(t, np.int32).reshape(-1, 3) for _, t in meshes]
    vo = np.concatenate([[0], np.cumsum([len(v) for v in vs])]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum([len(t) for t in ts])]).astype(np.int64)
    v = np.ascontiguousarray(np.concatenate(vs))
    t = np.ascontiguousarray(np.concatenate(ts))
    cap = 2 * len(t) + len(ts)
    n6 = np.zeros((cap, 6))
    i4 = np.zeros((cap, 4), np.int32)
    perm = np.zeros(len(t), np.int32)
    tri9 = np.zeros((len(t), 9))
    box6 = np.zeros((len(t), 6))
    depth = C.c_int32()
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    nn = lib.sdfo_debug_device_bvh_export(len(ts), p(v), p(vo), p(t), p(to), p(n6), p(i4), p(perm),
                                          p(tri9), p(box6), cap, C.byref(depth))
    check(nn if nn < 0 else 0)
    return n6[:nn].copy(), i4[:nn].copy(), perm, tri9, box6, depth.value


def check(rc: int) -> None:
    if rc != SDFO_OK:
        raise SdfoError(rc, lib.sdfo_last_error().decode(errors="replace"))


def version() -> str:
    return lib.sdfo_version().decode()


def default_build_params(**overrides) -> BuildParams:
    p = BuildParams()
    lib.sdfo_default_build_params(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown builder parameter {k!r}")
        setattr(p, k, v)
    return p


TUNE_OFF, TUNE_ON_FIRST_USE = 0, 1


def tuner_set_mode(mode: int):
    """SDFO_TUNE_OFF / SDFO_TUNE_ON_FIRST_USE (process-wide)."""
    check(lib.sdfo_tuner_set_mode(int(mode)))


def tuner_reset():
    check(lib.sdfo_tuner_reset())


def debug_tuner_pin(key: str, variant: int):
    """Test switch: make `variant` the tuner's decision for `key`."""
    check(lib.sdfo_debug_tuner_pin(key.encode(), int(variant)))


def tuner_report() -> str:
    n = lib.sdfo_tuner_report(None, 0)
    buf = C.create_string_buffer(int(n) + 1)
    lib.sdfo_tuner_report(buf, len(buf))
    return buf.value.decode()
