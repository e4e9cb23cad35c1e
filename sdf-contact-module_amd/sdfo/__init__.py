"""sdfo -- MI355X (gfx950) SdfOracle: GPU SDF build, query and contact reduction.

Python side of the C ABI in include/sdfo.h.  Device buffers are torch
tensors; the compute is the in-tree HIP library ``_lib/libsdfo.so`` (no CPU
fallback: import fails if the library is not built).
"""
from ._lib import MODES, SdfoError, default_build_params, version  # noqa: F401
from ._lib import tuner_report, tuner_reset, tuner_set_mode  # noqa: F401
from .grid import GridSpec, QueryPlan, SdfGrid, setup_grid  # noqa: F401
from .kdtree import KdTree  # noqa: F401

QueryMode = MODES
