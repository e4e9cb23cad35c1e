import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "sdf-contact-module_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def O():
    import oracle

    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def sdfo_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import sdfo

    return sdfo


def ograd(O, spec):
    return O.Grid(spec.bmin, spec.bmax, spec.voxel, spec.shape)


def f32(a):
    import numpy as np

    return None if a is None else np.asarray(a, np.float32).astype(np.float64)
