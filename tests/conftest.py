import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")

# Parity tolerance of the evaluator contract: max |x - x_ref| over the
# bounding-box diagonal of the reference output, FP32.
BBOX_TOL = 1e-5


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bbox_rel_err(x, ref):
    x = np.asarray(x, np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, np.float64).reshape(len(ref), -1)
    if ref.size == 0:
        return 0.0
    diag = np.linalg.norm(ref.max(axis=0) - ref.min(axis=0))
    diag = diag if diag > 0 else 1.0
    return float(np.abs(x - ref).max() / diag)


@pytest.fixture(scope="session")
def hip():
    """libosdhip + torch on cuda:0, for @pytest.mark.gpu tests."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from opensubdiv_amd import _lib
    return _lib.load_hip()
