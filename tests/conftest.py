import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_files(pattern="seq_small_*.npz"):
    return sorted(glob.glob(os.path.join(GOLDEN, pattern)))


def load_golden(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(params=golden_files(), ids=lambda p: os.path.basename(p)[:-4])
def golden(request):
    return load_golden(request.param)


def pm(vol_dm):
    """Disparity-major [D][H][W] -> the reference's pixel-major [H][W][D]."""
    return np.ascontiguousarray(np.moveaxis(vol_dm, 0, -1))


def dm(vol_pm):
    return np.ascontiguousarray(np.moveaxis(vol_pm, -1, 0))
