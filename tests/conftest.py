import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def golden_small():
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "bvh_small.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
