import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    d = os.path.join(ROOT, "tests", "golden")

    class G:
        @staticmethod
        def icos(N):
            return np.load(os.path.join(d, f"icos_{N}.npy"))

        md25 = staticmethod(lambda: np.load(os.path.join(d, "md25.npy")))
        md25_knn10 = staticmethod(lambda: np.load(os.path.join(d, "md25_knn10.npy")))
        path = d

    return G
