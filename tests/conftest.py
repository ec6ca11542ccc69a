import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "paint-sim-3d_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcanvaz_film.so")
    config.addinivalue_line("markers", "slow: long CPU restatement runs")


@pytest.fixture(scope="session")
def gpu():
    import torch  # noqa: F401  (initialises the HIP runtime like the product's callers do)
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return True
