import os
import sys

import pytest

# PyTorch (device buffers for the device-pointer entry points) carries its own HIP runtime; loading
# it before the engine library makes the process use one runtime for both, as bench.py does
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "xmvb-cpp_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device 0)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")
