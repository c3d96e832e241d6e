import os
import sys

import pytest

# PyTorch-ROCm bundles its own HIP runtime under the same soname as the
# system one.  Loading torch first makes libamdflow.so bind to that single
# runtime; loading libamdflow.so first would bring in a second HIP runtime,
# after which torch finds no device.
try:
    import torch  # noqa: F401
except Exception:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "sycl-optic-flow_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libamdflow.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
