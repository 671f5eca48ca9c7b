import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


_gpu_ok = None


def gpu_available() -> bool:
    global _gpu_ok
    if _gpu_ok is None:
        try:
            import torch
            _gpu_ok = bool(torch.cuda.is_available())
        except Exception:
            _gpu_ok = False
    return _gpu_ok


def pytest_collection_modifyitems(config, items):
    # GPU tests are skipped (not failed) only when no HIP device exists at
    # all, i.e. when someone runs the whole suite on a CPU-only machine.
    if any("gpu" in it.keywords for it in items) and not gpu_available():
        skip = pytest.mark.skip(reason="no HIP device visible")
        for it in items:
            if "gpu" in it.keywords:
                it.add_marker(skip)
