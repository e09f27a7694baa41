import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import binding  # noqa: E402  (oracle binding, test infrastructure)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def vv():
    return binding.load_package()


@pytest.fixture(scope="session")
def orc(vv):
    return binding.load(vv)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
