import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def mhdlib():
    """The HIP library; built in-tree if it is not there yet."""
    from resistivemhd_gpu_amd import build, mhd2d
    if not os.path.exists(mhd2d.LIB_PATH):
        build.build()
    return mhd2d.load_library()
