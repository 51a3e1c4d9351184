"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

REF_ROOT = Path("/root/reference")
CORNELL_LFS = REF_ROOT / "build_include/assets/cornell_box.lfs"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    from lavaframe_amd import lfh
    return lfh.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import native
    return native.lib()


def pytest_runtest_setup(item):
    if "gpu" in item.keywords and not gpu_available():
        pytest.skip("no GPU in this environment")
