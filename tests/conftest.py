import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvoxelcore.so)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def vc():
    """Session context on cuda:0 through the C ABI; fails loudly when the
    library or the GPU is missing (there is no CPU fallback)."""
    import voxel_ma_amd
    ctx = voxel_ma_amd.Context(0)
    yield ctx
    ctx.close()


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="session")
def simple_lut_cpu(oracle):
    return oracle.simple_lut()
