import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_dir():
    import kqed_amd
    return kqed_amd.table_dir("small")


@pytest.fixture(scope="session")
def full_dir():
    import kqed_amd
    return kqed_amd.table_dir("full")


@pytest.fixture(scope="session")
def oracle_small(small_dir):
    from oracle_ctypes import Oracle
    o = Oracle(small_dir)
    yield o
    o.close()


@pytest.fixture(scope="session")
def oracle_full(full_dir):
    from oracle_ctypes import Oracle
    o = Oracle(full_dir)
    yield o
    o.close()


@pytest.fixture(scope="session")
def gpu_small(small_dir):
    if not has_gpu():
        pytest.skip("no GPU")
    import kqed_amd
    t = kqed_amd.QEDKernelTemps.initialise(small_dir)
    yield t
    t.free()


@pytest.fixture(scope="session")
def gpu_full(full_dir):
    if not has_gpu():
        pytest.skip("no GPU")
    import kqed_amd
    t = kqed_amd.QEDKernelTemps.initialise(full_dir)
    yield t
    t.free()
