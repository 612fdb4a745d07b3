"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "cuda-n4_amd"))

GOLDEN = os.path.join(REPO, "tests", "golden", "n4py_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The product library on the GPU; raises (test error) when missing."""
    import n4amd
    lib = n4amd.load()
    print("loaded", lib.path, lib.version(), flush=True)
    return lib


@pytest.fixture(scope="session")
def emu():
    """Host emulation of the kernel bodies (CPU unit tests only)."""
    import n4amd
    return n4amd.load(emulate=True)
