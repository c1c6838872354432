"""Shared test setup: markers, import paths and the oracle/library handles.

`-m "not gpu"` runs the oracle against the golden fixtures, the host-side
model code and the library's exported symbols; `-m gpu` runs the HIP paths
through the C ABI and compares them with the oracle.
"""
import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (REPO, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("distributed-cube-average_amd")


@pytest.fixture(scope="session")
def H(pkg):
    """The ctypes bindings of libdc_hip.so."""
    pkg.dc_hip.load()
    return pkg.dc_hip


@pytest.fixture(scope="session")
def VM(pkg):
    return importlib.import_module("distributed-cube-average_amd.velocity_model")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
