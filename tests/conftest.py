"""Shared test setup: import paths for the product binding (cwp) and the oracle (pyoracle), the gpu marker, and
fixtures that build the scenes once per session."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda-wavefront-pathtracer_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the kernels")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
