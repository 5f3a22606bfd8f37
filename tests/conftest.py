"""Shared pytest setup: the `gpu` marker and import paths.

GPU tests call libfdfd through its C ABI (ctypes); CPU tests cover the
oracle, host logic and the ABI surface without launching kernels.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fdfd.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libfdfd kernels)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def engine():
    from fdfd.solve import Engine
    e = Engine(0)
    yield e
    e.close()
