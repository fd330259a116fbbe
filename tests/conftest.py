"""Test configuration: the `gpu` marker and import paths.

CPU tests (-m "not gpu") check the oracle against closed forms, brute force
and the golden fixtures, the host-side code, and that libpbf.so exports every
symbol of include/pbf.h.  GPU tests (-m gpu) compare libpbf's HIP path with
the oracle through the C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "webgpu-fluid-simulation_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpbf's HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def repo_root():
    return ROOT
