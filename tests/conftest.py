"""Shared set-up of the test suite.

Markers: `gpu` tests need an MI355X (they run the HIP kernels through the C
ABI and compare with the CPU oracle); everything else runs on the CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
for p in (ROOT, os.path.join(ROOT, "plane-sweep-cuda_amd", "python"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def views_small():
    from psw_amd import scene
    return scene.make_views(seed=3, width=160, height=120, nsrc=3)


@pytest.fixture(scope="session")
def views_tiny():
    from psw_amd import scene
    return scene.make_views(seed=5, width=64, height=48, nsrc=2)


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) without a device: a GPU test must run the kernels."""
    import psw_amd
    n = psw_amd.device_count()
    assert n > 0, "no HIP device visible"
    return psw_amd


def one_based(K):
    from psw_amd import scene
    return scene.one_based(K)


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
