"""Shared pytest setup: import paths, the `gpu` marker, and common fixtures."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("sci-algorithms_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def gpu():
    """Skip cleanly when no device is present; fail loudly if the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from pnp_sci import _lib
    return _lib.lib()
