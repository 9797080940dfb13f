"""Shared pytest configuration: the `gpu` marker and import paths."""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


# PyTorch ships its own HIP runtime library. Imported after libtgk.so has
# initialised the HIP runtime of /opt/rocm, torch finds no device ("No HIP
# GPUs are available"); imported first, both work. Tests that hand torch
# buffers to the library therefore import torch before any runtime exists.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libtgk.so")


@pytest.fixture(scope="session")
def runtime():
    """One HIP runtime for the GPU tests (fails loudly without a device)."""
    import tractography_amd as tg

    return tg.HipRuntime()


@pytest.fixture(scope="session")
def oracle_threads():
    return int(os.environ.get("TGO_THREADS", min(16, os.cpu_count() or 1)))
