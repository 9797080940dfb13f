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
