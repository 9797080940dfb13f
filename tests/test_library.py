"""libtgk.so loads and exports every entry point include/tgk.h declares
(no device calls)."""

import ctypes
import re
from pathlib import Path

import pytest

from tractography_amd import _native

HEADER = Path(__file__).resolve().parent.parent / "include" / "tgk.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tgk_\w+)\s*\(", text, re.M)))


def test_header_declares_the_bound_interface():
    assert declared_symbols() == sorted(_native.EXPORTED_SYMBOLS)


@pytest.mark.skipif(not _native.library_available(), reason="libtgk.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIBRARY_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert _native.lib().tgk_version() == 1
