"""Compatibility names for callers of the reference's opencl module.

``OpenCLRuntime`` is the HIP runtime (opencl.py:30-163 is replaced by
runtime.HipRuntime); code that constructs one keeps working.
"""

from enum import Enum

from .runtime import HipRuntime

OpenCLRuntime = HipRuntime


class BufferAccess(Enum):
    """Access modes named as in the reference (opencl.py:22-27)."""

    READ_ONLY = 1
    WRITE_ONLY = 2
    READ_WRITE = 3


__all__ = ["BufferAccess", "HipRuntime", "OpenCLRuntime"]
