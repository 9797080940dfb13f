"""The device runtime: one GPU, one HIP stream (replaces opencl.py).

``HipRuntime`` owns a ``tgk_runtime`` handle of libtgk.so. Fields and domains
upload their arrays once per runtime and keep the device handles, so repeated
tracking calls (and every batch of one call) reuse the volumes in HBM instead
of re-uploading them as the reference does per batch.
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _native

PATHS = {"plain": _native.PATH_PLAIN, "fast": _native.PATH_FAST}


class HipRuntime:
    """A GPU execution context for the tracking kernels.

    Args:
        device: HIP device index.
        path: ``"plain"`` (the reference's integrator, bit-exact with the FP32
            host restatement; default) or ``"fast"`` (persistent lanes with
            refill, transcendental-free model evaluation, FP64 arc length).
        precision: coefficient storage read by the fast path. Only
            ``"fp32"`` is served; ``"bf16x2"``, ``"fp16"`` and ``"bf16"`` are
            accepted names, but a fast-path call then raises
            NotImplementedError (FP16 or BF16 coefficients move tracks by
            1.2e-2 to 2.4e-2 voxel, past the 1e-3 budget; split BF16 uses 97%
            of the budget by itself and stores as many bytes as FP32).
    """

    def __init__(self, device: int = 0, *, path: str = "plain", precision: str = "fp32"):
        if path not in PATHS:
            raise ValueError(f"path must be one of {sorted(PATHS)}")
        if precision not in _native.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_native.PRECISIONS)}")
        lib = _native.lib()
        handle = ctypes.c_void_p()
        _native.check(lib.tgk_runtime_create(int(device), ctypes.byref(handle)))
        self._handle = handle
        self.device = int(device)
        self.path = path
        self.precision = precision
        self.fast_shape = 0
        # Handles made on this runtime: closing the runtime destroys them
        # first, since each one frees its device memory through it.
        self._children: weakref.WeakSet = weakref.WeakSet()
        self._finalizer = weakref.finalize(self, lib.tgk_runtime_destroy, handle)

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._finalizer.alive:
            raise RuntimeError("the runtime has been closed")
        return self._handle

    @property
    def path_id(self) -> int:
        return PATHS[self.path]

    def finish(self) -> None:
        """Wait for all work submitted to this runtime."""
        _native.check(_native.lib().tgk_runtime_synchronize(self._handle))

    def info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cus = ctypes.c_int32()
        total = ctypes.c_int64()
        _native.check(
            _native.lib().tgk_runtime_info(self._handle, name, 256, ctypes.byref(cus), ctypes.byref(total))
        )
        return {"name": name.value.decode(), "compute_units": cus.value, "total_bytes": total.value}

    def set_fast_grid(self, blocks: int) -> None:
        """Blocks of the fast path's persistent kernel (<= 0: occupancy)."""
        _native.check(_native.lib().tgk_runtime_set_fast_grid(self._handle, int(blocks)))
        self.fast_grid = int(blocks) if blocks > 0 else 0

    def set_fast_shape(self, shape: int) -> None:
        """Launch shape of the fast path's order-8 SH Transport and Diffusion
        kernels (see fast_shapes()); other kernels run shape 0."""
        _native.check(_native.lib().tgk_runtime_set_fast_shape(self._handle, int(shape)))
        self.fast_shape = int(shape)

    def set_fast_repack(self, min_streamlines: int) -> None:
        """Streamlines (half a generation's work) below which the fast path
        no longer cuts a generation short to repack its streamlines."""
        _native.check(_native.lib().tgk_runtime_set_fast_repack(self._handle, int(min_streamlines)))

    def fast_reads(self) -> dict:
        """Coefficient rows and mask values the last fast-path tracking launch
        read from memory (each lane reads them only when its cell changes)."""
        rows = ctypes.c_uint64()
        masks = ctypes.c_uint64()
        _native.check(_native.lib().tgk_runtime_fast_reads(self._handle, ctypes.byref(rows), ctypes.byref(masks)))
        return {"rows": rows.value, "mask_values": masks.value}

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _native.check(_native.lib().tgk_runtime_stream(self._handle, ctypes.byref(s)))
        return s.value or 0

    def close(self) -> None:
        """Free the runtime and every field and domain handle made on it
        (they become unusable)."""
        for child in list(self._children):
            child.release()
        self._finalizer()


def fast_shapes() -> list[dict]:
    """The fast path's launch shapes: lanes per block (streamlines per
    workgroup) and how many of a row's 45 coefficients stay in LDS between
    steps (the rest in registers). No device is needed."""
    shapes = []
    for shape in range(_native.FAST_SHAPES):
        lanes = ctypes.c_int32()
        lds = ctypes.c_int32()
        _native.check(_native.lib().tgk_fast_shape_info(shape, ctypes.byref(lanes), ctypes.byref(lds)))
        shapes.append({"shape": shape, "lanes_per_block": lanes.value, "lds_coefficients": lds.value})
    return shapes


class DeviceHandle:
    """A libtgk object (field or domain) owned by one runtime."""

    def __init__(self, runtime: HipRuntime, handle: ctypes.c_void_p, destroy):
        if not runtime._finalizer.alive:
            destroy(handle)
            raise RuntimeError("the runtime has been closed")
        self.runtime = runtime
        self._handle = handle
        self._finalizer = weakref.finalize(self, destroy, handle)
        runtime._children.add(self)

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._finalizer.alive:
            raise RuntimeError("the device handle has been released (its runtime was closed)")
        return self._handle

    def release(self) -> None:
        self._finalizer()


def as_float32_affine(matrix) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(matrix, dtype=np.float32).reshape(4, 4))
