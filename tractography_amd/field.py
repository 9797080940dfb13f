"""Fields of local-model parameters (field.py of the reference).

``Image`` is a 4-D image sampled at the nearest voxel (rint, ties to even).
Its volume is uploaded to HBM once per runtime and reused by every tracking
call.
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod
from typing import Any

import numpy as np

from . import _native
from .image import image_data
from .runtime import DeviceHandle, HipRuntime
from .utils import apply_affine


class BaseField(ABC):
    """A field of local-model parameters (field.py:15-49)."""

    def __init__(self, n_parameters: int):
        self.n_parameters = n_parameters

    @abstractmethod
    def evaluate(self, points) -> Any:
        """Evaluate the field at world points with NumPy."""

    @abstractmethod
    def upload(self, runtime: HipRuntime) -> DeviceHandle:
        """Device handle of this field on ``runtime`` (uploaded once)."""


class Image(BaseField):
    """A 4-D image field with nearest-voxel lookup (field.py:52-141)."""

    def __init__(self, image_data_in):
        if image_data_in.ndim != 4:
            raise ValueError("Only 4D images are supported for image fields.")
        super().__init__(image_data_in.shape[3])
        self.values = np.ascontiguousarray(image_data(image_data_in), dtype=np.float32)
        self.dims = np.ascontiguousarray(image_data_in.shape, dtype=np.uint32)
        affine = image_data_in.affine
        self.inverse_affine = np.ascontiguousarray(
            np.linalg.inv(affine if affine is not None else np.eye(4)), dtype=np.float32
        )
        self._device: dict[int, DeviceHandle] = {}

    def upload(self, runtime: HipRuntime) -> DeviceHandle:
        key = id(runtime)
        cached = self._device.get(key)
        if cached is not None and cached.runtime is runtime:
            return cached
        lib = _native.lib()
        handle = ctypes.c_void_p()
        _native.check(
            lib.tgk_field_create(
                runtime.handle, _native.ptr(self.values), _native.ptr(self.dims),
                _native.ptr(np.ascontiguousarray(self.inverse_affine, dtype=np.float32)), ctypes.byref(handle),
            )
        )
        device = DeviceHandle(runtime, handle, lib.tgk_field_destroy)
        self._device[key] = device
        return device

    def evaluate(self, points) -> np.ndarray:
        """NumPy twin (field.py:124-141): rint voxel, zeros outside."""
        points = np.asarray(points)
        voxel = np.round(apply_affine(self.inverse_affine, points)).astype(int)
        x, y, z = voxel[:, :3].T
        width, height, depth, n_parameters = self.values.shape
        in_bounds = (0 <= x) & (x < width) & (0 <= y) & (y < height) & (0 <= z) & (z < depth)
        result = np.zeros((points.shape[0], n_parameters), dtype=np.float32)
        result[in_bounds, :] = self.values[tuple(voxel[in_bounds].T)]
        return result


def evaluate(field: BaseField, points, runtime: HipRuntime | None = None) -> np.ndarray:
    """Sample the field on the GPU (field.py:144-197): (n, K) float32."""
    points = np.asarray(points)
    n_points = points.shape[0]
    if n_points == 0:
        return np.empty((0, field.n_parameters), dtype=np.float32)
    if runtime is None:
        runtime = HipRuntime()
    device = field.upload(runtime)
    pts = np.ascontiguousarray(points[:, :3], dtype=np.float32)
    out = np.empty((n_points, field.n_parameters), dtype=np.float32)
    _native.check(
        _native.lib().tgk_field_evaluate(runtime.handle, device.handle, _native.ptr(pts), n_points, _native.ptr(out))
    )
    return out
