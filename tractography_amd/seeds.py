"""Seed generation and validation (seeds.py of the reference).

Seeds are (n, 8) float32 rows: homogeneous position (x, y, z, 1) followed by
homogeneous direction (dx, dy, dz, 0).
"""

from __future__ import annotations

import numpy as np
import numpy.typing as npt

from . import _native, core
from . import local as local_models
from .image import image_data
from .runtime import HipRuntime
from .utils import apply_affine

METHODS = {"valu": _native.SEEDS_VALU, "mfma": _native.SEEDS_MFMA}


def _from_components(locations: npt.ArrayLike, orientations: npt.ArrayLike) -> npt.NDArray[np.float32]:
    """Canonical (n, 8) homogeneous seed array (seeds.py:16-32)."""
    locations = np.asarray(locations, dtype=np.float32)
    orientations = np.asarray(orientations, dtype=np.float32)
    if locations.ndim != 2 or locations.shape[1] != 3:
        raise ValueError("locations must have shape (n_seeds, 3).")
    if orientations.shape != locations.shape:
        raise ValueError("orientations must have the same shape as locations.")
    result = np.empty((len(locations), 8), dtype=np.float32)
    result[:, :3] = locations
    result[:, 3] = 1.0
    result[:, 4:7] = orientations
    result[:, 7] = 0.0
    return result


def validate_seed_array(seeds: npt.ArrayLike) -> npt.NDArray[np.float32]:
    """C-contiguous float32 copy of a valid seed array (seeds.py:35-54)."""
    result = np.ascontiguousarray(seeds, dtype=np.float32)
    if result.ndim != 2 or result.shape[1] != 8:
        raise ValueError("seeds must have shape (n_seeds, 8).")
    if not np.all(result[:, 3] == 1) or not np.all(result[:, 7] == 0):
        raise ValueError("seed position and orientation homogeneous values must be 1 and 0.")
    return result


def from_mask(mask, n_seeds: int) -> npt.NDArray[np.float32]:
    """Uniform seeds in the nonzero voxels of a mask, uniform directions
    (seeds.py:114-142; draws from NumPy's global legacy generator)."""
    if n_seeds == 0:
        return np.empty((0, 8), dtype=np.float32)
    voxels = np.array(list(zip(*np.nonzero(image_data(mask)))))
    indices = np.random.randint(len(voxels), size=n_seeds)
    locations_voxel = voxels[indices] + np.random.rand(n_seeds, 3) - [0.5, 0.5, 0.5]
    locations = apply_affine(mask.affine, locations_voxel)
    orientations = np.random.randn(n_seeds, 3)
    orientations /= np.linalg.norm(orientations, axis=1, keepdims=True)
    return _from_components(locations, orientations)


def sparse_fod(fod):
    """Nonzero-voxel coefficients (nnz, K) f32, voxel indices (nnz, 4) f32 with
    w = 1, and the forward affine as float32 (seeds.py:178-183)."""
    data = image_data(fod)
    mask = np.any(data != 0, axis=-1)
    indices = np.array(np.nonzero(mask), dtype=np.float32).T
    indices = np.hstack((indices, np.ones((len(indices), 1)))).astype(np.float32)
    inline = np.ascontiguousarray(data[mask], dtype=np.float32)
    affine = np.ascontiguousarray(np.asarray(fod.affine, dtype=np.float32))
    return inline, np.ascontiguousarray(indices), affine


def from_fod(
    fod,
    n_seeds: int,
    use_opencl: bool = True,
    *,
    runtime: HipRuntime | None = None,
    method: str = "valu",
    states: npt.NDArray[np.uint32] | None = None,
) -> npt.NDArray[np.float32]:
    """Seeds uniform in nonzero-FOD voxels, directions importance-sampled from
    the local FOD over 1000 Fibonacci directions (seeds.py:145-240).

    ``use_opencl`` keeps the reference's name for the device-vs-host switch:
    True runs the GPU kernel. ``method`` selects the kernel: ``"valu"``
    reproduces seeds.cl bit for bit; ``"mfma"`` evaluates the 1000 directions
    of every seed as one split-bfloat16 matrix product. ``states`` (n, 2)
    uint32 replaces the default draw from NumPy's global generator.
    """
    if n_seeds == 0:
        return np.empty((0, 8), dtype=np.float32)
    if use_opencl:
        if method not in METHODS:
            raise ValueError(f"method must be one of {sorted(METHODS)}")
        if runtime is None:
            runtime = HipRuntime()
        inline, indices, affine = sparse_fod(fod)
        model = local_models.from_shape(fod.shape)
        if states is None:
            states = np.random.randint(4294967295, size=(n_seeds, 2)).astype(np.uint32)
        states = np.array(states, dtype=np.uint32, copy=True, order="C")  # the device advances its copy
        if states.shape != (n_seeds, 2):
            raise ValueError("states must have shape (n_seeds, 2)")
        out = np.zeros((n_seeds, 8), dtype=np.float32)
        _native.check(
            _native.lib().tgk_seeds_from_fod(
                runtime.handle, model.model_id, model.n_parameters, _native.ptr(inline), _native.ptr(indices),
                len(indices), _native.ptr(affine), _native.ptr(states), n_seeds, METHODS[method], _native.ptr(out),
            )
        )
        return out

    data = image_data(fod)
    voxels = np.array(list(zip(*np.nonzero(np.any(data != 0, axis=-1)))))
    indices = np.random.randint(len(voxels), size=n_seeds)
    locations_voxel = voxels[indices] + np.random.rand(n_seeds, 3) - [0.5, 0.5, 0.5]
    locations = apply_affine(fod.affine, locations_voxel)
    vertices = core.fibonacci_sphere(1000)
    azimuths, colatitudes, _ = core.cart2sph(*vertices.T)
    ishtmtx, _ = core.ishtmtx(azimuths, colatitudes, fod.shape[-1])
    orientations = []
    for index in indices:
        values = np.dot(ishtmtx, data[tuple(voxels[index])])
        cumsum = np.cumsum(np.maximum(values, 0.0))
        i = np.searchsorted(cumsum, np.random.rand() * cumsum[-1])
        orientations.append(vertices[i])
    return _from_components(locations, orientations)
