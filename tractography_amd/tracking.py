"""Tracking driver (tracking.py of the reference).

``tractogram`` keeps the reference's signature and semantics: one streamline
per seed, points saved every ``save_at`` of arc length (back-interpolated to
the exact save distance), the seed always included, capacity
floor(maximum_streamline_length / save_at) points. The work runs in libtgk.so
on the GPU; volumes stay resident in HBM across calls and the dense per-seed
output of the reference is replaced by on-device compaction into one ragged
point array.
"""

from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass

import numpy as np

from . import _native
from .algorithm import BaseAlgorithm
from .domain import BaseDomain
from .field import BaseField
from .image import Streamlines, Tractogram
from .runtime import HipRuntime
from .seeds import validate_seed_array


@dataclass
class TrackResult:
    """Raw arrays of one tracking run (seed order).

    lengths: saved points per streamline (seed included), u32.
    steps: orientation updates per streamline, u32.
    hits: boundary triangle slot crossed (mesh domains), 0xffffffff if none.
    points: (sum(lengths), 3) float32 ragged points.
    """

    lengths: np.ndarray
    steps: np.ndarray
    hits: np.ndarray
    points: np.ndarray

    def streamlines(self) -> Streamlines:
        return Streamlines(self.points, self.lengths.astype(np.int64))


def point_capacity(maximum_streamline_length: float, save_at: float) -> int:
    """floor(maximum_streamline_length / save_at) (tracking.py:89)."""
    n_steps = int(np.floor(maximum_streamline_length / save_at))
    if n_steps <= 0:
        # The reference cannot allocate its zero-sized point buffer either
        # (opencl.py:254-262).
        raise ValueError("nbytes must be positive.")
    return n_steps


def track(
    algorithm: BaseAlgorithm,
    domain: BaseDomain,
    field: BaseField,
    seeds: np.ndarray,
    rng_seeds,
    batch_size: int,
    save_at: float,
    maximum_streamline_length: float,
    runtime: HipRuntime,
    path: str | None = None,
) -> TrackResult:
    """Track every seed in one device call.

    Seed g draws its noise from rng_init(rng_seeds[g // batch_size],
    g % batch_size), the stream the reference's per-batch calls give it.
    """
    n_seeds = len(seeds)
    n_points = point_capacity(maximum_streamline_length, save_at)
    rng = np.ascontiguousarray(np.asarray(rng_seeds, dtype=np.uint64) % 2**32, dtype=np.uint32)
    if len(rng) < (n_seeds + batch_size - 1) // batch_size:
        raise ValueError("one rng seed per batch is required")
    field_device = field.upload(runtime)
    domain_device = domain.upload(runtime)
    lib = _native.lib()
    path_id = runtime.path_id if path is None else {"plain": _native.PATH_PLAIN, "fast": _native.PATH_FAST}[path]
    if path_id == _native.PATH_FAST:
        _native.check(lib.tgk_field_set_fast_precision(field_device.handle, _native.PRECISIONS[runtime.precision]))
    handle = ctypes.c_void_p()
    _native.check(
        lib.tgk_track(
            runtime.handle, domain_device.handle, field_device.handle, algorithm.descriptor(), path_id,
            _native.ptr(seeds), n_seeds, _native.ptr(rng), int(batch_size), n_points, float(np.float32(save_at)),
            ctypes.byref(handle),
        )
    )
    owner = _TracksHandle(handle)
    n = ctypes.c_int64()
    total = ctypes.c_int64()
    _native.check(lib.tgk_tracks_size(handle, ctypes.byref(n), ctypes.byref(total)))
    lengths = np.empty(n.value, dtype=np.uint32)
    steps = np.empty(n.value, dtype=np.uint32)
    hits = np.empty(n.value, dtype=np.uint32)
    in_place = ctypes.c_void_p()
    _native.check(lib.tgk_tracks_points(handle, ctypes.byref(in_place)))
    if in_place.value and total.value > 0:
        # The library's page-locked buffer becomes the points array; the
        # buffer holds a reference to the handle, freed with the last view.
        buffer = (ctypes.c_float * (3 * total.value)).from_address(in_place.value)
        buffer._owner = owner
        points = np.ctypeslib.as_array(buffer).reshape(-1, 3)
        _native.check(lib.tgk_tracks_copy(handle, _native.ptr(lengths), _native.ptr(steps), _native.ptr(hits), None))
    else:
        points = np.empty((total.value, 3), dtype=np.float32)
        _native.check(
            lib.tgk_tracks_copy(handle, _native.ptr(lengths), _native.ptr(steps), _native.ptr(hits), _native.ptr(points))
        )
        owner.release()
    return TrackResult(lengths, steps, hits, points)


class _TracksHandle:
    """Owns a tgk_tracks handle: destroyed on release() or when the last
    array viewing its points is collected."""

    def __init__(self, handle: ctypes.c_void_p):
        self._finalizer = weakref.finalize(self, _native.lib().tgk_tracks_destroy, handle)

    def release(self) -> None:
        self._finalizer()


def tractogram(
    algorithm: BaseAlgorithm,
    domain: BaseDomain,
    field: BaseField,
    seeds: np.ndarray,
    save_at: float = 1.0,
    maximum_streamline_length: float = 300.0,
    runtime: HipRuntime | None = None,
) -> Tractogram:
    """Track streamlines from position-and-orientation seeds (tracking.py:22-178).

    Raises:
        ValueError: incompatible local model and field, or invalid seeds.

    An empty seed array returns an empty tractogram without creating a
    runtime.
    """
    seeds = validate_seed_array(seeds)
    if algorithm.local_model.n_parameters != field.n_parameters:
        raise ValueError(
            "Incompatible local model and field, they do not have the "
            "same number of parameters: "
            f"{algorithm.local_model.n_parameters} != {field.n_parameters}"
        )
    n_seeds = seeds.shape[0]
    if n_seeds == 0:
        return Tractogram([], affine_to_rasmm=np.eye(4))
    if runtime is None:
        runtime = HipRuntime()
    result = track(
        algorithm, domain, field, seeds, [algorithm.get_rng_seed()], n_seeds, save_at, maximum_streamline_length,
        runtime,
    )
    return Tractogram(result.streamlines(), affine_to_rasmm=np.eye(4))


def tod(
    algorithm: BaseAlgorithm,
    domain: BaseDomain,
    field: BaseField,
    image,
    n_seeds: int,
    maximum_streamline_length: float = 300.0,
    runtime: HipRuntime | None = None,
):
    """Track orientation distribution as SH coefficients (tracking.py:181-385).

    Each track starts at a uniform sample of the domain's boundary surface
    (directed inward) and runs for at most floor(maximum / step_size) steps
    while the domain contains its point and the field is defined. At every
    such step the 45-term SH basis of its orientation is added to the
    ``image`` grid cell of its point (when on the grid). Each nonempty cell
    is then normalised to unit integral (utils.normalize_odf).

    Returns an image of shape ``image.shape + (45,)`` on ``image``'s affine.

    Raises:
        ValueError: incompatible model and field, invalid seed count or
            maximum length, a non-3D output image, or a non-positive step.

    Zero seeds or zero steps return the zero image without creating a
    runtime.
    """
    from .image import float32_image
    from .utils import normalize_odf

    if algorithm.local_model.n_parameters != field.n_parameters:
        raise ValueError(
            "Incompatible local model and field, they do not have the "
            "same number of parameters: "
            f"{algorithm.local_model.n_parameters} != {field.n_parameters}"
        )
    if not isinstance(n_seeds, (int, np.integer)) or n_seeds < 0:
        raise ValueError("n_seeds must be a non-negative integer.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if len(image.shape) != 3:
        raise ValueError("The output image description must be 3D.")

    output_shape = tuple(int(size) for size in image.shape)
    affine = image.affine if image.affine is not None else np.eye(4)
    inverse_affine = np.ascontiguousarray(np.linalg.inv(affine), dtype=np.float32)
    step_size = float(algorithm.step_size)
    if not np.isfinite(step_size) or step_size <= 0:
        raise ValueError("algorithm.step_size must be finite and positive.")
    n_steps = int(np.floor(maximum_streamline_length / step_size))
    if n_seeds == 0 or n_steps == 0:
        raw = np.zeros(output_shape + (45,), dtype=np.float32)
        return float32_image(normalize_odf(raw), affine, like=image)

    if runtime is None:
        runtime = HipRuntime()
    domain_handle = domain.upload(runtime, with_boundary=True)
    field_handle = field.upload(runtime)
    out = np.empty(output_shape + (45,), dtype=np.float32)
    dims = np.asarray(output_shape, dtype=np.uint32)
    _native.check(
        _native.lib().tgk_tod(
            runtime.handle, domain_handle.handle, domain_handle.handle, field_handle.handle, algorithm.descriptor(),
            runtime.path_id, int(algorithm.get_rng_seed()), int(n_seeds), min(n_steps, 2**32 - 1), _native.ptr(dims),
            _native.ptr(inverse_affine), 1, _native.ptr(out),
        )
    )
    return float32_image(out, affine, like=image)


def connectome(
    algorithm: BaseAlgorithm,
    domain: BaseDomain,
    field: BaseField,
    n_seeds: int,
    maximum_streamline_length: float = 300.0,
    minimum_streamline_length: float = 0.0,
    runtime: HipRuntime | None = None,
):
    """Asymmetric, piecewise-linear boundary connectome (tracking.py:388-602).

    Each track starts at a uniform sample of the domain's boundary surface
    (directed inward) and runs for at most floor(maximum / step_size) steps
    while the domain contains its point and the field is defined. Its last
    segment is then tested against the boundary; a hit deposits the outer
    product of the start and hit barycentric weights onto the vertices of
    the two triangles, so an accepted track adds total weight one. Tracks
    shorter than ``minimum_streamline_length`` or without a hit add nothing.

    Returns a float32 (n_vertices, n_vertices) scipy CSR matrix over
    ``domain.boundary.vertices``.
    """
    from scipy.sparse import coo_matrix, csr_matrix

    if algorithm.local_model.n_parameters != field.n_parameters:
        raise ValueError(
            "Incompatible local model and field, they do not have the "
            "same number of parameters: "
            f"{algorithm.local_model.n_parameters} != {field.n_parameters}"
        )
    if not isinstance(n_seeds, (int, np.integer)) or n_seeds < 0:
        raise ValueError("n_seeds must be a non-negative integer.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if not np.isfinite(minimum_streamline_length) or minimum_streamline_length < 0:
        raise ValueError("minimum_streamline_length must be finite and non-negative.")
    if minimum_streamline_length > maximum_streamline_length:
        raise ValueError("minimum_streamline_length must not exceed maximum_streamline_length.")

    n_vertices = len(domain.boundary.vertices)
    if n_seeds == 0 or n_vertices == 0:
        return csr_matrix((n_vertices, n_vertices), dtype=np.float32)
    step_size = float(algorithm.step_size)
    if not np.isfinite(step_size) or step_size <= 0:
        raise ValueError("algorithm.step_size must be finite and positive.")
    n_steps = int(np.floor(maximum_streamline_length / step_size))
    if n_steps == 0:
        return csr_matrix((n_vertices, n_vertices), dtype=np.float32)

    from .domain import _mesh_handle_for

    if runtime is None:
        runtime = HipRuntime()
    domain_handle = domain.upload(runtime)
    boundary_handle = _mesh_handle_for(domain, runtime)
    field_handle = field.upload(runtime)
    rows = np.empty(9 * n_seeds, dtype=np.int32)
    cols = np.empty(9 * n_seeds, dtype=np.int32)
    weights = np.empty(9 * n_seeds, dtype=np.float32)
    _native.check(
        _native.lib().tgk_connectome(
            runtime.handle, domain_handle.handle, boundary_handle.handle, field_handle.handle,
            algorithm.descriptor(), runtime.path_id, int(algorithm.get_rng_seed()), int(n_seeds), n_steps,
            float(np.float32(minimum_streamline_length)), _native.ptr(rows), _native.ptr(cols), _native.ptr(weights),
        )
    )
    valid = rows >= 0
    if not np.any(valid):
        return csr_matrix((n_vertices, n_vertices), dtype=np.float32)
    matrix = coo_matrix((weights[valid], (rows[valid], cols[valid])), shape=(n_vertices, n_vertices),
                        dtype=np.float32).tocsr()
    matrix.sum_duplicates()
    return matrix
