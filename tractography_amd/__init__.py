"""MI355X-native streamline tracking, a drop-in for the OpenCL backend of the
reference ``tractography`` package on its tracking path.

``import tractography_amd as tg`` then ``tg.tractogram(...)``, ``tg.tod(...)``,
``tg.connectome(...)``, ``tg.algorithm.Transport/Diffusion/Probabilistic/
Deterministic``, ``tg.local.*``, ``tg.field.Image``, ``tg.domain.Image`` and
``tg.seeds.*`` with the reference's signatures, and ``tg.streaming`` for
batch-streamed TCK/TRX output. The device work runs in hand-written HIP
kernels (libtgk.so) on gfx950.
"""

from __future__ import annotations

import numpy as np
import numpy.typing as npt

from . import algorithm as algorithm_module
from . import autotune, core, local, seeds, streaming, tracking, utils
from . import domain as domain_module
from . import field as field_module
from .algorithm import BaseAlgorithm
from .image import Nifti1Image, Streamlines, Tractogram, is_image
from . import opencl
from .opencl import OpenCLRuntime
from .runtime import HipRuntime

algorithm = algorithm_module
domain = domain_module
field = field_module
seeds_module = seeds


def _tracking_components(fod, tracking_algorithm: BaseAlgorithm, domain_image=None):
    """Algorithm, image domain and image field for a 4-D image
    (__init__.py:19-37); the domain defaults to the image's nonzero voxels."""
    _validate_algorithm(tracking_algorithm)
    if tracking_algorithm.local_model.n_parameters != fod.shape[-1]:
        raise ValueError(
            "Algorithm local model parameter count does not match the FOD: "
            f"{tracking_algorithm.local_model.n_parameters} != {fod.shape[-1]}"
        )
    if domain_image is None:
        mask = np.any(np.asanyarray(fod.dataobj) != 0, axis=-1).astype(np.uint8)
        domain_image = Nifti1Image(mask, fod.affine)
    return tracking_algorithm, domain_module.Image(domain_image), field_module.Image(fod)


def _validate_algorithm(tracking_algorithm: BaseAlgorithm) -> None:
    if not isinstance(tracking_algorithm, BaseAlgorithm):
        raise TypeError("algorithm must be a BaseAlgorithm instance.")


def tractogram(
    field,
    seeds: npt.NDArray[np.float32] | int,
    domain=None,
    algorithm: BaseAlgorithm | None = None,
    *,
    save_at: float = 1.0,
    maximum_streamline_length: float = 300.0,
    minimum_streamline_length: float = 0.0,
    batch_size: int = 50000,
    runtime: HipRuntime | None = None,
) -> Tractogram:
    """Generate streamlines from a 4-D image or a prepared field
    (__init__.py:369-540).

    ``field`` is a 4-D image of 6 DTI or 45 SH coefficients per voxel (the
    local model and an image field are built from it) or a ``BaseField``.
    ``seeds`` is an (n, 8) array or a number of boundary seeds to sample.
    ``domain`` defaults to the nonzero voxels of an image field. Streamlines
    with ceil(minimum_streamline_length / save_at) points or fewer are
    dropped. ``batch_size`` only sets which seeds share a noise-stream draw
    (one ``algorithm.get_rng_seed()`` per batch, as the reference's batch
    loop); all seeds are tracked in one device call.
    """
    if not np.isfinite(save_at) or save_at <= 0:
        raise ValueError("save_at must be finite and positive.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if not np.isfinite(minimum_streamline_length) or minimum_streamline_length < 0:
        raise ValueError("minimum_streamline_length must be finite and non-negative.")
    if minimum_streamline_length > maximum_streamline_length:
        raise ValueError("minimum_streamline_length must not exceed maximum_streamline_length.")
    if not isinstance(batch_size, (int, np.integer)) or batch_size <= 0:
        raise ValueError("batch_size must be a positive integer.")
    if is_image(field):
        if field.ndim != 4:
            raise ValueError("The tracking image must be four-dimensional.")
        tracking_field = field_module.Image(field)
        domain_image = Nifti1Image(
            np.any(np.asanyarray(field.dataobj) != 0, axis=-1).astype(np.uint8), field.affine
        )
        local_model = local.from_shape(field.shape)
    elif isinstance(field, field_module.BaseField):
        tracking_field = field
        local_model = local.from_shape((field.n_parameters,)) if algorithm is None else None
        if isinstance(field, field_module.Image):
            affine = np.linalg.inv(field.inverse_affine)
            domain_image = Nifti1Image(np.any(field.values != 0, axis=-1).astype(np.uint8), affine)
        else:
            domain_image = None
    else:
        raise TypeError("field must be a Nifti1Image or a BaseField instance.")

    if algorithm is None:
        assert local_model is not None
        tracking_algorithm = algorithm_module.Transport(local_model)
    else:
        _validate_algorithm(algorithm)
        tracking_algorithm = algorithm
    if tracking_algorithm.local_model.n_parameters != tracking_field.n_parameters:
        raise ValueError(
            "Algorithm local model parameter count does not match the field: "
            f"{tracking_algorithm.local_model.n_parameters} != "
            f"{tracking_field.n_parameters}"
        )

    if isinstance(seeds, (int, np.integer)):
        if seeds < 0:
            raise ValueError("seeds must be a non-negative integer or an (n_seeds, 8) array.")
        n_seeds = int(seeds)
    elif isinstance(seeds, np.ndarray):
        seeds = seeds_module.validate_seed_array(seeds)
        n_seeds = len(seeds)
    else:
        raise TypeError("seeds must be a non-negative integer or an (n_seeds, 8) array.")

    if domain is not None and not (is_image(domain) or isinstance(domain, domain_module.BaseDomain)):
        raise TypeError("domain must be a BaseDomain, Nifti1Image, or None.")
    if n_seeds == 0:
        return Tractogram([], affine_to_rasmm=np.eye(4))

    if is_image(domain):
        tracking_domain = domain_module.Image(domain)
    elif domain is not None:
        tracking_domain = domain
    elif domain_image is not None:
        tracking_domain = domain_module.Image(domain_image)
    else:
        raise ValueError(
            "A domain is required when the field is not an image field and "
            "does not provide an image-derived nonzero domain."
        )

    if isinstance(seeds, (int, np.integer)):
        if runtime is None:
            runtime = HipRuntime()
        seeds = domain_module.seed(tracking_domain, n_seeds, runtime=runtime)
    if len(seeds) == 0:
        return Tractogram([], affine_to_rasmm=np.eye(4))
    if runtime is None:
        runtime = HipRuntime()

    n_batches = (len(seeds) + batch_size - 1) // batch_size
    rng_seeds = [tracking_algorithm.get_rng_seed() for _ in range(n_batches)]
    result = tracking.track(
        tracking_algorithm, tracking_domain, tracking_field, seeds, rng_seeds, int(batch_size), save_at,
        maximum_streamline_length, runtime,
    )
    lengths = result.lengths.astype(np.int64)
    keep = lengths > int(np.ceil(minimum_streamline_length / save_at))
    if np.all(keep):
        streamlines = result.streamlines()
    else:
        streamlines = Streamlines(result.points[np.repeat(keep, lengths)], lengths[keep])
    return Tractogram(streamlines, affine_to_rasmm=np.eye(4))


def tod(
    field,
    n_seeds: int = 1000000,
    domain=None,
    algorithm: BaseAlgorithm | None = None,
    *,
    maximum_streamline_length: float = 300.0,
    runtime: HipRuntime | None = None,
):
    """Track-orientation distribution of a field (__init__.py:217-366).

    ``field`` is a 4-D image (local model from its coefficient count) or a
    ``BaseField``; ``algorithm`` defaults to transport. Seeds are sampled on
    the boundary of ``domain`` (a domain object or a 3-D image; by default
    the nonzero voxels of an image field). The TOD is accumulated on the
    field's image grid, or on an image domain's grid for other fields.
    Returns an image of 45 normalised SH coefficients per voxel.
    """
    from .image import float32_image

    if not isinstance(n_seeds, (int, np.integer)) or n_seeds < 0:
        raise ValueError("n_seeds must be a non-negative integer.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if is_image(field):
        if field.ndim != 4:
            raise ValueError("The tracking image must be four-dimensional.")
        tracking_field = field_module.Image(field)
        domain_image = Nifti1Image(
            np.any(np.asanyarray(field.dataobj) != 0, axis=-1).astype(np.uint8), field.affine
        )
        local_model = local.from_shape(field.shape)
        grid_shape, grid_affine, grid_like = tuple(field.shape[:3]), field.affine, field
    elif isinstance(field, field_module.BaseField):
        tracking_field = field
        local_model = local.from_shape((field.n_parameters,)) if algorithm is None else None
        if isinstance(field, field_module.Image):
            affine = np.linalg.inv(field.inverse_affine)
            domain_image = Nifti1Image(np.any(field.values != 0, axis=-1).astype(np.uint8), affine)
            grid_shape, grid_affine, grid_like = tuple(int(s) for s in field.dims[:3]), affine, None
        else:
            domain_image = grid_shape = grid_affine = grid_like = None
    else:
        raise TypeError("field must be a Nifti1Image or a BaseField instance.")

    if algorithm is None:
        assert local_model is not None
        tracking_algorithm = algorithm_module.Transport(local_model)
    else:
        _validate_algorithm(algorithm)
        tracking_algorithm = algorithm
    if tracking_algorithm.local_model.n_parameters != tracking_field.n_parameters:
        raise ValueError(
            "Algorithm local model parameter count does not match the field: "
            f"{tracking_algorithm.local_model.n_parameters} != "
            f"{tracking_field.n_parameters}"
        )

    if domain is not None and not (is_image(domain) or isinstance(domain, domain_module.BaseDomain)):
        raise TypeError("domain must be a BaseDomain, Nifti1Image, or None.")
    tracking_domain = None
    if is_image(domain):
        if domain.ndim != 3:
            raise ValueError("Domain image must be three-dimensional.")
        domain_image = domain
        if grid_shape is None:
            grid_shape, grid_affine, grid_like = tuple(domain.shape), domain.affine, domain
    elif domain is not None:
        tracking_domain = domain
        if grid_shape is None and isinstance(domain, domain_module.Image):
            grid_shape = tuple(int(s) for s in domain.dims[:3])
            grid_affine = np.linalg.inv(domain.inverse_affine)
    elif domain_image is not None:
        domain = domain_image
    else:
        raise ValueError("A domain is required when the field is not an image field.")

    if grid_shape is None or grid_affine is None:
        raise ValueError("A field image or image-based domain is required to define the TOD grid.")
    image_grid = float32_image(np.zeros(grid_shape, dtype=np.float32), grid_affine, like=grid_like)

    if is_image(domain):
        if not np.any(np.asanyarray(domain.dataobj) != 0):
            return float32_image(np.zeros(tuple(grid_shape) + (45,), dtype=np.float32), grid_affine, like=grid_like)
        tracking_domain = domain_module.Image(domain)
    elif tracking_domain is None:
        raise ValueError("Could not infer a tracking domain from the field.")

    return tracking.tod(
        tracking_algorithm, tracking_domain, tracking_field, image_grid, n_seeds,
        maximum_streamline_length=maximum_streamline_length, runtime=runtime,
    )


def connectome(
    field,
    n_seeds: int = 1000000,
    domain=None,
    algorithm: BaseAlgorithm | None = None,
    segmentation=None,
    *,
    distance_upper_bound: float = 4.0,
    maximum_streamline_length: float = 300.0,
    minimum_streamline_length: float = 0.0,
    runtime: HipRuntime | None = None,
):
    """Boundary or region connectome (__init__.py:45-214 of the reference).

    ``field`` is a 4D image (its local model inferred from the coefficient
    count) or a prepared field; the algorithm defaults to Transport and the
    domain to the nonzero field voxels. Without ``segmentation`` the directed
    boundary-vertex CSR matrix is returned with vertex indices as labels. With
    one, each boundary vertex joins its nearest positive segmentation voxel
    within ``distance_upper_bound`` and the matrix is summed per region pair
    (dense, not symmetrized), returned with the sorted labels.
    """
    from scipy.sparse import csr_matrix
    from scipy.spatial import cKDTree

    if not isinstance(n_seeds, (int, np.integer)) or n_seeds < 0:
        raise ValueError("n_seeds must be a non-negative integer.")
    if not np.isfinite(distance_upper_bound) or distance_upper_bound < 0:
        raise ValueError("distance_upper_bound must be finite and non-negative.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if not np.isfinite(minimum_streamline_length) or minimum_streamline_length < 0:
        raise ValueError("minimum_streamline_length must be finite and non-negative.")
    if minimum_streamline_length > maximum_streamline_length:
        raise ValueError("minimum_streamline_length must not exceed maximum_streamline_length.")

    seg_data = None
    labels = None
    if segmentation is not None:
        if not is_image(segmentation):
            raise TypeError("segmentation must be a Nifti1Image or None.")
        seg_data = np.asanyarray(segmentation.dataobj)
        if seg_data.ndim != 3:
            raise ValueError("Segmentation image must be three-dimensional.")
        labels = np.unique(seg_data[seg_data > 0])
        labels = labels[labels > 0]
        if len(labels) == 0:
            return np.zeros((0, 0), dtype=np.float32), labels

    if is_image(field):
        if len(field.shape) != 4:
            raise ValueError("The tracking image must be four-dimensional.")
        tracking_field = field_module.Image(field)
        domain_image = Nifti1Image(np.any(np.asanyarray(field.dataobj) != 0, axis=-1).astype(np.uint8),
                                   field.affine)
        local_model = local.from_shape(field.shape)
    elif isinstance(field, field_module.BaseField):
        tracking_field = field
        local_model = local.from_shape((field.n_parameters,)) if algorithm is None else None
        if isinstance(field, field_module.Image):
            domain_image = Nifti1Image(np.any(field.values != 0, axis=-1).astype(np.uint8),
                                       np.linalg.inv(field.inverse_affine))
        else:
            domain_image = None
    else:
        raise TypeError("field must be a Nifti1Image or a BaseField instance.")

    if algorithm is None:
        tracking_algorithm = algorithm_module.Transport(local_model)
    else:
        _validate_algorithm(algorithm)
        tracking_algorithm = algorithm
    if tracking_algorithm.local_model.n_parameters != tracking_field.n_parameters:
        raise ValueError(
            "Algorithm local model parameter count does not match the field: "
            f"{tracking_algorithm.local_model.n_parameters} != {tracking_field.n_parameters}"
        )

    if domain is not None and not (is_image(domain) or isinstance(domain, domain_module.BaseDomain)):
        raise TypeError("domain must be a BaseDomain, Nifti1Image, or None.")
    if is_image(domain):
        tracking_domain = domain_module.Image(domain)
    elif domain is not None:
        tracking_domain = domain
    elif domain_image is not None:
        if not np.any(np.asanyarray(domain_image.dataobj) != 0):
            if labels is None:
                return csr_matrix((0, 0), dtype=np.float32), np.empty(0, dtype=np.intp)
            return np.zeros((len(labels), len(labels)), dtype=np.float32), labels
        tracking_domain = domain_module.Image(domain_image)
    else:
        raise ValueError("A domain is required when the field is not an image field.")

    boundary_matrix = tracking.connectome(
        tracking_algorithm, tracking_domain, tracking_field, n_seeds,
        maximum_streamline_length=maximum_streamline_length,
        minimum_streamline_length=minimum_streamline_length, runtime=runtime,
    )
    if labels is None:
        return boundary_matrix, np.arange(boundary_matrix.shape[0], dtype=np.intp)

    voxel_indices = np.argwhere(seg_data > 0)
    affine = segmentation.affine if segmentation.affine is not None else np.eye(4)
    voxel_centers = utils.apply_affine(affine, voxel_indices)
    distances, nearest = cKDTree(voxel_centers).query(tracking_domain.boundary.vertices,
                                                      distance_upper_bound=distance_upper_bound)
    distances = np.asarray(distances).reshape(-1)
    nearest = np.asarray(nearest, dtype=np.intp).reshape(-1)
    assigned = np.isfinite(distances)
    label_index = {label: index for index, label in enumerate(labels)}
    vertex_labels = np.zeros(len(tracking_domain.boundary.vertices), dtype=np.intp)
    assigned_voxels = voxel_indices[nearest[assigned]]
    assigned_labels = seg_data[assigned_voxels[:, 0], assigned_voxels[:, 1], assigned_voxels[:, 2]]
    vertex_labels[assigned] = np.fromiter((label_index[label] for label in assigned_labels), dtype=np.intp,
                                          count=int(np.count_nonzero(assigned)))
    selected = np.flatnonzero(assigned)
    assignment = csr_matrix((np.ones(len(selected), dtype=np.float32), (vertex_labels[selected], selected)),
                            shape=(len(labels), len(vertex_labels)))
    matrix = (assignment @ boundary_matrix @ assignment.T).tocsr()
    return matrix.toarray(), labels


__all__ = [
    "connectome",
    "HipRuntime",
    "Nifti1Image",
    "OpenCLRuntime",
    "Streamlines",
    "Tractogram",
    "algorithm",
    "core",
    "domain",
    "field",
    "opencl",
    "local",
    "seeds",
    "streaming",
    "tod",
    "tracking",
    "tractogram",
    "utils",
]
