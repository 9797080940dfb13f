"""Synthetic datasets and analytic answers of the reference's test suite.

Restated from test/data/__init__.py, test/data/tensor.py,
test/test_algorithms/exact_solutions.py and test/test_algorithm.py of the
reference repository.
"""

from __future__ import annotations

import numpy as np

import tractography_amd as tg

_BALL = [[5.0, 0, 0], [0, 5.0, 0], [0, 0, 5.0]]
_X_TENSOR = [[5.0, 0, 0], [0, 1.0, 0], [0, 0, 5.0]]
_Y_TENSOR = [[1.0, 0, 0], [0, 5.0, 0], [0, 0, 5.0]]


def _uniform_tensor_field(tensor, shape, bvectors):
    field = np.zeros(shape + (len(bvectors),))
    field[..., :] = np.exp(-np.diag(np.dot(np.dot(bvectors, tensor), bvectors.T)))
    return field


def _fit_spherical_harmonics(fod_values, bvectors, n_coefficients):
    azimuth, colatitude, _ = tg.core.cart2sph(*bvectors.T)
    ishtmtx, _ = tg.core.ishtmtx(azimuth, colatitude, n_coefficients)
    shtmtx = np.linalg.pinv(ishtmtx)
    fod = np.tensordot(fod_values, shtmtx, axes=(3, 1))
    return tg.utils.normalize_odf(fod)


def uniform_isotropic(shape=(10, 10, 10), n_coefficients=45):
    """test/data/__init__.py:27-31"""
    bvectors = tg.core.fibonacci_sphere(n_coefficients * 3)
    return _fit_spherical_harmonics(_uniform_tensor_field(_BALL, shape, bvectors), bvectors, n_coefficients)


def circle(shape=(10, 10, 1), n_coefficients=45, radius=2, width=1):
    """test/data/__init__.py:34-56"""
    bvectors = tg.core.fibonacci_sphere(n_coefficients * 3)
    fod_values = np.zeros(shape + (len(bvectors),))
    center = np.array(shape, dtype=float) / 2 - 0.5
    for i, j, k in np.ndindex(*shape):
        x, y, z = np.r_[i, j, k] - center
        r = np.sqrt(x * x + y * y)
        if radius - width < r < radius + width:
            eigvec = np.array([[y / r, x / r, 0], [-x / r, y / r, 0], [0, 0, 1.0]])
            tensor = eigvec.dot(np.diag([1.0, 5.0, 5.0])).dot(eigvec.T)
            fod_values[i, j, k] = np.exp(-np.diag(np.dot(np.dot(bvectors, tensor), bvectors.T)))
    return _fit_spherical_harmonics(fod_values, bvectors, n_coefficients)


def cross(shape=(10, 10, 1), n_coefficients=45):
    """test/data/__init__.py:59-94: (fod image, mask image, segmentation)."""
    bvectors = tg.core.fibonacci_sphere(n_coefficients * 3)
    fod_values = _uniform_tensor_field(_BALL, shape, bvectors)
    x_offset = shape[0] // 3
    x_size = shape[0] - 2 * x_offset
    fod_values[x_offset:-x_offset] = _uniform_tensor_field(_X_TENSOR, (x_size, shape[1], shape[2]), bvectors)
    y_offset = shape[1] // 3
    y_size = shape[1] - 2 * y_offset
    fod_values[:, y_offset:-y_offset] += _uniform_tensor_field(_Y_TENSOR, (shape[0], y_size, shape[2]), bvectors)
    fod = tg.Nifti1Image(_fit_spherical_harmonics(fod_values, bvectors, n_coefficients), np.eye(4))
    mask = np.zeros(shape, dtype=np.uint8)
    mask[x_offset:-x_offset] = 1
    mask[:, y_offset:-y_offset] = 1
    segmentation = np.zeros(shape, dtype=np.uint8)
    segmentation[x_offset:-x_offset, 0] = 1
    segmentation[x_offset:-x_offset, -1] = 2
    segmentation[0, y_offset:-y_offset] = 3
    segmentation[-1, y_offset:-y_offset] = 4
    return fod, tg.Nifti1Image(mask, np.eye(4)), tg.Nifti1Image(segmentation, np.eye(4))


def tensor_circle(shape=(10, 10, 1), radius=2, width=1):
    """test/data/tensor.py:25-56: (tensor image, wm image, None)."""
    fod_values = np.zeros(shape + (6,))
    fod_values[..., :3] = 0.003
    wm = np.zeros(shape, dtype=np.uint8)
    center = np.array(shape, dtype=float) / 2 - 0.5
    affine = np.eye(4)
    affine[:3, 3] = center
    for i, j, k in np.ndindex(*shape):
        x, y, z = np.r_[i, j, k] - center
        r = np.sqrt(x * x + y * y)
        if radius - width < r < radius + width:
            eigvec = np.array([[y / r, x / r, 0], [-x / r, y / r, 0], [0, 0, 1.0]])
            tensor = eigvec.dot(np.diag([1.5, 0.4, 0.4]) / 1000).dot(eigvec.T)
            fod_values[i, j, k] = tensor[[0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
            wm[i, j, k] = 1
    return tg.Nifti1Image(fod_values, affine), tg.Nifti1Image(wm, affine), None


def tensor_cross(shape=(20, 20, 1)):
    """test/data/tensor.py:59-90."""
    fod_values = np.zeros(shape + (6,))
    wm = np.zeros(shape, dtype=np.uint8)
    fod_values[..., :3] = 0.003
    x_offset = shape[0] // 3
    x_size = shape[0] - 2 * x_offset
    y_offset = shape[1] // 3
    y_size = shape[1] - 2 * y_offset
    fod_values[x_offset : x_offset + x_size, :, :, :3] = 0
    fod_values[:, y_offset : y_offset + y_size, :, :3] = 0
    center = np.array(shape, dtype=float) / 2 - 0.5
    affine = np.eye(4)
    affine[:3, 3] = center
    fod_values[x_offset : x_offset + x_size, :, :, :3] += [0.4e-3, 1.5e-3, 0.4e-3]
    wm[x_offset : x_offset + x_size, :, :] = 1
    fod_values[:, y_offset : y_offset + y_size, :, :3] += [1.5e-3, 0.4e-3, 0.4e-3]
    wm[:, y_offset : y_offset + y_size, :] = 1
    return tg.Nifti1Image(fod_values, affine), tg.Nifti1Image(wm, affine), None


def u0u3_coefficients(epsilon: float = 0.3, n_coefficients: int = 45) -> np.ndarray:
    """SH fit of 1 + epsilon u_x u_z (exact_solutions.py:6-14)."""
    directions = tg.core.fibonacci_sphere(5000)
    values = 1.0 + epsilon * directions[:, 0] * directions[:, 2]
    azimuths, colatitudes, _ = tg.core.cart2sph(*directions.T)
    basis, _ = tg.core.ishtmtx(azimuths, colatitudes, n_coefficients)
    return np.asarray(np.linalg.pinv(basis) @ values)


def u0u3_reference(step_size, gamma, theta_0, phi_0, epsilon, n_steps) -> np.ndarray:
    """Euler integration of the analytic drift (exact_solutions.py:17-37)."""
    orientation = np.asarray(tg.core.sph2cart(phi_0, theta_0, 1.0))
    orientations = [orientation.copy()]
    identity = np.eye(3)
    for _ in range(n_steps - 1):
        value = 1.0 + epsilon * orientation[0] * orientation[2]
        gradient = np.array([orientation[2], 0.0, orientation[0]])
        tangent = (identity - np.outer(orientation, orientation)) @ gradient
        tangent *= gamma * epsilon / value
        orientation = orientation + step_size * tangent
        orientation /= np.linalg.norm(orientation)
        orientations.append(orientation.copy())
    return np.asarray(orientations)


def quadratic_fod_coefficients(baseline, epsilon, n_coefficients=15):
    """Least-squares SH fit of baseline + epsilon x z (test_algorithm.py:7-13)."""
    directions = tg.core.fibonacci_sphere(1000)
    phi, theta, _ = tg.core.cart2sph(*directions.T)
    harmonics, _ = tg.core.ishtmtx(phi, theta, n_coefficients)
    values = baseline + epsilon * directions[:, 0] * directions[:, 2]
    return np.linalg.lstsq(harmonics, values, rcond=None)[0].astype(np.float32)


def expected_quadratic_update(orientation, baseline, epsilon, step_size, gamma):
    """One exponential-map step of the analytic softplus gradient
    (test_algorithm.py:16-30)."""
    raw_value = baseline + epsilon * orientation[0] * orientation[2]
    value = np.logaddexp(0.0, 100.0 * raw_value) / 100.0
    softmax_derivative = 1.0 / (1.0 + np.exp(-100.0 * raw_value))
    gradient = epsilon * np.array([orientation[2], 0.0, orientation[0]])
    tangent_gradient = gradient - orientation * np.dot(orientation, gradient)
    tangent = step_size * gamma * softmax_derivative / value * tangent_gradient
    tangent_norm = np.linalg.norm(tangent)
    if tangent_norm == 0:
        return orientation.copy()
    return np.cos(tangent_norm) * orientation + np.sin(tangent_norm) / tangent_norm * tangent


def shell_scene(shape=(32, 32, 32), outer=(13.0, 14.0, 12.0), inner=(5.0, 6.0, 5.0), n_rings=30, n_segments=36,
                amplitude=1.0, seed=0, n_seeds=1000):
    """A mesh domain bounded by two displaced ellipsoids about the volume
    centre (the shell workload at small size): (Mesh domain, seeds, radii)."""
    import bench

    rng = np.random.default_rng(seed)
    vertices, faces, centre, radii = bench.shell_domain(rng, shape, outer, inner, n_rings, n_segments, amplitude)
    seeds = bench.seeds_in_shell(rng, centre, radii, np.asarray(outer) + amplitude + 1.0, n_seeds)
    return tg.domain.Mesh(vertices, faces), seeds, centre, radii


def first_ray_hits(vertices, faces, origins, directions):
    """Brute-force closest triangle along each ray (FP64 Moller-Trumbore):
    (triangle index or -1, distance, smallest barycentric margin)."""
    tri = np.asarray(vertices, np.float64)[np.asarray(faces, np.int64)]
    a, e1, e2 = tri[:, 0], tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    best = np.full(len(origins), -1, np.int64)
    dist = np.full(len(origins), np.inf)
    margin = np.zeros(len(origins))
    for i, (o, d) in enumerate(zip(np.asarray(origins, np.float64), np.asarray(directions, np.float64))):
        p = np.cross(d[None, :], e2)
        det = np.einsum("ij,ij->i", e1, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / det
            s = o[None, :] - a
            u = np.einsum("ij,ij->i", s, p) * inv
            q = np.cross(s, e1)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", e2, q) * inv
        ok = (np.abs(det) > 1e-12) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 1e-9)
        if ok.any():
            j = np.flatnonzero(ok)[np.argmin(t[ok])]
            best[i], dist[i] = j, t[j]
            margin[i] = min(u[j], v[j], 1 - u[j] - v[j])
    return best, dist, margin
