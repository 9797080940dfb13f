"""FP64 geometry and the real symmetric SH basis (core.py of the reference).

These host functions define the spherical-harmonic convention the engine
tracks with (MRtrix order, index l(l+1)/2 + m) and are the closed form the
device basis is checked against.
"""

from __future__ import annotations

import numpy as np
import scipy.special


def cart2sph(x, y, z):
    """Cartesian to (azimuth in [0, 2pi), colatitude in [0, pi], radius).

    At the origin both angles and the radius are zero (core.py:6-26).
    """
    x, y, z = np.broadcast_arrays(x, y, z)
    radius = np.sqrt(x**2 + y**2 + z**2)
    nonzero = radius > 0
    cos_colatitude = np.divide(
        z,
        radius,
        out=np.zeros_like(radius, dtype=np.result_type(radius, np.float32)),
        where=nonzero,
    )
    azimuth = np.mod(np.arctan2(y, x), 2 * np.pi)
    colatitude = np.arccos(np.clip(cos_colatitude, -1.0, 1.0))
    azimuth = np.where(nonzero, azimuth, 0.0)
    colatitude = np.where(nonzero, colatitude, 0.0)
    return azimuth, colatitude, radius


def sph2cart(azimuth, colatitude, radius):
    """Spherical to Cartesian coordinates (core.py:29-33)."""
    x = radius * np.sin(colatitude) * np.cos(azimuth)
    y = radius * np.sin(colatitude) * np.sin(azimuth)
    z = radius * np.cos(colatitude)
    return x, y, z


def sh_order(n_coefficients: int) -> int:
    """Largest even degree of a symmetric SH expansion with this many terms."""
    lmax, count = 0, 1
    while count < n_coefficients:
        lmax += 2
        count += 2 * lmax + 1
    if count != n_coefficients:
        raise ValueError(f"{n_coefficients} is not a symmetric SH coefficient count (1, 6, 15, 28, 45, 66, 91, ...)")
    return lmax


def ishtmtx(azimuths, colatitudes, n_coefficients):
    """MRtrix-compatible inverse SH transform and its angular derivatives.

    Returns ``(matrix, derivatives)`` with ``matrix`` of shape (n, K) and
    ``derivatives`` of shape (2, n, K) holding d/dcolatitude and d/dazimuth
    (core.py:36-69). Coefficient i walks degrees l = 0, 2, 4, ... and orders
    m = -l..l; negative orders use sqrt(2) Im(Y), positive sqrt(2) Re(Y).
    """
    colatitudes = np.asarray(colatitudes, dtype=np.float64)
    azimuths = np.asarray(azimuths, dtype=np.float64)
    matrix = np.zeros((n_coefficients, len(colatitudes)))
    matrix_derivative = np.zeros((n_coefficients, len(colatitudes), 2))
    degree = 0
    order = 0
    for i in range(n_coefficients):
        if order < 0:
            m, d = scipy.special.sph_harm_y(degree, -order, colatitudes, azimuths, diff_n=1)
            matrix[i] = np.sqrt(2) * np.imag(m)
            matrix_derivative[i] = np.sqrt(2) * np.imag(d)
        elif order == 0:
            m, d = scipy.special.sph_harm_y(degree, order, colatitudes, azimuths, diff_n=1)
            matrix[i] = np.real(m)
            matrix_derivative[i] = np.real(d)
        else:
            m, d = scipy.special.sph_harm_y(degree, order, colatitudes, azimuths, diff_n=1)
            matrix[i] = np.sqrt(2) * np.real(m)
            matrix_derivative[i] = np.sqrt(2) * np.real(d)
        if degree == order:
            degree += 2
            order = -degree
        else:
            order += 1
    return matrix.T, matrix_derivative.T


def fibonacci_sphere(samples=1000):
    """Points on the unit sphere by the Fibonacci lattice (core.py:72-95)."""
    phi = (1 + np.sqrt(5)) / 2
    golden_angle = 2 * np.pi / phi
    indices = np.arange(samples)
    theta = golden_angle * indices
    z = 1 - (2 * indices / (samples - 1))
    radius = np.sqrt(1 - z**2)
    x = radius * np.cos(theta)
    y = radius * np.sin(theta)
    return np.vstack((x, y, z)).T
