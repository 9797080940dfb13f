"""FP64 geometry and the real symmetric SH basis.

These host functions fix the spherical-harmonic convention the engine tracks
with and are the closed form the device basis is checked against. Public
names and return shapes are those of the reference's ``core.py:6-95``.

Convention (reference ``core.py:36-69``): coefficient i belongs to the i-th
pair (l, m) of the walk l = 0, 2, 4, ..., m = -l..l (so i = l(l+1)/2 + m for
a complete expansion). Its real harmonic is Y_l^0 for m = 0, sqrt(2) Re
Y_l^m for m > 0 and sqrt(2) Im Y_l^|m| for m < 0, with Y_l^m the complex
harmonic of ``scipy.special.sph_harm_y`` (Condon-Shortley phase included).
"""

from __future__ import annotations

import numpy as np
import scipy.special

_SQRT2 = np.sqrt(2.0)


def cart2sph(x, y, z):
    """(azimuth in [0, 2 pi), colatitude in [0, pi], radius) of the points
    (x, y, z), arrays broadcast together. The origin maps to angles 0 and
    radius 0. Float32 input keeps float32 angles (reference ``core.py:6-26``)."""
    x, y, z = np.broadcast_arrays(x, y, z)
    radius = np.sqrt(x * x + y * y + z * z)
    away = radius > 0
    # Divide only away from the origin; the origin's cosine is never used.
    cosine = z / np.where(away, radius, 1)
    colatitude = np.where(away, np.arccos(np.clip(cosine, -1, 1)), 0)
    azimuth = np.where(away, np.mod(np.arctan2(y, x), 2 * np.pi), 0)
    return azimuth, colatitude, radius


def sph2cart(azimuth, colatitude, radius):
    """Inverse of :func:`cart2sph` (reference ``core.py:29-33``)."""
    ring = radius * np.sin(colatitude)
    return ring * np.cos(azimuth), ring * np.sin(azimuth), radius * np.cos(colatitude)


def sh_degrees_orders(n_coefficients: int):
    """Degree and order of each of the first ``n_coefficients`` terms of the
    symmetric expansion, as two int arrays."""
    degrees, orders = [], []
    degree = 0
    while len(degrees) < n_coefficients:
        span = np.arange(-degree, degree + 1)
        degrees.extend([degree] * span.size)
        orders.extend(span.tolist())
        degree += 2
    return np.asarray(degrees[:n_coefficients]), np.asarray(orders[:n_coefficients])


def sh_order(n_coefficients: int) -> int:
    """Largest even degree of a complete symmetric SH expansion with this many
    terms ((L+1)(L+2)/2 = 1, 6, 15, 28, 45, 66, 91, ...)."""
    degree = 0
    while (degree + 1) * (degree + 2) // 2 < n_coefficients:
        degree += 2
    if (degree + 1) * (degree + 2) // 2 != n_coefficients:
        raise ValueError(f"{n_coefficients} is not a symmetric SH coefficient count (1, 6, 15, 28, 45, 66, 91, ...)")
    return degree


def ishtmtx(azimuths, colatitudes, n_coefficients):
    """The real symmetric SH basis at n directions and its angular derivatives.

    Returns ``(matrix, derivatives)``: ``matrix`` of shape (n, K) and
    ``derivatives`` of shape (2, n, K), d/dcolatitude then d/dazimuth
    (reference ``core.py:36-69``). All K harmonics are evaluated in one
    broadcast call over the (direction, term) grid; each column then keeps
    the real or imaginary part its order selects.
    """
    colatitudes = np.atleast_1d(np.asarray(colatitudes, dtype=np.float64))
    azimuths = np.atleast_1d(np.asarray(azimuths, dtype=np.float64))
    degree, order = sh_degrees_orders(n_coefficients)
    value, jacobian = scipy.special.sph_harm_y(degree[None, :], np.abs(order)[None, :], colatitudes[:, None],
                                               azimuths[:, None], diff_n=1)
    weight = np.where(order == 0, 1.0, _SQRT2)
    imaginary = order < 0

    def real_part(z):
        # z: (n, K) complex -> the real basis column by column.
        return weight * np.where(imaginary, z.imag, z.real)

    derivatives = np.stack([real_part(jacobian[..., 0]), real_part(jacobian[..., 1])])
    return real_part(value), derivatives


def fibonacci_sphere(samples=1000):
    """``samples`` points of the golden-angle spiral on the unit sphere, from
    the north pole (z = 1) to the south pole in equal steps of z (reference
    ``core.py:72-95``). The points set the sample directions of FOD seeding,
    so the formula is the reference's: point k at z_k = 1 - 2k/(samples-1),
    azimuth k times the golden angle 2 pi / phi."""
    k = np.arange(samples)
    golden_angle = 2 * np.pi / ((1 + np.sqrt(5)) / 2)
    z = 1 - (2 * k / (samples - 1))
    turn = golden_angle * k
    ring = np.sqrt(1 - z * z)
    return np.stack([ring * np.cos(turn), ring * np.sin(turn), z], axis=1)
