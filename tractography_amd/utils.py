"""Host utilities kept from the reference (utils.py)."""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def normalize_odf(odf: NDArray):
    """Scale SH fields so every nonzero voxel has unit integral (utils.py:6-22).

    Each voxel is divided by sqrt(4 pi) times its l = 0 coefficient; voxels
    with a zero l = 0 coefficient become zero.
    """
    normalized = np.zeros_like(odf)
    np.divide(odf / np.sqrt(4 * np.pi), odf[..., :1], out=normalized, where=odf[..., :1] != 0)
    return normalized


def apply_affine(affine, points):
    """Apply a 4x4 homogeneous affine to (..., 3) points in float64."""
    affine = np.asarray(affine, dtype=np.float64)
    points = np.asarray(points, dtype=np.float64)
    return points @ affine[:3, :3].T + affine[:3, 3]
