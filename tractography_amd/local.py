"""Local models: real symmetric SH (with softplus) and DTI (local.py).

A model is a descriptor (model id + coefficient count) the kernels are
specialised on; ``evaluate`` methods are the reference's NumPy twins and
``tractography_amd.local.evaluate`` runs the device implementation.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from collections.abc import Sequence
from typing import Any

import numpy as np

from . import _native
from .core import cart2sph, ishtmtx
from .runtime import HipRuntime

Orientations = np.ndarray
LocalModelValues = np.ndarray

SOFTMAX_SCALE = 100.0


class BaseLocalModel(ABC):
    """A local model evaluated at orientations (local.py:20-71).

    ``evaluate`` returns (value, d/dtheta, (d/dphi)/sin(theta)) per row; the
    azimuth derivative is pre-divided by sin(theta).
    """

    opencl_model: str
    model_id: int

    def __init__(self, n_parameters: int):
        self.n_parameters = n_parameters

    @abstractmethod
    def evaluate(self, parameters: np.ndarray, orientations: Orientations) -> LocalModelValues:
        """Evaluate the model with NumPy (a host reference)."""


class RealSymmetricSphericalHarmonics(BaseLocalModel):
    """Real symmetric SH (MRtrix convention) passed through a softplus.

    The device kernels support up to 45 coefficients (l <= 8), as the
    reference's basis does (local.py:74-189).
    """

    opencl_model = "MODEL_SYMMETRIC_REAL_SPHERICAL_HARMONICS"
    model_id = _native.MODEL_SH

    def evaluate(self, parameters: np.ndarray, orientations: Orientations) -> LocalModelValues:
        """NumPy twin (local.py:123-189), with finite pole limits."""
        if orientations.shape[0] == 0:
            return np.empty((0, 3), dtype=np.float32)

        def softmax(x, k):
            return np.where(x * k < 30.0, np.log(1 + np.exp(np.minimum(k * x, 30.0))) / k, x)

        def dsoftmax(x, k):
            return 1.0 / (1 + np.exp(-k * x))

        phi, theta, _ = cart2sph(*np.asarray(orientations).T)
        mtx, mtx_derivatives = ishtmtx(phi, theta, self.n_parameters)
        value = np.einsum("ij,ij->i", mtx, parameters)
        d = dsoftmax(value, SOFTMAX_SCALE)
        value = softmax(value, SOFTMAX_SCALE)
        dphi_numerator = np.einsum("ij,ij->i", mtx_derivatives[1], parameters)
        sin_theta = np.sin(theta)
        at_pole = np.isclose(sin_theta, 0.0, atol=1e-7, rtol=0.0)
        dphi = np.divide(dphi_numerator, sin_theta, out=np.zeros_like(dphi_numerator), where=~at_pole)
        dtheta = np.einsum("ij,ij->i", mtx_derivatives[0], parameters)
        if np.any(at_pole):
            pole_theta = np.where(np.cos(theta[at_pole]) >= 0, 1e-6, np.pi - 1e-6)
            _, pole_derivatives = ishtmtx(phi[at_pole], pole_theta, self.n_parameters)
            pole_parameters = parameters[at_pole]
            dtheta[at_pole] = np.einsum("ij,ij->i", pole_derivatives[0], pole_parameters)
            dphi[at_pole] = np.einsum("ij,ij->i", pole_derivatives[1], pole_parameters) / np.sin(pole_theta)
        return np.stack([value, dtheta * d, dphi * d], axis=-1)


class DiffusionTensorImaging(BaseLocalModel):
    """u^T D u with D from (Dxx, Dyy, Dzz, Dxy, Dxz, Dyz) (local.py:192-262)."""

    opencl_model = "MODEL_DTI"
    model_id = _native.MODEL_DTI

    def __init__(self):
        super().__init__(6)

    def evaluate(self, parameters: np.ndarray, orientations: Orientations) -> LocalModelValues:
        n_orientations = orientations.shape[0]
        if n_orientations == 0:
            return np.empty((0, 3), dtype=np.float32)
        tensor = np.zeros((n_orientations, 3, 3), dtype=np.result_type(parameters, orientations))
        tensor[:, 0, 0] = parameters[:, 0]
        tensor[:, 1, 1] = parameters[:, 1]
        tensor[:, 2, 2] = parameters[:, 2]
        tensor[:, 0, 1] = tensor[:, 1, 0] = parameters[:, 3]
        tensor[:, 0, 2] = tensor[:, 2, 0] = parameters[:, 4]
        tensor[:, 1, 2] = tensor[:, 2, 1] = parameters[:, 5]
        tensor_dot_u = np.einsum("nij,nj->ni", tensor, orientations)
        value = np.einsum("ni,ni->n", orientations, tensor_dot_u)
        phi, theta, _ = cart2sph(*np.asarray(orientations).T)
        sin_theta, cos_theta = np.sin(theta), np.cos(theta)
        sin_phi, cos_phi = np.sin(phi), np.cos(phi)
        e_theta = np.stack([cos_theta * cos_phi, cos_theta * sin_phi, -sin_theta], axis=-1)
        e_phi = np.stack([-sin_phi, cos_phi, np.zeros(n_orientations)], axis=-1)
        dtheta = 2.0 * np.einsum("ni,ni->n", tensor_dot_u, e_theta)
        dphi = 2.0 * np.einsum("ni,ni->n", tensor_dot_u, e_phi)
        return np.stack([value, dtheta, dphi], axis=-1).astype(np.float32)


def from_shape(shape: Sequence[int]) -> BaseLocalModel:
    """Built-in model for a field's coefficient count (local.py:265-282)."""
    if len(shape) == 0:
        raise ValueError("A local-model shape must include a coefficient dimension.")
    n_parameters = shape[-1]
    if n_parameters == 6:
        return DiffusionTensorImaging()
    if n_parameters == 45:
        return RealSymmetricSphericalHarmonics(n_parameters)
    raise ValueError("For now, only DTI and fODF with 45 coefficients are supported (lmax=8).")


def evaluate(
    model: BaseLocalModel,
    parameters: np.ndarray,
    orientations: Orientations,
    runtime: HipRuntime | None = None,
) -> LocalModelValues:
    """Evaluate a local model on the GPU (local.py:285-362).

    Returns (n, 3) float32 rows of value, d/dtheta and (d/dphi)/sin(theta).
    """
    n_orientations = orientations.shape[0]
    if n_orientations == 0:
        return np.empty((0, 3), dtype=np.float32)
    if runtime is None:
        runtime = HipRuntime()
    params = np.ascontiguousarray(parameters, dtype=np.float32)
    if params.ndim != 2 or params.shape != (n_orientations, model.n_parameters):
        raise ValueError(
            f"Expected parameters to have shape ({n_orientations}, {model.n_parameters}), but got {params.shape}"
        )
    o = np.ascontiguousarray(np.asarray(orientations)[:, :3], dtype=np.float32)
    out = np.empty((n_orientations, 3), dtype=np.float32)
    _native.check(
        _native.lib().tgk_evaluate(
            runtime.handle, model.model_id, model.n_parameters, _native.ptr(params), _native.ptr(o),
            n_orientations, _native.ptr(out),
        )
    )
    return out


def _descriptor(model: BaseLocalModel) -> tuple[int, int]:
    if not isinstance(model, (RealSymmetricSphericalHarmonics, DiffusionTensorImaging)):
        raise TypeError(f"{type(model).__name__} has no device implementation")
    return model.model_id, model.n_parameters


def _unused(_: Any) -> None:  # keep typing imports used
    return None
