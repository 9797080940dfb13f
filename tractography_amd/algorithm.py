"""Orientation-update algorithms (algorithm.py of the reference).

``Transport`` (tangent drift along the gradient of the local model) and
``Diffusion`` (drift plus tangent-plane Gaussian noise) keep the reference's
constructors and NumPy twins; their device implementations are the HIP
kernels, selected by the algorithm descriptor.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from enum import Enum

import numpy as np

from . import _native
from .core import cart2sph
from .local import BaseLocalModel, Orientations, _descriptor
from .runtime import HipRuntime


class Algorithm(Enum):
    """The orientation-update algorithms (algorithm.py:18-27)."""

    DETERMINISTIC = "deterministic"
    PROBABILISTIC = "probabilistic"
    DIFFUSION = "diffusion"
    TRANSPORT = "transport"

    def __str__(self):
        return self.value


class BaseAlgorithm(ABC):
    """Interface of algorithms that update orientations (algorithm.py:30-53)."""

    algorithm_id: int

    def __init__(self, local_model: BaseLocalModel, step_size: float = 1.0):
        self.local_model = local_model
        self.step_size = step_size

    def get_rng_seed(self) -> int:
        """Seed for the per-seed device RNG streams of one tracking call."""
        return 0

    @abstractmethod
    def update_orientation(self, local_models: np.ndarray, orientations: Orientations) -> Orientations:
        """Update orientations with the NumPy implementation."""

    def descriptor(self) -> _native.tgk_algorithm:
        """The kernel-side description of this algorithm and its model."""
        model_id, n_parameters = _descriptor(self.local_model)
        return _native.algorithm_struct(
            model_id,
            self.algorithm_id,
            n_parameters,
            float(np.float32(self.step_size)),
            float(np.float32(getattr(self, "gamma", 1.0))),
            float(np.float32(getattr(self, "noise_variance", 0.0))),
        )


def _tangent_frame(orientations):
    phi, theta, _ = cart2sph(*np.asarray(orientations).T)
    sp, cp = np.sin(phi), np.cos(phi)
    st, ct = np.sin(theta), np.cos(theta)
    et = np.stack([ct * cp, ct * sp, -st], axis=-1)
    ep = np.stack([-sp, cp, np.zeros(len(phi))], axis=-1)
    return et, ep


class Transport(BaseAlgorithm):
    """Drift along the local model's tangent gradient (algorithm.py:56-156).

    Args:
        local_model: Model providing the value and its angular derivatives.
        step_size: Integration step (also the position step).
        gamma: Drift scaling factor.
    """

    algorithm_id = _native.ALGO_TRANSPORT

    def __init__(self, local_model: BaseLocalModel, step_size: float = 0.001, gamma: float = 1.0):
        super().__init__(local_model, step_size=step_size)
        self.gamma = gamma

    def update_orientation(self, local_models: np.ndarray, orientations: Orientations):
        model_values = self.local_model.evaluate(local_models, orientations)
        et, ep = _tangent_frame(orientations)
        dtheta = model_values[:, 1, np.newaxis]
        dphi = model_values[:, 2, np.newaxis]
        value = model_values[:, 0, np.newaxis]
        drift = (dtheta * et + dphi * ep) / np.maximum(value, 1e-6)
        tangent = self.step_size * self.gamma * drift
        return exps2(orientations, tangent, 1.0)


class Diffusion(BaseAlgorithm):
    """Drift plus tangent-plane Gaussian noise (algorithm.py:159-269).

    Args:
        local_model: Model providing the value and its angular derivatives.
        step_size: Integration time step.
        gamma: Drift and noise scaling factor.
        noise_variance: Variance of the stochastic tangent-plane term.
        seed: Seed of the NumPy generator that draws one device RNG seed per
            tracking call (and the NumPy twin's noise).
    """

    algorithm_id = _native.ALGO_DIFFUSION

    def __init__(
        self,
        local_model: BaseLocalModel,
        step_size: float = 0.001,
        gamma: float = 1.0,
        noise_variance: float = 0.1,
        seed: int | None = None,
    ):
        super().__init__(local_model, step_size=step_size)
        self.gamma = gamma
        self.noise_variance = noise_variance
        self._rng = np.random.default_rng(seed)

    def get_rng_seed(self) -> int:
        """One draw per tracking call (algorithm.py:239-241)."""
        return int(self._rng.integers(0, np.iinfo(np.uint32).max, dtype=np.uint32))

    def update_orientation(self, local_models: np.ndarray, orientations: Orientations):
        n_orientations = orientations.shape[0]
        model_values = self.local_model.evaluate(local_models, orientations)
        et, ep = _tangent_frame(orientations)
        dtheta = model_values[:, 1, np.newaxis]
        dphi = model_values[:, 2, np.newaxis]
        value = model_values[:, 0, np.newaxis]
        drift = (dtheta * et + dphi * ep) / np.maximum(value, 1e-6)
        noise = np.zeros_like(drift)
        if self.noise_variance > 0.0:
            samples = self._rng.standard_normal((n_orientations, 2))
            noise = samples[:, :1] * et + samples[:, 1:] * ep
        scale_drift = self.gamma * self.step_size
        scale_noise = np.sqrt(self.noise_variance * self.gamma * self.step_size)
        tangent = scale_drift * drift + scale_noise * noise
        return exps2(orientations, tangent, 1.0)


def exps2(p: np.ndarray, x: np.ndarray, t: float) -> np.ndarray:
    """Spherical exponential map of tangents x at points p (algorithm.py:680-698)."""
    n = np.linalg.norm(x, axis=1)
    c = np.cos(t * n)
    s = np.sin(t * n)
    scaled_sine = np.divide(s, n, out=np.zeros_like(n), where=n != 0)
    return c[:, np.newaxis] * p + x * scaled_sine[:, np.newaxis]


def update_orientation(
    algorithm: BaseAlgorithm,
    local_models: np.ndarray,
    orientations: Orientations,
    runtime: HipRuntime | None = None,
) -> Orientations:
    """One device orientation update per row (algorithm.py:701-790).

    Row i draws its noise from rng_init(algorithm.get_rng_seed(), i).
    """
    n_orientations = orientations.shape[0]
    if n_orientations == 0:
        return np.empty((0, 3), dtype=np.float32)
    if runtime is None:
        runtime = HipRuntime()
    if (
        local_models.ndim != 2
        or local_models.shape[1] != algorithm.local_model.n_parameters
        or local_models.shape[0] != n_orientations
    ):
        raise ValueError(
            f"Expected local_models to have shape ({n_orientations}, {algorithm.local_model.n_parameters}), "
            f"but got {local_models.shape}"
        )
    params = np.ascontiguousarray(local_models, dtype=np.float32)
    o = np.ascontiguousarray(np.asarray(orientations)[:, :3], dtype=np.float32)
    out = np.empty((n_orientations, 3), dtype=np.float32)
    descriptor = algorithm.descriptor()
    rng_seed = algorithm.get_rng_seed()
    _native.check(
        _native.lib().tgk_update_orientation(
            runtime.handle, descriptor, rng_seed, _native.ptr(params), _native.ptr(o), n_orientations,
            _native.ptr(out),
        )
    )
    return out
