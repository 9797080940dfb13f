"""Orientation-update algorithms (algorithm.py of the reference).

``Transport`` (tangent drift along the gradient of the local model),
``Diffusion`` (drift plus tangent-plane Gaussian noise), ``Probabilistic``
(a candidate drawn from a spherical cap with probability proportional to the
model value) and ``Deterministic`` (Newton ascent of the model on the sphere)
keep the reference's constructors and NumPy twins; their device
implementations are the HIP kernels, selected by the algorithm descriptor.
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
            int(getattr(self, "n_samples", 0)),
            float(np.float32(getattr(self, "max_angle", 0.0))),
            int(getattr(self, "max_iterations", 0)),
            float(np.float32(getattr(self, "tolerance", 0.0))),
            float(np.float32(getattr(self, "max_step_size", 0.0))),
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


class Probabilistic(BaseAlgorithm):
    """Sample a new orientation near the current one (algorithm.py:272-436).

    ``n_samples`` candidate directions are drawn uniformly from the spherical
    cap of radius ``max_angle`` around each orientation, and one is chosen
    with probability proportional to its non-negative local-model value.

    Args:
        local_model: Model whose non-negative value is the sampling density.
        n_samples: Candidates per update.
        max_angle: Cap radius in radians, in [0, pi].
        seed: Seed of the NumPy generator that draws the device RNG seeds
            (and the NumPy twin's samples).
        step_size: Tracking step; it does not affect sampling.
    """

    algorithm_id = _native.ALGO_PROBABILISTIC

    def __init__(
        self,
        local_model: BaseLocalModel,
        n_samples: int = 64,
        max_angle: float = np.pi / 12,
        seed: int | None = None,
        step_size: float = 0.25,
    ):
        super().__init__(local_model, step_size=step_size)
        if n_samples <= 0:
            raise ValueError("n_samples must be positive")
        if not np.isfinite(max_angle) or not 0.0 <= max_angle <= np.pi:
            raise ValueError("max_angle must be finite and between 0 and pi")
        self.n_samples = int(n_samples)
        self.max_angle = float(max_angle)
        self._rng = np.random.default_rng(seed)

    def get_rng_seed(self) -> int:
        """One draw per tracking batch (algorithm.py:382-384)."""
        return int(self._rng.integers(0, np.iinfo(np.uint32).max, dtype=np.uint32))

    def update_orientation(self, local_models: np.ndarray, orientations: Orientations):
        """NumPy twin (algorithm.py:386-436): its own NumPy random stream."""
        n_orientations = orientations.shape[0]
        result_dtype = np.result_type(orientations, np.float32)
        if n_orientations == 0:
            return np.empty((0, 3), dtype=result_dtype)
        units = np.asarray(orientations, dtype=result_dtype).copy()
        norms = np.linalg.norm(units, axis=1)
        invalid = norms <= 1e-8
        units[~invalid] /= norms[~invalid, np.newaxis]
        units[invalid] = (0.0, 0.0, 1.0)
        helpers = np.zeros_like(units)
        use_z = np.abs(units[:, 2]) < 0.9
        helpers[use_z, 2] = 1.0
        helpers[~use_z, 1] = 1.0
        tangent1 = np.cross(helpers, units)
        tangent1 /= np.linalg.norm(tangent1, axis=1, keepdims=True)
        tangent2 = np.cross(units, tangent1)
        draws = self._rng.random((n_orientations, self.n_samples, 2))
        cos_max = np.cos(self.max_angle)
        cos_angles = 1.0 - draws[:, :, 0] * (1.0 - cos_max)
        sin_angles = np.sqrt(np.maximum(0.0, 1.0 - cos_angles**2))
        azimuths = 2.0 * np.pi * draws[:, :, 1]
        candidates = cos_angles[:, :, None] * units[:, None, :] + sin_angles[:, :, None] * (
            np.cos(azimuths)[:, :, None] * tangent1[:, None, :] + np.sin(azimuths)[:, :, None] * tangent2[:, None, :]
        )
        repeated = np.repeat(local_models, self.n_samples, axis=0)
        values = self.local_model.evaluate(repeated, candidates.reshape(-1, 3))[:, 0].reshape(
            n_orientations, self.n_samples
        )
        weights = np.where(np.isfinite(values), np.maximum(values, 0.0), 0.0)
        totals = weights.sum(axis=1)
        result = units.copy()
        valid = totals > 0.0
        if np.any(valid):
            targets = self._rng.random(np.count_nonzero(valid)) * totals[valid]
            cumulative = np.cumsum(weights[valid], axis=1)
            selected = np.minimum(np.sum(cumulative <= targets[:, None], axis=1), self.n_samples - 1)
            result[valid] = candidates[valid, selected]
        return result.astype(result_dtype, copy=False)


class Deterministic(BaseAlgorithm):
    """Maximise the local model near each direction (algorithm.py:439-677).

    The device update is Newton ascent in the tangent frame: gradient from
    the model, Hessian from central differences of nearby gradients, a fixed
    small gradient step where that Hessian is not negative definite, steps
    capped at ``max_step_size`` and a backtracking line search. The NumPy
    twin maximises with SciPy's SLSQP under a unit-norm constraint, as the
    reference's does.

    Args:
        local_model: Model whose value is maximised.
        max_iterations: Iterations per update.
        tolerance: Gradient-norm threshold (device) / SLSQP ftol (twin).
        max_step_size: Largest device update angle, in radians.
        step_size: Tracking step; it does not enter the update.
    """

    algorithm_id = _native.ALGO_DETERMINISTIC

    def __init__(
        self,
        local_model: BaseLocalModel,
        max_iterations: int = 10,
        tolerance: float = 1e-5,
        max_step_size: float = 0.5,
        step_size: float = 0.1,
    ):
        super().__init__(local_model, step_size=step_size)
        self.max_iterations = max_iterations
        self.tolerance = tolerance
        self.max_step_size = max_step_size

    def update_orientation(self, local_models: np.ndarray, orientations: Orientations):
        """NumPy twin (algorithm.py:628-677): SciPy SLSQP per orientation."""
        from scipy.optimize import minimize

        result = np.array(orientations, dtype=np.result_type(orientations, np.float32), copy=True)
        for i, initial in enumerate(result):
            initial = np.asarray(initial, dtype=np.float64)
            initial /= np.linalg.norm(initial)
            parameters = local_models[i : i + 1]

            def objective(orientation, parameters=parameters):
                norm = np.linalg.norm(orientation)
                direction = orientation / norm
                value = self.local_model.evaluate(parameters, direction[None, :])[0]
                phi, theta, _ = cart2sph(*direction)
                st, ct = np.sin(theta), np.cos(theta)
                sp, cp = np.sin(phi), np.cos(phi)
                e1 = np.array([ct * cp, ct * sp, -st])
                e2 = np.array([-sp, cp, 0.0])
                gradient = value[1] * e1 + value[2] * e2
                gradient = (gradient - direction * np.dot(direction, gradient)) / norm
                return -value[0], -gradient

            if self.max_iterations <= 0:
                result[i] = initial
                continue
            optimisation = minimize(
                objective,
                initial,
                method="SLSQP",
                jac=True,
                constraints={
                    "type": "eq",
                    "fun": lambda orientation: np.dot(orientation, orientation) - 1.0,
                    "jac": lambda orientation: 2.0 * orientation,
                },
                options={"ftol": self.tolerance, "maxiter": self.max_iterations},
            )
            result[i] = optimisation.x / np.linalg.norm(optimisation.x)
        return result


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
