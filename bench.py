"""Streamline-tracking throughput on one MI355X.

Metric: streamline steps/s (one step = one orientation update of one
streamline, tracking.py:135) over the Diffusion workload on a 145x174x145
SH order-8 FOD with 1M seeds, FAST path (persistent lanes, transcendental-free
model evaluation, FP64 arc length) with the PLAIN FP32 path (the reference's
integrator, bit-exact with the host restatement) timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sh45] [--seeds N]

Prints one JSON line. Inputs (field, mask, seeds) are resident in HBM before
timing; each timed step is the tracking kernel plus the on-device compaction
of all saved points into one ragged array.

Other workloads (--workload): sphere32 (32^3 SH45 Transport, 10k seeds), dti
(145x174x145 tensors, Transport, 1M seeds), shell (the SH45 field bounded by
two displaced ellipsoid meshes of ~200k triangles each, Diffusion, 2M seeds).
--product tod / connectome times that product instead (tracks from boundary
seeds of the workload's mask, plain FP32 path) and prints its own line;
--product seeds times seeds.from_fod on the device (VALU and MFMA methods).
The module also holds the synthetic workload generators used by the tests.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

SHAPE_145 = (145, 174, 145)

# --------------------------------------------------------------------------
# Synthetic FOD fields
# --------------------------------------------------------------------------


def real_sh_basis(xyz: np.ndarray, lmax: int, transpose: bool = False) -> np.ndarray:
    """Real symmetric SH (MRtrix order, orthonormal) of unit vectors, FP64.

    Fully normalised associated Legendre functions by the standard three-term
    recurrence in l, Condon-Shortley phase included; index l(l+1)/2 + m over
    even l, m = -l..l, sqrt(2) Im for m < 0 and sqrt(2) Re for m > 0.
    Returns (n, K), or (K, n) with transpose=True.
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    s = np.sqrt(x * x + y * y)
    with np.errstate(divide="ignore", invalid="ignore"):
        e1 = np.where(s > 0, (x + 1j * y) / s, 1.0 + 0j)
    n = len(xyz)
    K = (lmax + 1) * (lmax + 2) // 2
    out = np.zeros((K, n))
    p_mm = np.full(n, 1.0 / math.sqrt(4 * math.pi))
    em = np.ones(n, dtype=np.complex128)
    for m in range(lmax + 1):
        if m > 0:
            p_mm = -math.sqrt((2 * m + 1) / (2 * m)) * s * p_mm
            em = em * e1
        p_prev2 = None
        p_prev = p_mm
        for l in range(m, lmax + 1):
            if l == m:
                p = p_mm
            elif l == m + 1:
                p = math.sqrt(2 * m + 3) * z * p_mm
            else:
                a = math.sqrt((4 * l * l - 1) / (l * l - m * m))
                b = math.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                p = a * (z * p_prev - b * p_prev2)
            if l > m:
                p_prev2, p_prev = p_prev, p
            if l % 2 == 0:
                centre = l * (l + 1) // 2
                if m == 0:
                    out[centre] = p
                else:
                    sp = math.sqrt(2) * p
                    out[centre + m] = sp * em.real
                    out[centre - m] = sp * em.imag
    return out if transpose else out.T


def axial_kernel_sh(lmax: int = 8, sharpness: float = 4.0) -> np.ndarray:
    """Funk-Hecke weights k_l = 2 pi int k(t) P_l(t) dt of k(t) = exp(s (t^2 - 1)):
    the SH coefficients of k(u . d) are k_l Y_lm(d)."""
    t, w = np.polynomial.legendre.leggauss(200)
    k = np.exp(sharpness * (t * t - 1.0))
    weights = np.zeros(lmax + 1)
    for l in range(0, lmax + 1, 2):
        weights[l] = 2 * math.pi * np.sum(w * k * np.polynomial.legendre.Legendre.basis(l)(t))
    return weights


def _coefficient_degrees(K: int) -> np.ndarray:
    degrees = []
    l = 0
    while len(degrees) < K:
        degrees.extend([l] * (2 * l + 1))
        l += 2
    return np.array(degrees[:K])


def fod_from_fibres(directions: np.ndarray, weights: np.ndarray, lmax: int = 8, chunk: int = 1 << 17) -> np.ndarray:
    """Sum of weighted axially symmetric fibre kernels as SH coefficients,
    normalised to unit integral per voxel (utils.normalize_odf).

    directions (V, F, 3) unit vectors, weights (V, F) -> (V, K) float32.
    """
    V, F, _ = directions.shape
    K = (lmax + 1) * (lmax + 2) // 2
    kernel = axial_kernel_sh(lmax)[_coefficient_degrees(K)][:, None]
    out = np.zeros((V, K), dtype=np.float32)

    def work(start):
        stop = min(V, start + chunk)
        acc = np.zeros((K, stop - start))
        for f in range(F):
            basis = real_sh_basis(directions[start:stop, f], lmax, transpose=True)
            basis *= kernel
            basis *= weights[start:stop, f][None, :]
            acc += basis
        c0 = acc[:1]
        scale = np.zeros_like(c0)
        np.divide(1.0 / np.sqrt(4 * np.pi), c0, out=scale, where=c0 != 0)
        out[start:stop] = (acc * scale).T

    workers = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        list(pool.map(work, range(0, V, chunk)))
    return out


def _random_fibres(rng, n_voxels: int, base_directions=None):
    """1-3 fibres per voxel with Dirichlet(1) weights."""
    n = rng.integers(1, 4, n_voxels)
    raw = rng.gamma(1.0, size=(n_voxels, 3))
    raw[np.arange(3)[None, :] >= n[:, None]] = 0
    weights = raw / raw.sum(axis=1, keepdims=True)
    if base_directions is None:
        d = rng.standard_normal((n_voxels, 3, 3))
    else:
        d = base_directions
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return d, weights


def smooth_direction_fields(rng, shape, sigma: float = 4.0, n_fields: int = 3) -> np.ndarray:
    """n_fields unit vector fields: three N(0,1) components per field,
    Gaussian-filtered (sigma voxels, reflect) and normalised. (V, n, 3)."""
    from scipy.ndimage import gaussian_filter

    fields = np.empty((int(np.prod(shape)), n_fields, 3), dtype=np.float64)
    for f in range(n_fields):
        for c in range(3):
            noise = rng.standard_normal(shape, dtype=np.float32)
            fields[:, f, c] = gaussian_filter(noise, sigma, mode="reflect").reshape(-1)
    norm = np.sqrt(np.einsum("vfc,vfc->vf", fields, fields))
    fields /= norm[..., None]
    return fields


def ellipsoid_mask(shape, semi_axes) -> np.ndarray:
    centre = (np.array(shape) - 1) / 2.0
    idx = np.indices(shape).reshape(3, -1).T.astype(np.float64)
    r = np.sum(((idx - centre) / np.array(semi_axes)) ** 2, axis=1)
    return (r <= 1.0).reshape(shape)


def seeds_in_mask(rng, mask: np.ndarray, n_seeds: int) -> np.ndarray:
    """from_mask semantics (seeds.py:114-142) with a seeded generator:
    uniform nonzero voxel, U[-0.5, 0.5)^3 jitter, normalised Gaussian
    direction; identity affine."""
    voxels = np.argwhere(mask)
    index = rng.integers(0, len(voxels), n_seeds)
    locations = voxels[index] + rng.random((n_seeds, 3)) - 0.5
    directions = rng.standard_normal((n_seeds, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    seeds = np.zeros((n_seeds, 8), dtype=np.float32)
    seeds[:, :3] = locations
    seeds[:, 3] = 1
    seeds[:, 4:7] = directions
    return seeds


def sphere32_transport(seed: int = 0, n_seeds: int = 10_000):
    """32^3 SH order-8 FOD from 1-3 random fibres per voxel, sphere mask of
    radius 14, seeds in the mask; Transport, step 0.1, save 0.25, max 300."""
    rng = np.random.default_rng(seed)
    shape = (32, 32, 32)
    d, w = _random_fibres(rng, int(np.prod(shape)))
    fod = fod_from_fibres(d, w).reshape(*shape, 45)
    idx = np.indices(shape).reshape(3, -1).T
    mask = (np.linalg.norm(idx - 15.5, axis=1) <= 14).reshape(shape)
    fod[~mask] = 0
    seeds = seeds_in_mask(rng, mask, n_seeds)
    return {"fod": fod, "mask": mask.astype(np.uint8), "seeds": seeds, "algorithm": "transport",
            "step_size": 0.1, "gamma": 1.0, "noise_variance": 0.0, "save_at": 0.25, "max_length": 300.0}


def brain145_diffusion(seed: int = 0, n_seeds: int = 1_000_000, shape=SHAPE_145):
    """145x174x145 SH order-8 FOD whose 1-3 fibres per voxel follow smooth
    random direction fields (sigma 4), ellipsoid mask 60/75/60; Diffusion
    gamma 1, noise variance 0.1, step 0.1, save 0.25, max 300."""
    rng = np.random.default_rng(seed)
    directions = smooth_direction_fields(rng, shape)
    _, weights = _random_fibres(rng, int(np.prod(shape)), base_directions=directions)
    fod = fod_from_fibres(directions, weights).reshape(*shape, 45)
    semi = np.array([60, 75, 60]) * np.array(shape) / np.array(SHAPE_145)
    mask = ellipsoid_mask(shape, semi)
    fod[~mask] = 0
    seeds = seeds_in_mask(rng, mask, n_seeds)
    return {"fod": fod, "mask": mask.astype(np.uint8), "seeds": seeds, "algorithm": "diffusion",
            "step_size": 0.1, "gamma": 1.0, "noise_variance": 0.1, "save_at": 0.25, "max_length": 300.0}


def _semi_axes(shape, semi) -> np.ndarray:
    return np.array(semi, dtype=np.float64) * np.array(shape) / np.array(SHAPE_145)


def brain145_dti_transport(seed: int = 0, n_seeds: int = 1_000_000, shape=SHAPE_145):
    """145x174x145 diffusion tensors D = R diag(1.7, 0.3, 0.3)e-3 R^T whose
    principal axis is the first smooth direction field of
    brain145_diffusion(seed) (sigma 4), 0.8e-3 I outside the ellipsoid
    60/75/60; Transport gamma 1, step 0.1, save 0.25, max 300."""
    rng = np.random.default_rng(seed)
    e = smooth_direction_fields(rng, shape, n_fields=1)[:, 0]
    mask = ellipsoid_mask(shape, _semi_axes(shape, (60, 75, 60)))
    inside = mask.reshape(-1)
    low, delta = 0.3e-3, 1.7e-3 - 0.3e-3
    tensors = np.zeros((len(e), 6), dtype=np.float64)
    for c in range(3):
        tensors[:, c] = low + delta * e[:, c] * e[:, c]
    tensors[:, 3] = delta * e[:, 0] * e[:, 1]
    tensors[:, 4] = delta * e[:, 0] * e[:, 2]
    tensors[:, 5] = delta * e[:, 1] * e[:, 2]
    tensors[~inside] = [0.8e-3, 0.8e-3, 0.8e-3, 0.0, 0.0, 0.0]
    seeds = seeds_in_mask(rng, mask, n_seeds)
    return {"fod": tensors.astype(np.float32).reshape(*shape, 6), "mask": mask.astype(np.uint8), "seeds": seeds,
            "model": "dti", "algorithm": "transport", "step_size": 0.1, "gamma": 1.0, "noise_variance": 0.0,
            "save_at": 0.25, "max_length": 300.0}


def _fibonacci(n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - 2.0 * i / (n - 1)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    t = i * (2.0 * math.pi / ((1.0 + math.sqrt(5.0)) / 2.0))
    return np.stack([r * np.cos(t), r * np.sin(t), z], axis=1)


def radial_noise(rng, amplitude: float, lmax: int = 4):
    """A smooth function on the sphere: random real SH coefficients of
    degree <= lmax, scaled to max |n| = amplitude. Returns n(u) for unit
    vectors u (m, 3)."""
    K = (lmax + 1) * (lmax + 2) // 2
    coefficients = rng.standard_normal(K)
    peak = np.abs(real_sh_basis(_fibonacci(20000), lmax) @ coefficients).max()
    scale = amplitude / peak

    def noise(u):
        return scale * (real_sh_basis(u, lmax) @ coefficients)

    return noise


def ellipsoid_radius(u: np.ndarray, semi_axes) -> np.ndarray:
    """Distance from the centre to the ellipsoid along unit vectors u."""
    return 1.0 / np.sqrt(np.sum((u / np.asarray(semi_axes, dtype=np.float64)) ** 2, axis=1))


def ellipsoid_surface(centre, semi_axes, n_rings: int, n_segments: int, noise=None, inward: bool = False):
    """Latitude-longitude triangulation of an ellipsoid whose vertices are
    moved radially by noise(u): poles plus n_rings - 1 rings of n_segments
    vertices, 2 n_segments (n_rings - 1) triangles, normals outward (or
    inward). Returns vertices (V, 3) float64 and faces (F, 3) int64."""
    theta = math.pi * np.arange(1, n_rings) / n_rings
    phi = 2.0 * math.pi * np.arange(n_segments) / n_segments
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    ring = np.stack([st * np.cos(phi)[None, :], st * np.sin(phi)[None, :],
                     np.broadcast_to(ct, (len(theta), n_segments))], axis=-1).reshape(-1, 3)
    u = np.vstack([[0.0, 0.0, 1.0], ring, [0.0, 0.0, -1.0]])
    r = ellipsoid_radius(u, semi_axes)
    if noise is not None:
        r = r + noise(u)
    vertices = np.asarray(centre, dtype=np.float64)[None, :] + r[:, None] * u
    north, south = 0, len(u) - 1
    j = np.arange(n_segments)
    jn = (j + 1) % n_segments
    faces = [np.stack([np.full(n_segments, north), 1 + j, 1 + jn], axis=1)]
    for i in range(n_rings - 2):
        a = 1 + i * n_segments + j
        b = 1 + (i + 1) * n_segments + j
        c = 1 + (i + 1) * n_segments + jn
        d = 1 + i * n_segments + jn
        faces.append(np.stack([a, b, c], axis=1))
        faces.append(np.stack([a, c, d], axis=1))
    last = 1 + (n_rings - 2) * n_segments
    faces.append(np.stack([np.full(n_segments, south), last + jn, last + j], axis=1))
    faces = np.vstack(faces).astype(np.int64)
    if inward:
        faces = faces[:, ::-1].copy()
    return vertices, faces


def shell_domain(rng, shape, outer, inner, n_rings: int = 315, n_segments: int = 318, amplitude: float = 2.0):
    """Two nested displaced ellipsoid surfaces (semi-axes in voxels) about
    the volume centre, the inner one facing inward, as one triangle list,
    with the radial functions that define them."""
    centre = (np.array(shape) - 1) / 2.0
    outer, inner = np.asarray(outer, dtype=np.float64), np.asarray(inner, dtype=np.float64)
    noise_out = radial_noise(rng, amplitude)
    noise_in = radial_noise(rng, amplitude)
    vo, fo = ellipsoid_surface(centre, outer, n_rings, n_segments, noise_out)
    vi, fi = ellipsoid_surface(centre, inner, n_rings, n_segments, noise_in, inward=True)
    vertices = np.vstack([vo, vi]).astype(np.float32)
    faces = np.vstack([fo, fi + len(vo)]).astype(np.uint32)

    def radii(u):
        return (ellipsoid_radius(u, inner) + noise_in(u), ellipsoid_radius(u, outer) + noise_out(u))

    return vertices, faces, centre, radii


def seeds_in_shell(rng, centre, radii, extent, n_seeds: int, margin: float = 0.5) -> np.ndarray:
    """Uniform seeds between the two surfaces (at least `margin` from each),
    by rejection in the box centre +- extent; normalised Gaussian directions."""
    kept = []
    count = 0
    while count < n_seeds:
        q = centre + rng.uniform(-1.0, 1.0, (max(4096, 3 * (n_seeds - count)), 3)) * extent
        d = q - centre
        r = np.linalg.norm(d, axis=1)
        u = d / np.maximum(r, 1e-12)[:, None]
        r_in, r_out = radii(u)
        q = q[(r > r_in + margin) & (r < r_out - margin)]
        kept.append(q)
        count += len(q)
    locations = np.vstack(kept)[:n_seeds]
    directions = rng.standard_normal((n_seeds, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    seeds = np.zeros((n_seeds, 8), dtype=np.float32)
    seeds[:, :3] = locations
    seeds[:, 3] = 1
    seeds[:, 4:7] = directions
    return seeds


def brain145_shell_diffusion(seed: int = 0, n_seeds: int = 2_000_000, shape=SHAPE_145, n_rings: int = 315,
                             n_segments: int = 318, amplitude: float = 2.0):
    """The brain145_diffusion field with the domain bounded by two nested
    closed meshes (ellipsoids 60/75/60 and 30/40/30, ~200k triangles each,
    vertices displaced radially by degree-4 noise of the given amplitude);
    seeds uniform between the surfaces; Diffusion gamma 1, noise variance
    0.1, step 0.1, save 0.25, max 300. Tracking stops on a boundary hit."""
    cfg = brain145_diffusion(seed, n_seeds=0, shape=shape)
    rng = np.random.default_rng(seed + 1)
    vertices, faces, centre, radii = shell_domain(rng, shape, _semi_axes(shape, (60, 75, 60)),
                                                  _semi_axes(shape, (30, 40, 30)), n_rings, n_segments, amplitude)
    extent = _semi_axes(shape, (60, 75, 60)) + amplitude + 1.0
    cfg["seeds"] = seeds_in_shell(rng, centre, radii, extent, n_seeds)
    cfg["mesh"] = (vertices, faces)
    # Culling cells of half a voxel certifying segments up to 1.5 steps.
    cfg["mesh_grid"] = (0.5, 1.5 * cfg["step_size"])
    return cfg


SH91_CHUNKS = (300_000, 700_000, 1_100_000, 1_500_000, 2_500_000, 1_900_000, 1_200_000, 800_000)


def brain145_sh91_fod_seeds(seed: int = 0, n_seeds: int = 10_000_000, shape=SHAPE_145):
    """brain145_diffusion's fibres projected to SH order 12 (91
    coefficients); the seeds are drawn on the device from the FOD
    (seeds.from_fod, split-bfloat16 MFMA sampler) in 8 uneven chunks summing
    to n_seeds. Diffusion gamma 1, noise variance 0.1, step 0.1, save 0.25,
    max 300."""
    rng = np.random.default_rng(seed)
    directions = smooth_direction_fields(rng, shape)
    _, weights = _random_fibres(rng, int(np.prod(shape)), base_directions=directions)
    fod = fod_from_fibres(directions, weights, lmax=12).reshape(*shape, 91)
    mask = ellipsoid_mask(shape, _semi_axes(shape, (60, 75, 60)))
    fod[~mask] = 0
    scale = n_seeds / sum(SH91_CHUNKS)
    chunks = [max(1, int(round(c * scale))) for c in SH91_CHUNKS]
    chunks[-1] = max(1, n_seeds - sum(chunks[:-1]))
    return {"fod": fod, "mask": mask.astype(np.uint8), "seeds": None, "n_seeds": int(sum(chunks)),
            "chunks": chunks, "algorithm": "diffusion", "step_size": 0.1, "gamma": 1.0, "noise_variance": 0.1,
            "save_at": 0.25, "max_length": 300.0, "state_seed": seed + 2}


# --------------------------------------------------------------------------
# Measurement
# --------------------------------------------------------------------------

MI355X_HBM_PEAK_GBS = 8000.0
# SIMD quad-cycles per second: 256 CUs x 4 SIMDs x 2.4 GHz / 4. A wave64 VALU
# instruction takes one quad-cycle of its SIMD; gfx950 issues two in one
# quad-cycle when their operands are all VGPRs (FP32 add, mul and fma measured
# at 2.4-2.6 cycles each, one per 4.2 cycles with an SGPR operand, FP64 fma
# one per 4.6), which SQ_ACTIVE_INST_VALU2 counts.
MI355X_VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def bytes_per_step(K: int, step: float, save_at: float, mask: bool = True) -> float:
    """Algorithmic HBM bytes per step: K f32 coefficients, one mask word (the
    reference's 4-byte int; mask domains only) and a 16-byte saved point
    times saves per step. BVH traversal of mesh domains is not counted."""
    return 4.0 * K + (4.0 if mask else 0.0) + 16.0 * step / save_at


def unique_bytes(K: int, reads: dict, saved_points: int, n_seeds: int) -> float:
    """Bytes a tracking launch with a per-lane cell cache must move at
    least: K f32 coefficients per row read and one mask byte per mask read
    (both counted by the kernel), 12 B per saved point (f32 x, y, z), the
    32-B seed rows and 12 B of per-seed counts (length, steps, hit)."""
    return 4.0 * K * reads["rows"] + reads["mask_values"] + 12.0 * saved_points + 44.0 * n_seeds


def make_algorithm(tg, cfg, seed: int = 0):
    if cfg.get("model") == "dti":
        model = tg.local.DiffusionTensorImaging()
    else:
        model = tg.local.RealSymmetricSphericalHarmonics(cfg["fod"].shape[-1])
    if cfg["algorithm"] == "probabilistic":
        # The reference's defaults (algorithm.py:288-295) at the workload's step.
        return tg.algorithm.Probabilistic(model, seed=seed, step_size=cfg["step_size"])
    if cfg["algorithm"] == "deterministic":
        return tg.algorithm.Deterministic(model, step_size=cfg["step_size"])
    if cfg["algorithm"] == "transport":
        return tg.algorithm.Transport(model, step_size=cfg["step_size"], gamma=cfg["gamma"])
    return tg.algorithm.Diffusion(model, step_size=cfg["step_size"], gamma=cfg["gamma"],
                                  noise_variance=cfg["noise_variance"], seed=seed)


def make_domain(tg, cfg):
    """The workload's tracking domain: its boundary meshes if it has any,
    else its mask image."""
    if cfg.get("mesh") is not None:
        cell_size, reach = cfg.get("mesh_grid", (1.0, 0.5))
        return tg.domain.Mesh(*cfg["mesh"], cell_size=cell_size, reach=reach)
    return tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))


class DeviceRun:
    """Device-resident inputs and outputs of one workload for repeated timing."""

    def __init__(self, tg, torch, runtime, cfg, batch_size: int = 50_000):
        from tractography_amd import _native

        self.tg, self.torch, self.runtime, self.cfg = tg, torch, runtime, cfg
        self.native = _native
        self.field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
        self.domain = make_domain(tg, cfg)
        self.algorithm = make_algorithm(tg, cfg)
        self.n_points = tg.tracking.point_capacity(cfg["max_length"], cfg["save_at"])
        seeds = np.ascontiguousarray(cfg["seeds"], dtype=np.float32)
        self.n = len(seeds)
        self.batch_size = batch_size
        n_batches = (self.n + batch_size - 1) // batch_size
        rng_seeds = np.array([self.algorithm.get_rng_seed() for _ in range(n_batches)], dtype=np.uint32)
        dev = torch.device("cuda", runtime.device)
        self.d_seeds = torch.from_numpy(seeds).to(dev)
        self.d_rng = torch.from_numpy(rng_seeds.view(np.int32)).to(dev)
        self.d_lengths = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.d_steps = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.d_hits = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.d_dense = torch.empty(self.n * self.n_points * 3, dtype=torch.float32, device=dev)
        self.d_offsets = torch.empty(self.n + 1, dtype=torch.int64, device=dev)
        self.d_points = torch.empty(self.n * self.n_points * 3, dtype=torch.float32, device=dev)
        self.field_handle = self.field.upload(runtime)
        self.domain_handle = self.domain.upload(runtime)
        torch.cuda.synchronize()
        self.stream = torch.cuda.ExternalStream(runtime.stream(), device=dev)

    def track(self, path: int):
        lib = self.native.lib()
        if path == self.native.PATH_FAST:
            self.native.check(lib.tgk_field_set_fast_precision(self.field_handle.handle, 0))
        self.native.check(lib.tgk_track_device(
            self.runtime.handle, self.domain_handle.handle, self.field_handle.handle, self.algorithm.descriptor(),
            path, self.d_seeds.data_ptr(), 0, self.n, self.d_rng.data_ptr(), self.batch_size, self.n_points,
            float(np.float32(self.cfg["save_at"])), self.d_lengths.data_ptr(), self.d_steps.data_ptr(),
            self.d_hits.data_ptr(), self.d_dense.data_ptr()))

    def compact(self, path: int):
        lib = self.native.lib()
        self.native.check(lib.tgk_compact_device(
            self.runtime.handle, path, self.d_dense.data_ptr(), self.n, self.n_points, self.d_lengths.data_ptr(),
            self.d_offsets.data_ptr(), self.d_points.data_ptr()))

    def timed(self, path: int, warmup: int, steps: int):
        """(ms per step, track-kernel ms per step, total orientation updates)."""
        torch = self.torch
        for _ in range(warmup):
            self.track(path)
            self.compact(path)
        self.runtime.finish()
        events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        self.runtime.finish()
        t0 = time.perf_counter()
        for start, stop in events:
            start.record(self.stream)
            self.track(path)
            stop.record(self.stream)
            self.compact(path)
        self.runtime.finish()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        kernel = float(np.mean([a.elapsed_time(b) for a, b in events]))
        total_steps = int(self.d_steps.to(torch.int64).sum().item())
        return wall, kernel, total_steps

    def results(self, n_first: int):
        """lengths, steps and ragged points of the first n_first seeds."""
        lengths = self.d_lengths[:n_first].cpu().numpy().view(np.uint32)
        steps = self.d_steps[:n_first].cpu().numpy().view(np.uint32)
        total = int(self.d_offsets[n_first].item())
        points = self.d_points[: total * 3].cpu().numpy().reshape(-1, 3)
        return lengths, steps, points


class GraphRun:
    """FOD seeding on the device, fast tracking and compaction over uneven
    chunks of seeds, recorded once into a HIP graph; one bench step is one
    replay of the whole graph. Chunks run in order on one stream and share
    one dense and one ragged point buffer sized for the largest chunk."""

    def __init__(self, tg, torch, runtime, cfg, method: int = 1, batch_size: int = 50_000):
        from tractography_amd import _native

        self.tg, self.torch, self.runtime, self.cfg = tg, torch, runtime, cfg
        self.native, self.lib = _native, _native.lib()
        self.method = method
        self.chunks = list(cfg["chunks"])
        self.n = sum(self.chunks)
        fod = cfg["fod"]
        self.K = fod.shape[-1]
        self.field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
        self.domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
        self.algorithm = make_algorithm(tg, cfg)
        self.n_points = tg.tracking.point_capacity(cfg["max_length"], cfg["save_at"])
        inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, np.eye(4)))
        self.nnz = len(voxels)
        dev = torch.device("cuda", runtime.device)
        self.d_fod = torch.from_numpy(inline).to(dev)
        self.d_voxels = torch.from_numpy(voxels).to(dev)
        self.d_affine = torch.from_numpy(affine).to(dev)
        del inline, voxels
        states = np.random.default_rng(cfg["state_seed"]).integers(0, 2**32 - 1, (self.n, 2), dtype=np.uint32)
        self.d_states = torch.from_numpy(states.view(np.int32)).to(dev)
        self.d_seeds = torch.zeros((self.n, 8), dtype=torch.float32, device=dev)
        n_batches = (self.n + batch_size - 1) // batch_size
        rng_seeds = np.array([self.algorithm.get_rng_seed() for _ in range(n_batches)], dtype=np.uint32)
        self.batch_size = batch_size
        self.d_rng = torch.from_numpy(rng_seeds.view(np.int32)).to(dev)
        self.d_lengths = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.d_steps = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.d_hits = torch.zeros(self.n, dtype=torch.int32, device=dev)
        largest = max(self.chunks)
        self.d_dense = torch.empty(largest * self.n_points * 3, dtype=torch.float32, device=dev)
        self.d_offsets = torch.zeros(largest + 1, dtype=torch.int64, device=dev)
        self.d_points = torch.empty(largest * self.n_points * 3, dtype=torch.float32, device=dev)
        self.field_handle = self.field.upload(runtime)
        self.domain_handle = self.domain.upload(runtime)
        torch.cuda.synchronize()
        self.native.check(self.lib.tgk_track_prepare(runtime.handle, self.domain_handle.handle,
                                                     self.field_handle.handle, self.algorithm.descriptor(),
                                                     _native.PATH_FAST, largest))
        self.native.check(self.lib.tgk_seeds_prepare(runtime.handle, 0, self.K, method))
        self.stream = torch.cuda.ExternalStream(runtime.stream(), device=dev)
        self.graph = None

    def enqueue(self, events=None, reads=None):
        """The chunk loop on the runtime stream; with `events`, a (start, stop)
        pair is recorded around each tracking launch; with `reads`, each
        launch's row and mask read counts are appended (synchronising)."""
        check, lib, h = self.native.check, self.lib, self.runtime.handle
        first = 0
        for count in self.chunks:
            seeds = self.d_seeds.data_ptr() + first * 32
            check(lib.tgk_seeds_from_fod_device(h, 0, self.K, self.d_fod.data_ptr(), self.d_voxels.data_ptr(),
                                                self.nnz, self.d_affine.data_ptr(),
                                                self.d_states.data_ptr() + first * 8, count, self.method, seeds))
            if events is not None:
                events.append((self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)))
                events[-1][0].record(self.stream)
            check(lib.tgk_track_device(h, self.domain_handle.handle, self.field_handle.handle,
                                       self.algorithm.descriptor(), self.native.PATH_FAST, seeds, first, count,
                                       self.d_rng.data_ptr(), self.batch_size, self.n_points,
                                       float(np.float32(self.cfg["save_at"])), self.d_lengths.data_ptr() + first * 4,
                                       self.d_steps.data_ptr() + first * 4, self.d_hits.data_ptr() + first * 4,
                                       self.d_dense.data_ptr()))
            if events is not None:
                events[-1][1].record(self.stream)
            if reads is not None:
                reads.append(self.runtime.fast_reads())
            check(lib.tgk_compact_device(h, self.native.PATH_FAST, self.d_dense.data_ptr(), count, self.n_points,
                                         self.d_lengths.data_ptr() + first * 4, self.d_offsets.data_ptr(),
                                         self.d_points.data_ptr()))
            first += count

    def record(self):
        import ctypes

        self.native.check(self.lib.tgk_graph_begin(self.runtime.handle))
        try:
            self.enqueue()
        finally:
            graph = ctypes.c_void_p()
            status = self.lib.tgk_graph_end(self.runtime.handle, ctypes.byref(graph))
        self.native.check(status)
        self.graph = graph

    def total_steps(self) -> int:
        self.runtime.finish()
        return int(self.d_steps.to(self.torch.int64).sum().item())

    def timed(self, warmup: int, steps: int):
        """(ms per graph replay, track-kernel ms per replay from an eager
        pass, orientation updates per replay)."""
        torch = self.torch
        events = []
        self.reads = []
        self.enqueue(events, self.reads)  # eager pass: tracking kernel time and reads per chunk
        self.runtime.finish()
        kernel = float(sum(a.elapsed_time(b) for a, b in events))
        self.saved_points = int(self.d_lengths.to(torch.int64).sum().item())
        self.record()
        for _ in range(warmup):
            self.native.check(self.lib.tgk_graph_launch(self.graph))
        self.runtime.finish()
        times, updates = [], 0
        for _ in range(steps):
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(self.stream)
            self.native.check(self.lib.tgk_graph_launch(self.graph))
            stop.record(self.stream)
            updates += self.total_steps()
            times.append(start.elapsed_time(stop))
        return float(np.mean(times)), kernel, updates // steps

    def close(self):
        if self.graph is not None:
            self.lib.tgk_graph_destroy(self.graph)
            self.graph = None


def host_leg(cfg, gpu_results: dict, seconds: float = 15.0, n_accuracy: int = 400):
    """The host restatement's leg of the benchmark.

    cpu_baseline: the FP32 restatement (OpenMP over seeds, all host cores) on
    a leading sample of the workload's seeds sized to ~`seconds` of wall time.
    accuracy: deviation of each GPU path from the FP64 restatement on the
    first `n_accuracy` seeds (same noise streams), measured up to the first
    possible voxel flip.
    """
    sys.path.insert(0, str(ROOT / "tests"))
    import deviation  # noqa: E402
    import oracle  # noqa: E402  (the host restatement)

    import tractography_amd as tg

    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = make_domain(tg, cfg)
    algorithm = make_algorithm(tg, cfg)
    rng_seeds = [algorithm.get_rng_seed() for _ in range((len(cfg["seeds"]) + 49_999) // 50_000)]
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    sample = max(64, threads * 8)
    while True:
        seeds = cfg["seeds"][:sample]
        t0 = time.perf_counter()
        r = oracle.track(domain, field, algorithm, seeds, rng_seeds, 50_000, cfg["save_at"], cfg["max_length"],
                         threads=threads)
        dt = time.perf_counter() - t0
        if dt > 0.6 * seconds or sample >= len(cfg["seeds"]):
            break
        # Steps/s is additive over seeds: size the next sample to ~`seconds`.
        sample = min(len(cfg["seeds"]), int(sample * min(64.0, max(2.0, seconds / max(dt, 1e-3)))))
    steps = int(r["steps"].astype(np.int64).sum())
    baseline = {"value": steps / dt, "unit": "streamline steps/s", "cores": threads, "kind": "port",
                "sample": f"first {len(seeds)} seeds of the workload, {steps} steps in {dt:.1f} s "
                          "(FP32 host restatement, OpenMP)"}
    accuracy = {}
    n_acc = min(n_accuracy, len(cfg["seeds"]))
    ref64 = oracle.track(domain, field, algorithm, cfg["seeds"][:n_acc], rng_seeds, 50_000, cfg["save_at"],
                         cfg["max_length"], precision=64, threads=threads)
    for name, (lengths, steps_, points) in gpu_results.items():
        rep = deviation.report(lengths[:n_acc], steps_[:n_acc], points, ref64)
        accuracy[name] = {"max_point_deviation_voxel": rep["max_deviation"],
                          "max_angular_error_rad": rep["max_angular_error_rad"],
                          "compared_fraction": rep["compared_fraction"],
                          "length_equal_fraction": rep["length_equal_fraction"],
                          "seeds": n_acc}
    return baseline, accuracy


Y0_F32 = float(np.float32(0.28209479177387814))  # the constant l = 0 basis value, as deposited
MI355X_BF16_DENSE_TFLOPS = 2500.0  # dense MFMA bf16 peak (no sparsity)


def seeds_leg(args, tg, torch, cfg, workload: str):
    """Time seeds.from_fod on the device (seeds.py:145-240): n seeds drawn
    from the workload's FOD inside its mask, states resident in HBM, for the
    VALU method (the reference's arithmetic) and the split-bf16 MFMA method.
    The MFMA rate is priced at its executed product flops, 2 passes x 1000
    directions x 2K flop x 3 split products per seed (an upper bound: the
    second pass stops once every lane of a wavefront has chosen)."""
    from tractography_amd import _native

    lib = _native.lib()
    runtime = tg.HipRuntime()
    fod = cfg["fod"] * cfg["mask"][..., None]
    K = fod.shape[-1]
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, np.eye(4)))
    n = int(args.seeds or 10_000_000)
    dev = torch.device("cuda", runtime.device)
    d_fod = torch.from_numpy(inline).to(dev)
    d_voxels = torch.from_numpy(voxels).to(dev)
    d_affine = torch.from_numpy(affine).to(dev)
    states = np.random.default_rng(0).integers(0, 2**32 - 1, (n, 2), dtype=np.uint32)
    d_states0 = torch.from_numpy(states.view(np.int32)).to(dev)
    d_states = d_states0.clone()
    d_seeds = torch.empty((n, 8), dtype=torch.float32, device=dev)
    stream = torch.cuda.ExternalStream(runtime.stream(), device=dev)
    methods = {"valu": _native.SEEDS_VALU, "mfma": _native.SEEDS_MFMA} if K <= 45 else {"mfma": _native.SEEDS_MFMA}
    results = {}
    for name, method in methods.items():
        _native.check(lib.tgk_seeds_prepare(runtime.handle, 0, K, method))
        times = []
        for i in range(args.warmup + args.steps):
            d_states.copy_(d_states0)
            torch.cuda.synchronize()
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(stream)
            _native.check(lib.tgk_seeds_from_fod_device(runtime.handle, 0, K, d_fod.data_ptr(), d_voxels.data_ptr(),
                                                        len(voxels), d_affine.data_ptr(), d_states.data_ptr(), n,
                                                        method, d_seeds.data_ptr()))
            stop.record(stream)
            runtime.finish()
            torch.cuda.synchronize()
            if i >= args.warmup:
                times.append(start.elapsed_time(stop))
        ms = float(np.mean(times))
        results[name] = {"ms": ms, "seeds_per_s": n / (ms * 1e-3)}
        print(f"[bench] seeds {name}: {ms:.1f} ms, {n / (ms * 1e-3):.3e} seeds/s", file=sys.stderr, flush=True)
    line = {"metric": "seeds.from_fod seeds/s on 1 MI355X", "unit": "seeds/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic (seeded numpy generators)",
            "config": {"workload": workload, "product": "seeds", "seeds": n, "K": K, "nnz": len(voxels)},
            "methods": results, "value": results["mfma"]["seeds_per_s"], "ms_per_step": results["mfma"]["ms"],
            "dtype": "bf16 (hi+lo split) MFMA, f32 accumulate"}
    # Algorithmic work: two passes over 1000 directions of a K-term dot
    # product per seed, 4000 K flop (the three split-bf16 products the
    # kernel issues per product are its cost, not the algorithm's). The
    # second pass ends early once every lane has chosen, so this is an
    # upper bound of the work the algorithm needs.
    flops = n * 4000 * K
    achieved = flops / (results["mfma"]["ms"] * 1e-3) / 1e12
    line["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": MI355X_BF16_DENSE_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / MI355X_BF16_DENSE_TFLOPS, "kernel": "seeds_mfma_kernel",
                        "algorithmic_flops_per_seed": 4000 * K,
                        "issued_mfma_flops_per_seed_upper_bound": 2 * 1000 * 2 * K * 3}
    print(json.dumps(line), flush=True)


def product_leg(args, tg, torch, cfg, workload: str):
    """Time one TOD or connectome pass over boundary seeds of the workload's
    mask domain (n_seeds tracks of at most max_length / step steps) on each
    path of --paths, and the FP32 host restatement (one core) on a bounded
    sample of the same tracks. Prints one JSON line."""
    from tractography_amd import _native
    from tractography_amd.domain import _mesh_handle_for

    lib = _native.lib()
    runtime = tg.HipRuntime()
    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
    algorithm = make_algorithm(tg, cfg)
    rng_seed = int(algorithm.get_rng_seed())
    n = int(args.seeds or len(cfg["seeds"]))
    n_steps = int(np.floor(cfg["max_length"] / cfg["step_size"]))
    field_handle = field.upload(runtime)
    domain_handle = domain.upload(runtime, with_boundary=True)
    dev = torch.device("cuda", runtime.device)
    stream = torch.cuda.ExternalStream(runtime.stream(), device=dev)
    dims = np.ascontiguousarray(cfg["mask"].shape, dtype=np.uint32)
    inverse = np.ascontiguousarray(np.eye(4), dtype=np.float32)
    path_ids = {"plain": _native.PATH_PLAIN, "fast": _native.PATH_FAST}
    if args.product == "tod":
        paths = [p for p in args.paths.split(",") if p in path_ids]
        out = torch.empty(int(np.prod(dims)) * 45, dtype=torch.float32, device=dev)

        def launch(path):
            _native.check(lib.tgk_tod_device(runtime.handle, domain_handle.handle, domain_handle.handle,
                                             field_handle.handle, algorithm.descriptor(), path_ids[path], rng_seed,
                                             n, n_steps, _native.ptr(dims), _native.ptr(inverse), 0,
                                             out.data_ptr()))
    else:
        paths = [p for p in args.paths.split(",") if p in path_ids]
        boundary = _mesh_handle_for(domain, runtime)
        # The host entry point copies its 9 entries per seed back; the timed
        # region is the whole call.
        rows_h = np.empty(9 * n, np.int32)
        cols_h = np.empty(9 * n, np.int32)
        weights_h = np.empty(9 * n, np.float32)

        def launch(path):
            _native.check(lib.tgk_connectome(runtime.handle, domain_handle.handle, boundary.handle,
                                             field_handle.handle, algorithm.descriptor(), path_ids[path], rng_seed,
                                             n, n_steps, 0.0, _native.ptr(rows_h), _native.ptr(cols_h),
                                             _native.ptr(weights_h)))
    results = {}
    for path in paths:
        for _ in range(args.warmup):
            launch(path)
        runtime.finish()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(stream)
            t0 = time.perf_counter()
            launch(path)
            stop.record(stream)
            runtime.finish()
            torch.cuda.synchronize()
            times.append(((time.perf_counter() - t0) * 1e3, start.elapsed_time(stop)))
        wall = float(np.mean([t[0] for t in times]))
        device_ms = float(np.mean([t[1] for t in times]))
        results[path] = {"ms_per_step": wall, "kernel_ms": device_ms}
        if args.product == "tod":
            raw = out.view(-1, 45)[:, 0].to(torch.float64)
            deposits = float(raw.sum().item()) / Y0_F32
            results[path].update(tracks_per_s=n / (device_ms * 1e-3), deposited_steps=deposits,
                                 steps_per_s=deposits / (device_ms * 1e-3))
        else:
            results[path].update(tracks_per_s=n / (wall * 1e-3), entries_set=int((rows_h >= 0).sum()))
        print(f"[bench] {args.product} {path}: {wall:.1f} ms per call, {results[path]['tracks_per_s']:.3e} tracks/s",
              file=sys.stderr, flush=True)
    headline = "fast" if "fast" in results else "plain"
    h = results[headline]
    line = {"metric": f"{args.product} tracks/s on 1 MI355X ({headline} path)", "unit": "tracks/s",
            "value": h["tracks_per_s"], "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded numpy generators)",
            "config": {"workload": workload, "product": args.product, "path": headline, "boundary_seeds": n,
                       "max_steps": n_steps},
            "paths": results}
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle  # noqa: E402  (the host restatement)

        sample = 64
        while True:
            t0 = time.perf_counter()
            if args.product == "tod":
                r, _, _ = oracle.tod(domain, field, algorithm, rng_seed, sample, n_steps, dims, inverse)
                work = float(r[..., 0].astype(np.float64).sum()) / Y0_F32
            else:
                oracle.connectome(domain, field, algorithm, rng_seed, sample, n_steps, threads=1)
                work = sample
            dt = time.perf_counter() - t0
            if dt > 5.0 or sample >= n:
                break
            sample = min(n, int(sample * min(16.0, max(2.0, 8.0 / max(dt, 1e-3)))))
        line["cpu_baseline"] = {"value": sample / dt, "unit": "tracks/s", "cores": 1, "kind": "port",
                                "sample": f"first {sample} boundary seeds in {dt:.1f} s (FP32 host restatement)"}
        if args.product == "tod":
            line["cpu_baseline"]["steps_per_s"] = work / dt
    print(json.dumps(line), flush=True)


WORKLOADS = {
    "sphere32": (sphere32_transport, 10_000,
                 "32^3 SH45 (lmax 8) Transport gamma 1, {n} seeds, step 0.1, save 0.25, max 300"),
    "sh45": (brain145_diffusion, 1_000_000,
             "145x174x145 SH45 (lmax 8) Diffusion gamma 1 sigma^2 0.1, {n} seeds, step 0.1, save 0.25, max 300"),
    "dti": (brain145_dti_transport, 1_000_000,
            "145x174x145 DTI Transport gamma 1, {n} seeds, step 0.1, save 0.25, max 300"),
    "shell": (brain145_shell_diffusion, 2_000_000,
              "145x174x145 SH45 Diffusion gamma 1 sigma^2 0.1 between two displaced ellipsoid meshes "
              "(2 x ~200k triangles, BVH termination), {n} seeds, step 0.1, save 0.25, max 300"),
    "sh91": (brain145_sh91_fod_seeds, 10_000_000,
             "145x174x145 SH91 (lmax 12) Diffusion gamma 1 sigma^2 0.1, {n} seeds drawn on the device from the "
             "FOD (split-bf16 MFMA sampler) in 8 uneven chunks, one HIP graph per replay, fast path, step 0.1, "
             "save 0.25, max 300"),
}


def main():
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--workload", default="sh45", choices=sorted(WORKLOADS))
    parser.add_argument("--seeds", type=int, default=None)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--paths", default="fast,plain")
    parser.add_argument("--fast-grid", type=int, default=0,
                        help="blocks of the fast kernel (default: occupancy-sized)")
    parser.add_argument("--fast-shape", type=int, default=0,
                        help="launch shape of the order-8 SH Transport / Diffusion fast kernels "
                             "(tractography_amd.runtime.fast_shapes(); default 0)")
    parser.add_argument("--autotune", action="store_true",
                        help="choose the fast kernel's launch shape, grid and repack threshold on up to 2M seeds of "
                             "the workload before timing")
    parser.add_argument("--repack", type=int, default=0,
                        help="fast path: streamlines below which a generation is no longer cut short "
                             "(default: the library's, 16384)")
    parser.add_argument("--no-api", dest="api", action="store_false",
                        help="skip timing tg.tractogram on host arrays (transfers and uploads included)")
    parser.add_argument("--api", dest="api", action="store_true", help=argparse.SUPPRESS)
    parser.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                        help="before timing, run one elementwise copy that reads and writes BYTES each "
                             "(reference dispatch for hardware byte counters)")
    parser.add_argument("--no-traffic", action="store_true",
                        help="skip the two hardware-counter runs that measure the tracking kernel's HBM bytes")
    parser.add_argument("--algorithm", default=None, choices=["transport", "diffusion", "probabilistic", "deterministic"],
                        help="replace the workload's algorithm (Probabilistic and Deterministic: plain path only, "
                             "the reference's default parameters)")
    parser.add_argument("--product", default="tractogram", choices=["tractogram", "tod", "connectome", "seeds"],
                        help="tod / connectome: time that product over boundary seeds of the workload's mask; "
                             "seeds: time seeds.from_fod on the device (VALU and MFMA methods)")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (seeds shard with no "
              "communication, see the multi-GPU section of the design notes)", file=sys.stderr)
        sys.exit(2)
    # Counter runs are child processes started before this process touches
    # the GPU.
    traffic = None
    if (not args.no_traffic and args.calibrate_copy == 0 and args.workload != "sh91" and "fast" in args.paths
            and args.product == "tractogram"):
        traffic = traffic_leg(args)

    import torch

    import tractography_amd as tg
    from tractography_amd import _native

    t0 = time.perf_counter()
    generator, default_seeds, description = WORKLOADS[args.workload]
    cfg = generator(n_seeds=args.seeds or default_seeds)
    n_seeds = cfg["n_seeds"] if cfg.get("seeds") is None else len(cfg["seeds"])
    workload = f"{args.workload}: {description.format(n=n_seeds)}"
    if args.algorithm is not None and args.algorithm != cfg["algorithm"]:
        cfg["algorithm"] = args.algorithm
        workload += f"; algorithm replaced by {args.algorithm}"
    print(f"[bench] workload generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if args.product == "seeds":
        seeds_leg(args, tg, torch, cfg, workload)
        return
    if args.product != "tractogram":
        if cfg.get("mask") is None:
            raise SystemExit(f"--product {args.product} needs a workload with a mask domain")
        product_leg(args, tg, torch, cfg, workload)
        return
    if cfg.get("chunks") is not None:
        run_graph_workload(args, cfg, workload)
        return

    if args.calibrate_copy > 0:
        calibration_copy(torch, args.calibrate_copy)
    runtime = tg.HipRuntime()
    tuned = None
    if args.autotune:
        # Kernel-only timing of the workload's seeds (up to 2M) for each
        # launch shape, grid and repack threshold (not cached: the bench
        # measures what it tunes). A sample smaller than a few grid fills
        # (about 262k lanes resident) would time a launch that never drains a
        # queue, and the grid and repack choices would not carry over.
        tuner = tg.HipRuntime(path="fast")
        sample = cfg["seeds"][:2_000_000]
        tuned = tg.autotune.tune_fast(make_algorithm(tg, cfg), make_domain(tg, cfg),
                                      tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4))), sample,
                                      runtime=tuner, save_at=cfg["save_at"],
                                      maximum_streamline_length=cfg["max_length"], cache=False)
        args.fast_grid = tuned["blocks"]
        args.repack = tuned["repack"]
        args.fast_shape = tuned["shape"]
        print(f"[bench] autotuned: {tuned}", file=sys.stderr, flush=True)
    runtime.set_fast_grid(args.fast_grid)
    runtime.set_fast_shape(args.fast_shape)
    if args.repack:
        runtime.set_fast_repack(args.repack)
    run = DeviceRun(tg, torch, runtime, cfg)
    paths = [p for p in args.paths.split(",") if p]
    results = {}
    gpu_results = {}
    for name in paths:
        path = _native.PATH_FAST if name == "fast" else _native.PATH_PLAIN
        try:
            wall, kernel, total = run.timed(path, args.warmup, args.steps)
        except NotImplementedError as e:
            print(f"[bench] {name} path unavailable: {e}", file=sys.stderr, flush=True)
            continue
        results[name] = {"ms_per_step": wall, "kernel_ms": kernel, "steps": total}
        if path == _native.PATH_FAST:
            results[name]["reads"] = runtime.fast_reads()
            results[name]["saved_points"] = int(run.d_lengths.to(torch.int64).sum().item())
        gpu_results[name] = run.results(min(400, run.n))
        print(f"[bench] {name}: {wall:.1f} ms/step, kernel {kernel:.1f} ms, {total} orientation updates, "
              f"{total / wall * 1e3:.3e} steps/s", file=sys.stderr, flush=True)
    headline = "fast" if "fast" in results else "plain"
    h = results[headline]
    K = cfg["fod"].shape[-1]
    bps = bytes_per_step(K, cfg["step_size"], cfg["save_at"], mask=cfg.get("mesh") is None)
    if "reads" in h:
        # The bytes the cell cache has to move: a row of K f32 coefficients
        # and a mask byte each time a lane's cell changes (counted in the
        # kernel), 12 B per saved point, the seed rows in and the per-seed
        # counts out. Every byte is needed by the kernel at least once, so
        # this is a lower bound of its memory traffic.
        unique = unique_bytes(K, h["reads"], h["saved_points"], len(cfg["seeds"]))
    else:
        unique = bps * h["steps"]
    achieved = unique / (h["kernel_ms"] * 1e-3) / 1e9
    line = {
        "metric": "streamline steps/s on 1 MI355X, mixed vs FP32 path, at <=1e-3 voxel max deviation",
        "value": h["steps"] / (h["ms_per_step"] * 1e-3),
        "unit": "streamline steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy generators)",
        "config": {"workload": workload, "path": headline, "seeds": len(cfg["seeds"]),
                   "orientation_updates_per_step": h["steps"],
                   "fast_launch": {"shape": tg.runtime.fast_shapes()[args.fast_shape],
                                   "grid_blocks": args.fast_grid or "occupancy",
                                   "repack": args.repack or "default",
                                   "autotuned": tuned is not None}},
        "paths": {k: {"steps_per_s": v["steps"] / (v["ms_per_step"] * 1e-3), "ms_per_step": v["ms_per_step"],
                      "kernel_ms": v["kernel_ms"]} for k, v in results.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic and headline == "fast" else None,
                     "kernel": f"track_{headline}", "algorithmic_bytes_per_launch": unique,
                     "basis": ("rows and mask values read on cell changes (kernel counters) x (4K, 1) B + 12 B "
                               "per saved point + 32 B per seed in + 12 B per seed out" if "reads" in h else
                               "survey bytes per step x steps"),
                     "survey_bytes_per_step": bps,
                     "survey_model_gbs": bps * h["steps"] / (h["kernel_ms"] * 1e-3) / 1e9},
    }
    if "reads" in h:
        line["roofline"]["reads"] = h["reads"]
        line["roofline"]["steps_per_row_read"] = h["steps"] / max(1, h["reads"]["rows"])
    if traffic and headline == "fast":
        line["roofline"]["traffic_detail"] = {k: v for k, v in traffic.items()
                                              if k not in ("valu_instructions", "valu_dual_quad_cycles", "l2")}
        if traffic.get("l2"):
            l2 = traffic["l2"]
            line["roofline"]["l2"] = {**l2, "hit_rate": l2["hits"] / max(1.0, l2["hits"] + l2["misses"]),
                                      "traffic_gbs": traffic["bytes_per_launch"] / (h["kernel_ms"] * 1e-3) / 1e9}
        # VALU occupancy: the quad-cycles the kernel's VALU instructions took
        # (instructions less the quad-cycles in which two issued together)
        # per second of the kernel, against all SIMDs' quad-cycles.
        seconds = h["kernel_ms"] * 1e-3
        instructions = traffic["valu_instructions"]
        dual = traffic.get("valu_dual_quad_cycles") or 0.0
        busy = (instructions - dual) / seconds
        line["roofline"]["issue"] = {
            "bound": "valu", "achieved": busy, "peak": MI355X_VALU_ISSUE_PEAK,
            "unit": "VALU-busy SIMD quad-cycles/s", "frac": busy / MI355X_VALU_ISSUE_PEAK,
            "basis": "SQ_INSTS_VALU less SQ_ACTIVE_INST_VALU2 (quad-cycles that issued two VALU instructions) "
                     "over the SIMDs' quad-cycles at 2.4 GHz",
            "wave_instructions_per_64_steps": instructions * 64.0 / h["steps"],
            "dual_issued_fraction": 2.0 * dual / max(1.0, instructions),
            "single_issue_frac": instructions / seconds / MI355X_VALU_ISSUE_PEAK}
    if "fast" in results and "plain" in results:
        line["speedup_fast_over_plain"] = (results["fast"]["steps"] / results["fast"]["ms_per_step"]) / (
            results["plain"]["steps"] / results["plain"]["ms_per_step"])
    if args.api:
        line["api_end_to_end"] = api_leg(tg, cfg, headline)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], line["accuracy_vs_fp64"] = host_leg(cfg, gpu_results)
    print(json.dumps(line), flush=True)


def api_leg(tg, cfg, path: str):
    """tg.tractogram on host arrays (seeds in, ragged streamlines out, so
    PCIe transfers and the host-side split are included), on a runtime
    warmed by one call on other objects. `cold_seconds`: the first call on
    fresh field and domain objects, which uploads the volumes (and, on the
    fast path, transforms the field); `seconds`: the same call again, with
    the volumes resident."""
    runtime = tg.HipRuntime(path=path)
    algorithm = make_algorithm(tg, cfg)
    kwargs = dict(save_at=cfg["save_at"], maximum_streamline_length=cfg["max_length"], runtime=runtime)
    tg.tractogram(tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4))), cfg["seeds"][:1000], make_domain(tg, cfg),
                  algorithm, **kwargs)
    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = make_domain(tg, cfg)
    if cfg.get("mesh") is not None:
        domain._ensure_boundary_data()  # host-side mesh tables are not part of the call
    t0 = time.perf_counter()
    result = tg.tractogram(field, cfg["seeds"], domain, algorithm, **kwargs)
    cold = time.perf_counter() - t0
    del result
    t0 = time.perf_counter()
    result = tg.tractogram(field, cfg["seeds"], domain, algorithm, **kwargs)
    seconds = time.perf_counter() - t0
    points = int(np.sum(result.streamlines.lengths))
    print(f"[bench] tg.tractogram on host arrays: {seconds:.2f} s ({cold:.2f} s with the uploads), "
          f"{len(result.streamlines)} streamlines, {points} points", file=sys.stderr, flush=True)
    return {"seconds": seconds, "cold_seconds": cold, "streamlines": len(result.streamlines), "points": points,
            "path": path}


TRAFFIC_CALIBRATION_BYTES = 1 << 30


def _counter_values(csv_path: Path, counter: str) -> dict:
    """Kernel name -> list of values of `counter`, one per dispatch."""
    import csv

    values = {}
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                values.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return values


def traffic_leg(args):
    """HBM bytes per launch of the fast tracking kernel from two rocprofv3
    counter runs of this script (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the counter reads for an elementwise copy
    that reads and writes 1 GiB in the same run, and the kernel's wavefront
    VALU instructions per launch and the quad-cycles that issued two of them
    from a third run (SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU2). Returns None if
    the profiler is unavailable or a run fails."""
    import shutil
    import subprocess
    import tempfile

    profiler = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not Path(profiler).exists():
        return None
    measured = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU SQ_ACTIVE_INST_VALU2", "TCC_HIT_sum TCC_MISS_sum"):
            out = Path(tmp) / counter.split()[0]
            cmd = [profiler, "--pmc", *counter.split(), "--output-format", "csv", "-d", str(out), "-o", "run", "--",
                   sys.executable, str(ROOT / "bench.py"), "--workload", args.workload, "--paths", "fast",
                   "--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--no-traffic", "--no-api",
                   "--calibrate-copy", str(TRAFFIC_CALIBRATION_BYTES)]
            if args.seeds:
                cmd += ["--seeds", str(args.seeds)]
            if args.fast_grid:
                cmd += ["--fast-grid", str(args.fast_grid)]
            if args.fast_shape:
                cmd += ["--fast-shape", str(args.fast_shape)]
            if args.repack:
                cmd += ["--repack", str(args.repack)]
            print(f"[bench] counter run {counter}", file=sys.stderr, flush=True)
            try:
                subprocess.run(cmd, cwd="/tmp", stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=240, check=True, env={**os.environ, "TMPDIR": "/tmp"})
            except (subprocess.SubprocessError, OSError) as e:
                print(f"[bench] counter run {counter} failed: {e}", file=sys.stderr, flush=True)
                return None
            files = sorted(out.rglob("*counter_collection.csv"))
            if not files:
                return None
            if counter.startswith("TCC_HIT"):
                # L2 hits and misses of the kernel's requests (gathers of
                # coefficient rows, mask values, point stores).
                hits = [v for k, v in _counter_values(files[0], "TCC_HIT_sum").items() if "track_fast_kernel" in k]
                misses = [v for k, v in _counter_values(files[0], "TCC_MISS_sum").items() if "track_fast_kernel" in k]
                if hits and misses:
                    measured["l2"] = {"hits": float(np.sum(hits[0])), "misses": float(np.sum(misses[0]))}
                continue
            # One tracking launch is several dispatches of the kernel (its
            # generations): per-launch figures are their sums.
            if counter.startswith("SQ_INSTS_VALU"):
                instructions = [v for k, v in _counter_values(files[0], "SQ_INSTS_VALU").items()
                                if "track_fast_kernel" in k]
                dual = [v for k, v in _counter_values(files[0], "SQ_ACTIVE_INST_VALU2").items()
                        if "track_fast_kernel" in k]
                if not instructions:
                    return None
                measured["SQ_INSTS_VALU"] = float(np.sum(instructions[0]))
                measured["SQ_ACTIVE_INST_VALU2"] = float(np.sum(dual[0])) if dual else None
                continue
            values = _counter_values(files[0], counter)
            kernel = [v for k, v in values.items() if "track_fast_kernel" in k]
            copy = [max(v) for k, v in values.items() if "CUDAFunctorOnSelf_add" in k]
            if not kernel or not copy or copy[0] <= 0:
                return None
            scale = TRAFFIC_CALIBRATION_BYTES / (copy[0] * 1024.0)
            measured[counter] = {"kib_per_launch": float(np.sum(kernel[0])), "dispatches": len(kernel[0]),
                                 "scale": scale, "bytes": float(np.sum(kernel[0])) * 1024.0 * scale}
    return {"bytes_per_launch": measured["FETCH_SIZE"]["bytes"] + measured["WRITE_SIZE"]["bytes"],
            "fetch_bytes": measured["FETCH_SIZE"]["bytes"], "write_bytes": measured["WRITE_SIZE"]["bytes"],
            "fetch_scale": measured["FETCH_SIZE"]["scale"], "write_scale": measured["WRITE_SIZE"]["scale"],
            "valu_instructions": measured["SQ_INSTS_VALU"], "valu_dual_quad_cycles": measured["SQ_ACTIVE_INST_VALU2"],
            "l2": measured.get("l2")}


def calibration_copy(torch, n_bytes: int):
    """out = src + 0 over n_bytes of float32: a kernel that reads and writes
    exactly n_bytes each, to scale hardware byte counters."""
    n = n_bytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.add(src, 0.0, out=dst)
    torch.cuda.synchronize()
    print(f"[bench] calibration copy: {n * 4} bytes read and written", file=sys.stderr, flush=True)
    del src, dst


def run_graph_workload(args, cfg, workload):
    import torch

    import tractography_amd as tg

    runtime = tg.HipRuntime()
    run = GraphRun(tg, torch, runtime, cfg)
    ms, kernel, updates = run.timed(args.warmup, args.steps)
    run.close()
    print(f"[bench] fast (graph): {ms:.1f} ms/replay, track kernels {kernel:.1f} ms, {updates} orientation updates, "
          f"{updates / ms * 1e3:.3e} steps/s", file=sys.stderr, flush=True)
    K = cfg["fod"].shape[-1]
    bps = bytes_per_step(K, cfg["step_size"], cfg["save_at"])
    reads = {"rows": sum(r["rows"] for r in run.reads), "mask_values": sum(r["mask_values"] for r in run.reads)}
    unique = unique_bytes(K, reads, run.saved_points, run.n)
    achieved = unique / (kernel * 1e-3) / 1e9
    line = {
        "metric": "streamline steps/s on 1 MI355X, mixed vs FP32 path, at <=1e-3 voxel max deviation",
        "value": updates / (ms * 1e-3),
        "unit": "streamline steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy generators; seeds drawn on the device)",
        "config": {"workload": workload, "path": "fast", "seeds": run.n, "chunks": run.chunks,
                   "orientation_updates_per_step": updates},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_PEAK_GBS, "traffic": None, "kernel": "track_fast",
                     "algorithmic_bytes_per_launch": unique, "reads": reads,
                     "basis": "rows and mask values read on cell changes (kernel counters, summed over the "
                              "chunks) x (4K, 1) B + 12 B per saved point + 44 B per seed",
                     "survey_bytes_per_step": bps,
                     "survey_model_gbs": bps * updates / (kernel * 1e-3) / 1e9},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
