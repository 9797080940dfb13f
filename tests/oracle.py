"""ctypes binding of the host restatement (oracle/liboracle.so).

Test infrastructure: the checker the HIP kernels are compared against.
"""

from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_float, c_int, c_int32, c_int64, c_uint32, c_void_p
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent.parent / "oracle"
LIBRARY = ORACLE_DIR / "liboracle.so"


class Algorithm(ctypes.Structure):
    _fields_ = [
        ("model", c_int32),
        ("algorithm", c_int32),
        ("n_parameters", c_int32),
        ("step_size", c_float),
        ("gamma", c_float),
        ("noise_variance", c_float),
        ("n_samples", c_int32),
        ("max_angle", c_float),
        ("max_iterations", c_int32),
        ("tolerance", c_float),
        ("max_step_size", c_float),
    ]


class Scene(ctypes.Structure):
    _fields_ = [
        ("domain_kind", c_int32),
        ("mask", c_void_p),
        ("mask_dims", c_uint32 * 3),
        ("mask_inverse_affine", c_float * 16),
        ("field", c_void_p),
        ("field_dims", c_uint32 * 4),
        ("field_inverse_affine", c_float * 16),
        ("vertices", c_void_p),
        ("triangles", c_void_p),
        ("bvh_min", c_void_p),
        ("bvh_max", c_void_p),
        ("bvh_nodes", c_void_p),
        ("bvh_order", c_void_p),
        ("n_bvh_nodes", c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            subprocess.run(["make", "-C", str(ORACLE_DIR), "-s"], check=True)
        _lib = ctypes.CDLL(str(LIBRARY))
    return _lib


def _p(a):
    return None if a is None else c_void_p(a.ctypes.data)


def algorithm_struct(algorithm) -> Algorithm:
    d = algorithm.descriptor()
    return Algorithm(d.model, d.algorithm, d.n_parameters, d.step_size, d.gamma, d.noise_variance, d.n_samples,
                     d.max_angle, d.max_iterations, d.tolerance, d.max_step_size)


class SceneHolder:
    """A tgo_scene over a domain and field, keeping its arrays alive."""

    def __init__(self, domain, field, boundary=False):
        from tractography_amd import domain as domain_module

        self.keep = []
        s = Scene()
        if boundary and not isinstance(domain, domain_module.Mesh):
            # A mask domain whose boundary surface and BVH are also needed.
            data = domain._ensure_boundary_data()
            for name in ("vertices", "indices", "bvh_min", "bvh_max", "bvh_nodes", "bvh_triangles"):
                self.keep.append(data[name])
            s.vertices = data["vertices"].ctypes.data
            s.triangles = data["indices"].ctypes.data
            s.bvh_min = data["bvh_min"].ctypes.data
            s.bvh_max = data["bvh_max"].ctypes.data
            s.bvh_nodes = data["bvh_nodes"].ctypes.data
            s.bvh_order = data["bvh_triangles"].ctypes.data
            s.n_bvh_nodes = len(data["bvh_nodes"])
        if isinstance(domain, domain_module.Mesh):
            s.domain_kind = 1
            data = domain._ensure_boundary_data()
            for name in ("vertices", "indices", "bvh_min", "bvh_max", "bvh_nodes", "bvh_triangles"):
                self.keep.append(data[name])
            s.vertices = data["vertices"].ctypes.data
            s.triangles = data["indices"].ctypes.data
            s.bvh_min = data["bvh_min"].ctypes.data
            s.bvh_max = data["bvh_max"].ctypes.data
            s.bvh_nodes = data["bvh_nodes"].ctypes.data
            s.bvh_order = data["bvh_triangles"].ctypes.data
            s.n_bvh_nodes = len(data["bvh_nodes"])
        elif domain is not None:
            mask = np.ascontiguousarray(domain.values != 0, dtype=np.uint8)
            self.keep.append(mask)
            s.domain_kind = 0
            s.mask = mask.ctypes.data
            for i in range(3):
                s.mask_dims[i] = mask.shape[i]
            inv = np.asarray(domain.inverse_affine, dtype=np.float32).ravel()
            for i in range(16):
                s.mask_inverse_affine[i] = inv[i]
        if field is not None:
            values = np.ascontiguousarray(field.values, dtype=np.float32)
            self.keep.append(values)
            s.field = values.ctypes.data
            for i in range(4):
                s.field_dims[i] = values.shape[i]
            inv = np.asarray(field.inverse_affine, dtype=np.float32).ravel()
            for i in range(16):
                s.field_inverse_affine[i] = inv[i]
        self.scene = s


def seed_streams(n_seeds: int, rng_seeds, batch_size: int, indices=None):
    """(rng seed, stream id) per seed for global indices (reference batching)."""
    idx = np.arange(n_seeds, dtype=np.int64) if indices is None else np.asarray(indices, dtype=np.int64)
    rng_seeds = np.asarray(rng_seeds, dtype=np.uint64).astype(np.uint32)
    return (
        np.ascontiguousarray(rng_seeds[idx // batch_size], dtype=np.uint32),
        np.ascontiguousarray(idx % batch_size, dtype=np.uint32),
    )


def track(domain, field, algorithm, seeds, rng_seeds, batch_size, save_at, max_length, *, precision=32,
          indices=None, threads=8):
    """Run the oracle on `seeds` (rows of the global seed array at `indices`).

    Returns dict(lengths, steps, hits, points (n, n_points, 3) dense[, margins]).
    """
    seeds = np.ascontiguousarray(seeds, dtype=np.float32)
    n = len(seeds)
    n_points = int(np.floor(max_length / save_at))
    seed_rng, seed_stream = seed_streams(n, rng_seeds, batch_size, indices)
    holder = SceneHolder(domain, field)
    alg = algorithm_struct(algorithm)
    lengths = np.zeros(n, np.uint32)
    steps = np.zeros(n, np.uint32)
    hits = np.zeros(n, np.uint32)
    save = float(np.float32(save_at))
    if precision == 32:
        points = np.zeros((n, n_points, 3), np.float32)
        lib().tgo32_track(ctypes.byref(holder.scene), ctypes.byref(alg), _p(seeds), c_int64(n), _p(seed_rng),
                          _p(seed_stream), c_uint32(n_points), c_float(save), _p(lengths), _p(steps), _p(hits),
                          _p(points), c_int(threads))
        return {"lengths": lengths, "steps": steps, "hits": hits, "points": points}
    points = np.zeros((n, n_points, 3), np.float64)
    margins = np.zeros((n, n_points), np.float64)
    lib().tgo64_track(ctypes.byref(holder.scene), ctypes.byref(alg), _p(seeds), c_int64(n), _p(seed_rng),
                      _p(seed_stream), c_uint32(n_points), c_float(save), _p(lengths), _p(steps), _p(hits),
                      _p(points), _p(margins), c_int(threads))
    return {"lengths": lengths, "steps": steps, "hits": hits, "points": points, "margins": margins}


def connectome(domain, field, algorithm, rng_seed, n_seeds, n_steps, min_length=0.0, *, precision=32,
               threads=8):
    """rows, cols, weights (9 per seed) of the connectome kernel."""
    holder = SceneHolder(domain, field, boundary=True)
    data = domain._ensure_boundary_data()
    alg = algorithm_struct(algorithm)
    rows = np.zeros(9 * n_seeds, np.int32)
    cols = np.zeros(9 * n_seeds, np.int32)
    weights = np.zeros(9 * n_seeds, np.float32)
    fn = lib().tgo32_connectome if precision == 32 else lib().tgo64_connectome
    fn(ctypes.byref(holder.scene), ctypes.byref(alg), _p(data["vertices"]), _p(data["indices"]),
       _p(data["cumulative"]), c_uint32(len(data["indices"])), c_uint32(rng_seed), c_int64(n_seeds),
       c_uint32(n_steps), c_float(min_length), _p(rows), _p(cols), _p(weights), c_int(threads))
    return rows, cols, weights


def tod(domain, field, algorithm, rng_seed, n_seeds, n_steps, out_dims, out_inverse_affine, *, precision=32):
    """Raw TOD sums (X, Y, Z, 45) in seed order, with the per-cell sum of
    |flushed values| (X, Y, Z, 45) and flush count (X, Y, Z)."""
    holder = SceneHolder(domain, field)
    data = domain._ensure_boundary_data()
    alg = algorithm_struct(algorithm)
    dims = np.ascontiguousarray(out_dims, dtype=np.uint32)
    inv = np.ascontiguousarray(out_inverse_affine, dtype=np.float32).reshape(16)
    shape = tuple(int(d) for d in dims)
    raw = np.zeros(shape + (45,), np.float32 if precision == 32 else np.float64)
    abs_sum = np.zeros(shape + (45,), np.float64)
    flushes = np.zeros(shape, np.uint32)
    fn = lib().tgo32_tod if precision == 32 else lib().tgo64_tod
    fn(ctypes.byref(holder.scene), ctypes.byref(alg), _p(data["vertices"]), _p(data["indices"]),
       _p(data["cumulative"]), c_uint32(len(data["indices"])), c_uint32(rng_seed), c_int64(n_seeds),
       c_uint32(n_steps), _p(dims), _p(inv), _p(raw), _p(abs_sum), _p(flushes))
    return raw, abs_sum, flushes


def normalize_odf(odf):
    """utils.py:6-22, in numpy as the reference writes it."""
    normalized = np.zeros_like(odf)
    np.divide(odf / np.sqrt(4 * np.pi), odf[..., :1], out=normalized, where=odf[..., :1] != 0)
    return normalized


def ragged(result):
    """Dense oracle points -> (total, 3) ragged array in seed order."""
    pts = result["points"]
    lengths = result["lengths"].astype(np.int64)
    mask = np.arange(pts.shape[1])[None, :] < lengths[:, None]
    return pts[mask]


def math32(fn: int, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = x if y is None else np.ascontiguousarray(y, dtype=np.float32)
    out = np.empty_like(x)
    lib().tgo_math32(c_int(fn), _p(x), _p(y), _p(out), c_int64(len(x)))
    return out


def rng_init(seed: int, stream: int) -> np.ndarray:
    state = np.zeros(2, np.uint32)
    lib().tgo_rng_init(c_uint32(seed), c_uint32(stream), _p(state))
    return state


def rng_draws(state, kind: int, n: int, aux: int = 0):
    state = np.ascontiguousarray(state, dtype=np.uint32)
    dtype = {0: np.uint32, 1: np.float32, 2: np.float32, 3: np.float64, 4: np.uint32}[kind]
    out = np.zeros(n, dtype)
    lib().tgo_rng_draws(_p(state), c_int(kind), c_uint32(aux), c_int64(n), _p(out))
    return out, state


def sh45(phi, theta, precision=32):
    dtype = np.float32 if precision == 32 else np.float64
    phi = np.ascontiguousarray(phi, dtype=dtype)
    theta = np.ascontiguousarray(theta, dtype=dtype)
    n = len(phi)
    ylm, dp, dt = (np.zeros((n, 45), dtype) for _ in range(3))
    fn = lib().tgo32_sh45 if precision == 32 else lib().tgo64_sh45
    fn(_p(phi), _p(theta), c_int64(n), _p(ylm), _p(dp), _p(dt))
    return ylm, dp, dt


def sh_basis64(lmax: int, phi, theta):
    """FP64 real SH of even order lmax with d/dtheta and (d/dphi)/sin(theta)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    K = (lmax + 1) * (lmax + 2) // 2
    ylm, dp, dt = (np.zeros((len(phi), K)) for _ in range(3))
    lib().tgo64_sh_basis(c_int(lmax), _p(phi), _p(theta), c_int64(len(phi)), _p(ylm), _p(dp), _p(dt))
    return ylm, dp, dt


def cart2sph32(xyz):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    out = np.zeros((len(xyz), 2), np.float32)
    lib().tgo32_cart2sph(_p(xyz), c_int64(len(xyz)), _p(out))
    return out


def evaluate(model_id: int, K: int, params, orientations, precision=32):
    dtype = np.float32 if precision == 32 else np.float64
    params = np.ascontiguousarray(params, dtype=dtype)
    o = np.ascontiguousarray(np.asarray(orientations)[:, :3], dtype=dtype)
    out = np.zeros((len(o), 3), dtype)
    fn = lib().tgo32_evaluate if precision == 32 else lib().tgo64_evaluate
    fn(c_int(model_id), c_int(K), _p(params), _p(o), c_int64(len(o)), _p(out))
    return out


def update(algorithm, rng_seed: int, params, orientations, precision=32):
    dtype = np.float32 if precision == 32 else np.float64
    params = np.ascontiguousarray(params, dtype=dtype)
    o = np.ascontiguousarray(np.asarray(orientations)[:, :3], dtype=dtype)
    out = np.zeros((len(o), 3), dtype)
    alg = algorithm_struct(algorithm)
    fn = lib().tgo32_update if precision == 32 else lib().tgo64_update
    fn(ctypes.byref(alg), c_uint32(rng_seed), _p(params), _p(o), c_int64(len(o)), _p(out))
    return out


def update_state(algorithm, state, params, orientation):
    """One FP32 update from an explicit noise state: (orientation, state)."""
    params = np.ascontiguousarray(params, dtype=np.float32).ravel()
    o = np.ascontiguousarray(np.asarray(orientation, dtype=np.float32).ravel()[:3])
    state = np.array(state, dtype=np.uint32)
    out = np.zeros(3, np.float32)
    alg = algorithm_struct(algorithm)
    lib().tgo32_update_state(ctypes.byref(alg), _p(params), _p(o), _p(state), _p(out))
    return out, state


def field_evaluate(field, points):
    holder = SceneHolder(None, field)
    pts = np.ascontiguousarray(np.asarray(points)[:, :3], dtype=np.float32)
    out = np.zeros((len(pts), field.n_parameters), np.float32)
    lib().tgo_field_evaluate(ctypes.byref(holder.scene), _p(pts), c_int64(len(pts)), _p(out))
    return out


def domain_contains(domain, points):
    holder = SceneHolder(domain, None)
    pts = np.ascontiguousarray(np.asarray(points)[:, :3], dtype=np.float32)
    out = np.zeros(len(pts), np.uint8)
    lib().tgo_domain_contains(ctypes.byref(holder.scene), _p(pts), c_int64(len(pts)), _p(out))
    return out.astype(bool)


def bvh_hit(domain, starts, ends):
    """Closest hit per segment; `domain` is any domain (its boundary BVH)."""
    data = domain._ensure_boundary_data()
    s = Scene()
    s.domain_kind = 1
    s.vertices = data["vertices"].ctypes.data
    s.triangles = data["indices"].ctypes.data
    s.bvh_min = data["bvh_min"].ctypes.data
    s.bvh_max = data["bvh_max"].ctypes.data
    s.bvh_nodes = data["bvh_nodes"].ctypes.data
    s.bvh_order = data["bvh_triangles"].ctypes.data
    s.n_bvh_nodes = len(data["bvh_nodes"])
    starts = np.ascontiguousarray(np.asarray(starts)[:, :3], dtype=np.float32)
    ends = np.ascontiguousarray(np.asarray(ends)[:, :3], dtype=np.float32)
    n = len(starts)
    found = np.zeros(n, np.uint8)
    hit = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    lib().tgo_bvh_segment_hit(ctypes.byref(s), _p(starts), _p(ends), c_int64(n), _p(found), _p(hit), _p(t))
    return found.astype(bool), hit, t


def domain_seed(domain, rng_seed: int, n: int):
    data = domain._ensure_boundary_data()
    out = np.zeros((n, 8), np.float32)
    lib().tgo_domain_seed(_p(data["vertices"]), _p(data["indices"]), _p(data["cumulative"]),
                          c_uint32(len(data["indices"])), c_uint32(rng_seed), c_int64(n), _p(out))
    return out


def seeds_from_fod(model_id: int, K: int, fod, voxels, affine, states, threads=8):
    fod = np.ascontiguousarray(fod, dtype=np.float32)
    voxels = np.ascontiguousarray(voxels, dtype=np.float32)
    affine = np.ascontiguousarray(affine, dtype=np.float32)
    states = np.array(states, dtype=np.uint32, copy=True, order="C")
    out = np.zeros((len(states), 8), np.float32)
    lib().tgo_seeds_from_fod(c_int(model_id), c_int(K), _p(fod), _p(voxels), c_uint32(len(voxels)), _p(affine),
                             _p(states), c_int64(len(states)), _p(out), c_int(threads))
    return out, states
