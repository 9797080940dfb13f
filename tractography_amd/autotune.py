"""Launch tuning of the fast tracking path.

The persistent fast kernel (one lane per streamline in flight, lanes refilled
from a device queue, run as generations that repack the streamlines still in
flight) has three launch settings that change its speed and never its
results:

* the launch shape of the order-8 SH Transport and Diffusion kernels: lanes
  per block (streamlines per workgroup: 128, 256 or 512) and how many of a
  row's 45 coefficients wait in LDS between steps (32, 40 or 44; the rest in
  registers), compiled as variants (``runtime.fast_shapes()``);
* the grid: enough blocks to fill every compute unit once, or a multiple;
* the repack threshold: how many streamlines a generation must still hold
  for it to be cut short and its streamlines resumed in fuller wavefronts.

``tune_fast`` times the real workload on a sample of its seeds, kernel only
(``tgk_track_timed``: HIP events around the tracking launches, seeds uploaded
beforehand, noise-stream seed 0, so the caller's algorithm draws nothing):
first every shape at the default grid and threshold, then every grid and
threshold with the fastest shape. It sets the fastest on the runtime and
keeps it in a JSON file keyed by device, model, coefficient count, algorithm
and domain kind, so later calls with the same key reuse it without timing.
None of this has a counterpart in the reference, whose OpenCL kernel runs one
work-item per seed (tracking.py:174).
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

import numpy as np

from . import _native
from .domain import Mesh
from .runtime import HipRuntime, fast_shapes
from .tracking import point_capacity

DEFAULT_GRID_FACTORS = (0.5, 1.0, 2.0)
DEFAULT_REPACKS = (4096, 16384, 65536, 1 << 30)
DEFAULT_SHAPES = tuple(range(_native.FAST_SHAPES))


def cache_path() -> Path:
    """Where tuned settings are kept: $TGK_AUTOTUNE_CACHE, else
    ~/.cache/tractography_amd/autotune.json."""
    env = os.environ.get("TGK_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "tractography_amd" / "autotune.json"


def cache_key(device_name: str, algorithm, domain) -> str:
    descriptor = algorithm.descriptor()
    kind = "mesh" if isinstance(domain, Mesh) else "mask"
    return (f"{device_name}|model={descriptor.model}|K={descriptor.n_parameters}|"
            f"algorithm={descriptor.algorithm}|domain={kind}")


def load_cache(path: Path | None = None) -> dict:
    path = cache_path() if path is None else Path(path)
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if isinstance(data, dict) else {}


def store_cache(key: str, entry: dict, path: Path | None = None) -> None:
    """Adds `entry` under `key`, rewriting the file atomically."""
    path = cache_path() if path is None else Path(path)
    data = load_cache(path)
    data[key] = entry
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(runtime: HipRuntime, entry: dict) -> None:
    runtime.set_fast_shape(int(entry.get("shape", 0)))
    runtime.set_fast_grid(int(entry["blocks"]))
    runtime.set_fast_repack(int(entry["repack"]))


def _grids(units: int, lanes_per_block: int, factors) -> list[int]:
    """0 (occupancy) and each factor x the grid of 1024 lanes per CU."""
    per_cu = max(1, 1024 // lanes_per_block)
    return [0] + sorted({max(1, int(round(f * units * per_cu))) for f in factors})


def time_fast(runtime: HipRuntime, algorithm, domain, field, seeds, *, save_at: float,
              maximum_streamline_length: float, repeats: int = 2) -> tuple[float, int]:
    """(milliseconds, orientation updates) of the fastest of `repeats`
    device tracking launches of `seeds` with the runtime's current settings."""
    seeds = np.ascontiguousarray(seeds, dtype=np.float32)
    field_device = field.upload(runtime)
    domain_device = domain.upload(runtime)
    ms = ctypes.c_float()
    steps = ctypes.c_uint64()
    _native.check(_native.lib().tgk_track_timed(
        runtime.handle, domain_device.handle, field_device.handle, algorithm.descriptor(), _native.PATH_FAST,
        _native.ptr(seeds), len(seeds), point_capacity(maximum_streamline_length, save_at),
        float(np.float32(save_at)), int(repeats), ctypes.byref(ms), ctypes.byref(steps)))
    return ms.value, steps.value


def tune_fast(algorithm, domain, field, seeds, *, runtime: HipRuntime, save_at: float,
              maximum_streamline_length: float, grid_factors=DEFAULT_GRID_FACTORS, repacks=DEFAULT_REPACKS,
              shapes=DEFAULT_SHAPES, repeats: int = 2, cache: bool | str | Path = True, force: bool = False) -> dict:
    """Times `seeds` (a sample of the workload, rows of (N, 8), several times
    the lanes the grid keeps resident: about 262k on a 256-CU device, so that
    the queue outlasts the first fill as in the full run) on the fast path: every launch shape at the occupancy-sized grid and the
    default repack threshold, then, with the fastest shape, every grid (0 =
    occupancy-sized, or factor x compute units x 1024 lanes) and repack
    threshold; sets the fastest on `runtime`. With `cache`, a stored entry
    for the same key is applied without timing unless `force`, and a new
    result is stored.

    Returns {"shape", "blocks", "repack", "milliseconds", "steps_per_s",
    "timings": {"shape/blocks/repack": ms}, "cached": bool}."""
    if runtime.path != "fast":
        raise ValueError("tune_fast needs a runtime created with path='fast'")
    path = None if cache is True else (Path(cache) if cache else None)
    info = runtime.info()
    key = cache_key(info["name"], algorithm, domain)
    if cache and not force:
        entry = load_cache(path).get(key)
        if entry is not None:
            apply(runtime, entry)
            return {**entry, "cached": True}
    units = int(info["compute_units"])
    table = {s["shape"]: s for s in fast_shapes()}
    default_repack = 16384
    timings = {}
    best = None

    def run(shape, blocks, repack):
        nonlocal best
        label = f"{shape}/{blocks}/{repack}"
        if label in timings:
            return
        runtime.set_fast_shape(shape)
        runtime.set_fast_grid(blocks)
        runtime.set_fast_repack(repack)
        ms, steps = time_fast(runtime, algorithm, domain, field, seeds, save_at=save_at,
                              maximum_streamline_length=maximum_streamline_length, repeats=repeats)
        timings[label] = ms
        if best is None or ms < best[0]:
            best = (ms, steps, shape, blocks, repack)

    for shape in shapes:
        run(int(shape), 0, default_repack)
    shape = best[2]
    for blocks in _grids(units, table[shape]["lanes_per_block"], grid_factors):
        for repack in repacks:
            run(shape, blocks, repack)
    ms, steps, shape, blocks, repack = best
    entry = {"shape": shape, "blocks": blocks, "repack": repack, "milliseconds": ms,
             "steps_per_s": steps / (ms * 1e-3) if ms > 0 else 0.0, "timings": timings}
    apply(runtime, entry)
    if cache:
        store_cache(key, entry, path)
    return {**entry, "cached": False}


def tune_fast_grid(algorithm, domain, field, seeds, *, runtime: HipRuntime, save_at: float,
                   maximum_streamline_length: float, factors=DEFAULT_GRID_FACTORS, repeats: int = 2) -> dict:
    """The grid alone (repack threshold left as set), kernel-only timing, no
    cache. Returns {"blocks": best, "timings": {blocks: ms}}."""
    if runtime.path != "fast":
        raise ValueError("tune_fast_grid needs a runtime created with path='fast'")
    units = int(runtime.info()["compute_units"])
    shape = getattr(runtime, "fast_shape", 0)
    candidates = _grids(units, fast_shapes()[shape]["lanes_per_block"], factors)
    timings = {}
    for blocks in candidates:
        runtime.set_fast_grid(blocks)
        timings[blocks], _ = time_fast(runtime, algorithm, domain, field, seeds, save_at=save_at,
                                       maximum_streamline_length=maximum_streamline_length, repeats=repeats)
    chosen = min(timings, key=timings.get)
    runtime.set_fast_grid(chosen)
    return {"blocks": chosen, "timings": timings}
