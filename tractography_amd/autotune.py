"""Launch-shape tuning of the fast tracking path.

The persistent fast kernel (one lane per streamline in flight, lanes refilled
from a device queue) needs a grid: enough blocks to fill every compute unit,
but more blocks than can be resident only add scheduling. ``tune_fast_grid``
times the real workload on a sample of its seeds for a few grid sizes around
the occupancy-derived one and keeps the fastest on the runtime. It has no
counterpart in the reference, whose OpenCL kernel runs one work-item per seed.
"""

from __future__ import annotations

import time

import numpy as np

from . import tracking
from .runtime import HipRuntime

DEFAULT_FACTORS = (0.5, 1.0, 1.5, 2.0, 4.0)


def tune_fast_grid(
    algorithm,
    domain,
    field,
    seeds,
    *,
    runtime: HipRuntime,
    save_at: float,
    maximum_streamline_length: float,
    factors=DEFAULT_FACTORS,
    repeats: int = 2,
) -> dict:
    """Times `seeds` (a sample of the workload, e.g. 100k rows of (N, 8)) on
    the fast path for grids of factor x (compute units x 3 blocks) and sets
    the fastest on `runtime`. Returns {"blocks": best, "timings": {blocks: s}}
    with blocks 0 meaning the occupancy-sized default."""
    if runtime.path != "fast":
        raise ValueError("tune_fast_grid needs a runtime created with path='fast'")
    seeds = np.ascontiguousarray(seeds, dtype=np.float32)
    rng_seeds = [algorithm.get_rng_seed()]
    units = int(runtime.info()["compute_units"])
    candidates = [0] + sorted({max(1, int(round(f * units * 3))) for f in factors})
    timings = {}
    for blocks in candidates:
        runtime.set_fast_grid(blocks)
        best = float("inf")
        for _ in range(max(1, repeats)):
            runtime.finish()
            t0 = time.perf_counter()
            tracking.track(algorithm, domain, field, seeds, rng_seeds, len(seeds), save_at,
                           maximum_streamline_length, runtime)
            best = min(best, time.perf_counter() - t0)
        timings[blocks] = best
    chosen = min(timings, key=timings.get)
    runtime.set_fast_grid(chosen)
    return {"blocks": chosen, "timings": timings}
