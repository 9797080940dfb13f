"""Writes rng_seed_draws.json: the first draws of
numpy.random.default_rng(seed).integers(0, 2**32 - 1, dtype=uint32), the
per-call RNG seed of the reference's Diffusion (algorithm.py:182, 239-241).
Recorded with numpy 2.2.6."""

import json
from pathlib import Path

import numpy as np

out = {}
for seed in (0, 1234):
    rng = np.random.default_rng(seed)
    out[str(seed)] = [int(rng.integers(0, np.iinfo(np.uint32).max, dtype=np.uint32)) for _ in range(8)]
Path(__file__).with_name("rng_seed_draws.json").write_text(json.dumps(out, indent=1) + "\n")
