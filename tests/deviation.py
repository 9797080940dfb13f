"""Trajectory deviation against the FP64 host restatement.

Nearest-voxel lookups make trajectories that differ by a rounding error
diverge by O(step) once one of them crosses a voxel face the other does not.
Saved points are therefore compared only while the running maximum of the
point deviation stays below the FP64 trajectory's running minimum distance
to a voxel face (the prefix over which both provably used the same voxels);
the report gives the maximum deviation over those prefixes, the largest
angle between corresponding saved segments of the two paths there, and how
much of the streamlines they cover.
"""

from __future__ import annotations

import numpy as np


def report(lengths, steps, points, ref64, jump: float | None = None) -> dict:
    """lengths/points: test path (ragged points in seed order); ref64: FP64
    oracle result (dense points, margins). With `jump`, a prefix also ends
    before the first point that deviates by more than `jump`: for sampling
    algorithms, where the two arithmetics may choose different candidates
    when a draw falls within rounding of a decision boundary."""
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(lengths)])
    max_dev = 0.0
    max_angle = 0.0
    compared = 0
    total = 0
    full = 0
    for i in range(len(lengths)):
        n_ref = int(ref64["lengths"][i])
        n = min(int(lengths[i]), n_ref)
        total += n_ref
        got = points[offsets[i] : offsets[i] + n].astype(np.float64)
        ref = ref64["points"][i, :n]
        dev = np.linalg.norm(got - ref, axis=1)
        running = np.maximum.accumulate(dev)
        ok = running < ref64["margins"][i, :n]
        if jump is not None:
            ok &= np.maximum.accumulate(dev > jump) == 0
        k = n if np.all(ok) else int(np.argmin(ok))
        if k > 0:
            max_dev = max(max_dev, float(running[k - 1]))
        if k > 1:
            # Angle between the saved segments of both paths over the prefix.
            a = np.diff(got[:k], axis=0)
            b = np.diff(ref[:k], axis=0)
            norms = np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1)
            live = norms > 0
            if np.any(live):
                cos = np.einsum("ij,ij->i", a[live], b[live]) / norms[live]
                max_angle = max(max_angle, float(np.max(np.arccos(np.clip(cos, -1.0, 1.0)))))
        compared += k
        if k == n and int(lengths[i]) == n_ref:
            full += 1
    return {
        "max_deviation": max_dev,
        "max_angular_error_rad": max_angle,
        "compared_fraction": compared / max(total, 1),
        "identical_length_and_voxels_fraction": full / max(len(lengths), 1),
        "length_equal_fraction": float(np.mean(lengths == ref64["lengths"])),
        "steps_equal_fraction": float(np.mean(np.asarray(steps) == ref64["steps"])),
    }
