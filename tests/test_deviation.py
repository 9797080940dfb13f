"""The deviation report against known trajectories (CPU)."""

import numpy as np

import deviation


def _straight(n, angle):
    t = 0.25 * np.arange(n)
    return np.stack([t * np.cos(angle), t * np.sin(angle), np.zeros(n)], axis=1)


def test_rotated_line_gives_the_rotation_angle_and_arc_deviation():
    n, angle = 40, 1e-3
    ref = {"lengths": np.array([n]), "steps": np.array([4 * n]), "points": _straight(n, 0.0)[None],
           "margins": np.full((1, n), 1.0)}
    got = _straight(n, angle).astype(np.float32)
    rep = deviation.report([n], [4 * n], got, ref)
    assert abs(rep["max_angular_error_rad"] - angle) < 2e-5
    chord = 2 * 0.25 * (n - 1) * np.sin(angle / 2)
    assert abs(rep["max_deviation"] - chord) < 1e-5
    assert rep["compared_fraction"] == 1.0


def test_prefix_stops_at_the_margin():
    n = 10
    ref_pts = _straight(n, 0.0)
    got = ref_pts.copy()
    got[6:, 1] += 0.5  # leaves the shared voxels at point 6
    ref = {"lengths": np.array([n]), "steps": np.array([4 * n]), "points": ref_pts[None],
           "margins": np.full((1, n), 0.1)}
    rep = deviation.report([n], [4 * n], got.astype(np.float32), ref)
    assert rep["compared_fraction"] == 0.6
    assert rep["max_deviation"] == 0.0
    assert rep["max_angular_error_rad"] < 1e-6
