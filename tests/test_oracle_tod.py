"""The TOD restatement (tracking.py:181-385) against closed forms: in an
isotropic field every track is a straight inward ray, so the deposit in each
cell is the number of steps taken there times the SH basis of the ray's
direction (scipy's spherical harmonics), one flush per track and cell."""

import numpy as np
import pytest

import fixtures
import oracle
import tractography_amd as tg


def _isotropic_scene(mask):
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), np.eye(4)))
    fod = np.zeros(mask.shape + (45,), dtype=np.float32)
    fod[..., 0] = 1.0
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    return domain, field


def _straight_tod(mask, seeds, step, n_steps, out_dims, out_inverse):
    """Steps per cell of straight float32 rays (contains, field, deposit,
    move), with the cell of each step from floor(inverse p + 0.5)."""
    inv = np.asarray(out_inverse, dtype=np.float32)
    counts = {}  # (seed, cell) -> steps
    order = []
    for g, s in enumerate(seeds):
        p = s[:3].astype(np.float32)
        o = s[4:7].astype(np.float32)
        for _ in range(n_steps):
            m = np.floor(p + np.float32(0.5))
            if np.any(m < 0) or np.any(m >= mask.shape) or not mask[tuple(m.astype(int))]:
                break
            f = np.rint(p)
            if np.any(f < 0) or np.any(f >= mask.shape):
                break
            v = np.array([((inv[i, 0] * p[0] + inv[i, 1] * p[1]) + inv[i, 2] * p[2]) + inv[i, 3] for i in range(3)],
                         dtype=np.float32)
            c = np.floor(v + np.float32(0.5))
            if np.all(c >= 0) and np.all(c < out_dims):
                key = (g, tuple(c.astype(int)))
                if key not in counts:
                    order.append(key)
                    counts[key] = 0
                counts[key] += 1
            p = p + np.float32(step) * o
    return counts, order


@pytest.mark.parametrize("out_scale,out_shift", [(1.0, 0.0), (0.5, 0.25)])
def test_tod_straight_rays_deposit_steps_times_the_basis(out_scale, out_shift):
    mask = np.zeros((7, 6, 5), dtype=bool)
    mask[1:6, 1:5, 1:4] = True
    domain, field = _isotropic_scene(mask)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=0.5)
    n_seeds, n_steps = 120, 80
    # Output grid: voxel size out_scale, shifted by out_shift (world = s i + t).
    affine = np.eye(4)
    affine[:3, :3] *= out_scale
    affine[:3, 3] = out_shift
    out_dims = np.array([int(np.ceil(7 / out_scale)), int(np.ceil(6 / out_scale)), int(np.ceil(5 / out_scale))])
    inverse = np.linalg.inv(affine).astype(np.float32)
    raw, abs_sum, flushes = oracle.tod(domain, field, algorithm, 17, n_seeds, n_steps, out_dims, inverse)

    seeds = oracle.domain_seed(domain, 17, n_seeds)
    counts, _ = _straight_tod(mask, seeds, 0.1, n_steps, out_dims, inverse)
    expected = np.zeros(tuple(out_dims) + (45,))
    expected_flushes = np.zeros(tuple(out_dims), dtype=np.uint32)
    xyz = seeds[:, 4:7].astype(np.float64)
    azimuth = np.mod(np.arctan2(xyz[:, 1], xyz[:, 0]), 2 * np.pi)
    colatitude = np.arccos(np.clip(xyz[:, 2], -1, 1))
    basis, _ = tg.core.ishtmtx(azimuth, colatitude, 45)
    for (g, cell), k in counts.items():
        expected[cell] += k * basis[g]
        expected_flushes[cell] += 1
    assert expected_flushes.sum() > 100
    np.testing.assert_array_equal(flushes, expected_flushes)
    np.testing.assert_allclose(raw, expected, rtol=1e-5, atol=2e-5 * np.abs(expected).max())
    # Y_0 is constant and positive: the l = 0 sums count steps exactly.
    steps = np.rint(raw[..., 0] / np.float32(0.28209479177387814)).astype(np.int64)
    per_cell = np.zeros(tuple(out_dims), dtype=np.int64)
    for (g, cell), k in counts.items():
        per_cell[cell] += k
    np.testing.assert_array_equal(steps, per_cell)
    np.testing.assert_allclose(abs_sum[..., 0], raw[..., 0], rtol=1e-6)


def test_tod_runs_flush_once_per_visit():
    # A track that leaves a cell and comes back flushes twice there: tracks
    # bounce in a narrow diffusive tube, and the flush count equals the
    # number of maximal runs of equal consecutive cells along each track.
    fod, wm, _ = fixtures.cross()
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.25, gamma=0.5,
                                       noise_variance=0.2, seed=5)
    seed = algorithm.get_rng_seed()
    n = 300
    dims = np.array(wm.shape)
    _, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, n, 60, dims, np.eye(4, dtype=np.float32))
    # The same tracks through the tracking restatement (boundary seeds,
    # rng_init(seed, g) after the sample drives the noise) give the cells.
    seeds = oracle.domain_seed(domain, seed, n)
    runs = np.zeros(wm.shape, dtype=np.uint32)
    steps_total = revisits = 0
    for g in range(n):
        cells = _cells_of_track(domain, field, algorithm, seed, g, seeds[g], 60)
        steps_total += len(cells)
        changes = sum(1 for a, b in zip(cells, cells[1:]) if a != b)
        revisits += changes + 1 - len(set(cells)) if cells else 0
        previous = None
        for c in cells:
            if c != previous:
                runs[c] += 1
            previous = c
    assert steps_total > 1000 and revisits > 5
    np.testing.assert_array_equal(flushes, runs)
    # Every deposit adds float32(Y_0) to the l = 0 sum.
    np.testing.assert_allclose(abs_sum[..., 0].sum() / np.float32(0.28209479177387814), steps_total, rtol=1e-6)


def _cells_of_track(domain, field, algorithm, rng_seed, g, seed, n_steps):
    """Cells of every deposited step of track g, from the one-op restatements
    (contains, field lookup, update) with the track's own noise stream."""
    mask = np.asarray(domain.values != 0)
    state = oracle.rng_init(rng_seed, g)
    # Skip the boundary sample's draws (triangle, two barycentric uniforms).
    _, state = oracle.rng_draws(state, 1, 3)
    p = seed[:3].astype(np.float32)
    o = seed[4:7].astype(np.float32)
    cells = []
    for _ in range(n_steps):
        m = np.floor(p + np.float32(0.5))
        if np.any(m < 0) or np.any(m >= mask.shape) or not mask[tuple(m.astype(int))]:
            break
        f = np.rint(p)
        if np.any(f < 0) or np.any(f >= field.values.shape[:3]):
            break
        cells.append(tuple(m.astype(int)))
        o, state = oracle.update_state(algorithm, state, field.values[tuple(f.astype(int))], o)
        if not np.all(np.isfinite(o)) or np.sqrt(np.float32((o[0] * o[0] + o[1] * o[1]) + o[2] * o[2])) < 0.5:
            break
        p = p + np.float32(algorithm.step_size) * o
    return cells


def test_tod_fp32_and_fp64_restatements_agree():
    fod, wm, _ = fixtures.cross()
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.25, gamma=0.5)
    dims = np.array(wm.shape)
    r32, _, f32 = oracle.tod(domain, field, algorithm, 3, 400, 60, dims, np.eye(4, dtype=np.float32))
    r64, _, f64 = oracle.tod(domain, field, algorithm, 3, 400, 60, dims, np.eye(4, dtype=np.float32), precision=64)
    same = f32 == f64
    assert same.mean() > 0.9 and f32.sum() > 100
    scale = np.abs(r64).max()
    np.testing.assert_allclose(r32[same], r64[same], atol=2e-3 * scale)


def test_normalize_odf_known_answers():
    # utils.py:6-22: unit integral per nonzero voxel, zero voxels stay zero.
    odf = np.zeros((3, 45), dtype=np.float32)
    odf[0, 0] = 2.0
    odf[0, 7] = -1.0
    odf[2, 0] = -0.5
    out = oracle.normalize_odf(odf)
    assert out.dtype == np.float32
    np.testing.assert_allclose(out[0, 0], 1 / np.sqrt(4 * np.pi), rtol=1e-7)
    np.testing.assert_allclose(out[0, 7], -0.5 / np.sqrt(4 * np.pi), rtol=1e-7)
    np.testing.assert_array_equal(out[1], 0.0)
    np.testing.assert_allclose(out[2, 0], 1 / np.sqrt(4 * np.pi), rtol=1e-7)
    np.testing.assert_array_equal(tg.utils.normalize_odf(odf), out)
