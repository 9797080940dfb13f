"""The Probabilistic and Deterministic restatements (algorithm.py:272-677)
against closed forms: the zero-radius cap, uniform cap sampling, the
selection rule rebuilt from the noise stream with float64 closed-form DTI
weights, analytic maxima (the reference's 1 + 0.3 x z test, DTI principal
axes), and straight tracks in an isotropic field."""

import numpy as np
import pytest
from scipy import stats

import fixtures
import oracle
import tractography_amd as tg


def _sh45():
    return tg.local.RealSymmetricSphericalHarmonics(45)


@pytest.mark.parametrize("precision", [32, 64])
def test_probabilistic_zero_angle_limit(precision):
    # test/test_algorithms/test_probabilistic.py:12-27
    algorithm = tg.algorithm.Probabilistic(_sh45(), max_angle=0.0, seed=0)
    initial = np.asarray([tg.core.sph2cart(0.5, 0.12, 1.0)])
    coefficients = fixtures.u0u3_coefficients()[np.newaxis, :]
    result = oracle.update(algorithm, 7, coefficients, initial, precision)
    np.testing.assert_allclose(result, initial, atol=1e-6)


def test_probabilistic_single_candidate_is_uniform_on_the_cap():
    # One candidate per update is always chosen (positive weight): the
    # output is uniform on the cap, so its cosine to the input is uniform on
    # [cos(max_angle), 1] and its azimuth in the tangent frame is uniform.
    max_angle = 0.6
    algorithm = tg.algorithm.Probabilistic(tg.local.DiffusionTensorImaging(), n_samples=1, max_angle=max_angle)
    n = 20000
    params = np.tile(np.array([1.0, 2.0, 3.0, 0.1, 0.2, 0.3], np.float32), (n, 1))
    o = np.tile(np.array([0.3, -0.4, 0.866], np.float32), (n, 1))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    out = oracle.update(algorithm, 11, params, o, 32).astype(np.float64)
    cosines = np.sum(out * o, axis=1) / np.linalg.norm(out, axis=1)
    assert cosines.min() >= np.cos(max_angle) - 1e-6
    assert stats.kstest((cosines - np.cos(max_angle)) / (1 - np.cos(max_angle)), "uniform").pvalue > 1e-3


def _candidates_from_draws(o, cos_max, draws):
    """Cap candidates from the uniform draws (cosine, azimuth per candidate),
    in float64: algorithm.py:330-343."""
    o = o / np.linalg.norm(o)
    helper = np.array([0.0, 0.0, 1.0]) if abs(o[2]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(helper, o)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(o, t1)
    c = 1.0 - draws[0::2] * (1.0 - cos_max)
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    az = 2.0 * np.pi * draws[1::2]
    return c[:, None] * o + s[:, None] * (np.cos(az)[:, None] * t1 + np.sin(az)[:, None] * t2)


def test_probabilistic_selection_rule_from_the_noise_stream():
    # Weights = DTI values u^T D u (closed form); target = the (2N+1)-th
    # uniform draw times the total; the chosen candidate is the first whose
    # running weight exceeds the target; afterwards the stream resumes after
    # the chosen candidate's two draws.
    rng = np.random.default_rng(3)
    n_samples = 16
    algorithm = tg.algorithm.Probabilistic(tg.local.DiffusionTensorImaging(), n_samples=n_samples,
                                           max_angle=1.2)
    cos_max = np.cos(np.float32(1.2))
    checked = 0
    for trial in range(300):
        a = rng.normal(size=(3, 3))
        D = a @ a.T + 0.1 * np.eye(3)
        params = np.array([D[0, 0], D[1, 1], D[2, 2], D[0, 1], D[0, 2], D[1, 2]], np.float32)
        D = np.array([[params[0], params[3], params[4]], [params[3], params[1], params[5]],
                      [params[4], params[5], params[2]]], dtype=np.float64)
        o = rng.normal(size=3).astype(np.float32)
        state = oracle.rng_init(int(rng.integers(0, 2**32 - 1)), trial)
        draws, _ = oracle.rng_draws(state.copy(), 1, 2 * n_samples + 1)
        cand = _candidates_from_draws(o.astype(np.float64), cos_max, draws[:2 * n_samples].astype(np.float64))
        weights = np.einsum("ni,ij,nj->n", cand, D, cand)
        cumulative = np.cumsum(weights)
        target = float(draws[-1]) * cumulative[-1]
        chosen = int(np.argmax(cumulative > target))
        margin = np.min(np.abs(cumulative - target)) / cumulative[-1]
        got, after = oracle.update_state(algorithm, state.copy(), params, o)
        if margin < 1e-4:
            continue  # a rounding-level tie; either neighbour is right
        checked += 1
        np.testing.assert_allclose(got, cand[chosen], atol=2e-5)
        _, expected_after = oracle.rng_draws(state.copy(), 1, 2 * (chosen + 1))
        np.testing.assert_array_equal(after, expected_after)
    assert checked > 250


@pytest.mark.parametrize("precision", [32, 64])
def test_deterministic_exact_cross_term_maximum(precision):
    # test/test_algorithms/test_deterministic.py:12-29: the maximum of
    # 1 + 0.3 x z on the sphere is (1, 0, 1) / sqrt(2).
    algorithm = tg.algorithm.Deterministic(_sh45(), max_iterations=30, tolerance=1e-6, max_step_size=0.5)
    initial = np.asarray([tg.core.sph2cart(0.5, 0.12, 1.0)])
    coefficients = fixtures.u0u3_coefficients()[np.newaxis, :]
    result = oracle.update(algorithm, 0, coefficients, initial, precision)
    np.testing.assert_allclose(result[0], np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0), atol=1e-3)


def test_deterministic_finds_the_dti_principal_axis():
    # u^T D u is maximal on the sphere at the principal eigenvector (sign of
    # the starting hemisphere). The line search accepts only gains above
    # 1e-7, so the ascent stops within about sqrt(1e-7 / (w1 - w2)) of the
    # axis: the reference's own test uses 1e-3.
    rng = np.random.default_rng(5)
    algorithm = tg.algorithm.Deterministic(tg.local.DiffusionTensorImaging(), max_iterations=50, tolerance=1e-7,
                                           max_step_size=0.5)
    for _ in range(50):
        w = np.array([3.0, 1.0, 0.5])
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        D = q @ np.diag(w) @ q.T
        params = np.array([D[0, 0], D[1, 1], D[2, 2], D[0, 1], D[0, 2], D[1, 2]], np.float32)
        axis = q[:, 0]
        start = axis + 0.4 * rng.normal(size=3)
        start *= np.sign(start @ axis)
        got = oracle.update(algorithm, 0, params[None, :], start[None, :].astype(np.float32), 64)[0]
        np.testing.assert_allclose(got, axis, atol=1e-3)


def test_deterministic_stops_at_once_when_there_is_nothing_to_climb():
    o = np.array([[0.3, 0.4, 2.0]], np.float32)
    unit = o / np.linalg.norm(o)
    isotropic = np.zeros((1, 45), np.float32)
    isotropic[0, 0] = 1.0
    climb = tg.algorithm.Deterministic(_sh45())
    np.testing.assert_allclose(oracle.update(climb, 0, isotropic, o, 32), unit, atol=1e-7)
    none = tg.algorithm.Deterministic(_sh45(), max_iterations=0)
    np.testing.assert_allclose(oracle.update(none, 0, fixtures.u0u3_coefficients()[None, :], o, 32), unit,
                               atol=1e-7)


@pytest.mark.parametrize("name", ["probabilistic", "deterministic"])
def test_plugin_tracks_in_an_isotropic_field(name, oracle_threads):
    # Deterministic: nothing to climb, so straight lines. Probabilistic: every
    # step turns by at most max_angle.
    fod = fixtures.uniform_isotropic()
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(np.any(fod != 0, axis=-1).astype(np.uint8), np.eye(4)))
    rng = np.random.default_rng(2)
    locations = rng.uniform(2, 7, (60, 3))
    directions = rng.standard_normal((60, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    seeds = tg.seeds._from_components(locations, directions)
    if name == "deterministic":
        algorithm = tg.algorithm.Deterministic(_sh45(), step_size=0.05)
    else:
        algorithm = tg.algorithm.Probabilistic(_sh45(), max_angle=0.2, step_size=0.05, seed=1)
    r = oracle.track(domain, field, algorithm, seeds, [algorithm.get_rng_seed()], 60, 0.05, 5.0,
                     threads=oracle_threads)
    assert r["steps"].sum() > 2000
    for i in range(len(seeds)):
        p = r["points"][i, :r["lengths"][i]].astype(np.float64)
        seg = np.diff(p, axis=0)
        seg = seg[np.linalg.norm(seg, axis=1) > 0.04]
        u = seg / np.linalg.norm(seg, axis=1, keepdims=True)
        cos_turn = np.sum(u[1:] * u[:-1], axis=1)
        if name == "deterministic":
            assert np.all(cos_turn > 1 - 1e-6)
        else:
            assert np.all(cos_turn >= np.cos(0.2) - 1e-5)
