"""SH basis, local models and orientation updates of the restatement against
closed forms (scipy's spherical harmonics) and the reference's known answers."""

import numpy as np
import pytest

import fixtures
import oracle
import tractography_amd as tg


def _grid(n=100):
    # test/test_src/test_utils.py:130-135
    phi = np.linspace(0, 2 * np.pi - 2 * np.pi / n, n)
    theta = np.linspace(0, np.pi - np.pi / n, n)
    return phi, theta


def test_sh45_fp64_matches_closed_form():
    phi, theta = _grid()
    theta = theta[1:]
    phi = phi[1:]
    ylm, dp, dt = oracle.sh45(phi, theta, 64)
    values, derivatives = tg.core.ishtmtx(phi, theta, 45)
    # The reference's normalisations carry 11-17 significant digits.
    np.testing.assert_allclose(ylm, values, atol=2e-10)
    np.testing.assert_allclose(dt, derivatives[0], atol=5e-10)
    np.testing.assert_allclose(dp * np.sin(theta)[:, None], derivatives[1], atol=5e-10)


def test_sh45_fp32_matches_closed_form_to_five_decimals():
    # test/test_src/test_utils.py:130-228 (decimal=5)
    phi, theta = _grid()
    phi32, theta32 = phi.astype(np.float32), theta.astype(np.float32)
    ylm, dp, dt = oracle.sh45(phi32, theta32, 32)
    values, derivatives = tg.core.ishtmtx(phi32, theta32, 45)
    np.testing.assert_array_almost_equal(ylm, values.astype(np.float32), 5)
    np.testing.assert_array_almost_equal(dt, derivatives[0].astype(np.float32), 5)
    np.testing.assert_array_almost_equal(dp * np.sin(theta32)[:, None], derivatives[1], 5)


def test_sh45_antipodal_symmetry():
    # test/test_src/test_utils.py:67-128
    directions = tg.core.fibonacci_sphere(20)
    a, c, _ = tg.core.cart2sph(*directions.T)
    a2, c2, _ = tg.core.cart2sph(*(-directions).T)
    ylm, _, _ = oracle.sh45(a, c, 32)
    ylm2, _, _ = oracle.sh45(a2, c2, 32)
    np.testing.assert_array_almost_equal(ylm, ylm2, 5)


def test_cart2sph_fp32_matches_float64():
    rng = np.random.default_rng(3)
    xyz = rng.standard_normal((10000, 3)).astype(np.float32)
    got = oracle.cart2sph32(xyz)
    az, col, _ = tg.core.cart2sph(*xyz.astype(np.float64).T)
    np.testing.assert_allclose(got[:, 0], az, atol=2e-6)
    np.testing.assert_allclose(got[:, 1], col, atol=2e-6)
    np.testing.assert_array_equal(oracle.cart2sph32(np.zeros((1, 3))), [[0, 0]])


@pytest.mark.parametrize("precision,tol", [(32, 1e-4), (64, 1e-7)])
def test_sh_model_matches_numpy_twin(precision, tol):
    # test/test_local.py:27-64
    n = 100
    params = np.ones((n, 45))
    phi = np.linspace(0, 2 * np.pi, n, dtype=np.float32)
    theta = np.linspace(0, np.pi, n, dtype=np.float32)[1:-1]
    phi = phi[1:-1]
    orientations = np.stack(tg.core.sph2cart(phi, theta, 1.0), axis=-1).astype(np.float64)
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    reference = model.evaluate(params[: len(phi)], orientations)
    got = oracle.evaluate(0, 45, params[: len(phi)], orientations, precision)
    np.testing.assert_allclose(got, reference, rtol=tol, atol=tol)
    rng = np.random.default_rng(42)
    p = rng.normal(size=(12, 45))
    o = np.stack(tg.core.sph2cart(np.linspace(0.1, 6, 12), np.linspace(0.1, 3, 12), 1.0), axis=-1)
    np.testing.assert_allclose(oracle.evaluate(0, 45, p, o, precision), model.evaluate(p, o), rtol=tol, atol=tol)


@pytest.mark.parametrize("precision,tol", [(32, 1e-5), (64, 1e-12)])
def test_dti_model_closed_form(precision, tol):
    # test/test_local.py:71-115
    params = np.array([[2.0, 1.5, 0.7, 0.3, -0.4, 0.2], [1.0, 2.0, 3.0, 0.0, 0.0, 0.0]], dtype=np.float32)
    phi = np.array([0.7, 1.2], dtype=np.float32)
    theta = np.array([1.1, 0.8], dtype=np.float32)
    orientations = np.stack(tg.core.sph2cart(phi.astype(np.float64), theta.astype(np.float64), 1.0), axis=-1)
    expected = np.empty((2, 3))
    for i, (c, o) in enumerate(zip(params.astype(np.float64), orientations)):
        t = np.array([[c[0], c[3], c[4]], [c[3], c[1], c[5]], [c[4], c[5], c[2]]])
        p, th, _ = tg.core.cart2sph(*o)
        et = np.array([np.cos(th) * np.cos(p), np.cos(th) * np.sin(p), -np.sin(th)])
        ep = np.array([-np.sin(p), np.cos(p), 0.0])
        tu = t @ o
        expected[i] = [o @ tu, 2 * tu @ et, 2 * tu @ ep]
    got = oracle.evaluate(1, 6, params, orientations, precision)
    np.testing.assert_allclose(got, expected, rtol=tol, atol=tol)


@pytest.mark.parametrize("precision,tol", [(32, 2e-4), (64, 2e-4)])
def test_transport_update_matches_analytic_step(precision, tol):
    # test/test_algorithm.py:54-74
    baseline, epsilon, step, gamma = 0.1, 0.3, 0.04, 1.7
    params = fixtures.quadratic_fod_coefficients(baseline, epsilon)[None, :]
    o = np.array([tg.core.sph2cart(0.5, 0.8, 1.0)]).reshape(1, 3)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=step, gamma=gamma)
    got = oracle.update(algorithm, 0, params, o, precision)
    expected = fixtures.expected_quadratic_update(o[0], baseline, epsilon, step, gamma)
    np.testing.assert_allclose(got[0], expected, rtol=tol, atol=tol)
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)
    np.testing.assert_allclose(got, algorithm.update_orientation(params, o), atol=1e-5)


def test_transport_update_batches_distinct_models():
    # test/test_algorithm.py:77-103
    fod = [(0.1, 0.2), (0.25, 0.4), (0.05, 0.55)]
    params = np.stack([fixtures.quadratic_fod_coefficients(b, e) for b, e in fod])
    o = np.stack(tg.core.sph2cart(np.array([0.3, 1.2, 2.4]), np.array([0.6, 1.1, 2.0]), 1.0), axis=-1)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.025, gamma=1.3)
    expected = np.stack(
        [fixtures.expected_quadratic_update(oi, b, e, 0.025, 1.3) for oi, (b, e) in zip(o, fod)]
    )
    np.testing.assert_allclose(oracle.update(algorithm, 0, params, o, 32), expected, rtol=3e-4, atol=3e-4)
    np.testing.assert_allclose(oracle.update(algorithm, 0, params, o, 64), expected, rtol=3e-4, atol=3e-4)


def test_update_at_and_near_poles_is_finite_and_normalised():
    # test/test_algorithm.py:133-150
    params = fixtures.quadratic_fod_coefficients(0.1, 0.3)[None, :]
    o = np.stack(
        tg.core.sph2cart(np.array([0.0, 1.1, 0.7, 2.0]), np.array([0.0, np.pi, 1e-4, np.pi - 1e-4]), 1.0), axis=-1
    )
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.01)
    got = oracle.update(algorithm, 0, np.repeat(params, 4, axis=0), o, 32)
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)
    assert np.all(np.linalg.norm(got - o, axis=1) > 0.0)


def test_exps2_known_answers():
    # test/test_algorithm.py:119-130 through a Transport update with a zero
    # model (zero tangent returns the input) and the NumPy twin.
    result = tg.algorithm.exps2(np.array([[0.0, 0, 1], [1, 0, 0]]), np.array([[0.0, 0, 0], [0, 0.3, 0]]), 1.0)
    np.testing.assert_allclose(result[1], [np.cos(0.3), np.sin(0.3), 0.0], atol=1e-7)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=1.0, gamma=0.0)
    o = np.array([[0.3, 0.4, 0.5]], np.float32)
    np.testing.assert_array_equal(oracle.update(algorithm, 0, np.zeros((1, 15)), o, 32), o)


def test_diffusion_without_noise_equals_transport_and_noise_is_reproducible():
    # test/test_algorithm.py:185-202
    model = tg.local.RealSymmetricSphericalHarmonics(15)
    params = np.zeros((3, 15), np.float32)
    params[:, 0] = 1.0
    o = np.eye(3, dtype=np.float32)
    quiet = tg.algorithm.Diffusion(model, step_size=0.01, noise_variance=0.0)
    transport = tg.algorithm.Transport(model, step_size=0.01)
    np.testing.assert_array_equal(oracle.update(quiet, 5, params, o, 32), oracle.update(transport, 5, params, o, 32))
    noisy = tg.algorithm.Diffusion(model, noise_variance=1.0)
    first = oracle.update(noisy, 1234, params, o, 32)
    second = oracle.update(noisy, 1234, params, o, 32)
    np.testing.assert_array_equal(first, second)
    np.testing.assert_allclose(np.linalg.norm(first, axis=1), 1.0, atol=1e-5)
    assert np.any(np.linalg.norm(first - o, axis=1) > 0.0)
    np.testing.assert_allclose(oracle.update(noisy, 1234, params, o, 64), first, atol=1e-5)


def test_generic_fp64_basis_matches_the_l8_basis():
    rng = np.random.default_rng(40)
    phi = rng.uniform(0, 2 * np.pi, 3000)
    theta = np.concatenate([rng.uniform(0, np.pi, 2997), [0.0, np.pi, 1e-9]])
    y8, dp8, dt8 = oracle.sh45(phi, theta, precision=64)
    y, dp, dt = oracle.sh_basis64(8, phi, theta)
    # The l <= 8 expressions carry the reference's 11-digit constants.
    np.testing.assert_allclose(y, y8, atol=1e-9)
    np.testing.assert_allclose(dp, dp8, atol=1e-8)
    np.testing.assert_allclose(dt, dt8, atol=1e-8)


@pytest.mark.parametrize("lmax", [10, 12])
def test_generic_fp64_basis_matches_ishtmtx(lmax):
    # Order 12 has no reference kernel; it is pinned by the closed form
    # (core.py:36-69, scipy's sph_harm_y) with its angular derivatives.
    rng = np.random.default_rng(lmax)
    phi = rng.uniform(0, 2 * np.pi, 2000)
    theta = rng.uniform(1e-3, np.pi - 1e-3, 2000)
    K = (lmax + 1) * (lmax + 2) // 2
    y, dp, dt = oracle.sh_basis64(lmax, phi, theta)
    matrix, derivatives = tg.core.ishtmtx(phi, theta, K)
    np.testing.assert_allclose(y, matrix, atol=1e-10)
    np.testing.assert_allclose(dt, derivatives[0], atol=1e-9)
    np.testing.assert_allclose(dp * np.sin(theta)[:, None], derivatives[1], atol=1e-9)


def test_fp64_model_of_order_12_matches_the_closed_form():
    rng = np.random.default_rng(41)
    params = rng.normal(size=(500, 91)) * 0.1
    o = rng.standard_normal((500, 3))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    got = oracle.evaluate(0, 91, params, o, 64)
    az, col, _ = tg.core.cart2sph(*o.T)
    matrix, derivatives = tg.core.ishtmtx(az, col, 91)
    v = np.einsum("nk,nk->n", matrix, params)
    d = 1 / (1 + np.exp(-100 * v))
    value = np.where(100 * v < 30, np.log1p(np.exp(np.minimum(100 * v, 30))) / 100, v)
    np.testing.assert_allclose(got[:, 0], value, atol=1e-12, rtol=1e-10)
    np.testing.assert_allclose(got[:, 1], d * np.einsum("nk,nk->n", derivatives[0], params), atol=1e-10)
    np.testing.assert_allclose(got[:, 2] * np.sin(col), d * np.einsum("nk,nk->n", derivatives[1], params),
                               atol=1e-10)
