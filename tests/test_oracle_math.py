"""The FP32 elementary-function contract (tgk_math.h) against float64.

These bounds pin the functions the kernels and the FP32 restatement share:
a wrong coefficient, reduction constant or quadrant rule fails them.
"""

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1)


def ulp_error(got, exact):
    exact32 = exact.astype(np.float32)
    spacing = np.spacing(np.abs(exact32)).astype(np.float64)
    spacing = np.where(spacing == 0, np.float64(np.spacing(np.float32(0))), spacing)
    return np.abs(got.astype(np.float64) - exact) / spacing


@pytest.mark.parametrize(
    "fn,reference,bound",
    [(0, np.sin, 1.5), (1, np.cos, 1.5)],
)
def test_sin_cos_ulp(fn, reference, bound):
    x = np.concatenate(
        [
            RNG.uniform(-20, 20, 1_000_000),
            RNG.uniform(-1e4, 1e4, 200_000),
            RNG.uniform(-1e-4, 1e-4, 50_000),
            np.arange(-64, 64) * (np.pi / 4),
        ]
    ).astype(np.float32)
    err = ulp_error(oracle.math32(fn, x), reference(x.astype(np.float64)))
    assert err.max() <= bound


def test_atan2_ulp_and_quadrants():
    y = (RNG.standard_normal(1_000_000) * np.exp(RNG.uniform(-20, 20, 1_000_000))).astype(np.float32)
    x = (RNG.standard_normal(1_000_000) * np.exp(RNG.uniform(-20, 20, 1_000_000))).astype(np.float32)
    err = ulp_error(oracle.math32(2, y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() <= 2.1


def test_acos_exp_log_ulp():
    a = np.concatenate([RNG.uniform(-1, 1, 1_000_000), 1 - RNG.uniform(0, 1e-3, 100_000), [1, -1, 0.5, -0.5, 0]])
    a = a.astype(np.float32)
    assert ulp_error(oracle.math32(3, a), np.arccos(a.astype(np.float64))).max() <= 1.1
    b = RNG.uniform(-103, 88.7, 1_000_000).astype(np.float32)
    assert ulp_error(oracle.math32(4, b), np.exp(b.astype(np.float64))).max() <= 1.1
    c = np.concatenate([np.exp(RNG.uniform(-100, 88, 1_000_000)), RNG.uniform(0.5, 2, 500_000), [1e-40, 1e-45]])
    c = c.astype(np.float32)
    assert ulp_error(oracle.math32(5, c), np.log(c.astype(np.float64))).max() <= 1.1


def test_sqrt_is_correctly_rounded():
    s = np.abs(RNG.standard_normal(1_000_000)).astype(np.float32)
    np.testing.assert_array_equal(oracle.math32(6, s), np.sqrt(s))


def test_special_values():
    inf, nan = np.inf, np.nan
    e = oracle.math32(4, np.float32([inf, -inf, 0, 89, -104, 88.72]))
    assert e[0] == inf and e[1] == 0 and e[2] == 1 and e[3] == inf and e[4] == 0 and np.isfinite(e[5])
    lg = oracle.math32(5, np.float32([inf, 0, -1, 1]))
    assert lg[0] == inf and lg[1] == -inf and np.isnan(lg[2]) and lg[3] == 0
    y = np.float32([0, -0.0, 0, -0.0, 1, inf, inf, -1])
    x = np.float32([-0.0, -0.0, 0, 0, 0, inf, -inf, -inf])
    got = oracle.math32(2, y, x)
    expected = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(got, expected)
    np.testing.assert_array_equal(np.signbit(got), np.signbit(expected))
    assert np.all(np.isnan(oracle.math32(0, np.float32([inf, nan]))))
    # The Cody-Waite reduction keeps sin(float(pi)) = pi - float(pi).
    assert oracle.math32(0, np.float32([np.pi]))[0] == np.float32(np.sin(np.float64(np.float32(np.pi))))
    big = oracle.math32(1, np.float32([1e30, 3e38, -2e20]))
    assert np.all(np.abs(big) <= 1)
