"""Fit the float32 polynomial coefficients used by ``tgk_math.h``.

Each elementary function in ``tgk_math.h`` is reduced to a short interval by
Cody & Waite style argument reduction (W. J. Cody and W. Waite, *Software
Manual for the Elementary Functions*, Prentice-Hall 1980) and then evaluated
with a polynomial whose coefficients come from this script:

* the target function is sampled in float64 (mpmath for the reference values)
  on a dense Chebyshev grid of the reduction interval;
* a weighted least-squares fit in the relative-error norm is iterated with
  Lawson's reweighting (C. L. Lawson, 1961), which converges to the minimax
  (equi-oscillating) solution;
* coefficients are then rounded to float32 one at a time, lowest order first,
  refitting the remaining higher-order ones after each rounding so that the
  rounding error of one coefficient is absorbed by the next.

Run ``python fit_math_poly.py`` to print the coefficient tables; the values in
``tgk_math.h`` were pasted from its output.
"""

import mpmath
import numpy as np

mpmath.mp.prec = 120


def _grid(a, b, n=4000):
    k = np.arange(n)
    x = 0.5 * (a + b) + 0.5 * (b - a) * np.cos(np.pi * (k + 0.5) / n)
    return np.sort(x)


def _lawson(basis, target, scale, iterations=200):
    """Minimise max |basis @ c - target| / scale by Lawson iterations."""
    a = basis / scale[:, None]
    b = target / scale
    weights = np.full(len(b), 1.0 / len(b))
    coefficients = None
    for _ in range(iterations):
        sw = np.sqrt(weights)
        coefficients = np.linalg.lstsq(a * sw[:, None], b * sw, rcond=None)[0]
        error = np.abs(a @ coefficients - b)
        weights = weights * error
        total = weights.sum()
        if total == 0:
            break
        weights /= total
    return coefficients


def _fit_rounded(basis, target, scale):
    """Fit, then round coefficients to float32 lowest order first."""
    n = basis.shape[1]
    fixed = []
    residual = target.copy()
    for j in range(n):
        c = _lawson(basis[:, j:], residual, scale)
        cj = float(np.float32(c[0]))
        fixed.append(cj)
        residual = residual - cj * basis[:, j]
    max_rel = np.max(np.abs(residual) / scale)
    return np.array(fixed, dtype=np.float32), max_rel


def _mp(f, xs):
    return np.array([float(f(mpmath.mpf(float(x)))) for x in xs])


def fit_sin(degree=4):
    # sin(r) = r + r^3 * (S1 + S2 w + ...), w = r^2, r in (0, pi/4].
    r = _grid(1e-6, np.pi / 4)
    exact = _mp(mpmath.sin, r)
    basis = np.stack([r ** (3 + 2 * j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - r, np.abs(exact))


def fit_cos(degree=5):
    # cos(r) = 1 + w * (C1 + C2 w + ...), w = r^2, r in [0, pi/4].
    r = _grid(1e-6, np.pi / 4)
    exact = _mp(mpmath.cos, r)
    basis = np.stack([r ** (2 + 2 * j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - 1.0, np.abs(exact))


def fit_atan(degree=7):
    # atan(t) = t + t^3 * P(t^2), t in (0, tan(pi/8)].
    t = _grid(1e-6, float(mpmath.tan(mpmath.pi / 8)))
    exact = _mp(mpmath.atan, t)
    basis = np.stack([t ** (3 + 2 * j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - t, np.abs(exact))


def fit_asin(degree=6):
    # asin(x) = x + x^3 * P(x^2), x in (0, 0.5].
    x = _grid(1e-6, 0.5)
    exact = _mp(mpmath.asin, x)
    basis = np.stack([x ** (3 + 2 * j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - x, np.abs(exact))


def fit_exp(degree=6):
    # exp(r) = 1 + r + r^2 * P(r), r in [-ln2/2, ln2/2].
    half = float(mpmath.log(2)) / 2
    r = _grid(-half, half)
    exact = _mp(mpmath.exp, r)
    basis = np.stack([r ** (2 + j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - 1.0 - r, np.abs(exact))


def fit_log(degree=4):
    # log(1 + f) = 2 s + s^3 * P(s^2) with s = f / (2 + f),
    # |s| <= (sqrt2 - 1) / (sqrt2 + 1).
    smax = float((mpmath.sqrt(2) - 1) / (mpmath.sqrt(2) + 1))
    s = _grid(1e-6, smax)
    exact = _mp(lambda v: 2 * mpmath.atanh(v), s)
    basis = np.stack([s ** (3 + 2 * j) for j in range(degree)], axis=1)
    return _fit_rounded(basis, exact - 2.0 * s, np.abs(exact))


def split_constant(value, parts=3):
    """Split a high-precision constant into float32 pieces that sum to it."""
    pieces = []
    remainder = mpmath.mpf(value)
    for _ in range(parts):
        p = float(np.float32(float(remainder)))
        pieces.append(p)
        remainder -= mpmath.mpf(p)
    return pieces


def main():
    np.set_printoptions(precision=9)
    for name, fit in [
        ("sin", fit_sin),
        ("cos", fit_cos),
        ("atan", fit_atan),
        ("asin", fit_asin),
        ("exp", fit_exp),
        ("log", fit_log),
    ]:
        coefficients, max_rel = fit()
        print(f"{name}: max relative polynomial error {max_rel:.3e}")
        for c in coefficients:
            print(f"    {float(c):.9e}f,")
    print("pi/2 pieces:", [f"{p:.9e}" for p in split_constant(mpmath.pi / 2)])
    print("pi pieces:", [f"{p:.9e}" for p in split_constant(mpmath.pi, 2)])
    print("pi/4 pieces:", [f"{p:.9e}" for p in split_constant(mpmath.pi / 4, 2)])
    print("ln2 pieces:", [f"{p:.9e}" for p in split_constant(mpmath.log(2), 2)])
    print("2/pi:", f"{float(np.float32(float(2 / mpmath.pi))):.9e}")
    print("log2(e):", f"{float(np.float32(float(1 / mpmath.log(2)))):.9e}")


if __name__ == "__main__":
    main()
