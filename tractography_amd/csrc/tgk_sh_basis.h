// tgk_sh_basis.h -- host-side FP64 real symmetric spherical harmonics, for
// the tables and transforms the library builds once per field or launch.
//
// Convention of the reference (core.py:36-69, spharm.cl:216-413), MRtrix
// order: even l only, index l(l+1)/2 + m, with for m > 0
//   Y_{l, m}  = sqrt(2) N_lm P_l^m(cos theta) cos(m phi)
//   Y_{l,-m}  = sqrt(2) N_lm P_l^m(cos theta) sin(m phi)
//   Y_{l, 0}  = N_l0 P_l(cos theta)
// where N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) and P_l^m includes the
// Condon-Shortley phase (-1)^m.
#pragma once

#include <cmath>
#include <vector>

namespace tgk {
namespace shb {

inline double factorial(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; i++)
        f *= i;
    return f;
}

// Legendre polynomial P_l as ascending monomial coefficients (Bonnet).
inline std::vector<double> legendre(int l)
{
    std::vector<double> prev(1, 1.0), cur(2, 0.0);
    cur[1] = 1.0;
    if (l == 0)
        return prev;
    for (int k = 1; k < l; k++) {
        std::vector<double> next(k + 2, 0.0);
        for (int j = 0; j <= k; j++)
            next[j + 1] += (2.0 * k + 1.0) * cur[j] / (k + 1.0);
        for (int j = 0; j < (int)prev.size(); j++)
            next[j] -= k * prev[j] / (k + 1.0);
        prev = cur;
        cur = next;
    }
    return cur;
}

// Number of coefficients of the even-order basis up to lmax.
inline int n_coefficients(int lmax) { return (lmax + 1) * (lmax + 2) / 2; }

// The basis at direction (x, y, z) (normalised here; the zero vector maps to
// the pole, as cart2sph does) into out[n_coefficients(lmax)].
inline void real_sh_basis(int lmax, double x, double y, double z, double *out)
{
    const double r = std::sqrt(x * x + y * y + z * z);
    double ct = 1.0, st = 0.0, phi = 0.0;
    if (r > 0.0) {
        ct = std::fmin(std::fmax(z / r, -1.0), 1.0);
        st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        phi = std::atan2(y, x);
    }
    // Associated Legendre P_l^m by the standard recurrence in l for each m:
    // P_m^m = (-1)^m (2m-1)!! s^m, P_{m+1}^m = (2m+1) z P_m^m,
    // P_l^m = ((2l-1) z P_{l-1}^m - (l+m-1) P_{l-2}^m) / (l-m).
    std::vector<double> P((size_t)(lmax + 1) * (lmax + 1), 0.0);
    double pmm = 1.0;
    for (int m = 0; m <= lmax; m++) {
        if (m > 0)
            pmm *= -(2.0 * m - 1.0) * st;
        P[(size_t)m * (lmax + 1) + m] = pmm;
        if (m + 1 <= lmax)
            P[(size_t)(m + 1) * (lmax + 1) + m] = (2.0 * m + 1.0) * ct * pmm;
        for (int l = m + 2; l <= lmax; l++)
            P[(size_t)l * (lmax + 1) + m] = ((2.0 * l - 1.0) * ct * P[(size_t)(l - 1) * (lmax + 1) + m] -
                                             (l + m - 1.0) * P[(size_t)(l - 2) * (lmax + 1) + m]) /
                                            (double)(l - m);
    }
    for (int l = 0; l <= lmax; l += 2) {
        const int centre = l * (l + 1) / 2;
        for (int m = 0; m <= l; m++) {
            const double N = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * factorial(l - m) / factorial(l + m));
            const double p = P[(size_t)l * (lmax + 1) + m];
            if (m == 0) {
                out[centre] = N * p;
            } else {
                out[centre + m] = std::sqrt(2.0) * N * p * std::cos(m * phi);
                out[centre - m] = std::sqrt(2.0) * N * p * std::sin(m * phi);
            }
        }
    }
}

}  // namespace shb
}  // namespace tgk
