"""The synthetic-workload generators against the closed-form basis: the FP64
real symmetric SH recurrence equals core.ishtmtx (scipy) to order 12, and
fibre FODs peak along their fibres with the unit-integral normalisation of
utils.normalize_odf."""

import numpy as np

import bench
import tractography_amd as tg


def test_real_sh_basis_matches_ishtmtx_to_order_12():
    rng = np.random.default_rng(0)
    xyz = rng.standard_normal((500, 3))
    xyz /= np.linalg.norm(xyz, axis=1, keepdims=True)
    xyz[:2] = [[0, 0, 1], [0, 0, -1]]
    azimuth, colatitude, _ = tg.core.cart2sph(*xyz.T)
    expected, _ = tg.core.ishtmtx(azimuth, colatitude, 91)
    np.testing.assert_allclose(bench.real_sh_basis(xyz, 12), expected, atol=1e-11)
    np.testing.assert_allclose(bench.real_sh_basis(xyz, 12, transpose=True), expected.T, atol=1e-11)


def test_fibre_fod_peaks_on_its_fibre():
    d = np.array([[[0.6, 0.0, 0.8]]])
    for lmax in (8, 12):
        fod = bench.fod_from_fibres(d, np.ones((1, 1)), lmax=lmax)[0].astype(np.float64)
        np.testing.assert_allclose(fod[0], 1 / np.sqrt(4 * np.pi), rtol=1e-6)
        sphere = tg.core.fibonacci_sphere(4000)
        values = bench.real_sh_basis(sphere, lmax) @ fod
        best = sphere[np.argmax(values)]
        assert abs(best @ d[0, 0]) > 0.999
