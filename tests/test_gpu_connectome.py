"""GPU parity of the boundary connectome (tracking.py:388-602): the device
kernel against the FP32 host restatement bit for bit, and the reference's
connectome tests through the public API."""

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import fixtures
import oracle
import tractography_amd as tg
from tractography_amd import _native

pytestmark = pytest.mark.gpu


def _device_connectome(runtime, domain, field, algorithm, rng_seed, n_seeds, n_steps, min_length=0.0):
    from tractography_amd.domain import _mesh_handle_for

    rows = np.empty(9 * n_seeds, np.int32)
    cols = np.empty(9 * n_seeds, np.int32)
    weights = np.empty(9 * n_seeds, np.float32)
    _native.check(_native.lib().tgk_connectome(
        runtime.handle, domain.upload(runtime).handle, _mesh_handle_for(domain, runtime).handle,
        field.upload(runtime).handle, algorithm.descriptor(), rng_seed, n_seeds, n_steps, float(np.float32(min_length)),
        _native.ptr(rows), _native.ptr(cols), _native.ptr(weights)))
    return rows, cols, weights


@pytest.mark.parametrize("case", ["sh45-transport", "sh45-diffusion", "sh15-diffusion", "dti-transport"])
def test_connectome_bit_exact(runtime, case, oracle_threads):
    if case.startswith("dti"):
        fod, wm, _ = fixtures.tensor_cross()
        model = tg.local.DiffusionTensorImaging()
        step, gamma = 0.05, 20.0
    else:
        K = 45 if case.startswith("sh45") else 15
        fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=K)
        model = tg.local.RealSymmetricSphericalHarmonics(K)
        step, gamma = 0.1, 1.0
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    if case.endswith("transport"):
        algorithm = tg.algorithm.Transport(model, step_size=step, gamma=gamma)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=step, gamma=gamma, noise_variance=0.1, seed=9)
    seed = algorithm.get_rng_seed()
    got = _device_connectome(runtime, domain, field, algorithm, seed, 20000, 300, 0.5)
    ref = oracle.connectome(domain, field, algorithm, seed, 20000, 300, 0.5, threads=oracle_threads)
    accepted = (got[0] >= 0).mean()
    print(f"[connectome {case}] {accepted:.3f} of entries set", flush=True)
    assert accepted > 0.2
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_connectome_returns_piecewise_linear_boundary_matrix(runtime):
    # test/test_tracking.py:183-209
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field_data = np.zeros((1, 1, 1, 15), dtype=np.float32)
    field_data[..., 0] = 1.0
    field = tg.field.Image(tg.Nifti1Image(field_data, np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.1, gamma=0.0)
    matrix = tg.tracking.connectome(algorithm, domain, field, n_seeds=16, maximum_streamline_length=2.0,
                                    runtime=runtime)
    assert isinstance(matrix, csr_matrix)
    assert matrix.shape == (len(domain.boundary.vertices),) * 2
    assert matrix.nnz > 0
    assert matrix.dtype == np.float32
    assert 0.0 < matrix.sum() <= 16.0 + 1e-5


def test_connectome_cross_dataset_connects_expected_regions(runtime):
    # test/test_tracking.py:227-255
    fod_image, mask_image, segmentation = fixtures.cross()
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(fod_image.shape[-1]),
                                       step_size=0.25, gamma=0.0)
    matrix, labels = tg.connectome(fod_image, n_seeds=10000, domain=mask_image, algorithm=algorithm,
                                   segmentation=segmentation, distance_upper_bound=2.0,
                                   maximum_streamline_length=12.0, runtime=runtime)
    np.testing.assert_array_equal(labels, [1, 2, 3, 4])
    undirected = matrix + matrix.T
    expected_partners = [1, 0, 3, 2]
    for region, partner in enumerate(expected_partners):
        competing_weights = np.delete(undirected[region], [region, partner])
        assert undirected[region, partner] > competing_weights.max()


def test_connectome_persistent_lanes_refill(runtime, oracle_threads):
    # Many more seeds than the grid has lanes, with short tracks: lanes claim
    # new seeds many times over, and every seed's entries still equal the
    # restatement's.
    fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=45)
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.25, gamma=1.0,
                                       noise_variance=0.1, seed=21)
    seed = algorithm.get_rng_seed()
    n = 600_000
    got = _device_connectome(runtime, domain, field, algorithm, seed, n, 24, 0.0)
    ref = oracle.connectome(domain, field, algorithm, seed, n, 24, 0.0, threads=oracle_threads)
    print(f"[connectome refill] {(got[0] >= 0).mean():.3f} of entries set", flush=True)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
