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


def _device_connectome(runtime, domain, field, algorithm, rng_seed, n_seeds, n_steps, min_length=0.0,
                       path=_native.PATH_PLAIN):
    from tractography_amd.domain import _mesh_handle_for

    rows = np.empty(9 * n_seeds, np.int32)
    cols = np.empty(9 * n_seeds, np.int32)
    weights = np.empty(9 * n_seeds, np.float32)
    _native.check(_native.lib().tgk_connectome(
        runtime.handle, domain.upload(runtime).handle, _mesh_handle_for(domain, runtime).handle,
        field.upload(runtime).handle, algorithm.descriptor(), path, rng_seed, n_seeds, n_steps,
        float(np.float32(min_length)), _native.ptr(rows), _native.ptr(cols), _native.ptr(weights)))
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


def _per_seed(entries):
    rows, cols, weights = (np.asarray(a).reshape(-1, 9) for a in entries)
    return rows, cols, weights


def _agreement(got, ref):
    """Per seed: the 9 (row, col) pairs equal the reference's; and, over the
    seeds where they do, the largest weight difference."""
    gr, gc, gw = _per_seed(got)
    rr, rc, rw = _per_seed(ref)
    same = np.all(gr == rr, axis=1) & np.all(gc == rc, axis=1)
    diff = np.abs(gw.astype(np.float64) - rw.astype(np.float64)).max(axis=1)
    return same, diff


@pytest.mark.parametrize("case", ["sh45-transport", "sh45-diffusion", "dti-transport", "sh45-probabilistic"])
def test_fast_connectome_tracks_fp64(runtime, case, oracle_threads):
    # The fast path's connectome against the FP64 restatement, with the
    # plain path's FP32 arithmetic (the reference's) measured the same way:
    # the same seeds must end on the same boundary triangles, and where both
    # FP32 paths agree with FP64 on the triangles the weights (products of
    # barycentrics, which move with the end point) must be as close.
    if case.startswith("dti"):
        fod, wm, _ = fixtures.tensor_cross()
        model = tg.local.DiffusionTensorImaging()
        step, gamma = 0.05, 20.0
    else:
        fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=45)
        model = tg.local.RealSymmetricSphericalHarmonics(45)
        step, gamma = 0.1, 1.0
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    if case.endswith("transport"):
        algorithm = tg.algorithm.Transport(model, step_size=step, gamma=gamma)
    elif case.endswith("diffusion"):
        algorithm = tg.algorithm.Diffusion(model, step_size=step, gamma=gamma, noise_variance=0.1, seed=9)
    else:
        algorithm = tg.algorithm.Probabilistic(model, step_size=step, seed=9)
    seed = algorithm.get_rng_seed()
    n = 20000
    fast = _device_connectome(runtime, domain, field, algorithm, seed, n, 300, 0.5, path=_native.PATH_FAST)
    plain = _device_connectome(runtime, domain, field, algorithm, seed, n, 300, 0.5)
    ref = oracle.connectome(domain, field, algorithm, seed, n, 300, 0.5, precision=64, threads=oracle_threads)
    same_fast, diff_fast = _agreement(fast, ref)
    same_plain, diff_plain = _agreement(plain, ref)
    both = same_fast & same_plain
    set_fast = (_per_seed(fast)[0][:, 0] >= 0).mean()
    print(f"[fast connectome {case}] same triangles as FP64: fast {same_fast.mean():.4f} plain "
          f"{same_plain.mean():.4f}; weight difference where both agree, max: fast {diff_fast[both].max():.2e} "
          f"plain {diff_plain[both].max():.2e}, 99.9th percentile: fast {np.quantile(diff_fast[both], 0.999):.2e} "
          f"plain {np.quantile(diff_plain[both], 0.999):.2e}; entries set {set_fast:.3f}", flush=True)
    assert set_fast > 0.2
    assert same_fast.mean() >= same_plain.mean() - 0.01
    # The end point's barycentrics move by about the point deviation over
    # the triangle size (voxel faces): tracks of up to 300 steps, compared
    # whole, not only up to a possible voxel flip. The reference's own FP32
    # arithmetic reaches 2.9e-2 on the Transport case.
    assert diff_fast[both].max() <= max(1e-2, 2.0 * diff_plain[both].max())
    # Accepted tracks carry total weight one each. On the one-voxel-thick DTI
    # slab, seeds start on the pole of an unstable equilibrium (f minimal
    # along z, gradient zero) with gamma 20, so whether a track turns before
    # it leaves depends on the rounding of the frame at the pole: FP32
    # (sin(pi) = -8.7e-8) accepts 76.5% of tracks, FP64 83.7%. The total is
    # held to 1%, or to twice the reference's own FP32 distance from FP64.
    total_fast, total_plain, total_ref = (float(_per_seed(x)[2].sum()) for x in (fast, plain, ref))
    print(f"[fast connectome {case}] total weight: fast {total_fast:.1f} plain {total_plain:.1f} "
          f"FP64 {total_ref:.1f}", flush=True)
    assert abs(total_fast - total_ref) <= max(0.01 * total_ref, 2.0 * abs(total_plain - total_ref))


def test_fast_connectome_is_deterministic_and_refills(runtime):
    # Many more seeds than the persistent grid has lanes: every seed's entries
    # are the same on a second call (which lane ran it does not matter).
    fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=45)
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.25, gamma=1.0,
                                       noise_variance=0.1, seed=21)
    seed = algorithm.get_rng_seed()
    first = _device_connectome(runtime, domain, field, algorithm, seed, 600_000, 24, 0.0, path=_native.PATH_FAST)
    second = _device_connectome(runtime, domain, field, algorithm, seed, 600_000, 24, 0.0, path=_native.PATH_FAST)
    for a, b in zip(first, second):
        np.testing.assert_array_equal(a, b)
    assert (first[0] >= 0).mean() > 0.05


def test_fast_connectome_through_the_public_api():
    # test/test_tracking.py:227-255 on the fast path.
    fod_image, mask_image, segmentation = fixtures.cross()
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(fod_image.shape[-1]),
                                       step_size=0.25, gamma=0.0)
    matrix, labels = tg.connectome(fod_image, n_seeds=10000, domain=mask_image, algorithm=algorithm,
                                   segmentation=segmentation, distance_upper_bound=2.0,
                                   maximum_streamline_length=12.0, runtime=tg.HipRuntime(path="fast"))
    np.testing.assert_array_equal(labels, [1, 2, 3, 4])
    undirected = matrix + matrix.T
    for region, partner in enumerate([1, 0, 3, 2]):
        assert undirected[region, partner] > np.delete(undirected[region], [region, partner]).max()
