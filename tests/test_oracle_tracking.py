"""The restated tracking loop (tracking.py:94-149) against the reference's
known answers, exact solutions and geometric invariants."""

import numpy as np
import pytest

import fixtures
import oracle
import tractography_amd as tg


def _run(domain, field, algorithm, seeds, save_at, max_length, precision=32, rng_seeds=(0,), batch=None):
    seeds = tg.seeds.validate_seed_array(seeds)
    batch = len(seeds) if batch is None else batch
    return oracle.track(domain, field, algorithm, seeds, list(rng_seeds), batch, save_at, max_length,
                        precision=precision)


def _streamline(result, i=0):
    return result["points"][i, : result["lengths"][i]]


@pytest.mark.parametrize("precision", [32, 64])
def test_tracks_seed_until_leaving_domain(precision):
    # test/test_tracking.py:28-53
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((1, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=1.0, gamma=0.0)
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], np.float32)
    r = _run(domain, field, algorithm, seeds, 1.0, 300.0, precision)
    np.testing.assert_allclose(_streamline(r), [[0, 0, 0], [1, 0, 0]], atol=1e-6)
    assert r["steps"][0] == 1


@pytest.mark.parametrize("precision", [32, 64])
def test_save_interval_and_maximum_length(precision):
    # test/test_tracking.py:71-100
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((20, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((20, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.25, gamma=0.0)
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], np.float32)
    r = _run(domain, field, algorithm, seeds, 0.5, 3.0, precision)
    s = _streamline(r)
    assert len(s) == 6
    np.testing.assert_allclose(s[:, 0], np.arange(0.0, 3.0, 0.5), atol=1e-6)
    np.testing.assert_allclose(s[:, 1:], 0.0, atol=1e-6)
    assert r["steps"][0] == 10


def test_diffusion_without_noise_and_drift():
    # test/test_tracking.py:103-127
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((1, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    algorithm = tg.algorithm.Diffusion(
        tg.local.RealSymmetricSphericalHarmonics(15), step_size=1.0, gamma=0.0, noise_variance=0.0, seed=1234
    )
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], np.float32)
    r = _run(domain, field, algorithm, seeds, 1.0, 300.0, rng_seeds=[algorithm.get_rng_seed()])
    np.testing.assert_allclose(_streamline(r), [[0, 0, 0], [1, 0, 0]], atol=1e-6)


def _u0u3_scene():
    epsilon = 0.3
    coefficients = fixtures.u0u3_coefficients(epsilon, 45)
    fod = np.broadcast_to(coefficients, (1, 1, 1, 45)).copy()
    affine = np.diag((1000.0, 1000.0, 1000.0, 1.0))
    affine[:3, 3] = -0.5
    image = tg.Nifti1Image(fod, affine)
    field = tg.field.Image(image)
    domain = tg.domain.Image(tg.Nifti1Image(np.any(fod != 0, axis=-1).astype(np.uint8), affine))
    return field, domain, epsilon


@pytest.mark.parametrize("algorithm_name", ["transport", "diffusion"])
@pytest.mark.parametrize("precision", [32, 64])
def test_u0u3_exact_solution(algorithm_name, precision):
    # test/test_algorithms/test_transport.py:117-233, test_diffusion.py:15-71
    field, domain, epsilon = _u0u3_scene()
    theta_0, phi_0 = 0.12, 0.5
    orientation = np.asarray(tg.core.sph2cart(phi_0, theta_0, 1.0))
    seeds = np.array([[10.0, 10.0, 10.0, 1.0, *orientation, 0.0]], np.float32)
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if algorithm_name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.001, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.001, gamma=1.0, noise_variance=0.0, seed=0)
    r = _run(domain, field, algorithm, seeds, 0.001, 1.0, precision, rng_seeds=[algorithm.get_rng_seed()])
    streamline = _streamline(r)
    u_ref = fixtures.u0u3_reference(0.001, 1.0, theta_0, phi_0, epsilon, len(streamline))
    expected = np.zeros((len(streamline), 3))
    expected[0] = [10.0, 10.0, 10.0]
    for k in range(1, len(streamline)):
        expected[k] = expected[k - 1] + 0.001 * u_ref[k]
    np.testing.assert_array_almost_equal(streamline, expected, decimal=3)


def test_circle_runs_to_maximum_length():
    # test/test_algorithms/test_transport.py:298-328
    shape, radius = (10, 10, 1), 2
    fod = fixtures.circle(shape=shape, radius=radius)
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image((fod[..., 0] > 0).astype(np.uint8), np.eye(4)))
    seeds = np.array([[(shape[0] - 1) / 2, (shape[1] - 1) / 2 + radius, 0, 1, 1, 0, 0, 0]], np.float32)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=1e-4, gamma=50.0)
    r = _run(domain, field, algorithm, seeds, 0.25, 100.0)
    assert r["lengths"][0] * 0.25 == pytest.approx(100.0)


def test_dti_circle_diffusion_runs_to_maximum_length():
    # test/test_algorithms/test_diffusion.py:127-158
    tensor, wm, _ = fixtures.tensor_circle((20, 20, 1), radius=5, width=2)
    field = tg.field.Image(tensor)
    domain = tg.domain.Image(tg.Nifti1Image(np.any(tensor.dataobj != 0, axis=-1).astype(np.uint8), tensor.affine))
    seeds = np.tile(np.array([[19.0, 22.5, 0, 1, -1, 0, 0, 0]], np.float32), (10, 1))
    algorithm = tg.algorithm.Diffusion(tg.local.DiffusionTensorImaging(), step_size=0.001, gamma=10.0,
                                       noise_variance=0.01, seed=3)
    r = _run(domain, field, algorithm, seeds, 0.25, 40.0, rng_seeds=[algorithm.get_rng_seed()])
    assert np.all(r["lengths"] * 0.25 == pytest.approx(40.0))


def test_straight_lines_in_isotropic_field():
    # test/test_algorithms/test_transport.py:244-260 (fewer, shorter tracks)
    fod = fixtures.uniform_isotropic()
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(np.any(fod != 0, axis=-1).astype(np.uint8), np.eye(4)))
    rng = np.random.default_rng(0)
    locations = rng.uniform(2, 7, (40, 3))
    directions = rng.standard_normal((40, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    seeds = tg.seeds._from_components(locations, directions)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.001, gamma=0.5)
    r = _run(domain, field, algorithm, seeds, 0.25, 10.0)
    for i in range(len(seeds)):
        d = np.diff(_streamline(r, i), n=2, axis=0)
        np.testing.assert_array_less(d, 1e-3)


def test_cross_coverage_equals_mask():
    # test/test_algorithms/test_transport.py:262-280 (2000 seeds)
    fod, wm, _ = fixtures.cross()
    masked = tg.Nifti1Image(fod.get_fdata() * wm.get_fdata()[..., None], fod.affine)
    field = tg.field.Image(masked)
    domain = tg.domain.Image(tg.Nifti1Image(np.any(masked.dataobj != 0, axis=-1).astype(np.uint8), np.eye(4)))
    np.random.seed(7)
    seeds = tg.seeds.from_mask(wm, 2000)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.001, gamma=0.5)
    r = _run(domain, field, algorithm, seeds, 0.25, 300.0)
    points = np.vstack([_streamline(r, i)[:-1] for i in range(len(seeds)) if r["lengths"][i] > 1])
    voxels = np.round(points).astype(int)
    mask = np.zeros(wm.shape, np.uint8)
    mask[tuple(voxels.T)] = 1
    np.testing.assert_array_equal(mask, wm.get_fdata().astype(np.uint8))


def test_fp32_and_fp64_restatements_agree_before_voxel_flips():
    # The FP32 restatement tracks the FP64 one to ~1e-5 voxel while no step
    # has come closer to a voxel face than the accumulated deviation.
    fod = fixtures.cross()[0]
    field = tg.field.Image(fod)
    domain = tg.domain.Image(tg.Nifti1Image(np.any(fod.dataobj != 0, axis=-1).astype(np.uint8), np.eye(4)))
    np.random.seed(3)
    seeds = tg.seeds.from_mask(domain_image := tg.Nifti1Image(domain.values, np.eye(4)), 64)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=11)
    rng = [algorithm.get_rng_seed()]
    r32 = _run(domain, field, algorithm, seeds, 0.25, 30.0, 32, rng_seeds=rng)
    r64 = _run(domain, field, algorithm, seeds, 0.25, 30.0, 64, rng_seeds=rng)
    compared = 0
    for i in range(len(seeds)):
        n = min(r32["lengths"][i], r64["lengths"][i])
        dev = np.linalg.norm(r32["points"][i, :n] - r64["points"][i, :n], axis=1)
        running = np.maximum.accumulate(dev)
        ok = running < r64["margins"][i, :n]
        k = int(np.argmin(ok)) if not np.all(ok) else n
        if k > 1:
            compared += k
            assert running[k - 1] < 1e-3
    assert compared > 500
    assert domain_image is not None


def test_mesh_domain_straight_rays_stop_at_the_first_triangle():
    # Isotropic field: orientations never change, so each streamline is the
    # ray from its seed; tracking stops after the step whose segment crosses
    # the first triangle along the ray (bvh.cl:83-131 per step segment).
    mesh, seeds, centre, radii = fixtures.shell_scene(n_seeds=300, seed=3)
    fod = fixtures.uniform_isotropic(shape=(32, 32, 32))
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0)
    r = _run(mesh, field, algorithm, seeds, 0.5, 200.0)
    data = mesh._ensure_boundary_data()
    faces = data["indices"][:, :3]
    expected, dist, margin = fixtures.first_ray_hits(data["vertices"][:, :3], faces, seeds[:, :3], seeds[:, 4:7])
    assert np.all(expected >= 0)
    ratio = dist / 0.1
    clear = (margin > 1e-4) & (np.abs(ratio - np.round(ratio)) > 2e-3)
    assert clear.mean() > 0.9
    np.testing.assert_array_equal(r["hits"][clear], expected[clear])
    np.testing.assert_array_equal(r["steps"][clear], np.ceil(ratio[clear]).astype(np.uint32))
    # Saved points stay in the shell, up to the last (crossing) step.
    for i in range(len(seeds)):
        p = _streamline(r, i).astype(np.float64) - centre
        rr = np.linalg.norm(p, axis=1)
        r_in, r_out = radii(p / rr[:, None])
        assert np.all(rr > r_in - 0.11) and np.all(rr < r_out + 0.11)


def test_connectome_restatement_known_answers():
    # test/test_tracking.py:183-209: a one-voxel domain with an isotropic
    # field and gamma 0: every track goes straight from a boundary sample to
    # the opposite face, and each accepted track deposits total weight one.
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field_data = np.zeros((1, 1, 1, 15), dtype=np.float32)
    field_data[..., 0] = 1.0
    field = tg.field.Image(tg.Nifti1Image(field_data, np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.1, gamma=0.0)
    rows, cols, weights = oracle.connectome(domain, field, algorithm, 0, 64, 20)
    per_seed = weights.reshape(64, 9).sum(axis=1)
    accepted = rows.reshape(64, 9)[:, 0] >= 0
    assert accepted.mean() > 0.9
    np.testing.assert_allclose(per_seed[accepted], 1.0, atol=1e-6)
    assert np.all(weights.reshape(64, 9)[~accepted] == 0)
    # Straight inward rays through a unit cube end on the face opposite
    # their start: the start and hit triangles never share a face plane.
    vertices = domain.boundary.vertices
    r = rows.reshape(64, 9)[accepted]
    c = cols.reshape(64, 9)[accepted]
    start_centre = vertices[r[:, [0, 3, 6]]].mean(axis=1)
    end_centre = vertices[c[:, :3]].mean(axis=1)
    assert np.all(np.abs(start_centre - end_centre).max(axis=1) > 0.5)


def test_connectome_restatement_fp32_and_fp64_agree():
    fod, wm, _ = fixtures.cross()
    domain = tg.domain.Image(wm)
    field = tg.field.Image(fod)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.25, gamma=0.5,
                                       noise_variance=0.05, seed=3)
    seed = algorithm.get_rng_seed()
    r32 = oracle.connectome(domain, field, algorithm, seed, 2000, 48)
    r64 = oracle.connectome(domain, field, algorithm, seed, 2000, 48, precision=64)
    same = np.all(r32[0].reshape(-1, 9) == r64[0].reshape(-1, 9), axis=1)
    assert same.mean() > 0.95
    np.testing.assert_allclose(r32[2].reshape(-1, 9)[same], r64[2].reshape(-1, 9)[same], atol=1e-4)
