"""GPU parity of the plain path and the one-op drivers: bit-exact against the
FP32 host restatement, and the reference's known answers through the public
API on the device."""

import os

import numpy as np
import pytest

import bench
import fixtures
import oracle
import tractography_amd as tg

pytestmark = pytest.mark.gpu


def _native_math(runtime, fn, x, y=None):
    from tractography_amd import _native

    x = np.ascontiguousarray(x, np.float32)
    y = x if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    _native.check(_native.lib().tgk_math(runtime.handle, fn, _native.ptr(x), _native.ptr(y), len(x), _native.ptr(out)))
    return out


def test_device_math_is_bit_identical_to_host(runtime):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-30, 30, 500_000), rng.uniform(-1e4, 1e4, 100_000),
                        [0.0, -0.0, np.inf, -np.inf, np.nan, 1e30]]).astype(np.float32)
    for fn in (0, 1):
        np.testing.assert_array_equal(_native_math(runtime, fn, x), oracle.math32(fn, x))
    y = rng.standard_normal(len(x)).astype(np.float32)
    np.testing.assert_array_equal(_native_math(runtime, 2, y, x), oracle.math32(2, y, x))
    a = np.concatenate([rng.uniform(-1, 1, 500_000), [1, -1, 0.5, -0.5, 2]]).astype(np.float32)
    np.testing.assert_array_equal(_native_math(runtime, 3, a), oracle.math32(3, a))
    b = rng.uniform(-110, 95, 500_000).astype(np.float32)
    np.testing.assert_array_equal(_native_math(runtime, 4, b), oracle.math32(4, b))
    c = np.exp(rng.uniform(-100, 88, 500_000)).astype(np.float32)
    c = np.concatenate([c, np.float32([0, -1, np.inf, 1e-42])])
    np.testing.assert_array_equal(_native_math(runtime, 5, c), oracle.math32(5, c))
    s = np.abs(rng.standard_normal(500_000)).astype(np.float32) * np.exp(rng.uniform(-40, 40, 500_000)).astype(np.float32)
    np.testing.assert_array_equal(_native_math(runtime, 6, s), np.sqrt(s))
    d = rng.standard_normal(500_000).astype(np.float32)
    e = rng.standard_normal(500_000).astype(np.float32)
    np.testing.assert_array_equal(_native_math(runtime, 7, d, e), d / e)


@pytest.mark.parametrize("K", [45, 15, 1])
def test_sh_evaluate_bit_exact(runtime, K):
    rng = np.random.default_rng(K)
    params = rng.normal(size=(20000, K)).astype(np.float32)
    o = rng.standard_normal((20000, 3)).astype(np.float32)
    o[:4] = [[0, 0, 1], [0, 0, -1], [1e-8, 0, 1], [0, 0, 0]]
    model = tg.local.RealSymmetricSphericalHarmonics(K)
    np.testing.assert_array_equal(tg.local.evaluate(model, params, o, runtime), oracle.evaluate(0, K, params, o, 32))


def test_dti_evaluate_bit_exact_and_closed_form(runtime):
    rng = np.random.default_rng(9)
    params = rng.normal(size=(20000, 6)).astype(np.float32)
    o = rng.standard_normal((20000, 3)).astype(np.float32)
    model = tg.local.DiffusionTensorImaging()
    got = tg.local.evaluate(model, params, o, runtime)
    np.testing.assert_array_equal(got, oracle.evaluate(1, 6, params, o, 32))
    np.testing.assert_allclose(got, oracle.evaluate(1, 6, params, o, 64), rtol=1e-4, atol=1e-4)


def test_local_model_matches_numpy_twin(runtime):
    # test/test_local.py:27-64 through the device
    n = 100
    params = np.ones((n, 45))
    phi = np.linspace(0, 2 * np.pi, n, dtype=np.float32)
    theta = np.linspace(0, np.pi, n, dtype=np.float32)
    o = np.stack(tg.core.sph2cart(phi, theta, 1.0), axis=-1)
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    np.testing.assert_allclose(tg.local.evaluate(model, params, o, runtime), model.evaluate(params, o),
                               rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
@pytest.mark.parametrize("K", [45, 15, 6])
def test_update_orientation_bit_exact(runtime, name, K):
    rng = np.random.default_rng(K + len(name))
    model = tg.local.DiffusionTensorImaging() if K == 6 else tg.local.RealSymmetricSphericalHarmonics(K)
    params = rng.normal(size=(10000, K)).astype(np.float32)
    if K != 6:
        params[:, 0] = np.abs(params[:, 0]) + 1
    o = rng.standard_normal((10000, 3)).astype(np.float32)
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.3)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=4)
    twin = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=4)
    got = tg.algorithm.update_orientation(algorithm, params, o, runtime)
    seed = twin.get_rng_seed() if name == "diffusion" else 0
    np.testing.assert_array_equal(got, oracle.update(algorithm, seed, params, o, 32))


def test_transport_update_analytic_step(runtime):
    # test/test_algorithm.py:54-74 on the device
    params = fixtures.quadratic_fod_coefficients(0.1, 0.3)[None, :]
    o = np.array([tg.core.sph2cart(0.5, 0.8, 1.0)]).reshape(1, 3)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.04, gamma=1.7)
    got = tg.algorithm.update_orientation(algorithm, params, o, runtime)
    np.testing.assert_allclose(got[0], fixtures.expected_quadratic_update(o[0], 0.1, 0.3, 0.04, 1.7),
                               rtol=2e-4, atol=2e-4)
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-5)


def test_field_and_domain_lookups(runtime):
    # test/test_field.py and test/test_domain.py known answers on the device
    data = np.arange(64, dtype=np.float32).reshape(4, 4, 4, 1)
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    pts = np.array([[0.5, 1.5, 2.5], [1.5, 0.5, 0.5], [-0.5, 2.5, 2.5], [-0.51, 0, 0], [0, 4, 0]], np.float32)
    np.testing.assert_array_equal(tg.field.evaluate(field, pts, runtime), oracle.field_evaluate(field, pts))
    np.testing.assert_array_equal(tg.field.evaluate(field, pts, runtime)[:3, 0], [10, 32, 10])
    cube = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    pts = np.array([[-0.5, 0, 0], [0.49, 0.49, 0.49], [0.5, 0, 0]], np.float32)
    np.testing.assert_array_equal(tg.domain.contains(cube, pts, runtime), [True, True, False])
    rng = np.random.default_rng(2)
    mask = rng.random((9, 10, 11)) > 0.5
    affine = np.diag([1.5, 2.0, 0.7, 1.0])
    affine[:3, 3] = [3, -2, 1]
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), affine))
    pts = rng.uniform(-5, 25, (50000, 3)).astype(np.float32)
    np.testing.assert_array_equal(tg.domain.contains(domain, pts, runtime), oracle.domain_contains(domain, pts))
    values = rng.normal(size=(9, 10, 11, 45)).astype(np.float32)
    field = tg.field.Image(tg.Nifti1Image(values, affine))
    np.testing.assert_array_equal(tg.field.evaluate(field, pts, runtime), oracle.field_evaluate(field, pts))


def test_boundary_seeds_and_bvh_bit_exact(runtime):
    mask = np.zeros((12, 12, 12), bool)
    idx = np.indices(mask.shape).reshape(3, -1).T
    mask.reshape(-1)[np.linalg.norm(idx - 5.5, axis=1) < 5] = True
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), np.eye(4)))
    seeds = tg.domain.seed(domain, 5000, rng_seed=77, runtime=runtime)
    np.testing.assert_array_equal(seeds, oracle.domain_seed(domain, 77, 5000))
    assert np.all(domain.contains(seeds[:, :3]))
    rng = np.random.default_rng(5)
    starts = rng.uniform(-1, 12, (20000, 3)).astype(np.float32)
    ends = (starts + rng.standard_normal((20000, 3)) * 2).astype(np.float32)
    found, hit, t = tg.domain.bvh_segment_hit(domain, starts, ends, runtime)
    ofound, ohit, ot = oracle.bvh_hit(domain, starts, ends)
    np.testing.assert_array_equal(found, ofound)
    np.testing.assert_array_equal(hit[found], ohit[found])
    np.testing.assert_array_equal(t[found], ot[found])
    # test/test_domain.py:183-260
    unit = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    assert tg.domain.bvh_segment_hit(unit, np.zeros((1, 3)), np.array([[-1.0, 0, 0]]), runtime)[0][0]
    s = tg.domain.seed(unit, 32, rng_seed=1234, runtime=runtime)
    assert np.all(tg.domain.bvh_segment_hit(unit, s[:, :3], s[:, :3] + 2 * s[:, 4:7], runtime)[0])


def _compare_tracks(got, ref):
    np.testing.assert_array_equal(got.lengths, ref["lengths"])
    np.testing.assert_array_equal(got.steps, ref["steps"])
    np.testing.assert_array_equal(got.hits, ref["hits"])
    np.testing.assert_array_equal(got.points, oracle.ragged(ref))


@pytest.fixture(scope="module")
def cfg0():
    return bench.sphere32_transport(seed=0, n_seeds=3000)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_plain_tracking_bit_exact_sh45(runtime, cfg0, name, oracle_threads):
    field = tg.field.Image(tg.Nifti1Image(cfg0["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg0["mask"], np.eye(4)))
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=5)
    rng = [algorithm.get_rng_seed() for _ in range(3)]
    got = tg.tracking.track(algorithm, domain, field, cfg0["seeds"], rng, 1000, 0.25, 60.0, runtime)
    ref = oracle.track(domain, field, algorithm, cfg0["seeds"], rng, 1000, 0.25, 60.0, threads=oracle_threads)
    _compare_tracks(got, ref)
    assert got.steps.sum() > 100_000


def test_plain_tracking_bit_exact_generic_k_and_dti(runtime, oracle_threads):
    fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=15)
    field = tg.field.Image(fod)
    domain = tg.domain.Image(wm)
    np.random.seed(1)
    seeds = tg.seeds.from_mask(wm, 2000)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.05, gamma=2.0,
                                       noise_variance=0.2, seed=8)
    rng = [algorithm.get_rng_seed() for _ in range(4)]
    got = tg.tracking.track(algorithm, domain, field, seeds, rng, 600, 0.5, 40.0, runtime)
    _compare_tracks(got, oracle.track(domain, field, algorithm, seeds, rng, 600, 0.5, 40.0, threads=oracle_threads))
    tensor, wm, _ = fixtures.tensor_cross()
    field = tg.field.Image(tensor)
    domain = tg.domain.Image(wm)
    seeds = tg.seeds.from_mask(wm, 2000)
    algorithm = tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=0.01, gamma=20.0)
    got = tg.tracking.track(algorithm, domain, field, seeds, [0], 2000, 0.25, 30.0, runtime)
    _compare_tracks(got, oracle.track(domain, field, algorithm, seeds, [0], 2000, 0.25, 30.0, threads=oracle_threads))


def test_plain_tracking_is_chunk_invariant(runtime, cfg0):
    field = tg.field.Image(tg.Nifti1Image(cfg0["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg0["mask"], np.eye(4)))
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1,
                                       noise_variance=0.1, seed=3)
    rng = [algorithm.get_rng_seed() for _ in range(2)]
    whole = tg.tracking.track(algorithm, domain, field, cfg0["seeds"], rng, 2000, 0.25, 30.0, runtime)
    os.environ["TGK_CHUNK_BYTES"] = str(120 * 12 * 333)
    try:
        chunked = tg.tracking.track(algorithm, domain, field, cfg0["seeds"], rng, 2000, 0.25, 30.0, runtime)
    finally:
        del os.environ["TGK_CHUNK_BYTES"]
    for a, b in [(whole.lengths, chunked.lengths), (whole.steps, chunked.steps), (whole.points, chunked.points)]:
        np.testing.assert_array_equal(a, b)


def test_public_api_known_answers_on_device(runtime):
    # test/test_tracking.py:28-127 through tg.tracking.tractogram on the GPU
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((1, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], np.float32)
    t = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=1.0, gamma=0.0)
    r = tg.tracking.tractogram(t, domain, field, seeds, runtime=runtime)
    np.testing.assert_allclose(np.asarray(r.streamlines[0]), [[0, 0, 0], [1, 0, 0]], atol=1e-6)
    d = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(15), step_size=1.0, gamma=0.0,
                               noise_variance=0.0, seed=1234)
    r = tg.tracking.tractogram(d, domain, field, seeds, runtime=runtime)
    np.testing.assert_allclose(np.asarray(r.streamlines[0]), [[0, 0, 0], [1, 0, 0]], atol=1e-6)
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((20, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((20, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    t = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.25, gamma=0.0)
    s = np.asarray(tg.tracking.tractogram(t, domain, field, seeds, save_at=0.5, maximum_streamline_length=3.0,
                                          runtime=runtime).streamlines[0])
    assert len(s) == 6
    np.testing.assert_allclose(s[:, 0], np.arange(0.0, 3.0, 0.5), atol=1e-6)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_u0u3_exact_solution_on_device(runtime, name):
    # test/test_algorithms/test_transport.py:117-233, test_diffusion.py:15-71
    coefficients = fixtures.u0u3_coefficients(0.3, 45)
    fod = np.broadcast_to(coefficients, (1, 1, 1, 45)).copy()
    affine = np.diag((1000.0, 1000.0, 1000.0, 1.0))
    affine[:3, 3] = -0.5
    image = tg.Nifti1Image(fod, affine)
    orientation = np.asarray(tg.core.sph2cart(0.5, 0.12, 1.0))
    seeds = np.array([[10.0, 10.0, 10.0, 1.0, *orientation, 0.0]], np.float32)
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.001, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.001, gamma=1.0, noise_variance=0.0, seed=0)
    streamline = np.asarray(tg.tractogram(image, seeds, None, algorithm, save_at=0.001, maximum_streamline_length=1.0,
                                          batch_size=1, runtime=runtime).streamlines[0])
    u_ref = fixtures.u0u3_reference(0.001, 1.0, 0.12, 0.5, 0.3, len(streamline))
    expected = np.zeros((len(streamline), 3))
    expected[0] = [10.0, 10.0, 10.0]
    for k in range(1, len(streamline)):
        expected[k] = expected[k - 1] + 0.001 * u_ref[k]
    np.testing.assert_array_almost_equal(streamline, expected, decimal=3)


def test_public_tractogram_batches_and_boundary_seeds(runtime):
    # test/test_tractography.py:67-104 and the batch RNG semantics
    fod, mask, _ = fixtures.cross()
    assert len(tg.tractogram(fod, 8, maximum_streamline_length=2.0, runtime=runtime)) == 8
    assert len(tg.tractogram(tg.field.Image(fod), 8, tg.domain.Image(mask), maximum_streamline_length=2.0,
                             runtime=runtime)) == 8
    np.random.seed(0)
    seeds = tg.seeds.from_mask(mask, 500)
    a = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.05, noise_variance=0.1,
                               seed=9)
    result = tg.tractogram(fod, seeds, None, a, save_at=0.25, maximum_streamline_length=20.0,
                           minimum_streamline_length=1.0, batch_size=128, runtime=runtime)
    b = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.05, noise_variance=0.1,
                               seed=9)
    rng = [b.get_rng_seed() for _ in range(4)]
    ref = oracle.track(tg.domain.Image(tg.Nifti1Image(np.any(fod.dataobj != 0, axis=-1).astype(np.uint8), np.eye(4))),
                       tg.field.Image(fod), b, seeds, rng, 128, 0.25, 20.0)
    keep = ref["lengths"] > 4
    expected = [ref["points"][i, : ref["lengths"][i]] for i in np.flatnonzero(keep)]
    assert len(result.streamlines) == len(expected)
    for got, exp in zip(result.streamlines, expected):
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_plain_mesh_domain_bit_exact(runtime, cfg0, name, oracle_threads):
    # Termination on a two-surface boundary mesh (BVH per step segment):
    # lengths, steps, hit triangles and points equal the FP32 restatement.
    mesh, seeds, _, _ = fixtures.shell_scene(n_seeds=3000, seed=11)
    field = tg.field.Image(tg.Nifti1Image(cfg0["fod"], np.eye(4)))
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=6)
    rng = [algorithm.get_rng_seed() for _ in range(3)]
    got = tg.tracking.track(algorithm, mesh, field, seeds, rng, 1000, 0.25, 60.0, runtime)
    ref = oracle.track(mesh, field, algorithm, seeds, rng, 1000, 0.25, 60.0, threads=oracle_threads)
    _compare_tracks(got, ref)
    hit = got.hits != 0xFFFFFFFF
    print(f"[mesh {name}] {hit.mean():.3f} of streamlines end on the boundary, {int(got.steps.sum())} steps")
    assert hit.mean() > 0.3


def _segment_hits(runtime, device, fn, starts, ends):
    from tractography_amd import _native

    starts = np.ascontiguousarray(starts, np.float32)
    ends = np.ascontiguousarray(ends, np.float32)
    n = len(starts)
    found = np.zeros(n, np.uint8)
    hit = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    _native.check(getattr(_native.lib(), fn)(runtime.handle, device.handle, _native.ptr(starts), _native.ptr(ends), n,
                                             _native.ptr(found), _native.ptr(hit), _native.ptr(t)))
    return found.astype(bool), hit, t


def test_candidate_lists_answer_as_the_bvh(runtime):
    # Segments near the surfaces (where the lists are used), longer than the
    # lists' reach (traversal fallback) and off the grid: identical answers.
    mesh, seeds, centre, radii = fixtures.shell_scene(n_seeds=10, seed=21)
    device = mesh.upload(runtime)
    data = mesh._ensure_boundary_data()
    rng = np.random.default_rng(5)
    tri = data["vertices"][data["indices"][:, :3].astype(np.int64), :3]
    pick = rng.integers(0, len(tri), 200_000)
    w = rng.dirichlet([1, 1, 1], len(pick))
    on_surface = np.einsum("nk,nkc->nc", w, tri[pick])
    direction = rng.standard_normal((len(pick), 3))
    direction /= np.linalg.norm(direction, axis=1, keepdims=True)
    length = np.concatenate([rng.uniform(0.001, mesh.reach, len(pick) - 20_000), rng.uniform(0.6, 3.0, 20_000)])
    starts = on_surface - direction * length[:, None] * rng.uniform(0, 1, (len(pick), 1))
    ends = starts + direction * length[:, None]
    starts[:1000] += 100.0  # off the grid
    ends[:1000] += 100.0
    bvh = _segment_hits(runtime, device, "tgk_bvh_segment_hit", starts, ends)
    grid = _segment_hits(runtime, device, "tgk_grid_segment_hit", starts, ends)
    print(f"[candidate lists] {bvh[0].mean():.3f} of segments hit", flush=True)
    assert bvh[0].mean() > 0.3
    np.testing.assert_array_equal(grid[0], bvh[0])
    np.testing.assert_array_equal(grid[1][bvh[0]], bvh[1][bvh[0]])
    np.testing.assert_array_equal(grid[2][bvh[0]], bvh[2][bvh[0]])


def test_candidate_lists_refused_beyond_their_limits(runtime):
    # Offsets that decrease, or a cell listing 2^26 candidates or more (a
    # wavefront sums 64 lists in 32 bits), are refused before any entry is read.
    from tractography_amd import _native

    mesh, _, _, _ = fixtures.shell_scene(n_seeds=10, seed=21)
    device = mesh.upload(runtime)
    near, _, _ = mesh.culling_grid()
    n_cells = int(np.prod(near.shape))
    entries = np.zeros(12, dtype=np.float32)
    lib = _native.lib()
    for big in [1 << 26, (1 << 26) + 5]:
        cell_start = np.zeros(n_cells + 1, dtype=np.uint32)
        cell_start[n_cells // 2 + 1:] = big
        with pytest.raises(ValueError, match="2\\^26"):
            _native.check(lib.tgk_domain_set_segment_grid(device.handle, _native.ptr(cell_start), n_cells,
                                                           _native.ptr(entries), big))
    cell_start = np.zeros(n_cells + 1, dtype=np.uint32)
    cell_start[1] = 3
    cell_start[-1] = 1
    with pytest.raises(ValueError, match="decrease"):
        _native.check(lib.tgk_domain_set_segment_grid(device.handle, _native.ptr(cell_start), n_cells,
                                                       _native.ptr(entries), 1))
