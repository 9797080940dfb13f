"""GPU accuracy of the fast path against the FP64 host restatement, with the
plain path's deviation measured the same way for comparison.

Tolerance (benchmark definition): saved points within 1e-3 voxel of the FP64
trajectory over every prefix on which both used the same voxels."""

import numpy as np
import pytest

import bench
import deviation
import fixtures
import oracle
import tractography_amd as tg

pytestmark = pytest.mark.gpu

TOLERANCE = 1e-3


@pytest.fixture(scope="module")
def scene():
    cfg = bench.sphere32_transport(seed=2, n_seeds=2000)
    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
    return cfg, field, domain


@pytest.fixture(scope="module")
def fast_runtime():
    return tg.HipRuntime(path="fast")


@pytest.fixture(scope="module")
def brain145():
    # The benchmark workload (smooth fibre fields on the 145x174x145 grid),
    # with 2000 of its seeds: the tolerance is defined on this field.
    cfg = bench.brain145_diffusion(seed=0, n_seeds=2000)
    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
    return cfg, field, domain


@pytest.fixture(scope="module")
def brain145_dti():
    # The DTI workload: tensors whose principal axis follows the first smooth
    # direction field, isotropic outside the ellipsoid; Transport gamma 1,
    # streamlines of 1 to 3000 steps.
    cfg = bench.brain145_dti_transport(seed=0, n_seeds=2000)
    field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
    return cfg, field, domain


def _compare_tracks(got, ref):
    np.testing.assert_array_equal(got.lengths, ref["lengths"])
    np.testing.assert_array_equal(got.steps, ref["steps"])
    np.testing.assert_array_equal(got.points, oracle.ragged(ref))


def test_dti_workload_plain_bit_exact_and_fast_within_tolerance(fast_runtime, runtime, brain145_dti, oracle_threads):
    cfg, field, domain = brain145_dti
    algorithm = tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=cfg["step_size"],
                                       gamma=cfg["gamma"])
    args = (cfg["seeds"], [0], 2000, cfg["save_at"], cfg["max_length"])
    plain = tg.tracking.track(algorithm, domain, field, *args, runtime)
    _compare_tracks(plain, oracle.track(domain, field, algorithm, *args, precision=32, threads=oracle_threads))
    # Uneven lengths: from a single step up to the 3000-step cap.
    assert plain.steps.min() <= 5 and plain.steps.max() >= 2900
    ref = oracle.track(domain, field, algorithm, *args, precision=64, threads=oracle_threads)
    fast = tg.tracking.track(algorithm, domain, field, *args, fast_runtime)
    rf = deviation.report(fast.lengths, fast.steps, fast.points, ref)
    rp = deviation.report(plain.lengths, plain.steps, plain.points, ref)
    print(f"\n[dti145] fast {rf}\n[dti145] plain {rp}")
    assert rf["max_deviation"] <= TOLERANCE
    assert rf["compared_fraction"] > 0.9 * rp["compared_fraction"]
    assert np.all(fast.lengths >= 1) and np.all(fast.lengths <= 1200)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_fast_sh45_tracks_fp64_oracle(fast_runtime, runtime, brain145, name, oracle_threads):
    cfg, field, domain = brain145
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=21)
    rng = [algorithm.get_rng_seed()]
    args = (cfg["seeds"], rng, 2000, 0.25, 300.0)
    ref = oracle.track(domain, field, algorithm, *args, precision=64, threads=oracle_threads)
    fast = tg.tracking.track(algorithm, domain, field, *args, fast_runtime)
    plain = tg.tracking.track(algorithm, domain, field, *args, runtime)
    # The plain path's contract on the benchmark field: every streamline's
    # saved-point count, step count and points equal the FP32 restatement's.
    _compare_tracks(plain, oracle.track(domain, field, algorithm, *args, precision=32, threads=oracle_threads))
    rf = deviation.report(fast.lengths, fast.steps, fast.points, ref)
    rp = deviation.report(plain.lengths, plain.steps, plain.points, ref)
    print(f"\n[{name}] fast {rf}\n[{name}] plain {rp}")
    assert rf["max_deviation"] <= TOLERANCE
    # Comparison stops where a trajectory may have used another voxel than
    # the FP64 one; over 300-voxel streamlines that leaves about half of the
    # points, as for the reference's FP32 arithmetic.
    assert rf["compared_fraction"] > max(0.3, 0.9 * rp["compared_fraction"])
    # Every streamline starts at its seed and saves at most n_points points.
    assert np.all(fast.lengths >= 1) and np.all(fast.lengths <= 1200)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_fast_path_on_a_chaotic_field(fast_runtime, runtime, scene, name, oracle_threads):
    # Independent random fibres in every voxel: where the FOD is negative
    # along the current direction, softplus makes f tiny and the drift
    # (grad f / f) large, and trajectories amplify any rounding. There the
    # fast path must stay as close to the FP64 trajectory as the reference's
    # own FP32 arithmetic does.
    cfg, field, domain = scene
    model = tg.local.RealSymmetricSphericalHarmonics(45)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=21)
    rng = [algorithm.get_rng_seed()]
    args = (cfg["seeds"], rng, 2000, 0.25, 80.0)
    ref = oracle.track(domain, field, algorithm, *args, precision=64, threads=oracle_threads)
    fast = tg.tracking.track(algorithm, domain, field, *args, fast_runtime)
    plain = tg.tracking.track(algorithm, domain, field, *args, runtime)
    rf = deviation.report(fast.lengths, fast.steps, fast.points, ref)
    rp = deviation.report(plain.lengths, plain.steps, plain.points, ref)
    print(f"\n[chaotic {name}] fast {rf}\n[chaotic {name}] plain {rp}")
    assert rf["max_deviation"] <= max(TOLERANCE, 1.25 * rp["max_deviation"])
    assert rf["compared_fraction"] > 0.9 * rp["compared_fraction"]


def test_fast_generic_k_and_dti(fast_runtime, oracle_threads):
    fod, wm, _ = fixtures.cross(shape=(14, 14, 3), n_coefficients=15)
    field = tg.field.Image(fod)
    domain = tg.domain.Image(wm)
    np.random.seed(4)
    seeds = tg.seeds.from_mask(wm, 1500)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.05, gamma=2.0,
                                       noise_variance=0.2, seed=8)
    rng = [algorithm.get_rng_seed()]
    got = tg.tracking.track(algorithm, domain, field, seeds, rng, 1500, 0.5, 40.0, fast_runtime)
    ref = oracle.track(domain, field, algorithm, seeds, rng, 1500, 0.5, 40.0, precision=64, threads=oracle_threads)
    r = deviation.report(got.lengths, got.steps, got.points, ref)
    print(f"\n[sh15] fast {r}")
    assert r["max_deviation"] <= TOLERANCE
    tensor, wm, _ = fixtures.tensor_cross()
    field = tg.field.Image(tensor)
    domain = tg.domain.Image(wm)
    seeds = tg.seeds.from_mask(wm, 1500)
    algorithm = tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=0.01, gamma=20.0)
    got = tg.tracking.track(algorithm, domain, field, seeds, [0], 1500, 0.25, 30.0, fast_runtime)
    ref = oracle.track(domain, field, algorithm, seeds, [0], 1500, 0.25, 30.0, precision=64, threads=oracle_threads)
    r = deviation.report(got.lengths, got.steps, got.points, ref)
    print(f"\n[dti] fast {r}")
    assert r["max_deviation"] <= TOLERANCE


def test_fast_known_answers(fast_runtime):
    # test/test_tracking.py:28-100 on the fast path
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((20, 1, 1), np.uint8), np.eye(4)))
    data = np.zeros((20, 1, 1, 15), np.float32)
    data[..., 0] = 1
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], np.float32)
    t = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.25, gamma=0.0)
    s = np.asarray(tg.tracking.tractogram(t, domain, field, seeds, save_at=0.5, maximum_streamline_length=3.0,
                                          runtime=fast_runtime).streamlines[0])
    assert len(s) == 6
    np.testing.assert_allclose(s[:, 0], np.arange(0.0, 3.0, 0.5), atol=1e-6)


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_fast_u0u3_exact_solution(fast_runtime, name):
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
                                          batch_size=1, runtime=fast_runtime).streamlines[0])
    u_ref = fixtures.u0u3_reference(0.001, 1.0, 0.12, 0.5, 0.3, len(streamline))
    expected = np.zeros((len(streamline), 3))
    expected[0] = [10.0, 10.0, 10.0]
    for k in range(1, len(streamline)):
        expected[k] = expected[k - 1] + 0.001 * u_ref[k]
    np.testing.assert_array_almost_equal(streamline, expected, decimal=3)


@pytest.mark.parametrize("name", ["diffusion", "dti"])
def test_fast_path_axis_aligned_orientations(fast_runtime, runtime, scene, name, oracle_threads):
    # Orientations on the coordinate axes with signed zeros, as boundary
    # seeds' inward normals have them: the reference's cart2sph gives the
    # axis the azimuth atan2(y, x) of those zeros (0 or pi), which sets the
    # sign of the frame the Diffusion noise is drawn in.
    cfg, field, domain = scene
    n = 1600
    signs = np.array([[sx, sy] for sx in (0.0, -0.0) for sy in (0.0, -0.0)], np.float32)
    o = np.zeros((n, 3), np.float32)
    o[:, :2] = signs[np.arange(n) % 4]
    o[:, 2] = np.where(np.arange(n) % 8 < 4, 1.0, -1.0)
    seeds = np.ascontiguousarray(np.concatenate([cfg["seeds"][:n, :4], o, np.zeros((n, 1), np.float32)], axis=1),
                                 dtype=np.float32)
    assert np.array_equal(np.signbit(seeds[:, 4:6]), np.signbit(o[:, :2]))
    if name == "dti":
        params = np.zeros(cfg["fod"].shape[:3] + (6,), np.float32)
        params[..., :3] = 1.0
        params[..., 3] = 0.3
        field = tg.field.Image(tg.Nifti1Image(params, np.eye(4)))
        model = tg.local.DiffusionTensorImaging()
    else:
        model = tg.local.RealSymmetricSphericalHarmonics(45)
    algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=4)
    rng = [algorithm.get_rng_seed()]
    args = (seeds, rng, n, 0.25, 20.0)
    ref = oracle.track(domain, field, algorithm, *args, precision=64, threads=oracle_threads)
    fast = tg.tracking.track(algorithm, domain, field, *args, fast_runtime)
    plain = tg.tracking.track(algorithm, domain, field, *args, runtime)
    rf = deviation.report(fast.lengths, fast.steps, fast.points, ref)
    rp = deviation.report(plain.lengths, plain.steps, plain.points, ref)
    print(f"\n[axis {name}] fast {rf}\n[axis {name}] plain {rp}")
    # A tensor's axis can be an unstable equilibrium of the drift, where the
    # reference's own FP32 rounding moves tracks by 4.5e-3 voxel from FP64;
    # there the bar is the chaotic field's.
    assert rf["max_deviation"] <= max(TOLERANCE, 1.25 * rp["max_deviation"])
    assert rf["compared_fraction"] > 0.9 * rp["compared_fraction"]


def test_fast_path_is_deterministic_and_chunk_invariant(fast_runtime, scene):
    import os

    cfg, field, domain = scene
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1,
                                       noise_variance=0.1, seed=3)
    rng = [algorithm.get_rng_seed() for _ in range(2)]
    a = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 1000, 0.25, 30.0, fast_runtime)
    b = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 1000, 0.25, 30.0, fast_runtime)
    os.environ["TGK_CHUNK_BYTES"] = str(120 * 12 * 333)
    try:
        c = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 1000, 0.25, 30.0, fast_runtime)
    finally:
        del os.environ["TGK_CHUNK_BYTES"]
    for x, y in [(a, b), (a, c)]:
        np.testing.assert_array_equal(x.lengths, y.lengths)
        np.testing.assert_array_equal(x.points, y.points)


def test_fast_mesh_domain_tracks_fp64_oracle(fast_runtime, scene, oracle_threads):
    cfg, field, _ = scene
    mesh, seeds, _, _ = fixtures.shell_scene(n_seeds=2000, seed=12)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=22)
    rng = [algorithm.get_rng_seed()]
    got = tg.tracking.track(algorithm, mesh, field, seeds, rng, 2000, 0.25, 80.0, fast_runtime)
    ref = oracle.track(mesh, field, algorithm, seeds, rng, 2000, 0.25, 80.0, precision=64, threads=oracle_threads)
    r = deviation.report(got.lengths, got.steps, got.points, ref)
    same = got.lengths == ref["lengths"]
    hits_equal = np.mean(got.hits[same] == ref["hits"][same])
    print(f"\n[mesh] fast {r}, hit ids equal on {hits_equal:.4f} of equal-length streamlines")
    assert r["max_deviation"] <= TOLERANCE
    assert hits_equal > 0.99


def test_fast_mesh_candidate_lists_equal_the_bvh_path(fast_runtime, scene):
    # The fast kernel answers a step's boundary test from its start cell's
    # candidate list, the lanes' lists spread over the wavefront; with no
    # reach (reach 0) every test traverses the BVH. Both must give the same
    # hits, hence the same tracks, bit for bit.
    import bench

    cfg, field, _ = scene
    outer, inner, amplitude = (13.0, 14.0, 12.0), (5.0, 6.0, 5.0), 1.0
    geometry = np.random.default_rng(13)
    vertices, faces, centre, radii = bench.shell_domain(geometry, (32, 32, 32), outer, inner, 30, 36, amplitude)
    seeds = bench.seeds_in_shell(geometry, centre, radii, np.asarray(outer) + amplitude + 1.0, 4000)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=23)
    rng = [algorithm.get_rng_seed()]
    results = []
    for cell_size, reach in [(0.5, 0.15), (1.0, 0.5), (1.0, 0.0)]:
        domain = tg.domain.Mesh(vertices, faces, cell_size=cell_size, reach=reach)
        results.append(tg.tracking.track(algorithm, domain, field, seeds, rng, 4000, 0.25, 80.0, fast_runtime))
    hits = np.mean(results[-1].hits != np.iinfo(np.uint32).max)
    print(f"\n[mesh lists] {hits:.3f} of streamlines end on a boundary hit")
    assert hits > 0.3
    for r in results[:-1]:
        np.testing.assert_array_equal(r.lengths, results[-1].lengths)
        np.testing.assert_array_equal(r.steps, results[-1].steps)
        np.testing.assert_array_equal(r.hits, results[-1].hits)
        np.testing.assert_array_equal(r.points, results[-1].points)


def _order12_scene(n_seeds=1500, seed=30):
    rng = np.random.default_rng(seed)
    shape = (24, 24, 24)
    directions = bench.smooth_direction_fields(rng, shape, sigma=3.0)
    _, weights = bench._random_fibres(rng, int(np.prod(shape)), base_directions=directions)
    fod = bench.fod_from_fibres(directions, weights, lmax=12).reshape(*shape, 91)
    idx = np.indices(shape).reshape(3, -1).T
    mask = (np.linalg.norm(idx - 11.5, axis=1) <= 10).reshape(shape)
    fod[~mask] = 0
    seeds = bench.seeds_in_mask(rng, mask, n_seeds)
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), np.eye(4)))
    return field, domain, seeds


@pytest.mark.parametrize("name", ["transport", "diffusion"])
def test_fast_sh_order12_tracks_fp64_oracle(fast_runtime, name, oracle_threads):
    # SH order 12 (91 coefficients) exists only on the fast path; the FP64
    # oracle evaluates the closed-form basis (pinned against core.ishtmtx).
    field, domain, seeds = _order12_scene()
    model = tg.local.RealSymmetricSphericalHarmonics(91)
    if name == "transport":
        algorithm = tg.algorithm.Transport(model, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=31)
    rng = [algorithm.get_rng_seed()]
    got = tg.tracking.track(algorithm, domain, field, seeds, rng, 1500, 0.25, 60.0, fast_runtime)
    ref = oracle.track(domain, field, algorithm, seeds, rng, 1500, 0.25, 60.0, precision=64, threads=oracle_threads)
    r = deviation.report(got.lengths, got.steps, got.points, ref)
    print(f"\n[sh91 {name}] fast {r}")
    assert r["max_deviation"] <= TOLERANCE
    assert r["compared_fraction"] > 0.5


def test_plain_path_refuses_order12(runtime):
    field, domain, seeds = _order12_scene(n_seeds=8)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(91), step_size=0.1, gamma=1.0)
    with pytest.raises(NotImplementedError, match="91"):
        tg.tracking.track(algorithm, domain, field, seeds, [0], 8, 0.25, 10.0, runtime)


def test_fast_results_do_not_depend_on_the_grid(scene):
    # Lanes pull streamlines from a queue, so the grid changes which lane
    # tracks a seed but never the result; the tuner only picks the fastest.
    cfg, field, domain = scene
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=41)
    rng = [algorithm.get_rng_seed()]
    runtime = tg.HipRuntime(path="fast")
    runtime.set_fast_grid(1)
    one_block = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 2000, 0.25, 80.0, runtime)
    tuned = tg.autotune.tune_fast_grid(algorithm, domain, field, cfg["seeds"][:500], runtime=runtime, save_at=0.25,
                                       maximum_streamline_length=80.0, repeats=1)
    assert tuned["blocks"] in tuned["timings"]
    got = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 2000, 0.25, 80.0, runtime)
    np.testing.assert_array_equal(got.lengths, one_block.lengths)
    np.testing.assert_array_equal(got.steps, one_block.steps)
    np.testing.assert_array_equal(got.points, one_block.points)


@pytest.mark.parametrize("precision", ["bf16x2", "fp16", "bf16"])
def test_fast_path_refuses_16_bit_coefficient_storage(scene, precision):
    cfg, field, domain = scene
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1)
    seeds = cfg["seeds"][:8]
    with pytest.raises(NotImplementedError, match="only FP32"):
        tg.tracking.tractogram(algorithm, domain, field, seeds, save_at=0.25, maximum_streamline_length=5.0,
                               runtime=tg.HipRuntime(path="fast", precision=precision))
    got = tg.tracking.tractogram(algorithm, domain, field, seeds, save_at=0.25, maximum_streamline_length=5.0,
                                 runtime=tg.HipRuntime(path="fast"))
    assert len(got.streamlines) == len(seeds)


@pytest.mark.parametrize("case", ["transport", "diffusion", "mesh"])
def test_fast_results_do_not_depend_on_the_launch_shape(scene, case):
    # Lanes per block and the coefficients kept in LDS (32, 40 or 44 of 45)
    # change where a lane's data waits, not its arithmetic: every shape
    # gives shape 0's results bit for bit.
    cfg, field, domain = scene
    seeds = cfg["seeds"]
    sh45 = tg.local.RealSymmetricSphericalHarmonics(45)
    if case == "transport":
        algorithm = tg.algorithm.Transport(sh45, step_size=0.1, gamma=1.0)
    else:
        algorithm = tg.algorithm.Diffusion(sh45, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=58)
    if case == "mesh":
        domain, seeds, _, _ = fixtures.shell_scene(n_seeds=2000, seed=13)
    rng = [algorithm.get_rng_seed()]
    results = []
    shapes = tg.runtime.fast_shapes() + [tg.runtime.fast_shapes()[0]]  # shape 0 again: a control
    for shape in shapes:
        runtime = tg.HipRuntime(path="fast")
        runtime.set_fast_shape(shape["shape"])
        results.append(tg.tracking.track(algorithm, domain, field, seeds, rng, len(seeds), 0.25, 80.0, runtime))
    first = results[0]
    assert first.steps.sum() > 10000
    differing = []
    for shape, got in zip(shapes[1:], results[1:]):
        seeds_differing = int(np.sum((got.lengths != first.lengths) | (got.steps != first.steps)))
        print(f"\n[shape] {case} {shape}: {seeds_differing} seeds differ from shape 0", flush=True)
        differing.append(seeds_differing)
    assert differing == [0] * len(differing)
    for got in results[1:]:
        np.testing.assert_array_equal(got.hits, first.hits)
        np.testing.assert_array_equal(got.points.view(np.uint32), first.points.view(np.uint32))


@pytest.mark.parametrize("case", ["sh45-diffusion", "sh45-probabilistic", "mesh", "dti", "sh91-diffusion"])
def test_fast_generations_change_no_result(scene, case):
    # A fast-path call runs as generations of the persistent kernel: when a
    # generation's work is all claimed and half of it has ended, the
    # streamlines in flight are saved and resumed by the next generation in
    # fuller wavefronts. A threshold of 8 makes every call here run all four
    # generations; the results must equal a single run to completion bit
    # for bit (noise streams, saved-point counters and arc-length clocks
    # carried across).
    cfg, field, domain = scene
    seeds = cfg["seeds"]
    sh45 = tg.local.RealSymmetricSphericalHarmonics(45)
    max_length = 80.0
    if case == "sh45-diffusion":
        algorithm = tg.algorithm.Diffusion(sh45, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=51)
    elif case == "sh45-probabilistic":
        algorithm = tg.algorithm.Probabilistic(sh45, step_size=0.1, seed=52)
        max_length = 20.0
    elif case == "mesh":
        domain, seeds, _, _ = fixtures.shell_scene(n_seeds=2000, seed=12)
        algorithm = tg.algorithm.Diffusion(sh45, step_size=0.1, gamma=1.0, noise_variance=0.1, seed=53)
    elif case == "dti":
        tensor, wm, _ = fixtures.tensor_cross()
        field = tg.field.Image(tensor)
        domain = tg.domain.Image(wm)
        algorithm = tg.algorithm.Diffusion(tg.local.DiffusionTensorImaging(), step_size=0.1, gamma=20.0,
                                           noise_variance=0.1, seed=54)
        seeds = tg.seeds.from_mask(wm, 2000)
    else:
        field, domain, seeds = _order12_scene(n_seeds=2000, seed=56)
        algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(91), step_size=0.1,
                                           gamma=1.0, noise_variance=0.1, seed=57)
    rng = [algorithm.get_rng_seed()]
    whole = tg.HipRuntime(path="fast")
    whole.set_fast_repack(1 << 30)
    repacked = tg.HipRuntime(path="fast")
    repacked.set_fast_repack(8)
    a = tg.tracking.track(algorithm, domain, field, seeds, rng, len(seeds), 0.25, max_length, whole)
    b = tg.tracking.track(algorithm, domain, field, seeds, rng, len(seeds), 0.25, max_length, repacked)
    print(f"\n[{case}] {int(a.steps.sum())} steps, {len(a.points)} points")
    np.testing.assert_array_equal(b.lengths, a.lengths)
    np.testing.assert_array_equal(b.steps, a.steps)
    np.testing.assert_array_equal(b.hits, a.hits)
    np.testing.assert_array_equal(b.points.view(np.uint32), a.points.view(np.uint32))


def test_tune_fast_caches_its_choice_and_changes_no_result(scene, tmp_path):
    cfg, field, domain = scene
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=61)
    rng = [algorithm.get_rng_seed()]
    default = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 2000, 0.25, 80.0,
                                tg.HipRuntime(path="fast"))
    runtime = tg.HipRuntime(path="fast")
    cache = tmp_path / "tune.json"
    kwargs = dict(runtime=runtime, save_at=0.25, maximum_streamline_length=80.0, grid_factors=(1.0,),
                  repacks=(64, 1 << 30), shapes=(0, 1, 4), repeats=1, cache=cache)
    tuned = tg.autotune.tune_fast(algorithm, domain, field, cfg["seeds"][:1000], **kwargs)
    print(f"\n[tune] {tuned}")
    # Three shapes, then two grids x two thresholds with the fastest.
    assert not tuned["cached"] and len(tuned["timings"]) == 7 and tuned["milliseconds"] > 0
    again = tg.autotune.tune_fast(algorithm, domain, field, cfg["seeds"][:1000], **kwargs)
    assert again["cached"] and (again["shape"], again["blocks"], again["repack"]) == (
        tuned["shape"], tuned["blocks"], tuned["repack"])
    assert runtime.fast_shape == tuned["shape"]
    got = tg.tracking.track(algorithm, domain, field, cfg["seeds"], rng, 2000, 0.25, 80.0, runtime)
    np.testing.assert_array_equal(got.lengths, default.lengths)
    np.testing.assert_array_equal(got.steps, default.steps)
    np.testing.assert_array_equal(got.points.view(np.uint32), default.points.view(np.uint32))
