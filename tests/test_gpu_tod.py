"""GPU parity of the track orientation distribution (tracking.py:181-385).

Each track's deposits are float32 sums computed in the same order on the
device and in the FP32 restatement, so a cell that receives one flush (or
none) matches bit for bit. A cell that receives m > 1 flushes sums the same
m float32 terms in the order the atomics land; any order of recursive
summation lies within gamma_{m-1} * sum|x_i| of the exact sum (Higham,
Accuracy and Stability of Numerical Algorithms, 2nd ed., eq. 4.4, with
gamma_k = k u / (1 - k u), u = 2^-24), so the device and the restatement
differ by at most twice that. The hardware float atomics flush subnormal
results, which adds at most one smallest normal (2^-126) per flush.
"""

import ctypes

import numpy as np
import pytest

import fixtures
import oracle
import tractography_amd as tg
from tractography_amd import _native

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TINY = 2.0**-126


def _device_tod(runtime, domain, field, algorithm, rng_seed, n_seeds, n_steps, dims, inverse, normalize=0,
                path=_native.PATH_PLAIN):
    handle = domain.upload(runtime, with_boundary=True)
    dims = np.ascontiguousarray(dims, dtype=np.uint32)
    inverse = np.ascontiguousarray(inverse, dtype=np.float32).reshape(16)
    out = np.full(tuple(int(d) for d in dims) + (45,), np.nan, dtype=np.float32)
    _native.check(_native.lib().tgk_tod(
        runtime.handle, handle.handle, handle.handle, field.upload(runtime).handle, algorithm.descriptor(), path,
        rng_seed, n_seeds, n_steps, _native.ptr(dims), _native.ptr(inverse), normalize, _native.ptr(out)))
    return out


def _reorder_bound(abs_sum, flushes):
    m = flushes[..., None].astype(np.float64)
    k = np.maximum(m - 1.0, 0.0)
    u = 2.0**-24
    return 2.0 * (k * u / (1.0 - k * u)) * abs_sum + m * TINY


def assert_same_sums(got, raw, abs_sum, flushes, label):
    """Bit-exact where a cell has at most one flush, within the reordering
    bound elsewhere; prints how much of the grid is exact."""
    single = flushes <= 1
    exact = got[single] == raw[single]
    subnormal = (np.abs(raw[single]) < TINY) & (got[single] == 0.0)
    assert np.all(exact | subnormal), f"{label}: single-flush cells differ"
    diff = np.abs(got.astype(np.float64) - raw.astype(np.float64))
    bound = _reorder_bound(abs_sum, flushes)
    worst = float(np.max(diff / np.maximum(bound, TINY)))
    print(f"[tod {label}] {single.mean():.3f} of cells single-flush; "
          f"{(got == raw).mean():.3f} of values bit-equal; worst diff/bound {worst:.3f}", flush=True)
    assert np.all(diff <= bound), f"{label}: reordered sums outside the bound"


def _scene(case):
    if case == "dti-volume":
        # A 24x28x24 volume of the DTI workload's tensors (smooth principal
        # axes in an ellipsoid): tracks leave the boundary at once, unlike
        # on the one-voxel-thick slabs below.
        import bench

        cfg = bench.brain145_dti_transport(seed=3, n_seeds=10, shape=(24, 28, 24))
        field = tg.field.Image(tg.Nifti1Image(cfg["fod"], np.eye(4)))
        domain = tg.domain.Image(tg.Nifti1Image(cfg["mask"], np.eye(4)))
        algorithm = tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=0.1, gamma=1.0)
        return domain, field, algorithm, cfg["mask"].shape
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
    elif case.endswith("probabilistic"):
        algorithm = tg.algorithm.Probabilistic(model, step_size=step, seed=9)
    elif case.endswith("deterministic"):
        algorithm = tg.algorithm.Deterministic(model, step_size=step)
    else:
        algorithm = tg.algorithm.Diffusion(model, step_size=step, gamma=gamma, noise_variance=0.1, seed=9)
    return domain, field, algorithm, wm.shape


@pytest.mark.parametrize("case", ["sh45-transport", "sh45-diffusion", "sh15-diffusion", "dti-transport"])
def test_tod_matches_the_restatement(runtime, case):
    domain, field, algorithm, shape = _scene(case)
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    got = _device_tod(runtime, domain, field, algorithm, seed, 4000, 200, dims, inverse)
    raw, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, 4000, 200, dims, inverse)
    assert flushes.sum() > 1000
    assert_same_sums(got, raw, abs_sum, flushes, case)


def test_tod_on_a_finer_shifted_grid(runtime):
    # Output cells half a domain voxel wide, offset by a quarter voxel, and
    # a grid that covers only part of the domain (tracks leave and re-enter).
    domain, field, algorithm, shape = _scene("sh45-diffusion")
    affine = np.diag([0.5, 0.5, 0.5, 1.0])
    affine[:3, 3] = [2.25, 1.75, -0.25]
    dims = np.array([16, 20, 7])
    inverse = np.linalg.inv(affine).astype(np.float32)
    seed = algorithm.get_rng_seed()
    got = _device_tod(runtime, domain, field, algorithm, seed, 3000, 200, dims, inverse)
    raw, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, 3000, 200, dims, inverse)
    assert flushes.sum() > 1000
    assert_same_sums(got, raw, abs_sum, flushes, "finer grid")


def test_tod_normalised_output(runtime):
    domain, field, algorithm, shape = _scene("sh45-transport")
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    got = _device_tod(runtime, domain, field, algorithm, seed, 3000, 200, dims, inverse, normalize=1)
    raw, _, flushes = oracle.tod(domain, field, algorithm, seed, 3000, 200, dims, inverse)
    populated = got[..., 0] != 0
    # Y_0 > 0, so a cell's l = 0 sum is positive exactly when it was flushed.
    np.testing.assert_array_equal(populated, flushes > 0)
    np.testing.assert_array_equal(got[..., 0][populated], np.float32(1.0 / np.sqrt(4.0 * np.pi)))
    np.testing.assert_array_equal(got[~populated], 0.0)
    np.testing.assert_allclose(got, oracle.normalize_odf(raw), rtol=1e-5, atol=1e-6)


def test_normalize_odf_bit_exact(runtime):
    rng = np.random.default_rng(4)
    odf = rng.standard_normal((500, 45)).astype(np.float32)
    odf[::7, 0] = 0.0
    odf[::11, 0] *= 1e-30
    odf[3, 5] = np.float32(1e-42)  # subnormal
    odf[5, 0] = np.nan
    odf[6, 0] = -0.0
    got = odf.copy()
    _native.check(_native.lib().tgk_normalize_odf(runtime.handle, len(got), _native.ptr(got)))
    np.testing.assert_array_equal(got, oracle.normalize_odf(odf))


def test_tod_device_entry_point_and_graph_replays(runtime):
    domain, field, algorithm, shape = _scene("sh45-diffusion")
    seed = algorithm.get_rng_seed()
    dims = np.ascontiguousarray(shape, dtype=np.uint32)
    inverse = np.ascontiguousarray(np.eye(4), dtype=np.float32).reshape(16)
    raw, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, 2000, 200, dims, inverse.reshape(4, 4))
    lib = _native.lib()
    handle = domain.upload(runtime, with_boundary=True)
    field_handle = field.upload(runtime)
    out = torch.full((int(np.prod(dims)) * 45,), float("nan"), dtype=torch.float32, device=f"cuda:{runtime.device}")
    torch.cuda.synchronize()

    def enqueue():
        return lib.tgk_tod_device(runtime.handle, handle.handle, handle.handle, field_handle.handle,
                                  algorithm.descriptor(), _native.PATH_PLAIN, seed, 2000, 200, _native.ptr(dims),
                                  _native.ptr(inverse), 0, out.data_ptr())

    _native.check(enqueue())
    runtime.finish()
    assert_same_sums(out.cpu().numpy().reshape(raw.shape), raw, abs_sum, flushes, "device entry point")
    _native.check(lib.tgk_graph_begin(runtime.handle))
    status = enqueue()
    graph = ctypes.c_void_p()
    end_status = lib.tgk_graph_end(runtime.handle, ctypes.byref(graph))
    _native.check(status)
    _native.check(end_status)
    try:
        for replay in range(3):
            out.fill_(float("nan"))  # eager work between replays
            torch.cuda.synchronize()
            _native.check(lib.tgk_graph_launch(graph))
            runtime.finish()
            got = out.cpu().numpy().reshape(raw.shape)
            assert_same_sums(got, raw, abs_sum, flushes, f"graph replay {replay}")
    finally:
        lib.tgk_graph_destroy(graph)


def test_tod_accumulates_boundary_seeded_track_orientations(runtime):
    # test/test_tracking.py:130-161
    description = tg.Nifti1Image(np.zeros((1, 1, 1), dtype=np.float32), np.eye(4))
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field_data = np.zeros((1, 1, 1, 15), dtype=np.float32)
    field_data[..., 0] = 1.0
    field = tg.field.Image(tg.Nifti1Image(field_data, np.eye(4)))
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.1, gamma=0.0,
                                       noise_variance=0.0, seed=1234)
    result = tg.tracking.tod(algorithm, domain, field, description, n_seeds=16, maximum_streamline_length=0.3,
                             runtime=runtime)
    assert result.shape == (1, 1, 1, 45)
    assert result.get_data_dtype() == np.dtype(np.float32)
    assert np.isfinite(result.get_fdata()).all()
    np.testing.assert_allclose(result.get_fdata()[0, 0, 0, 0], 1.0 / np.sqrt(4.0 * np.pi), atol=1e-6)


def test_tod_with_explicit_domain_and_algorithm(runtime):
    # test/test_tractography.py:18-38
    fod, seed_mask, _ = fixtures.cross()
    # test/test_algorithms/__init__.py:6-30 (make_algorithm's Diffusion defaults)
    algorithm = tg.algorithm.Diffusion(tg.local.from_shape(fod.shape), step_size=0.001, gamma=0.5,
                                       noise_variance=2.0)
    tod = tg.tod(fod, 32, seed_mask, algorithm, maximum_streamline_length=20.0, runtime=runtime)
    data = tod.get_fdata()
    assert tod.shape == fod.shape[:3] + (45,)
    assert np.isfinite(data).all()
    populated = data[..., 0] != 0
    assert np.any(populated)
    np.testing.assert_allclose(data[..., 0][populated], 1.0 / np.sqrt(4.0 * np.pi))


def test_tod_from_image_field_uses_defaults(runtime):
    # test/test_tractography.py:40-45
    fod, _, _ = fixtures.cross()
    result = tg.tod(tg.field.Image(fod), n_seeds=16, maximum_streamline_length=2.0, runtime=runtime)
    assert result.shape == fod.shape[:3] + (45,)


def test_tod_persistent_lanes_refill(runtime):
    # Many more seeds than the grid has lanes, with short tracks.
    domain, field, algorithm, shape = _scene("sh45-diffusion")
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    got = _device_tod(runtime, domain, field, algorithm, seed, 400_000, 12, dims, inverse)
    raw, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, 400_000, 12, dims, inverse)
    assert flushes.sum() > 400_000
    assert_same_sums(got, raw, abs_sum, flushes, "refill")


def test_tod_rejects_invalid_grids_and_domains(runtime):
    domain, field, algorithm, shape = _scene("sh45-transport")
    handle = domain.upload(runtime, with_boundary=True)
    lib = _native.lib()
    out = np.zeros(45, np.float32)

    def call(dims, inverse, dom=handle, path=_native.PATH_PLAIN):
        dims = np.ascontiguousarray(dims, dtype=np.uint32)
        inverse = np.ascontiguousarray(inverse, dtype=np.float32).reshape(16)
        return lib.tgk_tod(runtime.handle, dom.handle, dom.handle, field.upload(runtime).handle,
                           algorithm.descriptor(), path, 1, 10, 10, _native.ptr(dims),
                           _native.ptr(inverse), 0, _native.ptr(out))

    assert call([1, 1, 1], np.eye(4), path=7) == _native.ERR_INVALID
    assert b"unknown path" in lib.tgk_last_error()
    assert call([2**31, 2**31, 2**31], np.eye(4)) == _native.ERR_INVALID
    assert b"too large" in lib.tgk_last_error()
    bad = np.eye(4)
    bad[0, 3] = np.nan
    assert call([1, 1, 1], bad) == _native.ERR_INVALID
    assert b"finite" in lib.tgk_last_error()
    mesh, _, _, _ = fixtures.shell_scene(n_seeds=10, seed=1)
    assert call([1, 1, 1], np.eye(4), dom=mesh.upload(runtime)) == _native.ERR_INVALID
    assert b"mask domain" in lib.tgk_last_error()


Y00 = 1.0 / np.sqrt(4.0 * np.pi)


@pytest.mark.parametrize("case", ["sh45-transport", "sh45-diffusion", "sh15-diffusion", "dti-transport",
                                  "dti-volume", "sh45-probabilistic", "sh45-deterministic"])
def test_fast_tod_tracks_fp64(runtime, case):
    # The fast path's TOD against the FP64 restatement, with the plain
    # path's FP32 arithmetic (the reference's) measured the same way. Tracks
    # within ~1e-3 voxel of FP64 deposit the same steps in the same cells,
    # except for steps that close to a cell face or a track's end, so the
    # grids differ from FP64 by a small fraction of their mass; the fast
    # path is held to that fraction for the plain path, or 1e-3. A wrong
    # basis value, sign or index would move a large share of the mass. The
    # deposited steps (the l = 0 sums over Y_00) are held the same way. On
    # the one-voxel-thick DTI slab (dti-transport) boundary seeds start at
    # an unstable equilibrium, where rounding decides whether a track turns
    # (see the fast connectome test): both FP32 paths are far from FP64
    # there, and the plain path's distance is the yardstick.
    domain, field, algorithm, shape = _scene(case)
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    n, n_steps = 4000, 200
    fast = _device_tod(runtime, domain, field, algorithm, seed, n, n_steps, dims, inverse, path=_native.PATH_FAST)
    plain = _device_tod(runtime, domain, field, algorithm, seed, n, n_steps, dims, inverse)
    ref, _, flushes = oracle.tod(domain, field, algorithm, seed, n, n_steps, dims, inverse, precision=64)
    mass = float(np.abs(ref).sum())
    d_fast = float(np.abs(fast.astype(np.float64) - ref).sum()) / mass
    d_plain = float(np.abs(plain.astype(np.float64) - ref).sum()) / mass
    steps_ref, steps_fast, steps_plain = (float(x[..., 0].astype(np.float64).sum()) / Y00 for x in (ref, fast, plain))
    print(f"[fast tod {case}] L1 distance from FP64 over its mass: fast {d_fast:.2e} plain {d_plain:.2e}; "
          f"deposited steps FP64 {steps_ref:.0f} fast {steps_fast:.0f} plain {steps_plain:.0f}", flush=True)
    assert flushes.sum() > 1000
    assert np.isfinite(fast).all()
    assert d_fast <= max(1e-3, 2.0 * d_plain)
    assert abs(steps_fast - steps_ref) <= max(1e-3 * steps_ref, 2.0 * abs(steps_plain - steps_ref))


def test_fast_tod_basis_of_single_steps(runtime):
    # One deposit per track: tracks of one step (n_steps 1), so each cell
    # holds the basis at its tracks' start orientations, the inward normals
    # of the boundary faces. The fast path's angle-free basis must give the
    # FP64 basis there, summed over the same tracks (the starts are the same
    # on both paths), to FP32 rounding.
    domain, field, algorithm, shape = _scene("sh45-transport")
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    fast = _device_tod(runtime, domain, field, algorithm, seed, 20000, 1, dims, inverse, path=_native.PATH_FAST)
    ref, abs_sum, flushes = oracle.tod(domain, field, algorithm, seed, 20000, 1, dims, inverse, precision=64)
    assert flushes.sum() > 1000
    # Per cell: 1e-6 of its largest sum of magnitudes for the basis, plus
    # the float32 rounding of m summed deposits (m ulps). Relative to the
    # cell, not to each value: on the y axis FP64 gives cos(m pi / 2) =
    # 6e-17 where the exact value, and the fast basis's, is 0.
    scale = abs_sum.max(axis=-1, keepdims=True)
    bound = scale * (1e-6 + flushes[..., None] * 2.0**-23) + 1e-30
    err = np.abs(fast.astype(np.float64) - ref)
    print(f"[fast tod single steps] {int(flushes.sum())} deposits; worst error / bound {float(np.max(err / bound)):.3f}",
          flush=True)
    assert np.all(err <= bound)


def test_fast_tod_at_oblique_orientations(runtime):
    # A field whose single fibre runs along (1, 2, 3) / sqrt(14): tracks turn
    # off the axes at once, so every order m of the basis carries weight.
    shape = (12, 12, 12)
    rng = np.random.default_rng(5)
    direction = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    import bench

    fod = bench.fod_from_fibres(np.broadcast_to(direction, shape + (3,)).reshape(-1, 1, 3),
                                np.ones((int(np.prod(shape)), 1)), lmax=8).reshape(*shape, 45).astype(np.float32)
    mask = np.zeros(shape, np.uint8)
    mask[1:-1, 1:-1, 1:-1] = 1
    field = tg.field.Image(tg.Nifti1Image(fod, np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(mask, np.eye(4)))
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.05, seed=int(rng.integers(1, 1000)))
    seed = algorithm.get_rng_seed()
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    fast = _device_tod(runtime, domain, field, algorithm, seed, 3000, 150, dims, inverse, path=_native.PATH_FAST)
    plain = _device_tod(runtime, domain, field, algorithm, seed, 3000, 150, dims, inverse)
    ref, _, _ = oracle.tod(domain, field, algorithm, seed, 3000, 150, dims, inverse, precision=64)
    mass = float(np.abs(ref).sum())
    d_fast = float(np.abs(fast.astype(np.float64) - ref).sum()) / mass
    d_plain = float(np.abs(plain.astype(np.float64) - ref).sum()) / mass
    print(f"[fast tod oblique] L1 distance from FP64 over its mass: fast {d_fast:.2e} plain {d_plain:.2e}",
          flush=True)
    # Each order m's share of the mass, so that a basis error in any of them shows.
    for m in range(1, 9):
        idx = [l * (l + 1) // 2 + s * m for l in range(m + (m & 1), 9, 2) for s in (1, -1)]
        part = float(np.abs(ref[..., idx]).sum())
        assert part > 1e-4 * mass, f"order {m} carries no weight"
        d = float(np.abs(fast[..., idx].astype(np.float64) - ref[..., idx]).sum()) / part
        dp = float(np.abs(plain[..., idx].astype(np.float64) - ref[..., idx]).sum()) / part
        assert d <= max(2e-3, 2.0 * dp), f"order {m}: {d:.2e} (plain {dp:.2e})"


def test_fast_tod_zero_seeds_and_steps(runtime):
    # Nothing to track: the output is cleared (and normalised to zeros).
    domain, field, algorithm, shape = _scene("sh45-diffusion")
    dims = np.array(shape)
    inverse = np.eye(4, dtype=np.float32)
    for n, steps in ((0, 50), (100, 0)):
        got = _device_tod(runtime, domain, field, algorithm, 3, n, steps, dims, inverse, normalize=1,
                          path=_native.PATH_FAST)
        assert np.all(got == 0.0)


def test_fast_tod_through_the_public_api():
    fod, seed_mask, _ = fixtures.cross()
    algorithm = tg.algorithm.Transport(tg.local.from_shape(fod.shape), step_size=0.1, gamma=1.0)
    fast = tg.tod(fod, 2000, seed_mask, algorithm, maximum_streamline_length=20.0,
                  runtime=tg.HipRuntime(path="fast")).get_fdata()
    plain = tg.tod(fod, 2000, seed_mask, algorithm, maximum_streamline_length=20.0,
                   runtime=tg.HipRuntime()).get_fdata()
    populated = fast[..., 0] != 0
    assert np.any(populated)
    np.testing.assert_allclose(fast[..., 0][populated], Y00, rtol=1e-6)
    # Normalised cells: the same cells populated, shapes within tracking noise.
    assert (populated == (plain[..., 0] != 0)).mean() > 0.99
    both = populated & (plain[..., 0] != 0)
    assert float(np.abs(fast[both] - plain[both]).mean()) < 1e-2
