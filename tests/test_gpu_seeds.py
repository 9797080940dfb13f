"""GPU parity of seeding from FOD images (seeds.cl:12-118): the device VALU
kernel against the FP32 host restatement, bit for bit, on SH and DTI inputs,
and the reference's distributional checks (test/test_seeds.py:108-213)
through the public API."""

import numpy as np
import pytest

import bench
import fixtures
import oracle
import tractography_amd as tg

pytestmark = pytest.mark.gpu


def _states(n, seed):
    return np.random.default_rng(seed).integers(0, 2**32 - 1, (n, 2), dtype=np.uint32)


def _check_bit_exact(runtime, fod_image, model_id, K, n, seed):
    inline, voxels, affine = tg.seeds.sparse_fod(fod_image)
    states = _states(n, seed)
    got = tg.seeds.from_fod(fod_image, n, runtime=runtime, states=states)
    ref, ref_states = oracle.seeds_from_fod(model_id, K, inline, voxels, affine, states, threads=32)
    print(f"[seeds] model {model_id} K {K} nnz {len(voxels)} n {n}", flush=True)
    np.testing.assert_array_equal(got, ref)
    return got, ref_states


def test_sh45_seeds_bit_exact():
    runtime = tg.HipRuntime(0)
    cfg = bench.config0(seed=3, n_seeds=16)
    fod = cfg["fod"] * cfg["mask"][..., None]
    image = tg.Nifti1Image(fod, np.diag([1.5, 1.5, 2.0, 1.0]))
    _check_bit_exact(runtime, image, 0, 45, 30000, 11)


def _device_seeds(runtime, model_id, K, inline, voxels, affine, states):
    from tractography_amd import _native

    states = states.copy()
    out = np.zeros((len(states), 8), np.float32)
    _native.check(_native.lib().tgk_seeds_from_fod(
        runtime.handle, model_id, K, _native.ptr(inline), _native.ptr(voxels), len(voxels), _native.ptr(affine),
        _native.ptr(states), len(states), 0, _native.ptr(out)))
    return out, states


def test_sh15_seeds_bit_exact():
    # The public from_fod accepts 45 or 6 coefficients only (local.py:265-282);
    # the kernel takes any K <= 45.
    runtime = tg.HipRuntime(0)
    cfg = bench.config0(seed=4, n_seeds=16)
    fod = np.ascontiguousarray(cfg["fod"][..., :15] * cfg["mask"][..., None])
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, np.eye(4)))
    states = _states(10000, 12)
    got, got_states = _device_seeds(runtime, 0, 15, inline, voxels, affine, states)
    ref, ref_states = oracle.seeds_from_fod(0, 15, inline, voxels, affine, states, threads=32)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got_states, ref_states)


def test_dti_seeds_bit_exact():
    runtime = tg.HipRuntime(0)
    dti, mask, _ = fixtures.tensor_cross()
    fod = (dti.get_fdata() * mask.get_fdata()[..., None]).astype(np.float32)
    _check_bit_exact(runtime, tg.Nifti1Image(fod, dti.affine), 1, 6, 10000, 13)


def test_states_advance_as_on_the_host():
    runtime = tg.HipRuntime(0)
    sh, mask, _ = fixtures.cross()
    fod = (sh.get_fdata() * mask.get_fdata()[..., None]).astype(np.float32)
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, sh.affine))
    states = _states(5000, 14)
    got, got_states = _device_seeds(runtime, 0, 45, inline, voxels, affine, states)
    ref, ref_states = oracle.seeds_from_fod(0, 45, inline, voxels, affine, states, threads=32)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got_states, ref_states)
    assert not np.array_equal(got_states, states)


def test_sh_density_of_device_seeds():
    # test/test_seeds.py:163-187
    runtime = tg.HipRuntime(0)
    directions = tg.core.fibonacci_sphere(1000)
    sh_fod, _, _ = fixtures.cross()
    coefficients = sh_fod.get_fdata()[5, 1, 0]
    fod = np.zeros((1, 1, 1, 45), np.float32)
    fod[0, 0, 0] = coefficients
    seeds = tg.seeds.from_fod(tg.Nifti1Image(fod, np.eye(4)), 50000, runtime=runtime, states=_states(50000, 15))
    a, c, _ = tg.core.cart2sph(*directions.T)
    values = tg.core.ishtmtx(a, c, 45)[0] @ coefficients
    w = np.where(100 * values < 30, np.log1p(np.exp(np.clip(100 * values, -700, 30))) / 100, values)
    w = np.maximum(w, 0) / np.maximum(w, 0).sum()
    moment = np.einsum("ni,nj->ij", seeds[:, 4:7], seeds[:, 4:7]) / len(seeds)
    np.testing.assert_allclose(moment, np.einsum("n,ni,nj->ij", w, directions, directions), atol=2e-2)
    np.testing.assert_allclose(np.linalg.norm(seeds[:, 4:7], axis=1), 1, atol=1e-5)
    assert np.all(np.abs(seeds[:, :3]) <= 0.5)


def test_zero_fod_gives_the_default_direction():
    # A voxel whose model is nowhere positive has a zero total: no cumulative
    # sum exceeds the zero threshold and the seed keeps (0, 0, 1) (seeds.cl:56-61).
    runtime = tg.HipRuntime(0)
    fod = np.zeros((2, 1, 1, 6), np.float32)
    fod[0, 0, 0] = [-1e-3, -1e-3, -1e-3, 0, 0, 0]
    fod[1, 0, 0] = [1.7e-3, 3e-4, 3e-4, 0, 0, 0]
    image = tg.Nifti1Image(fod, np.eye(4))
    seeds = tg.seeds.from_fod(image, 4000, runtime=runtime, states=_states(4000, 16))
    first = seeds[:, 0] < 0.5
    assert first.any() and (~first).any()
    np.testing.assert_array_equal(seeds[first, 4:7], np.tile([0, 0, 1], (first.sum(), 1)))
    inline, voxels, affine = tg.seeds.sparse_fod(image)
    ref, _ = oracle.seeds_from_fod(1, 6, inline, voxels, affine, _states(4000, 16), threads=32)
    np.testing.assert_array_equal(seeds, ref)
