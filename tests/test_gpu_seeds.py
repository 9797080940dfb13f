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
    cfg = bench.sphere32_transport(seed=3, n_seeds=16)
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
    cfg = bench.sphere32_transport(seed=4, n_seeds=16)
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


# ------------------------------------------------------------ MFMA method ---

_MWC_A = np.uint64(4294883355)


def _randu_draws(states, n_draws):
    """randu (core.cl:216-225) draws 1..n_draws of every state, as float32."""
    x = states[:, 0].astype(np.uint64)
    c = states[:, 1].astype(np.uint64)
    out = np.empty((len(states), n_draws), np.float32)
    for d in range(n_draws):
        res = (x ^ c) & np.uint64(0xFFFFFFFF)
        t = x * _MWC_A + c
        x, c = t & np.uint64(0xFFFFFFFF), t >> np.uint64(32)
        out[:, d] = (res >> np.uint64(8)).astype(np.float32) * np.float32(2.0**-24)
    return out


def _fp64_weights(model_id, coefficients, directions):
    """Sampling weights max(f, 0) of every seed's model on the directions,
    FP64: softplus(k = 100) of the SH value, or the DTI quadratic form."""
    if model_id == 1:
        x, y, z = directions.T
        q = np.stack([x * x, y * y, z * z, 2 * x * y, 2 * x * z, 2 * y * z], 1)
        return np.maximum(coefficients.astype(np.float64) @ q.T, 0)
    K = coefficients.shape[1]
    lmax = 0
    while (lmax + 1) * (lmax + 2) // 2 < K:
        lmax += 2
    basis = bench.real_sh_basis(directions, lmax)[:, :K]
    v = coefficients.astype(np.float64) @ basis.T
    return np.maximum(np.where(100 * v < 30, np.log1p(np.exp(np.minimum(100 * v, 30))) / 100, v), 0)


def _check_selection_tolerance(model_id, inline, states, seeds, delta=1e-3):
    """Every selected direction j satisfies the reference's rule up to a
    relative tolerance on the FP64 cumulative weights S: S_j >= T - delta S_n
    and S_{j-1} <= T + delta S_n with T = u S_n (seeds.cl:46-61)."""
    directions = tg.core.fibonacci_sphere(1000)
    draws = _randu_draws(states, 5)
    index = np.minimum((draws[:, 0] * np.float32(len(inline))).astype(np.int64), len(inline) - 1)
    weights = _fp64_weights(model_id, inline[index], directions)
    S = np.cumsum(weights, axis=1)
    total = S[:, -1]
    chosen = np.argmax(seeds[:, 4:7].astype(np.float64) @ directions.T, axis=1)
    positive = total > 0
    np.testing.assert_array_equal(seeds[~positive, 4:7], np.tile([0, 0, 1], ((~positive).sum(), 1)))
    T = draws[:, 4].astype(np.float64) * total
    rows = np.arange(len(seeds))
    upper = S[rows, chosen]
    lower = np.where(chosen > 0, S[rows, np.maximum(chosen - 1, 0)], 0.0)
    ok = (upper >= T - delta * total) & (lower <= T + delta * total)
    bad = np.flatnonzero(positive & ~ok)
    assert bad.size == 0, f"{bad.size} selections outside the tolerance, first {bad[:5]}"


def _mfma_seeds(runtime, model_id, K, inline, voxels, affine, states):
    from tractography_amd import _native

    states = states.copy()
    out = np.zeros((len(states), 8), np.float32)
    _native.check(_native.lib().tgk_seeds_from_fod(
        runtime.handle, model_id, K, _native.ptr(inline), _native.ptr(voxels), len(voxels), _native.ptr(affine),
        _native.ptr(states), len(states), 1, _native.ptr(out)))
    return out, states


@pytest.mark.parametrize("case", ["sh45", "dti"])
def test_mfma_seeds_against_the_host(case):
    runtime = tg.HipRuntime(0)
    if case == "sh45":
        cfg = bench.sphere32_transport(seed=5, n_seeds=16)
        fod = cfg["fod"] * cfg["mask"][..., None]
        image, model_id, K = tg.Nifti1Image(fod, np.diag([1.5, 1.5, 2.0, 1.0])), 0, 45
    else:
        dti, mask, _ = fixtures.tensor_cross()
        fod = (dti.get_fdata() * mask.get_fdata()[..., None]).astype(np.float32)
        image, model_id, K = tg.Nifti1Image(fod, dti.affine), 1, 6
    inline, voxels, affine = tg.seeds.sparse_fod(image)
    states = _states(20011, 21)  # not a multiple of the 64-seed wavefront
    got, got_states = _mfma_seeds(runtime, model_id, K, inline, voxels, affine, states)
    ref, ref_states = oracle.seeds_from_fod(model_id, K, inline, voxels, affine, states, threads=32)
    # Voxel draws, jitter and the state sequence do not depend on the weights.
    np.testing.assert_array_equal(got[:, :4], ref[:, :4])
    np.testing.assert_array_equal(got_states, ref_states)
    same = np.all(got[:, 4:] == ref[:, 4:], axis=1)
    print(f"[seeds mfma] {case}: {same.mean():.5f} of directions identical to the FP32 host", flush=True)
    assert same.mean() >= 0.99
    _check_selection_tolerance(model_id, inline, states, got)


def test_mfma_seeds_order12():
    # SH order 12 (91 coefficients): no FP32 reference kernel exists; the
    # selections are checked against the FP64 cumulative weights.
    runtime = tg.HipRuntime(0)
    rng = np.random.default_rng(22)
    directions = rng.standard_normal((216, 2, 3))
    directions /= np.linalg.norm(directions, axis=-1, keepdims=True)
    weights = rng.dirichlet([1, 1], 216)
    fod = bench.fod_from_fibres(directions, weights, lmax=12).reshape(6, 6, 6, 91)
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, np.eye(4)))
    states = _states(12000, 23)
    got, got_states = _mfma_seeds(runtime, 0, 91, inline, voxels, affine, states)
    draws = _randu_draws(states, 4)
    index = np.minimum((draws[:, 0] * np.float32(len(inline))).astype(np.int64), len(inline) - 1)
    np.testing.assert_array_equal(got[:, :3], voxels[index, :3] + (draws[:, 1:4] - np.float32(0.5)))
    np.testing.assert_allclose(np.linalg.norm(got[:, 4:7], axis=1), 1, atol=1e-5)
    _check_selection_tolerance(0, inline, states, got)
