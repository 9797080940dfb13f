"""The autotuner's JSON cache (host logic only): keys separate device,
model, coefficient count, algorithm and domain kind; entries survive a round
trip; a missing or malformed file reads as empty."""

import json

import numpy as np

import tractography_amd as tg
from tractography_amd import autotune


def test_cache_keys_separate_what_changes_the_kernel():
    sh45 = tg.local.RealSymmetricSphericalHarmonics(45)
    sh91 = tg.local.RealSymmetricSphericalHarmonics(91)
    mask = tg.domain.Image(tg.Nifti1Image(np.ones((2, 2, 2), dtype=np.uint8), np.eye(4)))
    vertices = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float32)
    faces = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int64)
    mesh = tg.domain.Mesh(vertices, faces)
    keys = {
        autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh45, step_size=0.1), mesh),
        autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-b", tg.algorithm.Diffusion(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh91, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Transport(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=0.1), mask),
    }
    assert len(keys) == 6
    # Step size and noise variance are not part of the key.
    assert autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh45, step_size=0.2, noise_variance=0.3), mask) in keys


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    assert autotune.load_cache(path) == {}
    autotune.store_cache("k1", {"blocks": 1024, "repack": 16384}, path)
    autotune.store_cache("k2", {"blocks": 0, "repack": 4096}, path)
    assert autotune.load_cache(path) == {"k1": {"blocks": 1024, "repack": 16384}, "k2": {"blocks": 0, "repack": 4096}}
    autotune.store_cache("k1", {"blocks": 2048, "repack": 65536}, path)
    assert autotune.load_cache(path)["k1"] == {"blocks": 2048, "repack": 65536}
    assert not list(path.parent.glob("*.tmp"))
    monkeypatch.setenv("TGK_AUTOTUNE_CACHE", str(path))
    assert autotune.cache_path() == path
    assert autotune.load_cache()["k2"]["repack"] == 4096


def test_malformed_cache_reads_as_empty(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("not json")
    assert autotune.load_cache(path) == {}
    path.write_text(json.dumps([1, 2, 3]))
    assert autotune.load_cache(path) == {}


def test_launch_shapes_and_grid_candidates():
    # The shapes the library compiles (lanes per block, coefficients in LDS)
    # and the grids tried for each: the same lanes per CU whatever the block.
    shapes = tg.runtime.fast_shapes()
    assert [(s["lanes_per_block"], s["lds_coefficients"]) for s in shapes] == [
        (256, 40), (128, 40), (512, 40), (256, 32), (256, 44)]
    assert autotune.DEFAULT_SHAPES == tuple(range(len(shapes)))
    from tractography_amd import _native

    assert _native.lib().tgk_fast_shape_info(len(shapes), None, None) == _native.ERR_INVALID
    assert _native.lib().tgk_fast_shape_info(-1, None, None) == _native.ERR_INVALID
    for s in shapes:
        grids = autotune._grids(256, s["lanes_per_block"], (0.5, 1.0, 2.0))
        assert grids[0] == 0
        assert [g * s["lanes_per_block"] for g in grids[1:]] == [131072, 262144, 524288]
