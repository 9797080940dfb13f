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
    keys = {
        autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-b", tg.algorithm.Diffusion(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Diffusion(sh91, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Transport(sh45, step_size=0.1), mask),
        autotune.cache_key("gpu-a", tg.algorithm.Transport(tg.local.DiffusionTensorImaging(), step_size=0.1), mask),
    }
    assert len(keys) == 5
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
