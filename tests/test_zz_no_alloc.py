"""tgk_compact_device_noalloc (the allocation-free compaction) against
tgk_compact_device through the C ABI: the same offsets and the same ragged
points, bit for bit, on calls in a row that share the library's fixed
scratch array (a larger call, then a smaller one, then the largest call the
scratch admits), and the size limit refused before anything runs."""

import numpy as np
import pytest

from tractography_amd import _native

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SENTINEL = -7.25


def _buffers(dev, n, n_points, seed):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(0, n_points + 1, n).astype(np.uint32)
    lengths[:4] = [0, n_points, 1, 0]
    dense = rng.standard_normal(n * n_points * 3).astype(np.float32)
    return (torch.from_numpy(dense).to(dev), torch.from_numpy(lengths.view(np.int32)).to(dev))


def _compact(fn, runtime, d_dense, d_lengths, n, n_points):
    dev = d_dense.device
    d_offsets = torch.full((n + 1,), -1, dtype=torch.int64, device=dev)
    d_points = torch.full((n * n_points * 3,), SENTINEL, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    _native.check(fn(runtime.handle, _native.PATH_FAST, d_dense.data_ptr(), n, n_points, d_lengths.data_ptr(),
                     d_offsets.data_ptr(), d_points.data_ptr()))
    runtime.finish()
    return d_offsets.cpu().numpy(), d_points.cpu().numpy()


def _check_same(runtime, dev, n, n_points, seed):
    lib = _native.lib()
    print(f"[no_alloc] n_seeds {n}, n_points {n_points}, seed {seed}", flush=True)
    d_dense, d_lengths = _buffers(dev, n, n_points, seed)
    off_free, pts_free = _compact(lib.tgk_compact_device_noalloc, runtime, d_dense, d_lengths, n, n_points)
    off_ref, pts_ref = _compact(lib.tgk_compact_device, runtime, d_dense, d_lengths, n, n_points)
    np.testing.assert_array_equal(off_free, off_ref)
    np.testing.assert_array_equal(pts_free.view(np.uint32), pts_ref.view(np.uint32))
    lengths = d_lengths.cpu().numpy().view(np.uint32).astype(np.int64)
    assert off_ref[-1] == lengths.sum()


def test_noalloc_compaction_equals_compact_device(runtime):
    dev = torch.device("cuda", runtime.device)
    # A larger call, then a smaller one: the second reuses the scratch the
    # first left behind.
    _check_same(runtime, dev, 600_001, 23, seed=1)
    _check_same(runtime, dev, 3001, 57, seed=2)
    # The largest call the 64 KiB scratch admits (one point per seed).
    _check_same(runtime, dev, _native.COMPACT_NOALLOC_MAX_SEEDS, 1, seed=3)


def test_noalloc_compaction_refuses_sizes_past_its_scratch(runtime):
    dev = torch.device("cuda", runtime.device)
    # Valid but small buffers: the size check refuses the call before any
    # kernel could read or write them.
    tiny = torch.zeros(16, dtype=torch.float32, device=dev)
    lib = _native.lib()
    status = lib.tgk_compact_device_noalloc(runtime.handle, _native.PATH_FAST, tiny.data_ptr(),
                                            _native.COMPACT_NOALLOC_MAX_SEEDS + 1, 1, tiny.data_ptr(),
                                            tiny.data_ptr(), tiny.data_ptr())
    assert status == _native.ERR_INVALID
    assert b"maximum" in lib.tgk_last_error()
    # Zero seeds: nothing to do, no error.
    assert lib.tgk_compact_device_noalloc(runtime.handle, _native.PATH_FAST, 0, 0, 1, 0, 0, 0) == _native.OK
