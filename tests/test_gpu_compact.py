"""Device compaction (tgk_compact_device) of both dense layouts against a
NumPy restatement of its contract: offsets are the exclusive scan of the
lengths; seed g's first min(lengths[g], n_points) points land at offsets[g];
nothing is written outside the n_seeds * n_points * 3 floats of the output,
whatever the lengths hold."""

import numpy as np
import pytest

import tractography_amd as tg
from tractography_amd import _native

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SENTINEL = np.float32(-7.25)


def _expected(dense_seed_major, lengths, n_points):
    n = len(lengths)
    capacity = n * n_points * 3
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths.astype(np.int64), out=offsets[1:])
    out = np.full(capacity, SENTINEL, dtype=np.float32)
    for g in range(n):
        count = min(int(lengths[g]), n_points) * 3
        first = int(offsets[g]) * 3
        hi = min(count, capacity - first)
        if hi > 0:
            out[first:first + hi] = dense_seed_major[g].reshape(-1)[:hi]
    return offsets, out


@pytest.mark.parametrize("path", [_native.PATH_PLAIN, _native.PATH_FAST])
@pytest.mark.parametrize("case", ["ragged", "oversize"])
def test_compaction_matches_its_contract(runtime, path, case):
    rng = np.random.default_rng(17 + path)
    n, n_points = 3001, 57
    if case == "ragged":
        lengths = rng.integers(0, n_points + 1, n).astype(np.uint32)
        lengths[:5] = [0, n_points, 1, 0, n_points]
    else:  # lengths past the rows: later seeds' offsets run past the output
        lengths = rng.integers(n_points, 3 * n_points, n).astype(np.uint32)
    seed_major = rng.standard_normal((n, n_points, 3)).astype(np.float32)
    dense = seed_major  # both paths write seed-major dense rows
    want_offsets, want_points = _expected(seed_major, lengths, n_points)
    dev = torch.device("cuda", runtime.device)
    d_dense = torch.from_numpy(dense.reshape(-1)).to(dev)
    d_lengths = torch.from_numpy(lengths.view(np.int32)).to(dev)
    d_offsets = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    d_points = torch.full((n * n_points * 3,), float(SENTINEL), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    _native.check(_native.lib().tgk_compact_device(runtime.handle, path, d_dense.data_ptr(), n, n_points,
                                                   d_lengths.data_ptr(), d_offsets.data_ptr(),
                                                   d_points.data_ptr()))
    runtime.finish()
    np.testing.assert_array_equal(d_offsets.cpu().numpy(), want_offsets)
    got = d_points.cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint32), want_points.view(np.uint32))
