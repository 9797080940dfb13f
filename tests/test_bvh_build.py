"""Native host builders of libtgk.so against their NumPy restatements,
array for array (CPU): the BVH (domain.py:72-116) and a mesh domain's
culling grid and per-cell candidate lists."""

from __future__ import annotations

import ctypes
import sys
import time
from pathlib import Path

import numpy as np
import pytest

import tractography_amd as tg
from tractography_amd import _native

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import bvh_build  # noqa: E402  (NumPy restatement of the build)
import mesh_grid  # noqa: E402  (NumPy restatement of the candidate tables)

pytestmark = pytest.mark.skipif(not _native.library_available(), reason="libtgk.so not built")


def _boxes(rng, n, ties=False):
    centre = rng.normal(size=(n, 3)) * 20
    if ties:  # many exactly equal centroids along every axis
        centre = np.round(centre / 8) * 8
    half = rng.uniform(0.0, 1.0, size=(n, 3))
    return centre - half, centre + half


def _check_equal(tmin, tmax, leaf):
    got = tg.domain.build_bvh(tmin, tmax, leaf)
    want = bvh_build.build_bvh(tmin, tmax, leaf)
    for g, w, name in zip(got, want, ("bounds_min", "bounds_max", "nodes", "ordered")):
        assert g.dtype == w.dtype and g.shape == w.shape, name
        np.testing.assert_array_equal(g, w, err_msg=name)
    return got


@pytest.mark.parametrize("n", [1, 2, 4, 5, 9, 64, 1000, 4099])
@pytest.mark.parametrize("leaf", [1, 4, 7])
def test_native_build_equals_restatement(n, leaf):
    rng = np.random.default_rng(n * 31 + leaf)
    _check_equal(*_boxes(rng, n), leaf)


@pytest.mark.parametrize("n", [17, 3000])
def test_native_build_equal_centroids_keep_slot_order(n):
    rng = np.random.default_rng(n)
    _check_equal(*_boxes(rng, n, ties=True), 4)
    point = np.zeros((n, 3))  # every box the same point: all keys tie
    _check_equal(point, point, 4)


def test_native_build_nan_follows_numpy():
    rng = np.random.default_rng(5)
    tmin, tmax = _boxes(rng, 50)
    tmin[7, 1] = np.nan
    tmax[20, 2] = np.nan
    _check_equal(tmin, tmax, 4)


def test_native_build_invariants():
    rng = np.random.default_rng(2)
    tmin, tmax = _boxes(rng, 777)
    bmin, bmax, nodes, ordered = tg.domain.build_bvh(tmin, tmax, 4)
    assert sorted(ordered.tolist()) == list(range(777))
    for i, (left, right, first, count) in enumerate(nodes):
        if left < 0:
            assert right < 0 and 1 <= count <= 4
            slots = ordered[first : first + count]
            assert np.all(bmin[i, :3] == tmin[slots].min(axis=0).astype(np.float32))
        else:
            assert left == i + 1 and right > left  # pre-order, left first
            assert np.all(bmin[i, :3] <= np.minimum(bmin[left, :3], bmin[right, :3]))
            assert np.all(bmax[i, :3] >= np.maximum(bmax[left, :3], bmax[right, :3]))


def test_native_build_rejects_bad_shapes():
    with pytest.raises(ValueError):
        tg.domain.build_bvh(np.zeros((4, 2)), np.zeros((4, 2)))
    with pytest.raises(ValueError, match="empty BVH"):
        tg.domain.build_bvh(np.zeros((0, 3)), np.zeros((0, 3)))


def test_native_build_is_fast_on_a_large_mesh():
    rng = np.random.default_rng(3)
    tmin, tmax = _boxes(rng, 400_000)
    t0 = time.perf_counter()
    tg.domain.build_bvh(tmin, tmax, 4)
    assert time.perf_counter() - t0 < 3.0


@pytest.mark.parametrize("cell_size,reach", [(1.0, 0.5), (0.5, 0.75), (2.0, 0.25)])
def test_native_candidate_tables_equal_restatement(cell_size, reach):
    import bench

    rng = np.random.default_rng(11)
    shape = (40, 46, 40)
    vertices, faces, _, _ = bench.shell_domain(rng, shape, bench._semi_axes(shape, (15, 19, 15)),
                                               bench._semi_axes(shape, (7, 9, 7)), n_rings=24, n_segments=30,
                                               amplitude=1.0)
    domain = tg.domain.Mesh(vertices, faces, cell_size=cell_size, reach=reach)
    data = domain._ensure_boundary_data()
    dims, _, lo, hi = domain._grid_cells()
    near, _, got_reach = domain.culling_grid()
    np.testing.assert_array_equal(near, mesh_grid.culling_grid(dims, lo, hi))
    assert got_reach == np.float32(reach)
    cell_start, entries = domain.segment_grid()
    want_start, want_entries = mesh_grid.segment_grid(dims, lo, hi, data["bvh_triangles"], data["vertices"],
                                                      data["indices"])
    np.testing.assert_array_equal(cell_start, want_start)
    assert entries.shape == want_entries.shape
    np.testing.assert_array_equal(entries.view(np.uint32), want_entries.view(np.uint32))


def test_native_candidate_tables_reject_a_box_off_the_grid():
    from tractography_amd import _native as nat

    dims = np.array([2, 2, 2], dtype=np.int64)
    lo = np.array([[0, 0, 0]], dtype=np.int64)
    hi = np.array([[2, 0, 0]], dtype=np.int64)
    order = np.zeros(1, dtype=np.uint32)
    start = np.empty(9, dtype=np.uint32)
    n = ctypes.c_uint64()
    with pytest.raises(ValueError, match="leaves the grid"):
        nat.check(nat.lib().tgk_segment_grid_count(nat.ptr(dims), nat.ptr(lo), nat.ptr(hi), nat.ptr(order), 1,
                                                    nat.ptr(start), ctypes.byref(n)))
