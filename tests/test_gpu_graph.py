"""Recorded HIP graphs of the device pipeline (FOD seeding, tracking,
compaction): every replay must equal the eager run bit for bit, with eager
work and synchronisations between replays, and recording must refuse work
that would allocate."""

import ctypes

import numpy as np
import pytest

import bench
import tractography_amd as tg
from tractography_amd import _native

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def sphere32_scene():
    """The 32^3 SH45 sphere workload: FOD, mask, Diffusion."""
    cfg = bench.sphere32_transport(seed=7, n_seeds=16)
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=3)
    return cfg["fod"], cfg["mask"], algorithm


def order12_scene():
    """A 24^3 field of order 12 (91 coefficients) built as the order-12
    benchmark field is (smooth fibre directions, Funk-Hecke projection), in a
    sphere of radius 10; Diffusion."""
    rng = np.random.default_rng(30)
    shape = (24, 24, 24)
    directions = bench.smooth_direction_fields(rng, shape, sigma=3.0)
    _, weights = bench._random_fibres(rng, int(np.prod(shape)), base_directions=directions)
    fod = bench.fod_from_fibres(directions, weights, lmax=12).reshape(*shape, 91).astype(np.float32)
    idx = np.indices(shape).reshape(3, -1).T
    mask = (np.linalg.norm(idx - 11.5, axis=1) <= 10).reshape(shape).astype(np.uint8)
    fod[mask == 0] = 0
    algorithm = tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(91), step_size=0.1, gamma=1.0,
                                       noise_variance=0.1, seed=33)
    return fod, mask, algorithm


class Pipeline:
    """Device buffers for seeding `counts` chunks of seeds from a sparse FOD
    and tracking them (one dense/points region per chunk)."""

    def __init__(self, runtime, counts, path, method, n_points=120, save_at=0.25, scene=sphere32_scene):
        fod_data, mask, algorithm = scene()
        self.runtime = runtime
        self.lib = _native.lib()
        self.path, self.method = path, method
        self.counts = list(counts)
        self.n_points, self.save_at = n_points, save_at
        self.K = fod_data.shape[-1]
        fod = tg.Nifti1Image(fod_data * mask[..., None], np.eye(4))
        inline, voxels, affine = tg.seeds.sparse_fod(fod)
        self.field = tg.field.Image(tg.Nifti1Image(fod_data, np.eye(4)))
        self.domain = tg.domain.Image(tg.Nifti1Image(mask, np.eye(4)))
        self.algorithm = algorithm
        self.field_handle = self.field.upload(runtime)
        self.domain_handle = self.domain.upload(runtime)
        dev = torch.device("cuda", runtime.device)
        n = sum(self.counts)
        self.n = n
        self.nnz = len(voxels)
        self.d_fod = torch.from_numpy(inline).to(dev)
        self.d_voxels = torch.from_numpy(voxels).to(dev)
        self.d_affine = torch.from_numpy(affine).to(dev)
        states = np.random.default_rng(8).integers(0, 2**32 - 1, (n, 2), dtype=np.uint32)
        self.d_states0 = torch.from_numpy(states.view(np.int32)).to(dev)
        self.d_states = self.d_states0.clone()
        self.d_seeds = torch.zeros((n, 8), dtype=torch.float32, device=dev)
        rng = np.array([self.algorithm.get_rng_seed() for _ in range((n + 49_999) // 50_000)], dtype=np.uint32)
        self.rng_seeds = rng
        self.d_rng = torch.from_numpy(rng.view(np.int32)).to(dev)
        self.d_lengths = torch.zeros(n, dtype=torch.int32, device=dev)
        self.d_steps = torch.zeros(n, dtype=torch.int32, device=dev)
        self.d_hits = torch.zeros(n, dtype=torch.int32, device=dev)
        self.d_dense = torch.zeros(n * n_points * 3, dtype=torch.float32, device=dev)
        self.d_offsets = torch.zeros(n + len(self.counts), dtype=torch.int64, device=dev)
        self.d_points = torch.zeros(n * n_points * 3, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()

    def prepare(self):
        _native.check(self.lib.tgk_track_prepare(self.runtime.handle, self.domain_handle.handle,
                                                 self.field_handle.handle, self.algorithm.descriptor(), self.path,
                                                 max(self.counts)))
        _native.check(self.lib.tgk_seeds_prepare(self.runtime.handle, 0, self.K, self.method))

    def enqueue(self):
        """Seeding, tracking and compaction of every chunk on the runtime
        stream; returns the first failing status (0 if none)."""
        first = 0
        for c, count in enumerate(self.counts):
            f4 = 4  # bytes per float / int32
            seeds = self.d_seeds.data_ptr() + first * 8 * f4
            status = self.lib.tgk_seeds_from_fod_device(
                self.runtime.handle, 0, self.K, self.d_fod.data_ptr(), self.d_voxels.data_ptr(), self.nnz,
                self.d_affine.data_ptr(), self.d_states.data_ptr() + first * 2 * f4, count, self.method, seeds)
            if status:
                return status
            dense = self.d_dense.data_ptr() + first * self.n_points * 3 * f4
            status = self.lib.tgk_track_device(
                self.runtime.handle, self.domain_handle.handle, self.field_handle.handle,
                self.algorithm.descriptor(), self.path, seeds, first, count, self.d_rng.data_ptr(), 50_000,
                self.n_points, self.save_at, self.d_lengths.data_ptr() + first * f4,
                self.d_steps.data_ptr() + first * f4, self.d_hits.data_ptr() + first * f4, dense)
            if status:
                return status
            status = self.lib.tgk_compact_device(
                self.runtime.handle, self.path, dense, count, self.n_points, self.d_lengths.data_ptr() + first * f4,
                self.d_offsets.data_ptr() + (first + c) * 8, self.d_points.data_ptr() + first * self.n_points * 3 * f4)
            if status:
                return status
            first += count
        return 0

    def reset(self):
        """Eager work between runs: fresh states, cleared outputs."""
        self.d_states.copy_(self.d_states0)
        for t in (self.d_seeds, self.d_lengths, self.d_steps, self.d_hits, self.d_offsets, self.d_points):
            t.zero_()
        torch.cuda.synchronize()

    def results(self):
        self.runtime.finish()
        torch.cuda.synchronize()
        out = {name: getattr(self, "d_" + name).cpu().numpy().copy()
               for name in ("seeds", "states", "lengths", "steps", "hits", "offsets")}
        first, points = 0, []
        for c, count in enumerate(self.counts):
            total = int(out["offsets"][first + c + count])
            base = first * self.n_points * 3
            points.append(self.d_points[base:base + total * 3].cpu().numpy())
            first += count
        out["points"] = np.concatenate(points)
        return out


@pytest.mark.parametrize("path,method", [(1, 1), (0, 0)])
def test_graph_replays_equal_the_eager_run(path, method):
    runtime = tg.HipRuntime(0)
    pipe = Pipeline(runtime, [1500, 700, 2300], path, method)
    pipe.prepare()
    pipe.reset()
    _native.check(pipe.enqueue())
    eager = pipe.results()
    assert eager["steps"].sum() > 10_000
    pipe.reset()
    _native.check(pipe.lib.tgk_graph_begin(runtime.handle))
    status = pipe.enqueue()
    graph = ctypes.c_void_p()
    end_status = pipe.lib.tgk_graph_end(runtime.handle, ctypes.byref(graph))
    _native.check(status)
    _native.check(end_status)
    try:
        for replay in range(3):
            pipe.reset()
            _native.check(pipe.lib.tgk_graph_launch(graph))
            got = pipe.results()
            print(f"[graph path {path} method {method}] replay {replay}", flush=True)
            for name, value in eager.items():
                np.testing.assert_array_equal(got[name], value, err_msg=f"replay {replay}: {name}")
    finally:
        pipe.lib.tgk_graph_destroy(graph)


def test_recording_refuses_unprepared_work():
    runtime = tg.HipRuntime(0)
    pipe = Pipeline(runtime, [300], 1, 1)
    # Nothing prepared: the seeding table would have to be built while
    # recording, so the call fails and the runtime stays usable.
    _native.check(pipe.lib.tgk_graph_begin(runtime.handle))
    status = pipe.enqueue()
    graph = ctypes.c_void_p()
    pipe.lib.tgk_graph_end(runtime.handle, ctypes.byref(graph))
    pipe.lib.tgk_graph_destroy(graph)
    assert status == _native.ERR_INVALID
    assert b"prepared" in pipe.lib.tgk_last_error()
    pipe.reset()
    _native.check(pipe.enqueue())
    assert pipe.results()["steps"].sum() > 0


def test_order12_chunked_pipeline(oracle_threads):
    # The order-12 workload's pipeline at test size: MFMA seeding from the
    # order-12 FOD, fast tracking and compaction over 8 chunks of uneven
    # size (the workload's 0.3M-2.5M proportions, /1000), recorded once.
    # Every replay equals the eager run bit for bit, and the eager tracks
    # of a sample of seeds from every chunk follow the FP64 oracle, started
    # from the device's seeds, within the 1e-3 voxel budget.
    import deviation
    import oracle

    counts = [300, 700, 1100, 1500, 2500, 1900, 1200, 800]
    runtime = tg.HipRuntime(0)
    pipe = Pipeline(runtime, counts, _native.PATH_FAST, _native.SEEDS_MFMA, n_points=200, save_at=0.25,
                    scene=order12_scene)
    pipe.prepare()
    pipe.reset()
    _native.check(pipe.enqueue())
    eager = pipe.results()
    assert np.all(eager["lengths"] >= 1) and eager["steps"].sum() > 100_000
    pipe.reset()
    _native.check(pipe.lib.tgk_graph_begin(runtime.handle))
    status = pipe.enqueue()
    graph = ctypes.c_void_p()
    end_status = pipe.lib.tgk_graph_end(runtime.handle, ctypes.byref(graph))
    _native.check(status)
    _native.check(end_status)
    try:
        for replay in range(3):
            pipe.reset()
            _native.check(pipe.lib.tgk_graph_launch(graph))
            got = pipe.results()
            print(f"[order-12 graph] replay {replay}", flush=True)
            for name, value in eager.items():
                np.testing.assert_array_equal(got[name], value, err_msg=f"replay {replay}: {name}")
    finally:
        pipe.lib.tgk_graph_destroy(graph)
    # 40 seeds from every chunk against the FP64 oracle.
    starts = np.concatenate([[0], np.cumsum(counts)])
    idx = np.concatenate([np.arange(a, a + 40) for a in starts[:-1]])
    # results() concatenates the chunks' ragged points in seed order.
    offsets = np.concatenate([[0], np.cumsum(eager["lengths"].astype(np.int64))])
    lengths = eager["lengths"].view(np.uint32)[idx]
    points = np.concatenate([eager["points"].reshape(-1, 3)[offsets[g]: offsets[g] + eager["lengths"][g]]
                             for g in idx])
    ref = oracle.track(pipe.domain, pipe.field, pipe.algorithm, eager["seeds"][idx], pipe.rng_seeds, 50_000,
                       pipe.save_at, pipe.n_points * pipe.save_at, precision=64, indices=idx,
                       threads=oracle_threads)
    r = deviation.report(lengths, eager["steps"].view(np.uint32)[idx], points, ref)
    print(f"[order-12 graph] sample vs FP64: {r}", flush=True)
    assert r["max_deviation"] <= 1e-3
    assert r["compared_fraction"] > 0.5
