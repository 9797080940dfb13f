"""Streaming tractogram writers (cli/tractogram.py:28-139): TCK and TRX
files read back by readers written from the formats' specifications, and
the batch loop (test/test_cli/test_tractogram.py:125-166)."""

import json
import zipfile

import numpy as np
import pytest

import tractography_amd as tg
from tractography_amd import streaming


def read_tck(path):
    """MRtrix3 .tck: header lines up to END; `file: . <offset>` locates
    float32 data; NaN triplets separate streamlines, an Inf triplet ends."""
    raw = path.read_bytes()
    end = raw.index(b"\nEND\n") + 5
    lines = raw[:end].decode("ascii").splitlines()
    assert lines[0] == "mrtrix tracks"
    fields = dict(line.split(": ", 1) for line in lines[1:-1])
    assert fields["datatype"] == "Float32LE"
    offset = int(fields["file"].split()[1])
    assert offset == end
    data = np.frombuffer(raw[offset:], dtype="<f4").reshape(-1, 3)
    assert np.all(np.isinf(data[-1]))
    streamlines, current = [], []
    for row in data[:-1]:
        if np.all(np.isnan(row)):
            streamlines.append(np.array(current, dtype=np.float32).reshape(-1, 3))
            current = []
        else:
            current.append(row)
    assert current == []
    assert int(fields["count"]) == len(streamlines)
    return streamlines


def read_trx(path):
    with zipfile.ZipFile(path) as archive:
        header = json.loads(archive.read("header.json"))
        positions = np.frombuffer(archive.read("positions.3.float32"), dtype="<f4").reshape(-1, 3)
        offsets = np.frombuffer(archive.read("offsets.uint64"), dtype="<u8")
        assert all(info.compress_type == zipfile.ZIP_STORED for info in archive.infolist())
    assert header["NB_VERTICES"] == len(positions)
    assert header["NB_STREAMLINES"] == len(offsets)
    bounds = list(offsets.astype(np.int64)) + [len(positions)]
    return header, [positions[bounds[i] : bounds[i + 1]] for i in range(len(offsets))]


def _random_batches(rng, n_batches):
    batches = []
    for _ in range(n_batches):
        lengths = rng.integers(1, 9, rng.integers(0, 6))
        batches.append((rng.normal(size=(int(lengths.sum()), 3)).astype(np.float32), lengths))
    return batches


def test_tck_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    batches = _random_batches(rng, 7)
    writer = streaming.TckWriter(tmp_path / "t.tck")
    for points, lengths in batches:
        writer.append(points, lengths)
    writer.close()
    expected = [s for points, lengths in batches for s in np.split(points, np.cumsum(lengths)[:-1])
                if len(lengths)]
    got = read_tck(tmp_path / "t.tck")
    assert len(got) == len(expected) == sum(len(l) for _, l in batches)
    for a, b in zip(got, expected):
        np.testing.assert_array_equal(a, b)


def test_tck_empty_and_large_count_header(tmp_path):
    writer = streaming.TckWriter(tmp_path / "e.tck")
    writer.close()
    assert read_tck(tmp_path / "e.tck") == []


def test_trx_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    batches = _random_batches(rng, 5)
    affine = np.diag([2.0, 2.0, 2.0, 1.0])
    writer = streaming.TrxWriter(tmp_path / "t.trx", affine=affine, dimensions=(4, 5, 6))
    for points, lengths in batches:
        writer.append(points, lengths)
    writer.close()
    header, got = read_trx(tmp_path / "t.trx")
    assert header["DIMENSIONS"] == [4, 5, 6]
    np.testing.assert_array_equal(header["VOXEL_TO_RASMM"], affine)
    expected = [s for points, lengths in batches for s in np.split(points, np.cumsum(lengths)[:-1])
                if len(lengths)]
    assert len(got) == len(expected)
    for a, b in zip(got, expected):
        np.testing.assert_array_equal(a, b)
    # Temporary streams beside the output are removed.
    assert sorted(p.name for p in tmp_path.iterdir()) == ["t.trx"]


def test_trx_output_is_written_incrementally(monkeypatch, tmp_path):
    # test/test_cli/test_tractogram.py:125-166
    fod = tg.Nifti1Image(np.zeros((2, 2, 2, 1), dtype=np.float32), np.eye(4))
    mask = tg.Nifti1Image(np.ones((2, 2, 2), dtype=np.uint8), np.eye(4))
    seeds = np.array([[i, 0, 0, 1, 1, 0, 0, 0] for i in range(3)], dtype=np.float32)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(1))
    monkeypatch.setattr(tg, "_tracking_components", lambda *args, **kwargs: (algorithm, object(), object()))
    monkeypatch.setattr(streaming, "HipRuntime", lambda *args, **kwargs: object())
    batch_sizes = []

    def fake_track(_algorithm, _domain, _field, batch, *args, **kwargs):
        batch_sizes.append(len(batch))
        points = np.concatenate([[[i, 0, 0], [i + 1, 0, 0]] for i in range(len(batch))]).astype(np.float32)
        return tg.tracking.TrackResult(np.full(len(batch), 2, np.uint32), np.ones(len(batch), np.uint32),
                                       np.zeros(len(batch), np.uint32), points)

    monkeypatch.setattr(tg.tracking, "track", fake_track)
    output_path = tmp_path / "streamlines.trx"
    written = streaming.write_tractogram(fod, seeds, output_path, mask, algorithm, save_at=1.0,
                                         maximum_streamline_length=5.0, batch_size=2)
    header, streamlines = read_trx(output_path)
    assert batch_sizes == [2, 1]
    assert written == 3
    assert header["NB_STREAMLINES"] == 3
    assert header["NB_VERTICES"] == 6


def test_write_tractogram_argument_errors(tmp_path):
    fod = tg.Nifti1Image(np.ones((2, 2, 2, 45), dtype=np.float32), np.eye(4))
    seeds = np.array([[0, 0, 0, 1, 1, 0, 0, 0]], dtype=np.float32)
    with pytest.raises(ValueError, match="save_at"):
        streaming.write_tractogram(fod, seeds, tmp_path / "x.tck", save_at=0.0)
    with pytest.raises(ValueError, match="must not exceed"):
        streaming.write_tractogram(fod, seeds, tmp_path / "x.tck", maximum_streamline_length=1.0,
                                   minimum_streamline_length=2.0)
    with pytest.raises(ValueError, match="batch_size"):
        streaming.write_tractogram(fod, seeds, tmp_path / "x.tck", batch_size=0)
    with pytest.raises(ValueError, match="unsupported"):
        streaming.write_tractogram(fod, seeds, tmp_path / "x.trk")


@pytest.mark.gpu
@pytest.mark.parametrize("suffix", [".tck", ".trx"])
def test_streamed_file_equals_the_in_memory_tractogram(runtime, tmp_path, suffix):
    import fixtures

    fod, wm, _ = fixtures.cross()
    np.random.seed(4)
    seeds = tg.seeds.from_mask(wm, 1000)

    def algorithm():
        return tg.algorithm.Diffusion(tg.local.RealSymmetricSphericalHarmonics(45), step_size=0.1, gamma=1.0,
                                      noise_variance=0.1, seed=12)

    expected = tg.tractogram(fod, seeds, wm, algorithm(), save_at=0.5, maximum_streamline_length=30.0,
                             minimum_streamline_length=2.0, batch_size=300, runtime=runtime)
    path = tmp_path / ("tracks" + suffix)
    written = streaming.write_tractogram(fod, seeds, path, wm, algorithm(), save_at=0.5,
                                         maximum_streamline_length=30.0, minimum_streamline_length=2.0,
                                         batch_size=300, runtime=runtime)
    got = read_tck(path) if suffix == ".tck" else read_trx(path)[1]
    assert written == len(expected) == len(got)
    assert 0 < written < 1000
    for a, b in zip(got, expected.streamlines):
        np.testing.assert_array_equal(a, np.asarray(b))
