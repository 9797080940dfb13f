"""Streaming tractogram output (the batch writers of cli/tractogram.py:28-139).

``write_tractogram`` tracks seeds in batches and appends each batch's ragged
streamlines to a ``.tck`` or ``.trx`` file while the GPU tracks the next
batch: a writer thread takes a batch's points as soon as the device call
returns them (the GIL is released during the call), so file output overlaps
the tracking. No tractogram-sized host buffer exists; one batch's ragged
points are held at a time, plus the one being written.

Formats:

- TCK (MRtrix3): a text header ending in ``END``, whose ``file: . <offset>``
  gives the start of the float32 little-endian data, and ``count`` the
  number of streamlines; each streamline's points, then a NaN triplet; an
  Inf triplet ends the data. ``count`` is written with fixed width and
  rewritten on close.
- TRX: a zip archive (stored, no compression) with ``header.json``
  (``VOXEL_TO_RASMM``, ``DIMENSIONS``, ``NB_VERTICES``, ``NB_STREAMLINES``),
  ``positions.3.float32`` (vertices, row-major) and ``offsets.uint64`` (the
  index of each streamline's first vertex). Positions and offsets stream to
  temporary files beside the output and are zipped on close.
"""

from __future__ import annotations

import json
import os
import queue
import shutil
import tempfile
import threading
import zipfile
from pathlib import Path

import numpy as np

from .algorithm import BaseAlgorithm
from .runtime import HipRuntime

_TCK_COUNT_WIDTH = 10
_TCK_SEPARATOR = np.full(3, np.nan, dtype="<f4").tobytes()
# Buffers per writev call (below the usual IOV_MAX of 1024).
_IOV_BATCH = 1024


class TckWriter:
    """Append-only MRtrix .tck writer (float32 little-endian)."""

    def __init__(self, path, header: dict | None = None):
        self.path = Path(path)
        self.count = 0
        fields = {"datatype": "Float32LE"}
        fields.update({k: str(v) for k, v in (header or {}).items() if k not in ("count", "file", "datatype")})
        lines = ["mrtrix tracks"] + [f"{k}: {v}" for k, v in fields.items()]
        lines.append("count: " + "0" * _TCK_COUNT_WIDTH)
        body = "\n".join(lines) + "\n"
        # "file: . <offset>" names the offset of the data, which includes its
        # own digits.
        offset = len(body) + len("file: . \nEND\n")
        while len(str(offset)) + len(body) + len("file: . \nEND\n") != offset:
            offset = len(str(offset)) + len(body) + len("file: . \nEND\n")
        text = body + f"file: . {offset}\nEND\n"
        self._count_at = text.index("count: ") + len("count: ")
        # Unbuffered: the data goes out with os.writev, straight from the
        # point array.
        self._file = open(self.path, "wb", buffering=0)
        self._write_all([text.encode("ascii")])
        assert self._file.tell() == offset

    def _write_all(self, buffers) -> None:
        """Writes the buffers in order, gathering them into one system call
        per _IOV_BATCH (no copy of the point data)."""
        fd = self._file.fileno()
        for start in range(0, len(buffers), _IOV_BATCH):
            views = [memoryview(b).cast("B") for b in buffers[start : start + _IOV_BATCH]]
            written = os.writev(fd, views) if hasattr(os, "writev") else 0
            # A short write (or no writev): finish the rest one buffer at a time.
            for view in views:
                if written >= len(view):
                    written -= len(view)
                    continue
                rest = view[written:]
                written = 0
                while len(rest):
                    rest = rest[os.write(fd, rest) :]

    def append(self, points: np.ndarray, lengths: np.ndarray) -> None:
        """Streamlines given as ragged points (sum(lengths), 3) and lengths."""
        lengths = np.asarray(lengths, dtype=np.int64)
        if len(lengths) == 0:
            return
        points = np.ascontiguousarray(points, dtype="<f4").reshape(-1, 3)
        if points.shape[0] != int(lengths.sum()):
            raise ValueError("points and lengths disagree")
        # Each streamline's rows, then a NaN triplet: views of the point
        # array interleaved with one separator buffer.
        ends = np.cumsum(lengths)
        starts = ends - lengths
        buffers = []
        for a, b in zip(starts.tolist(), ends.tolist()):
            buffers.append(points[a:b])
            buffers.append(_TCK_SEPARATOR)
        self._write_all(buffers)
        self.count += len(lengths)

    def close(self) -> None:
        if self._file is None:
            return
        self._write_all([np.full(3, np.inf, dtype="<f4").tobytes()])
        if self.count >= 10**_TCK_COUNT_WIDTH:
            raise ValueError("too many streamlines for the count field")
        self._file.seek(self._count_at)
        self._write_all([f"{self.count:0{_TCK_COUNT_WIDTH}d}".encode("ascii")])
        self._file.close()
        self._file = None


class TrxWriter:
    """Append-only TRX writer (zip of header.json, positions, offsets)."""

    def __init__(self, path, affine=None, dimensions=None):
        self.path = Path(path)
        self.affine = np.eye(4) if affine is None else np.asarray(affine, dtype=np.float64)
        self.dimensions = [1, 1, 1] if dimensions is None else [int(d) for d in dimensions]
        self._dir = Path(tempfile.mkdtemp(prefix=self.path.name + ".", dir=self.path.parent))
        self._positions = open(self._dir / "positions.3.float32", "wb")
        self._offsets = open(self._dir / "offsets.uint64", "wb")
        self.n_vertices = 0
        self.n_streamlines = 0

    def append(self, points: np.ndarray, lengths: np.ndarray) -> None:
        lengths = np.asarray(lengths, dtype=np.int64)
        if len(lengths) == 0:
            return
        points = np.ascontiguousarray(points, dtype="<f4").reshape(-1, 3)
        if points.shape[0] != int(lengths.sum()):
            raise ValueError("points and lengths disagree")
        starts = self.n_vertices + np.concatenate([[0], np.cumsum(lengths)[:-1]])
        self._offsets.write(starts.astype("<u8").tobytes())
        self._positions.write(points.tobytes())
        self.n_vertices += points.shape[0]
        self.n_streamlines += len(lengths)

    def close(self) -> None:
        if self._positions is None:
            return
        self._positions.close()
        self._offsets.close()
        self._positions = self._offsets = None
        header = {
            "VOXEL_TO_RASMM": self.affine.tolist(),
            "DIMENSIONS": self.dimensions,
            "NB_VERTICES": int(self.n_vertices),
            "NB_STREAMLINES": int(self.n_streamlines),
        }
        try:
            with zipfile.ZipFile(self.path, "w", compression=zipfile.ZIP_STORED) as archive:
                archive.writestr("header.json", json.dumps(header))
                archive.write(self._dir / "positions.3.float32", "positions.3.float32")
                archive.write(self._dir / "offsets.uint64", "offsets.uint64")
        finally:
            shutil.rmtree(self._dir, ignore_errors=True)


def _writer_for(path: Path, field_image) -> TckWriter | TrxWriter:
    suffix = path.suffix.lower()
    affine = getattr(field_image, "affine", None)
    if suffix == ".tck":
        return TckWriter(path)
    if suffix == ".trx":
        shape = getattr(field_image, "shape", None)
        return TrxWriter(path, affine=affine, dimensions=None if shape is None else shape[:3])
    raise ValueError(f"unsupported tractogram format {suffix!r}: .tck or .trx")


def write_tractogram(
    field,
    seeds: np.ndarray,
    tractogram_path,
    domain=None,
    algorithm: BaseAlgorithm | None = None,
    *,
    save_at: float = 1.0,
    maximum_streamline_length: float = 300.0,
    minimum_streamline_length: float = 0.0,
    batch_size: int = 50_000,
    runtime: HipRuntime | None = None,
) -> int:
    """Track `seeds` through a 4-D field image in batches and stream the
    streamlines to a .tck or .trx file (cli/tractogram.py:28-139).

    Streamlines with ceil(minimum_streamline_length / save_at) points or
    fewer are dropped, as the reference's batch writer does. Batch b draws
    its noise streams from the algorithm's b-th RNG seed. Returns the number
    of streamlines written.
    """
    import tractography_amd as tg

    if not np.isfinite(save_at) or save_at <= 0:
        raise ValueError("save_at must be finite and positive.")
    if not np.isfinite(maximum_streamline_length) or maximum_streamline_length < 0:
        raise ValueError("maximum_streamline_length must be finite and non-negative.")
    if not np.isfinite(minimum_streamline_length) or minimum_streamline_length < 0:
        raise ValueError("minimum_streamline_length must be finite and non-negative.")
    if minimum_streamline_length > maximum_streamline_length:
        raise ValueError("minimum_streamline_length must not exceed maximum_streamline_length.")
    if not isinstance(batch_size, (int, np.integer)) or batch_size <= 0:
        raise ValueError("batch_size must be a positive integer.")
    seeds = tg.seeds.validate_seed_array(seeds)
    if algorithm is None:
        algorithm = tg.algorithm.Transport(tg.local.from_shape(field.shape))
    tracking_algorithm, tracking_domain, tracking_field = tg._tracking_components(
        field, algorithm, domain_image=domain
    )
    path = Path(tractogram_path)
    writer = _writer_for(path, field)
    minimum_points = int(np.ceil(minimum_streamline_length / save_at))
    if runtime is None and len(seeds) > 0:
        runtime = HipRuntime()

    # The writer thread appends batch k while the caller's thread tracks
    # batch k + 1; at most two batches are alive.
    pending: queue.Queue = queue.Queue(maxsize=1)
    failure: list[BaseException] = []

    def drain():
        while True:
            item = pending.get()
            if item is None:
                return
            try:
                if not failure:
                    writer.append(*item)
            except BaseException as error:  # noqa: BLE001 - handed to the caller
                failure.append(error)

    thread = threading.Thread(target=drain, name="tractogram-writer", daemon=True)
    thread.start()
    written = 0
    try:
        for start in range(0, len(seeds), batch_size):
            batch = seeds[start : start + batch_size]
            result = tg.tracking.track(
                tracking_algorithm, tracking_domain, tracking_field, batch, [tracking_algorithm.get_rng_seed()],
                len(batch), save_at, maximum_streamline_length, runtime,
            )
            lengths = result.lengths.astype(np.int64)
            keep = lengths > minimum_points
            points = result.points if np.all(keep) else result.points[np.repeat(keep, lengths)]
            written += int(keep.sum())
            pending.put((points, lengths[keep]))
            if failure:
                break
    finally:
        pending.put(None)
        thread.join()
        writer.close()
    if failure:
        os.unlink(path)
        raise failure[0]
    return written
