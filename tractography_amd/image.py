"""Image and tractogram containers for environments without NiBabel.

The reference API takes ``nibabel`` images and returns a ``nibabel``
``Tractogram``. When NiBabel is importable its classes are used; otherwise
these minimal stand-ins provide the attributes the tracking path reads
(``shape``, ``ndim``, ``affine``, ``dataobj``, ``get_fdata``) and returns
(``streamlines``, ``affine_to_rasmm``).
"""

from __future__ import annotations

from typing import Iterator

import numpy as np

try:  # pragma: no cover - NiBabel is optional
    from nibabel.nifti1 import Nifti1Image as _NibabelImage
    from nibabel.streamlines.tractogram import Tractogram as _NibabelTractogram
except Exception:  # noqa: BLE001
    _NibabelImage = None
    _NibabelTractogram = None


class Nifti1Image:
    """Array plus voxel-to-world affine, read like a NiBabel image."""

    def __init__(self, dataobj, affine=None):
        self.dataobj = np.asanyarray(dataobj)
        self.affine = np.eye(4) if affine is None else np.asarray(affine, dtype=np.float64)

    @property
    def shape(self):
        return self.dataobj.shape

    @property
    def ndim(self):
        return self.dataobj.ndim

    def get_fdata(self):
        return np.asarray(self.dataobj, dtype=np.float64)

    def get_data_dtype(self):
        return self.dataobj.dtype


def float32_image(data, affine, like=None):
    """An image of float32 ``data`` on ``affine``: a NiBabel image with a copy
    of ``like``'s header (data type set to float32) when ``like`` carries a
    NiBabel header, else the stand-in."""
    data = np.asarray(data, dtype=np.float32)
    header = getattr(like, "header", None)
    if _NibabelImage is not None and header is not None and hasattr(header, "set_data_dtype"):
        header = header.copy()
        header.set_data_dtype(np.float32)
        return _NibabelImage(data, affine, header=header)
    return Nifti1Image(data, affine)


def is_image(value) -> bool:
    """True for NiBabel images and for the stand-in above."""
    if isinstance(value, Nifti1Image):
        return True
    if _NibabelImage is not None and isinstance(value, _NibabelImage):
        return True
    return False


def image_data(image) -> np.ndarray:
    """float64 voxel data of an image (NiBabel's get_fdata)."""
    return np.asarray(image.get_fdata(), dtype=np.float64)


class Streamlines:
    """Ragged sequence of (n_i, 3) float32 streamlines over one flat array."""

    def __init__(self, points: np.ndarray | None = None, lengths: np.ndarray | None = None):
        if points is None:
            points = np.empty((0, 3), dtype=np.float32)
            lengths = np.empty((0,), dtype=np.int64)
        self._data = np.asarray(points, dtype=np.float32).reshape(-1, 3)
        self._lengths = np.asarray(lengths, dtype=np.int64)
        self._offsets = np.zeros(len(self._lengths) + 1, dtype=np.int64)
        np.cumsum(self._lengths, out=self._offsets[1:])

    @classmethod
    def from_list(cls, arrays) -> "Streamlines":
        arrays = [np.asarray(a, dtype=np.float32).reshape(-1, 3) for a in arrays]
        lengths = np.array([len(a) for a in arrays], dtype=np.int64)
        data = np.concatenate(arrays) if arrays else np.empty((0, 3), dtype=np.float32)
        return cls(data, lengths)

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def lengths(self) -> np.ndarray:
        return self._lengths

    @property
    def offsets(self) -> np.ndarray:
        return self._offsets[:-1]

    def __len__(self) -> int:
        return len(self._lengths)

    def __getitem__(self, index):
        if isinstance(index, (int, np.integer)):
            if index < 0:
                index += len(self)
            if not 0 <= index < len(self):
                raise IndexError(index)
            return self._data[self._offsets[index] : self._offsets[index + 1]]
        selected = np.arange(len(self))[index]
        return Streamlines.from_list([self[int(i)] for i in selected])

    def __iter__(self) -> Iterator[np.ndarray]:
        for i in range(len(self)):
            yield self[i]


class Tractogram:
    """Streamlines with the affine to RAS+ mm (NiBabel's Tractogram subset)."""

    def __init__(self, streamlines=None, affine_to_rasmm=None):
        if streamlines is None:
            streamlines = Streamlines()
        elif not isinstance(streamlines, Streamlines):
            streamlines = Streamlines.from_list(list(streamlines))
        self.streamlines = streamlines
        self.affine_to_rasmm = np.eye(4) if affine_to_rasmm is None else np.asarray(affine_to_rasmm)

    def __len__(self) -> int:
        return len(self.streamlines)

    def __iter__(self):
        return iter(self.streamlines)
