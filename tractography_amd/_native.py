"""ctypes binding of libtgk.so (the C ABI declared in include/tgk.h).

The library is loaded on first use from the package directory. There is no
host fallback: without libtgk.so every device entry point raises.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().with_name("libtgk.so")

OK = 0
ERR_INVALID = -1
ERR_NO_MEMORY = -2
ERR_DEVICE = -3
ERR_UNSUPPORTED = -4

MODEL_SH = 0
MODEL_DTI = 1
ALGO_TRANSPORT = 0
ALGO_DIFFUSION = 1
ALGO_PROBABILISTIC = 2
ALGO_DETERMINISTIC = 3
PATH_PLAIN = 0
PATH_FAST = 1
FAST_SHAPES = 5
PRECISIONS = {"fp32": 0, "bf16x2": 1, "fp16": 2, "bf16": 3}
SEEDS_VALU = 0
SEEDS_MFMA = 1

NO_HIT = 0xFFFFFFFF
# Largest n_seeds tgk_compact_device_noalloc takes (TGK_COMPACT_NOALLOC_MAX_SEEDS).
COMPACT_NOALLOC_MAX_SEEDS = 16775168


class tgk_algorithm(ctypes.Structure):
    _fields_ = [
        ("model", c_int32),
        ("algorithm", c_int32),
        ("n_parameters", c_int32),
        ("step_size", c_float),
        ("gamma", c_float),
        ("noise_variance", c_float),
        ("n_samples", c_int32),
        ("max_angle", c_float),
        ("max_iterations", c_int32),
        ("tolerance", c_float),
        ("max_step_size", c_float),
    ]


_P = c_void_p  # every array and handle crosses the boundary as a plain pointer

_SIGNATURES = {
    "tgk_last_error": (c_char_p, []),
    "tgk_version": (c_int, []),
    "tgk_runtime_create": (c_int, [c_int32, POINTER(c_void_p)]),
    "tgk_runtime_set_fast_grid": (c_int, [_P, c_int32]),
    "tgk_runtime_set_fast_shape": (c_int, [_P, c_int32]),
    "tgk_fast_shape_info": (c_int, [c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "tgk_runtime_set_fast_repack": (c_int, [_P, c_int32]),
    "tgk_runtime_fast_reads": (c_int, [_P, POINTER(c_uint64), POINTER(c_uint64)]),
    "tgk_track_timed": (c_int, [_P, _P, _P, POINTER(tgk_algorithm), c_int32, _P, c_int64, c_uint32, c_float, c_int32,
                                POINTER(c_float), POINTER(c_uint64)]),
    "tgk_runtime_destroy": (c_int, [_P]),
    "tgk_runtime_synchronize": (c_int, [_P]),
    "tgk_runtime_info": (c_int, [_P, c_char_p, c_int32, POINTER(c_int32), POINTER(c_int64)]),
    "tgk_runtime_stream": (c_int, [_P, POINTER(c_void_p)]),
    "tgk_field_create": (c_int, [_P, _P, _P, _P, POINTER(c_void_p)]),
    "tgk_field_destroy": (c_int, [_P]),
    "tgk_field_set_fast_precision": (c_int, [_P, c_int32]),
    "tgk_mask_domain_create": (c_int, [_P, _P, _P, _P, POINTER(c_void_p)]),
    "tgk_mesh_domain_create": (
        c_int,
        [_P, _P, c_uint32, _P, _P, c_uint32, _P, _P, _P, c_uint32, _P, c_uint32, _P, _P, _P, c_float,
         POINTER(c_void_p)],
    ),
    "tgk_domain_set_boundary": (c_int, [_P, _P, c_uint32, _P, _P, c_uint32]),
    "tgk_domain_destroy": (c_int, [_P]),
    "tgk_track": (
        c_int,
        [_P, _P, _P, POINTER(tgk_algorithm), c_int32, _P, c_int64, _P, c_int64, c_uint32, c_float,
         POINTER(c_void_p)],
    ),
    "tgk_tracks_size": (c_int, [_P, POINTER(c_int64), POINTER(c_int64)]),
    "tgk_tracks_copy": (c_int, [_P, _P, _P, _P, _P]),
    "tgk_tracks_points": (c_int, [_P, POINTER(c_void_p)]),
    "tgk_tracks_destroy": (c_int, [_P]),
    "tgk_track_device": (
        c_int,
        [_P, _P, _P, POINTER(tgk_algorithm), c_int32, _P, c_int64, c_int64, _P, c_int64, c_uint32, c_float,
         _P, _P, _P, _P],
    ),
    "tgk_compact_device": (c_int, [_P, c_int32, _P, c_int64, c_uint32, _P, _P, _P]),
    "tgk_compact_device_noalloc": (c_int, [_P, c_int32, _P, c_int64, c_uint32, _P, _P, _P]),
    "tgk_domain_seed": (c_int, [_P, _P, c_uint32, c_int64, _P]),
    "tgk_connectome": (c_int, [_P, _P, _P, _P, POINTER(tgk_algorithm), c_int32, c_uint32, c_int64, c_uint32, c_float,
                               _P, _P, _P]),
    "tgk_tod": (c_int, [_P, _P, _P, _P, POINTER(tgk_algorithm), c_int32, c_uint32, c_int64, c_uint32, _P, _P, c_int32,
                        _P]),
    "tgk_tod_device": (c_int, [_P, _P, _P, _P, POINTER(tgk_algorithm), c_int32, c_uint32, c_int64, c_uint32, _P, _P,
                               c_int32, _P]),
    "tgk_normalize_odf": (c_int, [_P, c_int64, _P]),
    "tgk_seeds_from_fod": (c_int, [_P, c_int32, c_int32, _P, _P, c_uint32, _P, _P, c_int64, c_int32, _P]),
    "tgk_seeds_prepare": (c_int, [_P, c_int32, c_int32, c_int32]),
    "tgk_seeds_from_fod_device": (c_int, [_P, c_int32, c_int32, _P, _P, c_uint32, _P, _P, c_int64, c_int32, _P]),
    "tgk_track_prepare": (c_int, [_P, _P, _P, POINTER(tgk_algorithm), c_int32, c_int64]),
    "tgk_graph_begin": (c_int, [_P]),
    "tgk_graph_end": (c_int, [_P, POINTER(c_void_p)]),
    "tgk_graph_launch": (c_int, [_P]),
    "tgk_graph_destroy": (c_int, [_P]),
    "tgk_evaluate": (c_int, [_P, c_int32, c_int32, _P, _P, c_int64, _P]),
    "tgk_update_orientation": (c_int, [_P, POINTER(tgk_algorithm), c_uint32, _P, _P, c_int64, _P]),
    "tgk_field_evaluate": (c_int, [_P, _P, _P, c_int64, _P]),
    "tgk_domain_contains": (c_int, [_P, _P, _P, c_int64, _P]),
    "tgk_bvh_segment_hit": (c_int, [_P, _P, _P, _P, c_int64, _P, _P, _P]),
    "tgk_bvh_node_count": (c_int, [c_uint32, c_uint32, POINTER(c_uint32)]),
    "tgk_bvh_build": (c_int, [_P, _P, c_uint32, c_uint32, _P, _P, _P, c_uint32, _P]),
    "tgk_segment_grid_count": (c_int, [_P, _P, _P, _P, c_uint32, _P, POINTER(c_uint64)]),
    "tgk_segment_grid_fill": (c_int, [_P, _P, _P, _P, c_uint32, _P, c_uint32, _P, _P, _P, c_uint64]),
    "tgk_grid_segment_hit": (c_int, [_P, _P, _P, _P, c_int64, _P, _P, _P]),
    "tgk_domain_set_segment_grid": (c_int, [_P, _P, c_uint64, _P, c_uint64]),
    "tgk_math": (c_int, [_P, c_int32, _P, _P, c_int64, _P]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def library_available() -> bool:
    return LIBRARY_PATH.exists()


def lib() -> ctypes.CDLL:
    """The loaded libtgk.so; raises RuntimeError if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build the HIP extension with "
                "`python -m tractography_amd.build` (there is no host fallback)"
            )
        handle = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


class NativeError(RuntimeError):
    """A device-side failure reported by libtgk.so."""


def check(status: int) -> None:
    """Map a TGK status to the exception the reference API would raise."""
    if status == OK:
        return
    message = lib().tgk_last_error().decode("utf-8", "replace")
    if status == ERR_INVALID:
        raise ValueError(message)
    if status == ERR_NO_MEMORY:
        raise MemoryError(message)
    if status == ERR_UNSUPPORTED:
        raise NotImplementedError(message)
    raise NativeError(message)


def ptr(array: np.ndarray | None) -> int | None:
    """Address of a C-contiguous numpy array (None passes NULL)."""
    if array is None:
        return None
    if not array.flags.c_contiguous:
        raise ValueError("arrays passed to libtgk.so must be C-contiguous")
    return array.ctypes.data


def algorithm_struct(model: int, algorithm: int, n_parameters: int, step_size: float, gamma: float,
                     noise_variance: float, n_samples: int = 0, max_angle: float = 0.0, max_iterations: int = 0,
                     tolerance: float = 0.0, max_step_size: float = 0.0) -> tgk_algorithm:
    return tgk_algorithm(model, algorithm, n_parameters, step_size, gamma, noise_variance, n_samples, max_angle,
                         max_iterations, tolerance, max_step_size)
