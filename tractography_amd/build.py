"""Build libtgk.so (the HIP kernels and the C ABI) in-tree for gfx950.

Usage: ``python -m tractography_amd.build`` (or ``build()`` from Python).

Every translation unit is compiled by hipcc for gfx950 only. The plain path,
the one-op drivers and the C ABI are compiled with ``-ffp-contract=off`` so
their arithmetic rounds exactly as written (the bit-exactness contract with
the FP32 host restatement); the fast path is compiled with
``-ffp-contract=on``: a product and a sum in one source expression become an
fma, decided by the source alone, so that its variants compiled for
different launch shapes round alike. Packed FP32 instructions are disabled
for all code.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
INCLUDE = PACKAGE_DIR.parent / "include"
BUILD_DIR = PACKAGE_DIR / "_build"
LIBRARY = PACKAGE_DIR / "libtgk.so"
ARCH = "gfx950"

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
    f"-I{CSRC}",
    f"-I{INCLUDE}",
]

# (source, contraction policy)
SOURCES = [
    ("tgk_plain.hip", "off"),
    ("tgk_capi.hip", "off"),
    ("tgk_seeds.hip", "off"),
    ("tgk_scan.hip", "off"),
    ("tgk_bvh.hip", "off"),
    ("tgk_fast.hip", "on"),
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    candidate = Path(rocm) / "bin" / "hipcc"
    if candidate.exists():
        return str(candidate)
    found = shutil.which("hipcc")
    if not found:
        raise RuntimeError("hipcc not found: a ROCm installation is required to build libtgk.so")
    return found


def _compile(source: str, contract: str) -> Path:
    obj = BUILD_DIR / (Path(source).stem + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS, f"-ffp-contract={contract}", "-c", str(CSRC / source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed for {source}:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every HIP translation unit and link ``libtgk.so``."""
    BUILD_DIR.mkdir(exist_ok=True)
    sources = [(s, c) for s, c in SOURCES if (CSRC / s).exists()]
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(sources)) as pool:
        objects = list(pool.map(lambda sc: _compile(*sc), sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
