"""Build ``libpsx.so`` (the HIP kernels and the C ABI) in-tree for gfx950 with hipcc.

``python -m pyspectral_amd.build`` or ``__graft_entry__.build()``.  Translation units are
compiled in parallel into ``pyspectral_amd/_build`` and linked into ``pyspectral_amd/libpsx.so``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpsx.so"
SOURCES = ["psx_k2d.hip", "psx_k2d_ct.hip", "psx_k3d.hip", "psx_k3d_ct.hip", "psx_kutil.hip",
           "psx_kdrv.hip", "psx_kens.hip", "psx_kles.hip", "psx_kcnn.hip", "psx_kfft.hip",
           "psx_api.cpp", "psx_ens.cpp", "psx_les.cpp"]
ARCH = os.environ.get("PSX_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions stay disabled for every device compile.
HIPFLAGS = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [_hipcc(), *HIPFLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd = [_hipcc(), *HIPFLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build_library(verbose: bool = True) -> Path:
    """Compile every translation unit and link ``libpsx.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build_library()
    sys.exit(0)
