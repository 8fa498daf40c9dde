"""Builds libcorr_gpu.so (HIP, gfx950) in-tree.

Every ``corr_amd/csrc/*.hip`` file is compiled with hipcc for gfx950 and the
objects are linked into ``corr_amd/libcorr_gpu.so``.  Packed-FP32 instruction
selection is disabled for all device code.  Run ``python -m corr_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libcorr_gpu.so"
DROPIN_LIB = PKG / "libcorr_dropin.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep per-thread register arrays in registers (not LDS)
    "-mllvm", "-disable-promote-alloca-to-lds",
]
LINK_LIBS = ["-lrocsolver", "-lrocblas"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcorr_gpu.so")


def _compile(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-8000:]}")
    return out


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = min(len(sources), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           "-L/opt/rocm/lib", *LINK_LIBS]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-8000:]}")
    os.replace(tmp, LIB)
    build_dropin()
    if verbose:
        print(f"built {LIB} from {len(sources)} sources and {DROPIN_LIB.name}")
    return LIB


def _run(cmd: list[str], what: str) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{res.stderr[-8000:]}")


def build_dropin() -> None:
    """Host-only C++ drop-in layer (reference signatures) over libcorr_gpu.so."""
    cxx = os.environ.get("CXX", "g++")
    src = PKG / "dropin"
    tmp = DROPIN_LIB.with_suffix(".so.tmp")
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", str(src / "corr_dropin.cpp"), "-o", str(tmp),
          f"-L{PKG}", "-lcorr_gpu", "-Wl,-rpath,$ORIGIN"], "g++ (libcorr_dropin.so)")
    os.replace(tmp, DROPIN_LIB)


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
