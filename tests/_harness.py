"""Builds the host-only check program tests/dropin_check.cpp (g++, no device
code) into tests/_bin/, optionally with the address and undefined-behaviour
sanitizers.  The program runs the reference's post-HF sequence
(src/main.cc:130-175) through libcorr_dropin.so."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "corr_amd"
BIN = Path(__file__).resolve().parent / "_bin"
SRC = Path(__file__).resolve().parent / "dropin_check.cpp"
DROPIN_SRC = PKG / "dropin" / "corr_dropin.cpp"
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def _stale(out: Path, *deps: Path) -> bool:
    return not out.exists() or any(out.stat().st_mtime < d.stat().st_mtime for d in deps)


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{' '.join(map(str, cmd))} failed:\n{res.stderr[-4000:]}")


def dropin_check(sanitize: bool = False) -> Path:
    """Path of the check program, built if missing or older than its sources.
    sanitize=True compiles the drop-in layer into the program, both with the
    address and undefined-behaviour sanitizers (runtimes linked statically);
    libcorr_gpu.so stays uninstrumented."""
    if not (PKG / "libcorr_gpu.so").exists():
        raise RuntimeError("corr_amd/libcorr_gpu.so is missing: run python -m corr_amd.build first")
    BIN.mkdir(exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    hdrs = (ROOT / "include" / "corr_dropin.h", ROOT / "include" / "corr_abi.h")
    if not sanitize:
        out = BIN / "corr_dropin_check"
        if _stale(out, SRC, PKG / "libcorr_dropin.so", *hdrs):
            _run([cxx, "-O2", "-std=c++17", "-Wall", str(SRC), "-o", str(out), f"-L{PKG}", "-lcorr_dropin",
                  "-lcorr_gpu", f"-Wl,-rpath,{PKG}"])
        return out
    out = BIN / "corr_dropin_check_san"
    if _stale(out, SRC, DROPIN_SRC, *hdrs):
        _run([cxx, "-O1", "-g", "-std=c++17", "-Wall", *SAN, "-static-libasan", "-static-libubsan", str(SRC),
              str(DROPIN_SRC), "-o", str(out), f"-L{PKG}", "-lcorr_gpu", f"-Wl,-rpath,{PKG}"])
    return out
