"""CPU test of the BF16x3 energy kernel's tile list (corr_amd/csrc/wide_tiles.hpp,
the functions the device builder runs): tests/wide_tiles_check.cpp is built
with g++ and the address and undefined-behaviour sanitizers and checks counts,
field ranges, the diagonal tiles' place and weighted coverage of every
occupied pair i <= j and virtual block pair A <= B (restating the pair loop of
run_canonical_mp2_ri, can_mp2.cc:82-112)."""
import os
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "wide_tiles_check.cpp"
HDR = HERE.parent / "corr_amd" / "csrc" / "wide_tiles.hpp"
OUT = HERE / "_bin" / "wide_tiles_check"

# (o, v): one occupied or virtual block, odd and even block counts, ragged
# edges, the test shapes of the GPU suite and the (H2O)32 layout
SHAPES = [(1, 1), (3, 30), (4, 32), (5, 33), (7, 64), (9, 150), (13, 70), (16, 224), (17, 96), (20, 76),
          (21, 608), (31, 130), (33, 300), (40, 152), (41, 97), (61, 200), (160, 608)]


@pytest.fixture(scope="module")
def checker():
    OUT.parent.mkdir(exist_ok=True)
    if not OUT.exists() or OUT.stat().st_mtime < max(SRC.stat().st_mtime, HDR.stat().st_mtime):
        cxx = os.environ.get("CXX", "g++")
        res = subprocess.run([cxx, "-O1", "-g", "-std=c++17", "-Wall", "-fsanitize=address,undefined",
                              "-fno-sanitize-recover=undefined", f"-I{HDR.parent}", str(SRC), "-o", str(OUT)],
                             capture_output=True, text=True)
        if res.returncode != 0:
            pytest.fail(res.stderr[-3000:])
    return OUT


def test_tile_list_covers_every_pair_once(checker):
    args = [str(x) for s in SHAPES for x in s]
    res = subprocess.run([str(checker), *args], capture_output=True, text=True, timeout=300,
                         env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert res.returncode == 0, res.stderr[-3000:]
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr
    lines = res.stdout.strip().splitlines()
    assert len(lines) == len(SHAPES)
    # (H2O)32: 410 occupied units x 181 virtual tiles, 10 diagonal tiles per unit
    assert lines[-1] == "o=160 v=608 nIB=40 nAC=19 tiles=74210 (diagonal 4100)"
