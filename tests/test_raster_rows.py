"""The render kernel's closed-form row emitter (blend2d_amd/csrc/raster_device.h,
compiled here for the host) against the oracle's restatement of the
reference rasterizer (oracle/orc_raster.c, pinned by the reference-rendered
golden canvases): tests/native/raster_rows_check.cpp compares the coverage
prefix sum of every pixel of every row of random and boundary-hugging
segments (vertical, near-vertical, flat, on pixel boundaries, both
directions and signs, 45-degree and very steep ones) for tiles at many
offsets.  Two planted mistakes in a
copy of the emitter (the rounding term of the row-boundary x, and the
modulo-256 crossing of steep edges) must make it fail."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "raster_rows_check.cpp")
EMITTER = os.path.join(ROOT, "blend2d_amd", "csrc", "raster_device.h")
ORC = os.path.join(ROOT, "oracle", "_build")


def _build(tmp_path, emitter_text=None):
    if not shutil.which("g++") or not os.path.exists(os.path.join(ORC, "liborc.so")):
        pytest.skip("g++ or the oracle library is missing")
    src_dir = tmp_path / "tests" / "native"
    inc_dir = tmp_path / "blend2d_amd" / "csrc"
    src_dir.mkdir(parents=True)
    inc_dir.mkdir(parents=True)
    shutil.copy(SRC, src_dir / "check.cpp")
    (inc_dir / "raster_device.h").write_text(emitter_text if emitter_text is not None else open(EMITTER).read())
    exe = tmp_path / "check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", str(exe), str(src_dir / "check.cpp"), "-L" + ORC, "-lorc",
                    "-Wl,-rpath," + ORC], check=True)
    return exe


def _run(exe, n, seed):
    res = subprocess.run([str(exe), str(n), str(seed)], capture_output=True, text=True)
    m = re.search(r"mismatches (\d+)", res.stdout)
    assert m, res.stdout + res.stderr
    return int(m.group(1)), res


@pytest.mark.parametrize("seed", [11, 12])
def test_row_emitter_matches_the_rasterizer(tmp_path, seed):
    exe = _build(tmp_path)
    bad, res = _run(exe, 60000, seed)
    assert bad == 0 and res.returncode == 0, res.stderr


def test_planted_mistakes_are_caught(tmp_path):
    text = open(EMITTER).read()
    rounding = text.replace("(uint64_t)(e.dy - (e.dy >> 1))", "(uint64_t)(e.dy >> 1)")
    modulo = text.replace("return e.steep ? (y & 255) : y - row_y;", "return y - row_y;")
    assert rounding != text and modulo != text
    bad, _ = _run(_build(tmp_path / "a", rounding), 20000, 1)
    assert bad > 0
    bad, _ = _run(_build(tmp_path / "b", modulo), 300000, 1)
    assert bad > 0
