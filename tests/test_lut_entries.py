"""Gradient tables computed entry by entry on the device
(blend2d_amd/csrc/lut_device.h, compiled here for the host) against the
oracle's span-by-span restatement of the reference's table builder
(oracle/orc_scene.c orc_gradient_lut and orc_gradient_lut64 -- the 64-bit
tables of dithered gradients -- pinned by the reference-rendered gradient
goldens): tests/native/lut_check.cpp compares every entry of random stop
lists.  Planted mistakes in a copy of the header -- the earlier span
winning a shared entry, the tail taking the first colour, the half-step
rounding bias of the interpolation dropped -- must make it fail."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "lut_check.cpp")
HEADER = os.path.join(ROOT, "blend2d_amd", "csrc", "lut_device.h")
ORC = os.path.join(ROOT, "oracle", "_build")


def _build(tmp_path, header_text=None):
    if not shutil.which("g++") or not os.path.exists(os.path.join(ORC, "liborc.so")):
        pytest.skip("g++ or the oracle library is missing")
    src_dir = tmp_path / "tests" / "native"
    inc_dir = tmp_path / "blend2d_amd" / "csrc"
    src_dir.mkdir(parents=True)
    inc_dir.mkdir(parents=True)
    shutil.copy(SRC, src_dir / "check.cpp")
    (inc_dir / "lut_device.h").write_text(header_text if header_text is not None else open(HEADER).read())
    exe = tmp_path / "check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", str(exe), str(src_dir / "check.cpp"), "-L" + ORC, "-lorc",
                    "-Wl,-rpath," + ORC], check=True)
    return exe


def _run(exe, n, seed):
    res = subprocess.run([str(exe), str(n), str(seed)], capture_output=True, text=True)
    m = re.search(r"mismatches (\d+)", res.stdout)
    assert m, res.stdout + res.stderr
    return int(m.group(1)), res


@pytest.mark.parametrize("seed", [21, 22])
def test_entries_match_the_span_builder(tmp_path, seed):
    bad, res = _run(_build(tmp_path), 20000, seed)
    assert bad == 0 and res.returncode == 0, res.stdout


@pytest.mark.parametrize("old,new", [
    ("if (start(mid) <= e) lo = mid;", "if (start(mid) < e) lo = mid;"),
    ("if (e > last_end) return premultiplied(st[n - 1u].rgba64);", "if (e > last_end) return premultiplied(st[0].rgba64);"),
    ("const uint32_t cx = (uint32_t)(ca << 23) + (1u << 22);", "const uint32_t cx = (uint32_t)(ca << 23);"),
    # 64-bit tables: an empty span interpolating from the stop before, the
    # half-step bias dropped, both colours taken as opaque when either is
    ("const uint64_t c0 = (q == p) ? c1 : st[lo == 0 ? 0u : lo - 1u].rgba64;",
     "const uint64_t c0 = st[lo == 0 ? 0u : lo - 1u].rgba64;"),
    ("return a0 + (1u << 14) + inc * j;", "return a0 + inc * j;"),
    ("if (((c0 & c1) >> 48) == 0xFFFFu)", "if (((c0 | c1) >> 48) == 0xFFFFu)"),
])
def test_planted_mistakes_are_caught(tmp_path, old, new):
    text = open(HEADER).read()
    wrong = text.replace(old, new)
    assert wrong != text
    bad, _ = _run(_build(tmp_path, wrong), 4000, 3)
    assert bad > 0
