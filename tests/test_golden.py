"""The golden fixtures (tests/golden/golden.json, reference renderings of the
scenes in tests/golden/cases.py) against the CPU oracle, and against the
MI355X backend on a GPU.  Bit-exact: sha256 of every canvas."""
import json
import os
import sys

import pytest

import checkers

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))
CASES = cases.cases()


def _first_bad_band(got, want):
    for i, (a, b) in enumerate(zip(got["band_crc32"], want["band_crc32"])):
        if a != b:
            return i * 64
    return None


def test_every_case_has_a_fixture():
    assert sorted(GOLDEN) == sorted(CASES)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    scenes = CASES[name]()
    assert len(scenes) == len(GOLDEN[name])
    for k, (sc, want) in enumerate(zip(scenes, GOLDEN[name])):
        px, err = checkers.render_oracle(sc)
        got = cases.digest(px)
        assert err == want["error"], (name, k)
        assert got["sha256"] == want["sha256"], f"{name}[{k}]: first differing row band at y={_first_bad_band(got, want)}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_golden(name):
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    scenes = CASES[name]()
    if len(scenes) > 1:
        outs, errs = blend2d_amd.render_scenes_batched(scenes)
    else:
        out, err = blend2d_amd.render_scene(scenes[0])
        outs, errs = [out], [err]
    for k, (px, err, want) in enumerate(zip(outs, errs, GOLDEN[name])):
        got = cases.digest(px)
        assert err == want["error"], (name, k)
        assert got["sha256"] == want["sha256"], f"{name}[{k}]: first differing row band at y={_first_bad_band(got, want)}"


def test_small_canvas_scene_renders_in_the_oracle():
    # the scene of __graft_entry__.smoke(): boxes larger than the canvas
    from blend2d_amd import scene as S
    sc = S.cubic_gradient_scene(seed=7, width=256, height=192, count=40)
    px, err = checkers.render_oracle(sc)
    assert err == 0 and px.shape == (192, 256)
