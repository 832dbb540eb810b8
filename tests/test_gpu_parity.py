"""GPU parity: the MI355X backend against the CPU oracle (and the reference
library when it is built) on seeded scenes, through the C ABI.

Tolerances (see the module docstrings of the kernels): coverage, solid fills,
linear gradients, nearest and bilinear patterns and every integer comp op are
compared bit-exact; radial and conic gradients are computed with the same
float operation order and are also expected bit-exact, with at most 1 LSB
allowed per channel."""
import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


def _cmp(amd, sc, tol=0):
    got, err = amd.render_scene(sc)
    want, oerr = checkers.render_oracle(sc)
    assert err == oerr
    mx, mean, n = checkers.diff_info(got, want)
    assert mx <= tol, f"max diff {mx}, mean {mean}, {n} channels differ"
    return got


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_polygons_exact(amd, seed):
    _cmp(amd, S.polygons_scene(seed, 512, 512, 200))


@pytest.mark.parametrize("seed", [1, 2])
def test_cubic_gradients(amd, seed):
    _cmp(amd, S.cubic_gradient_scene(seed, 640, 360, 200), tol=1)


@pytest.mark.parametrize("seed", [1, 2])
def test_patterns_mixed_ops(amd, seed):
    _cmp(amd, S.pattern_shapes_scene(seed, 800, 600, 80, pattern_size=96))


def test_glyph_canvases_a8_and_prgb(amd):
    for sc in S.glyph_canvases(3, 12, 30):
        _cmp(amd, sc)


def test_batched_equals_single(amd):
    scenes = S.glyph_canvases(5, 24, 20)
    outs, errs = amd.render_scenes_batched(scenes)
    for sc, out, err in zip(scenes, outs, errs):
        want, oerr = checkers.render_oracle(sc)
        assert err == oerr
        assert checkers.diff_info(out, want)[0] == 0
