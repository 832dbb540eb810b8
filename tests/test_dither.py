"""Dithered gradients (BL_CONTEXT_HINT_GRADIENT_QUALITY = DITHER:
pipeline/reference/fetchgeneric_p.h:899-934, pixelops/interpolation.cpp:
161-275).  The reference's renderings are pinned in tests/golden
(dither_*) and the 64-bit tables in tests/test_lut_entries.py.

CPU: what the hint fixes without the reference -- A8 targets never dither;
a style keeps the hint of the moment it was set; SMOOTH renders as NEAREST;
a two-stop opaque ramp dithers within one level of the undithered pixels,
alpha untouched; out-of-range hints are BL_ERROR_INVALID_VALUE.  GPU: batched frames that mix dithered (64-bit) and
plain tables, and single canvases, against the oracle."""
import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

INVALID_VALUE = 0x10001


def _ramp(q, fmt=S.FMT_PRGB32, late_hint=None):
    sc = S.Scene(256, 64, fmt, 0xFF000000 if fmt != S.FMT_A8 else 0)
    sc.set_gradient_quality(q)
    sc.set_fill_style_gradient(S.GRADIENT_LINEAR, [0.0, 0.0, 256.0, 0.0], [(0.0, 0xFF000000), (1.0, 0xFF3060FF)], 0, None)
    if late_hint is not None:
        sc.set_gradient_quality(late_hint)
    sc.fill_all()
    px, err = checkers.render_oracle(sc)
    assert err == 0
    return px


def test_a8_targets_never_dither():
    assert np.array_equal(_ramp(2, S.FMT_A8), _ramp(0, S.FMT_A8))


def test_smooth_renders_as_nearest():
    assert np.array_equal(_ramp(1), _ramp(0))


def test_the_style_keeps_the_hint_it_was_set_with():
    assert np.array_equal(_ramp(2, late_hint=0), _ramp(2))
    assert np.array_equal(_ramp(0, late_hint=2), _ramp(0))
    assert not np.array_equal(_ramp(2), _ramp(0))


def test_dithered_ramp_stays_within_a_level():
    d, n = _ramp(2), _ramp(0)
    assert not np.array_equal(d, n)
    for sh in (0, 8, 16, 24):
        dc = ((d >> sh) & 0xFF).astype(np.int32)
        nc = ((n >> sh) & 0xFF).astype(np.int32)
        assert np.abs(dc - nc).max() <= 1
    # opaque stops: alpha is never dithered
    assert np.all((d >> 24) == 0xFF)


def test_out_of_range_hint_is_invalid_and_ignored():
    sc = S.Scene(32, 16, S.FMT_PRGB32, 0xFF000000)
    sc.set_gradient_quality(2)
    sc.set_gradient_quality(3)
    sc.set_fill_style_gradient(S.GRADIENT_LINEAR, [0.0, 0.0, 32.0, 0.0], [(0.0, 0xFF000000), (1.0, 0xFFFFFFFF)], 0, None)
    sc.fill_all()
    px, err = checkers.render_oracle(sc)
    assert err == INVALID_VALUE
    sc2 = S.Scene(32, 16, S.FMT_PRGB32, 0xFF000000)
    sc2.set_gradient_quality(2)
    sc2.set_fill_style_gradient(S.GRADIENT_LINEAR, [0.0, 0.0, 32.0, 0.0], [(0.0, 0xFF000000), (1.0, 0xFFFFFFFF)], 0, None)
    sc2.fill_all()
    assert np.array_equal(px, checkers.render_oracle(sc2)[0])


def _device():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_gpu_batched_dither_scenes_match_the_oracle(fmt):
    """Canvases of one batched frame mixing 64-bit (dithered) and 32-bit
    tables: table offsets are rebased per canvas in 32-bit words."""
    b = _device()
    scenes = [S.dither_scene(seed, 140 + 11 * seed, 90 + 9 * seed, fmt, 60) for seed in (41, 42, 43, 44)]
    outs, errs = b.render_scenes_batched(scenes)
    for sc, px, err in zip(scenes, outs, errs):
        want, werr = checkers.render_oracle(sc)
        assert err == werr
        assert np.array_equal(px, want)
