"""fill_mask (raster/rastercontext.cpp:3590-3780).

CPU: the oracle's mask fills against what the operation fixes -- a constant
mask v composites as a box fill at alpha v (fill alpha multiplied in with
udiv255 below full alpha), an all-zero mask changes nothing, a 32-bit image
used as a mask is the A8 mask of its row bytes (the reference reads a mask
as bytes whatever its format), a fractional start or a rotation is
BL_ERROR_NOT_IMPLEMENTED unless the fill is a NOP, an area outside the mask
is BL_ERROR_INVALID_VALUE.  The reference's own renderings of mask scenes
are pinned in tests/golden (masks_*).

GPU: batched mask scenes (the mask's image index is rebased per canvas)
and single ones against the oracle."""
import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

INVALID_VALUE, NOT_IMPLEMENTED = 0x10001, 0x10007


def _scene(fmt=S.FMT_PRGB32, w=64, h=48):
    return S.Scene(w, h, fmt, 0xFF203040 if fmt != S.FMT_A8 else 0x30)


def _div255(x):
    x = np.asarray(x, dtype=np.uint32) + 128
    return (x + (x >> 8)) >> 8


@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
@pytest.mark.parametrize("v", [0, 1, 77, 128, 254, 255])
@pytest.mark.parametrize("fill_alpha", [1.0, 0.6])
def test_constant_mask_is_a_box_fill_at_that_alpha(fmt, v, fill_alpha):
    outs = []
    fa_i = int(np.floor(fill_alpha * 255 + 0.5))
    m = v if fa_i >= 255 else int(_div255(v * fa_i))
    for masked in (True, False):
        sc = _scene(fmt)
        k = sc.add_image(np.full((13, 17), v, dtype=np.uint8), S.FMT_A8)
        sc.set_fill_style_rgba32(0xC0508020)
        if masked:
            sc.set_fill_alpha(fill_alpha)
            sc.fill_mask(S.MASK_I, k, (5, 9))
        else:
            sc.set_global_alpha(m / 255.0)
            sc.fill_rect_i(5, 9, 17, 13)
        px, err = checkers.render_oracle(sc)
        assert err == 0
        outs.append(px)
    assert np.array_equal(outs[0], outs[1])


def test_zero_mask_changes_nothing():
    sc = _scene()
    k = sc.add_image(np.zeros((20, 20), dtype=np.uint8), S.FMT_A8)
    sc.set_comp_op(S.COMP_OP_SRC_COPY)
    sc.fill_mask(S.MASK_D, k, (3.0, 4.0))
    px, err = checkers.render_oracle(sc)
    assert err == 0 and np.all(px == 0xFF203040)


def test_a_32bit_mask_is_read_as_bytes():
    """Area x = 3 of a 12-pixel-wide 32-bit mask starts at byte 12 of its
    row; a 9-pixel-wide fill reads bytes 12 .. 20 (the reference's VMaskA8
    command over the row), the same as an A8 mask holding those bytes."""
    img = S.premultiplied_noise(np.random.default_rng(3), 12, 9)
    rows = img.view(np.uint8).reshape(9, 48)
    outs = []
    for mask, area in ((img, (3, 0, 9, 9)), (np.ascontiguousarray(rows[:, 12:21]), None),
                       (np.ascontiguousarray(rows[:, 3:12]), None)):
        sc = _scene()
        k = sc.add_image(mask) if mask.dtype == np.uint32 else sc.add_image(mask, S.FMT_A8)
        sc.set_fill_style_rgba32(0xFF3366CC)
        sc.fill_mask(S.MASK_I, k, (20, 30), area)
        px, err = checkers.render_oracle(sc)
        assert err == 0
        outs.append(px)
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])


def test_unrendered_placements_and_errors():
    mask = np.full((10, 10), 200, dtype=np.uint8)
    cases = [
        ([1.0, 0.0, 0.0, 1.0, 0.0, 0.0], S.MASK_D, (3.5, 4.0), None, 1.0, NOT_IMPLEMENTED),
        ([1.0, 0.0, 0.0, 1.0, 0.25, 0.0], S.MASK_I, (3, 4), None, 1.0, NOT_IMPLEMENTED),
        ([1.0, 0.0, 0.0, 1.0, 0.5, 0.0], S.MASK_D, (3.5, 4.0), None, 1.0, 0),       # aligned again
        ([2.0, 0.0, 0.0, 2.0, 0.0, 0.0], S.MASK_I, (3, 4), None, 1.0, NOT_IMPLEMENTED),
        ([2.0, 0.0, 0.0, 2.0, 0.0, 0.0], S.MASK_I, (3, 4), None, 0.0, 0),           # a NOP
        ([1.0, 0.0, 0.0, 1.0, 0.0, 0.0], S.MASK_D, (500.5, 4.0), None, 1.0, 0),     # clipped out first
        ([1.0, 0.0, 0.0, 1.0, 0.0, 0.0], S.MASK_I, (3, 4), (5, 5, 6, 1), 1.0, INVALID_VALUE),
        ([1.0, 0.0, 0.0, 1.0, 0.0, 0.0], S.MASK_I, (3, 4), (5, 5, 6, 1), 0.0, INVALID_VALUE),
    ]
    for m, kind, dst, area, ga, want in cases:
        sc = _scene()
        k = sc.add_image(mask, S.FMT_A8)
        sc.set_transform(m)
        sc.set_global_alpha(ga)
        sc.fill_mask(kind, k, dst, area)
        px, err = checkers.render_oracle(sc)
        assert err == want, (m, kind, dst, area, ga)
        if want:
            assert np.all(px == 0xFF203040)


# ---- GPU -----------------------------------------------------------------

def _device():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_gpu_batched_mask_scenes_match_the_oracle(fmt):
    """Several mask canvases in one batched frame: each command's mask image
    index is rebased to its canvas's images."""
    b = _device()
    scenes = [S.mask_scene(seed, 150 + 13 * seed, 97 + 7 * seed, fmt, 70) for seed in (21, 22, 23, 24)]
    outs, errs = b.render_scenes_batched(scenes)
    for sc, px, err in zip(scenes, outs, errs):
        want, werr = checkers.render_oracle(sc)
        assert err == werr
        assert np.array_equal(px, want)


@pytest.mark.gpu
def test_gpu_mask_errors_match_the_oracle():
    b = _device()
    for seed in (31, 32):
        sc = S.mask_scene(seed, 211, 150, S.FMT_PRGB32, 90, unaligned=True)
        want, werr = checkers.render_oracle(sc)
        got, gerr = b.render_scene(sc)
        assert gerr == werr
        assert np.array_equal(got, want)
