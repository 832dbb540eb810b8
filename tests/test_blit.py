"""blit_image / blit_scaled_image (raster/rastercontext.cpp:3938-4150).

CPU: the oracle's blits against what the operation fixes without the
reference -- an integral SRC_COPY blit writes the image's pixels, an area is
the cropped image, a scaled blit of the area's own size is a plain blit, an
integral user translation moves an integer blit, global alpha 0 and an empty
area render nothing, an area outside the image is BL_ERROR_INVALID_VALUE.
The reference's own renderings of the blit scenes are pinned in
tests/golden (the blits_* and blit_clip_robustness_* cases).

GPU: the same properties through the C ABI, and the case the reference
cannot render (a zero-height area under a rotation, where its pattern fetch
divides by zero): no pixels, no error."""
import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

INVALID_VALUE = 0x10001


def _noise(seed, w, h):
    return S.premultiplied_noise(np.random.default_rng(seed), w, h)


def _scene(fmt=S.FMT_PRGB32, w=64, h=48, clear=None):
    if clear is None:
        clear = 0xFF203040 if fmt != S.FMT_A8 else 0x30
    return S.Scene(w, h, fmt, clear)


def test_integer_src_copy_blit_writes_the_image():
    img = _noise(1, 13, 9)
    sc = _scene()
    k = sc.add_image(img)
    sc.set_comp_op(S.COMP_OP_SRC_COPY)
    sc.blit_image(S.BLIT_I, k, (5, 7))
    px, err = checkers.render_oracle(sc)
    assert err == 0
    want = np.full((48, 64), 0xFF203040, dtype=np.uint32)
    want[7:16, 5:18] = img
    assert np.array_equal(px, want)


def test_integer_blit_is_clipped_at_the_canvas():
    img = _noise(2, 20, 20)
    sc = _scene()
    k = sc.add_image(img)
    sc.set_comp_op(S.COMP_OP_SRC_COPY)
    sc.blit_image(S.BLIT_I, k, (-6, 40))
    sc.blit_image(S.BLIT_D, k, (55.0, -11.0))
    px, err = checkers.render_oracle(sc)
    assert err == 0
    want = np.full((48, 64), 0xFF203040, dtype=np.uint32)
    want[40:48, 0:14] = img[0:8, 6:20]
    want[0:9, 55:64] = img[11:20, 0:9]
    assert np.array_equal(px, want)


@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
@pytest.mark.parametrize("transform", ["identity", "fraction", "rotation", "scale"])
def test_area_is_the_cropped_image(fmt, transform):
    """A blit of an area samples exactly what a blit of the cropped image
    does: the area is the pattern's whole image (its REFLECT extend included)."""
    img = _noise(3, 31, 23)
    area = (4, 3, 17, 11)
    crop = np.ascontiguousarray(img[3:14, 4:21])
    m = {"identity": [1, 0, 0, 1, 0, 0], "fraction": [1, 0, 0, 1, 0.375, -0.8125],
         "rotation": S.rotation_about(0.4, 30, 20), "scale": [1.7, 0, 0, 0.6, 2.5, 1.0]}[transform]
    out = []
    for use_area in (True, False):
        sc = _scene(fmt, 80, 60)
        k = sc.add_image(img if use_area else crop)
        sc.set_transform([float(v) for v in m])
        for q in (0, 1):
            sc.set_pattern_quality(q)
            sc.set_global_alpha(0.75)
            sc.blit_image(S.BLIT_D, k, (9.25 + 30 * q, 5.5), area if use_area else None)
            sc.blit_image(S.BLIT_SCALED_D, k, (2.0, 30.0 + 5 * q, 45.5, 21.0), area if use_area else None)
            sc.blit_image(S.BLIT_I, k, (50, 40 - 20 * q), area if use_area else None)
        px, err = checkers.render_oracle(sc)
        assert err == 0
        out.append(px)
    assert np.array_equal(out[0], out[1])


def test_scaled_blit_of_the_area_size_is_a_plain_blit():
    img = _noise(4, 19, 14)
    outs = []
    for scaled in (False, True):
        sc = _scene()
        k = sc.add_image(img)
        sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.3, 0.7])
        if scaled:
            sc.blit_image(S.BLIT_SCALED_D, k, (7.25, 3.5, 19.0, 14.0))
            sc.blit_image(S.BLIT_SCALED_I, k, (30, 20, 19, 14))
        else:
            sc.blit_image(S.BLIT_D, k, (7.25, 3.5))
            sc.blit_image(S.BLIT_I, k, (30, 20))
        outs.append(checkers.render_oracle(sc)[0])
    assert np.array_equal(outs[0], outs[1])


def test_integral_translation_moves_an_integer_blit():
    img = _noise(5, 11, 17)
    outs = []
    for moved in (False, True):
        sc = _scene()
        k = sc.add_image(img)
        if moved:
            sc.set_transform([1.0, 0.0, 0.0, 1.0, 9.0, -4.0])
            sc.blit_image(S.BLIT_I, k, (3, 12))
        else:
            sc.blit_image(S.BLIT_I, k, (12, 8))
        outs.append(checkers.render_oracle(sc)[0])
    assert np.array_equal(outs[0], outs[1])


def test_nops_and_invalid_areas():
    img = _noise(6, 10, 10)
    sc = _scene()
    k = sc.add_image(img)
    sc.set_global_alpha(0.0)
    sc.blit_image(S.BLIT_I, k, (1, 1))
    sc.blit_image(S.BLIT_SCALED_D, k, (1.0, 1.0, 30.0, 30.0))
    sc.set_global_alpha(1.0)
    sc.set_fill_alpha(0.0)   # fill alpha does not apply to blits
    sc.blit_image(S.BLIT_D, k, (40.0, 30.0), (2, 2, 0, 5))     # empty area
    sc.blit_image(S.BLIT_SCALED_D, k, (2.0, 2.0, 0.0, 9.0))    # empty rectangle
    px, err = checkers.render_oracle(sc)
    assert err == 0
    assert np.all(px == 0xFF203040)
    for bad in [(-1, 0, 5, 5), (0, 0, 11, 10), (6, 6, 5, 4), (0, 0, -1, 3), (10, 0, 1, 1)]:
        sc = _scene()
        k = sc.add_image(img)
        sc.set_global_alpha(0.0)   # an invalid area is reported even by a NOP blit
        sc.blit_image(S.BLIT_I, k, (0, 0), bad)
        px, err = checkers.render_oracle(sc)
        assert err == INVALID_VALUE, bad
        assert np.all(px == 0xFF203040)


def test_fill_alpha_is_ignored_and_global_alpha_applies():
    img = np.full((8, 8), 0xFFFFFFFF, dtype=np.uint32)
    sc = _scene(clear=0xFF000000)
    k = sc.add_image(img)
    sc.set_fill_alpha(0.25)
    sc.set_global_alpha(128 / 255)
    sc.blit_image(S.BLIT_I, k, (0, 0))
    px, _ = checkers.render_oracle(sc)
    # SRC_OVER of opaque white at alpha 128 over opaque black: 128 per channel
    assert np.all(px[0:8, 0:8] == 0xFF808080)


# ---- GPU -----------------------------------------------------------------

def _device():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
def test_gpu_blit_properties_match_the_oracle():
    b = _device()
    rng = np.random.default_rng(7)
    for fmt in (S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8):
        sc = _scene(fmt, 120, 90)
        imgs = [sc.add_image(_noise(8, 37, 29)), sc.add_image(rng.integers(0, 256, (13, 21), dtype=np.uint8), S.FMT_A8)]
        for i in range(40):
            sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.integers(-3, 4)), 0.0] if i % 3 else
                             S.rotation_about(float(rng.uniform(0, 6)), 60, 45))
            sc.set_pattern_quality(i % 2)
            sc.set_global_alpha(float(rng.uniform(0.3, 1.0)))
            k = imgs[i % 2]
            sc.blit_image(i % 4, k, (float(rng.uniform(-20, 110)), float(rng.uniform(-20, 80)),
                                     float(rng.uniform(-30, 90)), float(rng.uniform(1, 70))),
                          (1, 2, 9, 7) if i % 5 == 0 else None)
        want, werr = checkers.render_oracle(sc)
        got, gerr = b.render_scene(sc)
        assert gerr == werr
        assert np.array_equal(got, want), fmt


@pytest.mark.gpu
def test_gpu_zero_height_area_renders_nothing():
    """A zero-height area under a rotation or a scale: the reference's
    pattern fetch divides by the area height there (SIGFPE); this backend
    renders nothing and reports no error, while a solid override (an
    XRGB32 image onto A8, whose alpha is always opaque) still fills its
    destination box, as in the reference, which sets no pattern up then."""
    b = _device()
    img = _noise(9, 20, 20)
    sc = _scene(S.FMT_PRGB32, 50, 40)
    k = sc.add_image(img)
    sc.set_transform(S.rotation_about(0.3, 25, 20))
    sc.blit_image(S.BLIT_D, k, (3.0, 4.0), (2, 5, 10, 0))
    sc.set_transform([2.0, 0.0, 0.0, 2.0, 0.0, 0.0])
    sc.blit_image(S.BLIT_SCALED_D, k, (3.0, 4.0, 20.0, 10.0), (2, 5, 10, 0))
    px, err = b.render_scene(sc)
    assert err == 0 and np.all(px == 0xFF203040)

    sc = _scene(S.FMT_A8, 50, 40, clear=0)
    k = sc.add_image(np.full((20, 20), 0x00123456, dtype=np.uint32), S.FMT_XRGB32)
    sc.set_transform([2.0, 0.0, 0.0, 2.0, 0.0, 0.0])
    sc.blit_image(S.BLIT_SCALED_I, k, (3, 4, 10, 5), (2, 5, 10, 0))
    px, err = b.render_scene(sc)
    want = np.zeros((40, 50), dtype=np.uint8)
    want[8:18, 6:26] = 255
    assert err == 0 and np.array_equal(px, want)
