"""Object lifetimes at the C ABI: images stay allocated while a context,
a fill style or a replayable frame uses them (the reference defers the free
of an image that still has users, core/image.cpp:171-175), and a batch
refuses contexts whose canvases would race (one context twice, two contexts
on one image; the reference renders each context's calls in order)."""
import gc

import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

pytestmark = pytest.mark.gpu

E_INVALID_VALUE = 0x10001


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


def _pattern_scene():
    rng = np.random.default_rng(4)
    sc = S.Scene(96, 64, S.FMT_PRGB32, 0xFF000000)
    idx = sc.add_image(S.premultiplied_noise(rng, 33, 17))
    sc.set_fill_style_pattern(idx, S.EXTEND_REPEAT, (1.5, 0.2, -0.3, 1.2, 5.0, 3.0))
    sc.fill_polygon([(5.0, 5.0), (90.0, 12.0), (60.0, 60.0), (8.0, 50.0)])
    return sc


def test_pattern_image_outlives_its_python_object(amd):
    sc = _pattern_scene()
    want, _ = checkers.render_oracle(sc)
    pat = amd.BLImage(33, 17, S.FMT_PRGB32)
    pat.upload(sc.images[0][1])
    target = amd.BLImage(sc.width, sc.height, sc.format)
    target.fill_raw(sc.clear_value)
    ctx = amd.BLContext(target)
    ctx.set_fill_style(amd.BLPattern(pat, S.EXTEND_REPEAT, (1.5, 0.2, -0.3, 1.2, 5.0, 3.0)))
    ctx.flush()                  # nothing recorded yet; the style keeps the image
    pat.close()                  # the caller's reference is gone
    del pat
    gc.collect()
    # a fill recorded after the flush still samples the pattern
    ctx.fill_polygon(np.array([(5.0, 5.0), (90.0, 12.0), (60.0, 60.0), (8.0, 50.0)]))
    ctx.flush()
    assert (target.download() == want).all()
    # the resident frame can still be replayed
    target.fill_raw(sc.clear_value)
    ctx.replay()
    assert (target.download() == want).all()
    ctx.end()


def test_target_outlives_its_python_object_while_attached(amd):
    sc = S.polygons_scene(2, 64, 48, 20)
    want, _ = checkers.render_oracle(sc)
    target = amd.BLImage(sc.width, sc.height, sc.format)
    target.fill_raw(sc.clear_value)
    handle = target.handle.value
    ctx = amd.BLContext(target)
    amd.replay_into(ctx, sc, [])
    ctx.flush()
    target.close()               # the attached context still holds the pixels
    ctx.replay()                 # composites the frame a second time into them
    # read the pixels through the C ABI with the raw handle, still valid
    import ctypes
    out = np.empty((sc.height, sc.width), dtype=np.uint32)
    assert amd.lib().bl_amd_image_download(ctypes.c_void_p(handle), out.ctypes.data, sc.width * 4) == 0
    twice = S.polygons_scene(2, 64, 48, 20)
    twice.cmds = twice.cmds + twice.cmds
    want2, _ = checkers.render_oracle(twice)
    assert (out == want2).all()
    ctx.end()
    ctx.close()


def test_batch_rejects_a_context_twice_and_a_shared_target(amd):
    img = amd.BLImage(32, 32, S.FMT_PRGB32)
    other = amd.BLImage(32, 32, S.FMT_PRGB32)
    a, b, c = amd.BLContext(img), amd.BLContext(img), amd.BLContext(other)
    batch = amd.BLBatch()
    batch.add(a)
    with pytest.raises(amd.BLError) as e:
        batch.add(a)
    assert e.value.code == E_INVALID_VALUE
    with pytest.raises(amd.BLError) as e:
        batch.add(b)
    assert e.value.code == E_INVALID_VALUE
    batch.add(c)
    # the refused context renders by its own flush, in order after `a`
    img.fill_raw(0xFF000000)
    a.fill_rect(2, 2, 20, 20, 0x80FF0000)
    b.fill_rect(10, 10, 20, 20, 0x8000FF00)
    c.fill_rect(0, 0, 8, 8, 0xFFFFFFFF)
    batch.flush()
    b.flush()
    sc = S.Scene(32, 32, S.FMT_PRGB32, 0xFF000000)
    sc.fill_rect_i_rgba32(2, 2, 20, 20, 0x80FF0000)
    sc.fill_rect_i_rgba32(10, 10, 20, 20, 0x8000FF00)
    want, _ = checkers.render_oracle(sc)
    assert (img.download() == want).all()
    for ctx in (a, b, c):
        ctx.end()


def test_reset_renders_pending_calls_and_detaches(amd):
    img = amd.BLImage(40, 30, S.FMT_PRGB32)
    img.fill_raw(0xFF000000)
    ctx = amd.BLContext(img)
    ctx.fill_rect(5, 5, 20, 10, 0xFF00FF00)
    ctx.reset()
    sc = S.Scene(40, 30, S.FMT_PRGB32, 0xFF000000)
    sc.fill_rect_i_rgba32(5, 5, 20, 10, 0xFF00FF00)
    want, _ = checkers.render_oracle(sc)
    assert (img.download() == want).all()
    # detached: rendering calls fail until the next begin()
    with pytest.raises(amd.BLError) as e:
        ctx.fill_all()
    assert e.value.code == 0x10002
    ctx.begin(img)
    ctx.end()


def test_explicit_style_rejects_an_unknown_kind(amd):
    img = amd.BLImage(16, 16, S.FMT_PRGB32)
    ctx = amd.BLContext(img)
    st = amd.BLAmdStyle()
    st.kind = 9
    assert amd.lib().bl_amd_context_fill_all_ext(ctx.handle, st) == E_INVALID_VALUE
    ctx.end()
