"""bl_amd_context_replay_noalloc: the replay whose intermediate buffers are
slices of the library's fixed per-device scratch region.  Called through the
C ABI and compared bit-exact with bl_amd_context_replay (the entry point it
varies) on two calls in a row, the second with a smaller frame (the region
keeps the first call's contents), and a frame too large for the region must
fail with BL_AMD_ERROR_DATA_TOO_LARGE without touching the target."""
import numpy as np
import pytest

from blend2d_amd import scene as S

pytestmark = pytest.mark.gpu

E_DATA_TOO_LARGE = 0x10032


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


def _resident(amd, sc):
    images = []
    for fmt, px in sc.images:
        im = amd.BLImage(px.shape[1], px.shape[0], fmt)
        im.upload(px)
        images.append(im)
    img = amd.BLImage(sc.width, sc.height, sc.format)
    img.fill_raw(sc.clear_value)
    ctx = amd.BLContext(img)
    amd.replay_into(ctx, sc, images)
    ctx.flush()
    return ctx, img, images


def _both(amd, sc):
    """(pixels, stats) of replay and of replay_noalloc on a re-cleared target."""
    ctx, img, keep = _resident(amd, sc)
    out = []
    for name in ("replay", "replay_noalloc"):
        img.fill_raw(sc.clear_value)
        r = getattr(amd.lib(), "bl_amd_context_" + name)(ctx.handle)
        assert r == 0, (name, hex(r))
        out.append((img.download(), ctx.last_stats()))
    ctx.end()
    return out


def test_noalloc_replay_matches_replay_twice_second_smaller(amd):
    big = S.cubic_gradient_scene(3, 320, 200, 120)
    small = S.polygons_scene(4, 90, 70, 25)
    for sc in (big, small):
        (p_ref, st_ref), (p_na, st_na) = _both(amd, sc)
        assert p_ref.shape == p_na.shape
        assert (p_ref == p_na).all()
        for k in ("commands", "segments", "binned", "pixels"):
            assert st_ref[k] == st_na[k], k


def test_noalloc_replay_rejects_a_frame_larger_than_its_region(amd):
    sc = S.cubic_gradient_scene(5, 1920, 1080, 10000)   # ~100 MB of intermediate buffers
    ctx, img, keep = _resident(amd, sc)
    img.fill_raw(0x12345678)
    r = amd.lib().bl_amd_context_replay_noalloc(ctx.handle)
    assert r == E_DATA_TOO_LARGE
    assert (img.download() == 0x12345678).all()
    ctx.end()
