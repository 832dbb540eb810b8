"""Fills whose geometry the reference rejects (NaN reaching its line
clipper: BL_ERROR_INVALID_GEOMETRY, edges reverted, raster/edgebuilder_p.h:
1494-1496 and 1584-1586, raster/rastercontextops.cpp:38-45).  The device
finds these while it builds edges, so the fill call returns success and the
flush reports the failing call (bl_amd_context_get_deferred_error) and the
accumulated BLContextErrorFlags; the pixels must equal the oracle's, which
restates the reference (and is pinned against it by the invalid_geometry_*
golden cases)."""
import numpy as np
import pytest

import checkers
from blend2d_amd import scene as S

NAN = float("nan")
INVALID_GEOMETRY = 0x10034
FLAG_INVALID_GEOMETRY = 0x4


def _scene(fmt=S.FMT_PRGB32):
    """Rendering calls: 0 valid polygon, 1 polygon with a NaN second vertex,
    2 valid cubic path, 3 cubic path with a NaN control point, 4 valid rect."""
    sc = S.Scene(120, 90, fmt, 0xFF102030 if fmt != S.FMT_A8 else 0x10)
    sc.set_fill_style_rgba32(0xC0FF8040)
    sc.fill_polygon([(10.0, 10.0), (70.0, 15.0), (40.0, 60.0)])
    sc.set_fill_style_rgba32(0xC040FF80)
    sc.fill_polygon([(50.0, 20.0), (NAN, 30.0), (100.0, 80.0), (60.0, 85.0)])
    sc.fill_path([S.CMD_MOVE, S.CMD_CUBIC, S.CMD_CUBIC, S.CMD_CUBIC, S.CMD_CLOSE],
                 [(5.0, 80.0), (30.0, 20.0), (80.0, 120.0), (110.0, 40.0), (5.0, 80.0)])
    sc.fill_path([S.CMD_MOVE, S.CMD_CUBIC, S.CMD_CUBIC, S.CMD_CUBIC, S.CMD_CLOSE],
                 [(20.0, 20.0), (NAN, 10.0), (90.0, 70.0), (100.0, 20.0), (20.0, 20.0)])
    sc.fill_rect_d(30.5, 40.25, 50.0, 20.0)
    return sc


def test_oracle_reports_the_first_failing_call_and_draws_the_rest():
    sc = _scene()
    px, err = checkers.render_oracle(sc)
    assert err == INVALID_GEOMETRY
    # the same scene without the two failing fills renders the same pixels
    ok = _scene()
    ok.cmds = [c for k, c in enumerate(ok.cmds) if k not in (3, 5)]
    px2, err2 = checkers.render_oracle(ok)
    assert err2 == 0
    assert (px == px2).all()


def test_oracle_nan_line_matches_reference_when_the_reference_is_available():
    if not checkers.have_ref():
        pytest.skip("reference library not built")
    sc = _scene()
    sc.cmds = sc.cmds[:5]   # the cubic with a NaN control point has no defined reference result
    r, re, _ = checkers.render_reference(sc)
    o, oe = checkers.render_oracle(sc)
    assert re == oe == INVALID_GEOMETRY
    assert (r == o).all()


def _gpu():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_A8])
def test_gpu_deferred_error_names_the_failing_call(fmt):
    amd = _gpu()
    sc = _scene(fmt)
    img = amd.BLImage(sc.width, sc.height, sc.format)
    img.fill_raw(sc.clear_value)
    ctx = amd.BLContext(img)
    errs = amd.replay_into(ctx, sc, [])
    assert errs == []          # the calls themselves succeed; the device finds the failures
    ctx.flush()
    assert ctx.deferred_error() == (1, INVALID_GEOMETRY)
    assert ctx.error_flags() == FLAG_INVALID_GEOMETRY
    ctx.end()
    want, werr = checkers.render_oracle(sc)
    assert werr == INVALID_GEOMETRY
    got = img.download()
    assert (got == want).all()


@pytest.mark.gpu
def test_gpu_batched_contexts_get_their_own_failures():
    amd = _gpu()
    scenes = [_scene(), S.polygons_scene(3, 64, 48, 20), _scene(S.FMT_A8)]
    outs, errs = amd.render_scenes_batched(scenes)
    assert errs == [INVALID_GEOMETRY, 0, INVALID_GEOMETRY]
    for sc, got in zip(scenes, outs):
        want, _ = checkers.render_oracle(sc)
        assert (got == want).all()


@pytest.mark.gpu
def test_gpu_error_flags_accumulate_until_begin():
    amd = _gpu()
    img = amd.BLImage(64, 64, S.FMT_PRGB32)
    img.fill_raw(0)
    ctx = amd.BLContext(img)
    ctx.fill_polygon(np.array([(5.0, 5.0), (NAN, 9.0), (40.0, 50.0)]))
    ctx.flush()
    assert ctx.deferred_error() == (0, INVALID_GEOMETRY)
    ctx.fill_rect(1.0, 1.0, 10.0, 10.0)
    ctx.flush()
    assert ctx.deferred_error() is None                 # per flush
    assert ctx.error_flags() == FLAG_INVALID_GEOMETRY   # accumulated
    ctx.end()
    ctx.begin(img)
    assert ctx.error_flags() == 0
    ctx.end()
