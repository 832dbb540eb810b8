"""Tile schedules of the render stage on the GPU, through the C ABI: one
workgroup per tile (GRID), a persistent grid draining a device work queue in
band order (QUEUE), the queue sorted longest first by a segment-count
heuristic (QUEUE_HEURISTIC) or by the tile cost model (QUEUE_LEARNED), the
autotuner that picks among them, and tile-cost logging.  A schedule only
decides which workgroup renders which tile, when; tiles own disjoint pixels,
so every schedule must give the oracle's pixels (tolerance 0)."""
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
    yield blend2d_amd
    blend2d_amd.set_render_schedule(blend2d_amd.SCHEDULE_AUTO)
    blend2d_amd.tile_model_set(None)
    blend2d_amd.autotune_reset()


def _random_model(seed=3):
    from blend2d_amd import train_cost_model as TC
    return TC.init_params(seed)


def _check_single(amd, sc):
    got, err, stats = amd.render_scene(sc, return_context=True)
    want, oerr = checkers.render_oracle(sc)
    assert err == oerr
    mx, mean, n = checkers.diff_info(got, want)
    assert mx == 0, f"max diff {mx}, mean {mean}, {n} channels differ"
    return stats


def _check_batched(amd, scenes):
    outs, errs = amd.render_scenes_batched(scenes)
    for sc, out, err in zip(scenes, outs, errs):
        want, oerr = checkers.render_oracle(sc)
        assert err == oerr
        mx, mean, n = checkers.diff_info(out, want)
        assert mx == 0, f"max diff {mx}, mean {mean}, {n} channels differ"


SCHEDULES = [(0, 0), (1, 4), (2, 2), (2, 3), (2, 4), (3, 3), (3, 4)]


@pytest.mark.parametrize("schedule,wgs", SCHEDULES)
def test_every_schedule_gives_the_oracles_pixels(amd, schedule, wgs):
    amd.tile_model_set(_random_model() if schedule == amd.SCHEDULE_QUEUE_LEARNED else None)
    amd.set_render_schedule(schedule, wgs)
    try:
        stats = _check_single(amd, S.cubic_gradient_scene(4, 640, 360, 250))
        assert stats["schedule"] == schedule
        assert stats["queue_wgs"] == (0 if schedule == 0 else wgs)
        _check_batched(amd, S.glyph_canvases(6, 20, 30))
        _check_batched(amd, S.cost_canvases(5, 4, 128, 384, 16))
        _check_single(amd, S.pattern_shapes_scene(2, 400, 300, 40, pattern_size=64))
    finally:
        amd.set_render_schedule(amd.SCHEDULE_AUTO)
        amd.tile_model_set(None)


def test_learned_schedule_without_a_model_runs_the_heuristic(amd):
    amd.tile_model_set(None)
    amd.set_render_schedule(amd.SCHEDULE_QUEUE_LEARNED, 4)
    try:
        stats = _check_single(amd, S.polygons_scene(3, 256, 256, 120))
        assert stats["schedule"] == amd.SCHEDULE_QUEUE_HEURISTIC
    finally:
        amd.set_render_schedule(amd.SCHEDULE_AUTO)


def test_queue_with_learned_sizing(amd):
    from blend2d_amd import train_cost_model as TC
    params = TC.load()
    if params is None:
        pytest.skip("no trained sizing model")
    amd.cost_model_set(params)
    amd.set_sizing(amd.SIZING_LEARNED)
    amd.tile_model_set(_random_model(5))
    amd.set_render_schedule(amd.SCHEDULE_QUEUE_LEARNED, 4)
    try:
        _check_batched(amd, S.glyph_canvases(8, 16, 40))
    finally:
        amd.set_sizing(amd.SIZING_EXACT)
        amd.cost_model_set(None)
        amd.tile_model_set(None)
        amd.set_render_schedule(amd.SCHEDULE_AUTO)


def _resident_batch(amd, scenes):
    keep = []
    batch = amd.BLBatch()
    for sc in scenes:
        img = amd.BLImage(sc.width, sc.height, sc.format)
        img.fill_raw(sc.clear_value)
        ctx = amd.BLContext(img)
        amd.replay_into(ctx, sc, [])
        batch.add(ctx)
        keep.append((img, ctx))
    batch.flush()
    return batch, keep


def test_autotune_records_a_schedule_that_auto_then_runs(amd):
    amd.autotune_reset()
    amd.tile_model_set(_random_model())
    scenes = S.glyph_canvases(9, 24, 40)
    batch, keep = _resident_batch(amd, scenes)
    try:
        tune = amd.autotune(batch, reps=2)
        assert tune["candidates"] == 8
        assert 0 <= tune["schedule"] <= 3
        assert tune["best_ms"] <= tune["grid_ms"]
        # a replay onto re-cleared canvases runs the tuned schedule and
        # gives the oracle's pixels
        for (img, ctx), sc in zip(keep, scenes):
            img.fill_raw(sc.clear_value)
        batch.replay()
        st = batch.last_stats()
        assert st["schedule"] == tune["schedule"]
        log = amd.autotune_log()
        assert log.count("\n") == 1 and "grid/0" in log and "queue_learned/4" in log
        for (img, ctx), sc in zip(keep, scenes):
            want, _ = checkers.render_oracle(sc)
            assert checkers.diff_info(img.download(), want)[0] == 0
    finally:
        for img, ctx in keep:
            ctx.end()
        batch.close()
        amd.autotune_reset()
        amd.tile_model_set(None)


def test_tile_costs_are_logged_per_analytic_tile(amd):
    sc = S.cubic_gradient_scene(6, 512, 256, 150)
    img = amd.BLImage(sc.width, sc.height, sc.format)
    img.fill_raw(sc.clear_value)
    ctx = amd.BLContext(img)
    amd.replay_into(ctx, sc, [])
    ctx.flush()
    try:
        X, T = amd.collect_tile_costs(ctx)
        stats = ctx.last_stats()
        tiles_x, tiles_y = (sc.width + 63) // 64, (sc.height + 15) // 16
        assert 0 < len(X) <= tiles_x * tiles_y
        assert np.isfinite(X).all() and np.isfinite(T).all()
        assert (T[:, 0] > 0).all()
        # composited pixels of the analytic tiles add up to at most the frame's
        px = np.round(np.exp2(T[:, 1].astype(np.float64)) - 1.0).sum()
        assert 0 < px <= stats["pixels"]
        assert (X[:, 12] > 0).all() and (X[:, 12] <= 1).all() and (X[:, 13] <= 1).all()
        # the logging replay rendered the frame once more: onto a re-cleared
        # canvas it gives the oracle's pixels
        img.fill_raw(sc.clear_value)
        X2, _ = amd.collect_tile_costs(ctx)
        assert np.array_equal(X, X2)
        want, _ = checkers.render_oracle(sc)
        assert checkers.diff_info(img.download(), want)[0] == 0
    finally:
        ctx.end()
