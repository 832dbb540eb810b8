"""Graph replays on the GPU, through the C ABI: with graph replay on, a
flush runs with exact sizing and records its counts, the first replay of the
resident frame captures its device stages into one HIP graph and later
replays launch it.  Every replay -- with eager work (canvas
clears, downloads) and synchronizations between them, a schedule change
that forces a new capture, and a new upload -- must give the eager flush's
pixels and pixel count."""
import pytest

import checkers
from blend2d_amd import scene as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    blend2d_amd.set_graph_replay(True)
    yield blend2d_amd
    blend2d_amd.set_graph_replay(False)
    blend2d_amd.set_render_schedule(blend2d_amd.SCHEDULE_AUTO)


def _frame(amd, scenes):
    keep = []
    images_all = []
    batch = amd.BLBatch() if len(scenes) > 1 else None
    for sc in scenes:
        images = []
        for fmt, px in sc.images:
            im = amd.BLImage(px.shape[1], px.shape[0], fmt)
            im.upload(px)
            images.append(im)
        images_all.append(images)
        img = amd.BLImage(sc.width, sc.height, sc.format)
        img.fill_raw(sc.clear_value)
        ctx = amd.BLContext(img)
        amd.replay_into(ctx, sc, images)
        if batch is not None:
            batch.add(ctx)
        keep.append((img, ctx))
    return batch, keep, images_all


def _replays(amd, scenes, n=4, schedule_switch=None):
    batch, keep, images = _frame(amd, scenes)
    target = batch if batch is not None else keep[0][1]
    try:
        target.flush()
        first = [img.download() for img, _ in keep]
        st0 = target.last_stats()
        graphs = []
        for k in range(n):
            if schedule_switch is not None and k == schedule_switch:
                amd.set_render_schedule(amd.SCHEDULE_QUEUE_HEURISTIC, 4)
            for (img, _), sc in zip(keep, scenes):
                img.fill_raw(sc.clear_value)
            target.replay()
            st = target.last_stats()
            graphs.append(st["graph"])
            assert st["pixels"] == st0["pixels"], k
            for (img, _), want in zip(keep, first):
                assert checkers.diff_info(img.download(), want)[0] == 0, k
        return first, graphs
    finally:
        amd.set_render_schedule(amd.SCHEDULE_AUTO)
        for img, ctx in keep:
            ctx.end()
        if batch is not None:
            batch.close()


def test_single_context_graph_replays(amd):
    sc = S.cubic_gradient_scene(11, 480, 270, 200)
    first, graphs = _replays(amd, [sc])
    assert graphs == [1, 1, 1, 1]
    want, _ = checkers.render_oracle(sc)
    assert checkers.diff_info(first[0], want)[0] == 0


def test_batched_graph_replays_with_a_schedule_change(amd):
    scenes = S.glyph_canvases(12, 24, 40) + S.cost_canvases(13, 3, 96, 300, 12)
    first, graphs = _replays(amd, scenes, n=5, schedule_switch=3)
    assert graphs == [1, 1, 1, 1, 1]
    for sc, out in zip(scenes, first):
        want, _ = checkers.render_oracle(sc)
        assert checkers.diff_info(out, want)[0] == 0


def test_graph_replay_of_patterns_and_boxes(amd):
    scenes = [S.pattern_shapes_scene(5, 300, 200, 30, pattern_size=48),
              S.srcover_layers_scene(6, 300, 200, 2)]
    first, graphs = _replays(amd, scenes, n=3)
    assert graphs == [1, 1, 1]
    for sc, out in zip(scenes, first):
        want, _ = checkers.render_oracle(sc)
        assert checkers.diff_info(out, want)[0] == 0


def test_a_new_upload_is_not_replayed_from_the_old_graph(amd):
    a = S.polygons_scene(21, 256, 256, 60)
    b = S.polygons_scene(22, 256, 256, 90)
    img = amd.BLImage(256, 256, a.format)
    ctx = amd.BLContext(img)
    try:
        for sc in (a, b):
            img.fill_raw(sc.clear_value)
            amd.replay_into(ctx, sc, [])
            ctx.flush()
            want = img.download()
            for k in range(3):
                img.fill_raw(sc.clear_value)
                ctx.replay()
                assert ctx.last_stats()["graph"] == 1
                assert checkers.diff_info(img.download(), want)[0] == 0
    finally:
        ctx.end()
