"""GPU parity: the MI355X backend against the CPU oracle (and the reference
library when it is built) on seeded scenes, through the C ABI.

Tolerance: 0 everywhere.  Coverage, solid fills, linear gradients, nearest
and bilinear patterns (affine and translated) and every integer comp op are
integer arithmetic; radial and conic gradients run the oracle's float
operation order (unfused, correctly rounded division and square root), so
they are bit-exact against the oracle too.  (The +-1 LSB that the north star
allows radial / conic and bilinear output applies against the reference's
JIT pipeline, not against this oracle.)"""
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
    _cmp(amd, S.cubic_gradient_scene(seed, 640, 360, 200))


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


@pytest.mark.parametrize("seed", [5, 6])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_mixed_all_comp_ops(amd, seed, fmt):
    # PLUS / MULTIPLY / SCREEN have no runnable reference (JIT only): the
    # oracle's restatement is the checker here
    ops = (S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY, S.COMP_OP_PLUS, S.COMP_OP_MULTIPLY, S.COMP_OP_SCREEN)
    _cmp(amd, S.mixed_scene(seed, 389, 301, 160, fmt, comp_ops=ops))


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_box_only_bands(amd, seed, fmt):
    # every band holds box fills only: the tile-group compositor path, with
    # canvas widths that are not multiples of a tile group
    ops = (S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY, S.COMP_OP_PLUS, S.COMP_OP_MULTIPLY, S.COMP_OP_SCREEN)
    if fmt != S.FMT_PRGB32:
        ops = ops[:3]
    _cmp(amd, S.box_scene(seed, 333 + 257 * seed, 77, 80, fmt, comp_ops=ops))


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_translated_patterns(amd, seed, fmt):
    # pure-translation pattern transforms: aligned and fractional (FxFy)
    # fetchers, every extend mode, tiny and 1-pixel images, bit-exact
    ops = (S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY, S.COMP_OP_PLUS, S.COMP_OP_MULTIPLY)
    if fmt != S.FMT_PRGB32:
        ops = ops[:3]
    _cmp(amd, S.translated_patterns_scene(seed, 301 + 64 * seed, 211, 120, fmt, comp_ops=ops))


@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_A8])
def test_cancelling_pattern_transforms(amd, fmt):
    # user and pattern transforms that cancel to a translation
    _cmp(amd, S.cancelling_pattern_transforms_scene(93, 200, 150, 40, fmt))


def test_box_and_analytic_canvases_batched(amd):
    scenes = [S.box_scene(3, 255, 33, 30), S.polygons_scene(9, 70, 40, 50), S.box_scene(4, 1, 1, 5, S.FMT_A8),
              S.box_scene(5, 64, 16, 20, S.FMT_XRGB32)]
    outs, errs = amd.render_scenes_batched(scenes)
    for sc, out, err in zip(scenes, outs, errs):
        want, oerr = checkers.render_oracle(sc)
        assert err == oerr
        assert checkers.diff_info(out, want)[0] == 0


def test_pareto_mix_canvases_batched(amd):
    # the many-canvas mixed workload at test size: Pareto-distributed
    # segment counts of lines / quads / cubics, every fetcher and four comp
    # ops, both fill rules, one batched frame
    scenes = S.cost_canvases(7, 6, 96, 400, 16, 3000)
    outs, errs = amd.render_scenes_batched(scenes)
    for sc, out, err in zip(scenes, outs, errs):
        want, oerr = checkers.render_oracle(sc)
        assert err == oerr
        assert checkers.diff_info(out, want)[0] == 0


def test_edge_cases_exact(amd):
    for fmt in (S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8):
        _cmp(amd, S.edge_case_scene(fmt=fmt))


def test_replay_reproduces_the_frame(amd):
    # a replay re-executes the resident frame: on a re-cleared target it
    # must produce the flushed pixels again, and the same statistics
    sc = S.mixed_scene(8, 300, 200, 80, S.FMT_PRGB32)
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
    first = img.download()
    st1 = ctx.last_stats()
    img.fill_raw(sc.clear_value)
    ctx.replay()
    second = img.download()
    st2 = ctx.last_stats()
    ctx.end()
    assert (first == second).all()
    for k in ("commands", "segments", "binned", "pixels"):
        assert st1[k] == st2[k], k
    assert st1["pixels"] > 0
    want, _ = checkers.render_oracle(sc)
    assert checkers.diff_info(first, want)[0] == 0


def _star(n, cx, cy, r0, r1, phase=0.0):
    a = np.arange(n) * (2 * np.pi / n) + phase
    r = np.where(np.arange(n) % 2 == 0, r0, r1)
    return np.stack([cx + np.cos(a) * r, cy + np.sin(a) * r], 1)


def scan_boundary_scenes():
    """Frames whose scanned arrays have exactly 1024 * k elements (the scan
    block): (band, command) matrices of 1024 and 2048 cells and source
    segment counts of 1024, 2048 and 4096 -- the sizes at which the grand
    total of a scan is written by the last thread of the last block."""
    out = []
    rng = np.random.default_rng(77)
    for h, n in ((128, 128), (256, 128)):              # 8 or 16 bands x 128 commands
        sc = S.Scene(64, h, S.FMT_PRGB32, 0xFF000000)
        for _ in range(n):
            sc.set_fill_style_rgba32(int(rng.integers(0x40000000, 1 << 32)))
            sc.fill_polygon(rng.random((5, 2)) * np.array([64, h]))
        out.append(sc)
    for n in (1024, 2048, 4096):                         # one polygon: n line sources
        sc = S.Scene(300, 200, S.FMT_PRGB32, 0xFF102030)
        sc.set_fill_style_rgba32(0xC08040FF)
        sc.fill_polygon(_star(n, 150, 100, 90, 40))
        out.append(sc)
    return out


def test_scan_block_boundaries(amd):
    for sc in scan_boundary_scenes():
        _cmp(amd, sc)


def _box_scene(seed, w, h, fmt, count, fractional=False):
    """Aligned (or, with `fractional`, unaligned) solid boxes with SRC_OVER /
    SRC_COPY, random colours and global alphas, many overlapping one band."""
    rng = np.random.default_rng(seed)
    sc = S.Scene(w, h, fmt, 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != S.FMT_A8 else int(rng.integers(0, 256)))
    for _ in range(count):
        r = rng.random()
        if r < 0.1:
            sc.set_comp_op(int(rng.choice([S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY])))
        elif r < 0.2:
            sc.set_global_alpha(float(rng.choice([1.0, rng.uniform(0.05, 1.0)])))
        x, y = int(rng.integers(-20, w)), int(rng.integers(-20, h))
        bw, bh = int(rng.integers(1, w // 2 + 2)), int(rng.integers(1, h // 2 + 2))
        col = int(rng.integers(0, 1 << 32))
        if fractional:
            sc.fill_rect_d_rgba32(x + float(rng.random()), y + float(rng.random()), bw + float(rng.random()), bh, col)
        else:
            sc.fill_rect_i_rgba32(x, y, bw, bh, col)
    return sc


@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32])
@pytest.mark.parametrize("seed,w,h,count", [(1, 333, 70, 40), (2, 1000, 48, 300), (3, 64, 16, 5), (4, 257, 33, 600)])
def test_aligned_solid_boxes_take_the_fast_kernel(amd, fmt, seed, w, h, count):
    """Bands of aligned solid SRC_OVER / SRC_COPY boxes only are rendered by
    k_render_box_fast (more than 256 commands in a band take several record
    chunks; ragged widths and heights leave partial groups)."""
    sc = _box_scene(seed, w, h, fmt, count)
    got, err, st = amd.render_scene(sc, return_context=True)
    assert st["render_kernels"] & 8
    want, oerr = checkers.render_oracle(sc)
    assert err == oerr
    mx, mean, n = checkers.diff_info(got, want)
    assert mx == 0, f"max diff {mx}, {n} channels differ"


@pytest.mark.parametrize("fmt,fractional", [(S.FMT_A8, False), (S.FMT_PRGB32, True)])
def test_boxes_outside_the_fast_kernel_still_match(amd, fmt, fractional):
    """A8 canvases and unaligned boxes stay on the general box kernels."""
    sc = _box_scene(5, 300, 64, fmt, 200, fractional)
    got, err, st = amd.render_scene(sc, return_context=True)
    want, oerr = checkers.render_oracle(sc)
    assert err == oerr
    mx, mean, n = checkers.diff_info(got, want)
    assert mx == 0, f"max diff {mx}, {n} channels differ"
