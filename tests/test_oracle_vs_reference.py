"""The CPU oracle against the reference library itself (compiled from the
reference sources into oracle/_ref by oracle/ref.mk) on seeds beyond the
golden fixtures.  Skipped where that build is absent."""
import pytest

import checkers
from blend2d_amd import scene as S

pytestmark = [pytest.mark.ref,
              pytest.mark.skipif(not checkers.have_ref(), reason="reference library not built (oracle/_ref)")]


def _same(sc, threads=0):
    want, rerr, _ = checkers.render_reference(sc, threads=threads)
    got, oerr = checkers.render_oracle(sc)
    assert oerr == rerr
    mx, mean, n = checkers.diff_info(got, want)
    assert mx == 0, f"max {mx} mean {mean} channels {n}"


@pytest.mark.parametrize("seed", [11, 12, 13])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_mixed(seed, fmt):
    _same(S.mixed_scene(seed, 301, 257, 200, fmt))


@pytest.mark.parametrize("seed", [21, 22])
def test_cubic_gradients(seed):
    _same(S.cubic_gradient_scene(seed, 800, 450, 400))


def test_patterns_src_over_copy():
    _same(S.pattern_shapes_scene(31, 640, 480, 60, comp_ops=(S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY), pattern_size=100))


def test_wide_canvas_band_height():
    # 4096 px wide: the reference band height drops to 16 rows, which moves
    # the conic fetcher's row origin
    _same(S.mixed_scene(41, 4100, 96, 80, S.FMT_PRGB32))


def test_worker_threads_do_not_change_pixels():
    # reference with worker threads (its own mt test: tolerance 0) vs oracle
    _same(S.mixed_scene(51, 513, 513, 150, S.FMT_PRGB32), threads=4)


@pytest.mark.parametrize("seed", [61, 62, 63, 64])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_frontend_calls(seed, fmt):
    # fill_geometry, explicit colours, clears, fill alpha, 64-bit / float
    # colours, origins, non-invertible transforms
    _same(S.frontend_scene(seed, 300, 220, fmt, 150))


@pytest.mark.parametrize("seed", [71, 72])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_box_fills(seed, fmt):
    # rects, fill_all and clears only, under translate / scale transforms
    _same(S.box_scene(seed, 333, 77, 80, fmt))


@pytest.mark.parametrize("seed", [81, 82, 83])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_translated_patterns(seed, fmt):
    # pure-translation pattern transforms: aligned and fractional fetchers
    # (the reference's portable pipeline implements SRC_OVER and SRC_COPY only)
    _same(S.translated_patterns_scene(seed, 301, 211, 120, fmt, comp_ops=(S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY)))


@pytest.mark.parametrize("seed", [91, 92])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_A8])
def test_cancelling_pattern_transforms(seed, fmt):
    # user and pattern transforms that cancel: the near-identity test
    _same(S.cancelling_pattern_transforms_scene(seed, 200, 150, 40, fmt))


@pytest.mark.parametrize("seed", [81, 82, 83, 84])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_blits(seed, fmt):
    # blit_image / blit_scaled_image in all four forms, areas, transforms
    _same(S.blit_scene(seed, 257, 181, fmt, 140, bad_areas=seed == 84))


@pytest.mark.parametrize("seed", [91, 92, 93])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_mask_fills(seed, fmt):
    # fill_mask in all forms; seed 93 adds the placements the reference
    # reports as NOT_IMPLEMENTED and invalid areas
    _same(S.mask_scene(seed, 241, 173, fmt, 120, unaligned=seed == 93))


@pytest.mark.parametrize("seed", [101, 102, 103])
@pytest.mark.parametrize("fmt", [S.FMT_PRGB32, S.FMT_XRGB32, S.FMT_A8])
def test_dithered_gradients(seed, fmt):
    # the gradient quality hint: nearest, smooth and dither, styles and
    # explicit styles, every gradient type and extend mode
    _same(S.dither_scene(seed, 223, 167, fmt, 120, bad_hint=seed == 103))
