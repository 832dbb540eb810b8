"""Named seeded scenes whose reference renderings are pinned in
golden.json (written by make_golden.py from the reference library)."""
from blend2d_amd import scene as S


def cases():
    out = {}
    for seed in (1, 2, 3):
        out[f"polygons_512_s{seed}"] = lambda seed=seed: [S.polygons_scene(seed, 512, 512, 1000)]
    for seed in (1, 2):
        out[f"cubic_gradients_640x360_s{seed}"] = lambda seed=seed: [S.cubic_gradient_scene(seed, 640, 360, 300)]
    out["pattern_shapes_800x600_s1"] = lambda: [S.pattern_shapes_scene(
        1, 800, 600, 100, comp_ops=(S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY), pattern_size=128)]
    out["glyph_canvases_s1"] = lambda: S.glyph_canvases(1, 16, 60)
    for seed in (1, 2, 3, 4):
        for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
            out[f"mixed_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.mixed_scene(seed, 513, 513, 150, fmt)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        out[f"edge_cases_{fname}"] = lambda fmt=fmt: [S.edge_case_scene(fmt=fmt)]
        for seed in (1, 2, 3):
            out[f"frontend_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.frontend_scene(seed, fmt=fmt)]
        out[f"translated_patterns_{fname}_s1"] = lambda fmt=fmt: [S.translated_patterns_scene(
            1, 301, 211, 120, fmt, comp_ops=(S.COMP_OP_SRC_OVER, S.COMP_OP_SRC_COPY))]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2):
            out[f"comp_op_folds_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.comp_op_folds_scene(seed, fmt=fmt)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2):
            out[f"explicit_styles_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.explicit_styles_scene(seed, fmt=fmt)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_A8, "a8")):
        out[f"clip_robustness_{fname}"] = lambda fmt=fmt: S.clip_robustness_scenes(fmt=fmt)
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2):
            out[f"invalid_geometry_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.invalid_geometry_scene(seed, fmt=fmt)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2, 3):
            out[f"blits_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.blit_scene(seed, fmt=fmt, bad_areas=seed == 3)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2, 3):
            out[f"masks_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.mask_scene(seed, fmt=fmt, unaligned=seed == 3)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_XRGB32, "xrgb32"), (S.FMT_A8, "a8")):
        for seed in (1, 2, 3):
            out[f"dither_{fname}_s{seed}"] = lambda seed=seed, fmt=fmt: [S.dither_scene(seed, fmt=fmt, bad_hint=seed == 3)]
    for fmt, fname in ((S.FMT_PRGB32, "prgb32"), (S.FMT_A8, "a8")):
        out[f"blit_clip_robustness_{fname}"] = lambda fmt=fmt: S.blit_clip_robustness_scenes(fmt=fmt)
    return out


def digest(pixels):
    """sha256 of the packed canvas and crc32 of each 64-row band."""
    import hashlib
    import zlib
    raw = pixels.tobytes()
    row = len(raw) // pixels.shape[0]
    bands = [zlib.crc32(raw[y * row:(y + 64) * row]) for y in range(0, pixels.shape[0], 64)]
    return {"sha256": hashlib.sha256(raw).hexdigest(), "band_crc32": bands}
