"""Replays a recorded Scene (blend2d_amd.scene) through the MI355X backend."""
from __future__ import annotations

import struct

import numpy as np

from . import scene as S


def _images(ctx_cls, sc):
    from . import BLImage
    out = []
    for fmt, px in sc.images:
        h, w = px.shape
        img = BLImage(w, h, fmt)
        img.upload(px)
        out.append(img)
    return out


# Scene ops that are rendering calls (fill_* / clear_*): the context counts
# them, and an error found on the device names its call by that count.
RENDER_OPS = {S.OP_FILL_PATH, S.OP_FILL_POLYGON, S.OP_FILL_RECT_D, S.OP_FILL_RECT_I, S.OP_FILL_ALL,
              S.OP_CLEAR_ALL, S.OP_CLEAR_RECT_I, S.OP_CLEAR_RECT_D, S.OP_FILL_GEOMETRY, S.OP_FILL_ALL_RGBA32,
              S.OP_FILL_RECT_I_RGBA32, S.OP_FILL_RECT_D_RGBA32, S.OP_FILL_PATH_EX, S.OP_FILL_EXPLICIT,
              S.OP_BLIT_IMAGE, S.OP_FILL_MASK}


def _mask(pl, at=0):
    """Decodes a SceneMask: (integer, image index, dst, area)."""
    kind, index, has_area, _, ax, ay, aw, ah, dx, dy = struct.unpack_from("<4I4i2d", pl, at)
    return kind == S.MASK_I, index, (dx, dy), ((ax, ay, aw, ah) if has_area else None)


def _explicit(pl, images):
    """Decodes a FILL_EXPLICIT payload: (target, style object, geometry)."""
    from . import BLGradient, BLPattern, BLRgba64
    target, kind, gtype, count, rgba = struct.unpack_from("<IIIIQ", pl)
    at = 24
    if kind == S.STYLE_RGBA32:
        style = int(rgba) & 0xFFFFFFFF
    elif kind == S.STYLE_RGBA64:
        style = BLRgba64(rgba)
    elif kind == S.STYLE_GRADIENT:
        h = struct.unpack_from("<IIII6d6d", pl, at)
        n = h[2]
        stops = [struct.unpack_from("<dI", pl, at + 112 + 16 * k) for k in range(n)]
        style = BLGradient(h[0], list(h[4:10]), h[1], stops, list(h[10:16]) if h[3] else None)
        at += 112 + 16 * n
    else:
        h = struct.unpack_from("<IIII6d", pl, at)
        style = BLPattern(images[h[0]], h[1], list(h[4:10]) if h[2] else None)
        at += 64
    if target == S.TARGET_RECT_I:
        geom = [int(v) for v in struct.unpack_from("<4i", pl, at)]
    elif target == S.TARGET_RECT_D:
        geom = [float(v) for v in struct.unpack_from("<4d", pl, at)]
    elif target == S.TARGET_PATH:
        cmds = np.frombuffer(pl, dtype=np.uint8, count=count, offset=at)
        vtx = np.frombuffer(pl, dtype=np.float64, count=2 * count, offset=at + ((count + 7) & ~7)).reshape(-1, 2)
        geom = (cmds, vtx)
    elif target == S.TARGET_MASK:
        geom = _mask(pl, at)
    elif target == S.TARGET_GEOMETRY:
        if gtype in S.GEOMETRY_ARRAY:
            fmt = S.GEOMETRY_ARRAY[gtype]
            size = struct.calcsize(fmt)
            geom = (gtype, [struct.unpack_from(fmt, pl, at + k * size) for k in range(count)])
        else:
            geom = (gtype, struct.unpack_from(S.GEOMETRY_STRUCT[gtype], pl, at))
    else:
        geom = None
    return target, style, geom


def replay_into(ctx, sc: "S.Scene", images) -> list:
    """Issues the scene's calls on `ctx`; returns the BLResult errors seen,
    each as (rendering-call index, code) (calls keep going after an error, as
    the reference driver does).  State calls that fail get the index of the
    next rendering call."""
    from . import BLError, BLGradient, BLPattern, BLRgba, BLRgba64
    errors = []
    n_render = 0

    def call(fn, *args):
        try:
            fn(*args)
        except BLError as e:
            errors.append((n_render, e.code))

    for op, pl in sc.cmds:
        if op in RENDER_OPS:
            n_render += 1
        if op == S.OP_SET_COMP_OP:
            call(ctx.set_comp_op, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_SET_FILL_RULE:
            call(ctx.set_fill_rule, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_SET_TRANSFORM:
            call(ctx.set_transform, struct.unpack_from("<6d", pl))
        elif op == S.OP_SET_GLOBAL_ALPHA:
            call(ctx.set_global_alpha, struct.unpack_from("<d", pl)[0])
        elif op == S.OP_SET_PATTERN_QUALITY:
            call(ctx.set_pattern_quality, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_SET_GRADIENT_QUALITY:
            call(ctx.set_gradient_quality, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_SET_STYLE_SOLID:
            call(ctx.set_fill_style, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_SET_STYLE_GRADIENT:
            h = struct.unpack_from("<IIII6d6d", pl)
            gtype, extend, n, has_tr = h[:4]
            values, tr = list(h[4:10]), list(h[10:16])
            stops = [struct.unpack_from("<dI", pl, 112 + 16 * k) for k in range(n)]
            call(ctx.set_fill_style, BLGradient(gtype, values, extend, stops, tr if has_tr else None))
        elif op == S.OP_SET_STYLE_PATTERN:
            h = struct.unpack_from("<IIII6d", pl)
            idx, extend, has_tr = h[0], h[1], h[2]
            call(ctx.set_fill_style, BLPattern(images[idx], extend, list(h[4:10]) if has_tr else None))
        elif op == S.OP_FILL_PATH:
            n = struct.unpack_from("<I", pl)[0]
            cmds = np.frombuffer(pl, dtype=np.uint8, count=n, offset=8)
            vtx = np.frombuffer(pl, dtype=np.float64, count=2 * n, offset=8 + ((n + 7) & ~7)).reshape(-1, 2)
            call(ctx.fill_path, (cmds, vtx))
        elif op == S.OP_FILL_POLYGON:
            n = struct.unpack_from("<I", pl)[0]
            call(ctx.fill_polygon, np.frombuffer(pl, dtype=np.float64, count=2 * n, offset=8).reshape(-1, 2))
        elif op == S.OP_FILL_RECT_D:
            call(ctx.fill_rect, *[float(v) for v in struct.unpack_from("<4d", pl)])
        elif op == S.OP_FILL_RECT_I:
            call(ctx.fill_rect, *[int(v) for v in struct.unpack_from("<4i", pl)])
        elif op == S.OP_FILL_ALL:
            call(ctx.fill_all)
        elif op == S.OP_SET_FILL_ALPHA:
            call(ctx.set_fill_alpha, struct.unpack_from("<d", pl)[0])
        elif op == S.OP_SET_STYLE_RGBA64:
            call(ctx.set_fill_style, BLRgba64(struct.unpack_from("<Q", pl)[0]))
        elif op == S.OP_SET_STYLE_RGBA:
            call(ctx.set_fill_style, BLRgba(*struct.unpack_from("<4f", pl)))
        elif op == S.OP_CLEAR_ALL:
            call(ctx.clear_all)
        elif op == S.OP_CLEAR_RECT_I:
            call(ctx.clear_rect, *[int(v) for v in struct.unpack_from("<4i", pl)])
        elif op == S.OP_CLEAR_RECT_D:
            call(ctx.clear_rect, *[float(v) for v in struct.unpack_from("<4d", pl)])
        elif op == S.OP_FILL_ALL_RGBA32:
            call(ctx.fill_all, struct.unpack_from("<I", pl)[0])
        elif op == S.OP_FILL_RECT_I_RGBA32:
            v = struct.unpack_from("<4iI", pl)
            call(ctx.fill_rect, *[int(x) for x in v[:4]], v[4])
        elif op == S.OP_FILL_RECT_D_RGBA32:
            v = struct.unpack_from("<4dI", pl)
            call(ctx.fill_rect, *[float(x) for x in v[:4]], v[4])
        elif op == S.OP_FILL_GEOMETRY:
            gtype, has_col, col, count = struct.unpack_from("<4I", pl)
            if gtype in S.GEOMETRY_ARRAY:
                fmt = S.GEOMETRY_ARRAY[gtype]
                size = struct.calcsize(fmt)
                data = [struct.unpack_from(fmt, pl, 16 + k * size) for k in range(count)]
            elif gtype in S.GEOMETRY_STRUCT:
                data = struct.unpack_from(S.GEOMETRY_STRUCT[gtype], pl, 16)
            else:
                data = ()
            call(ctx.fill_geometry, gtype, data, col if has_col else None)
        elif op == S.OP_FILL_EXPLICIT:
            target, style, geom = _explicit(pl, images)
            if target == S.TARGET_ALL:
                call(lambda: ctx.fill_all(style=style))
            elif target in (S.TARGET_RECT_I, S.TARGET_RECT_D):
                call(lambda: ctx.fill_rect(*geom, style=style))
            elif target == S.TARGET_PATH:
                call(lambda: ctx.fill_path(geom, style=style))
            elif target == S.TARGET_MASK:
                integer, index, dst, area = geom
                call(lambda: ctx.fill_mask(dst, images[index], area, integer, style=style))
            else:
                call(lambda: ctx.fill_geometry(geom[0], geom[1], style=style))
        elif op == S.OP_BLIT_IMAGE:
            kind, index, has_area, _, ax, ay, aw, ah, dx, dy, dw, dh = struct.unpack_from("<4I4i4d", pl)
            area = (ax, ay, aw, ah) if has_area else None
            dst = (dx, dy) if kind in (S.BLIT_I, S.BLIT_D) else (dx, dy, dw, dh)
            integer = kind in (S.BLIT_I, S.BLIT_SCALED_I)
            call(ctx.blit_image, dst, images[index], area, integer)
        elif op == S.OP_FILL_MASK:
            integer, index, dst, area = _mask(pl)
            call(ctx.fill_mask, dst, images[index], area, integer)
        elif op == S.OP_FILL_PATH_EX:
            n, flags, col, _, ox, oy = struct.unpack_from("<4I2d", pl)
            cmds = np.frombuffer(pl, dtype=np.uint8, count=n, offset=32)
            vtx = np.frombuffer(pl, dtype=np.float64, count=2 * n, offset=32 + ((n + 7) & ~7)).reshape(-1, 2)
            call(ctx.fill_path, (cmds, vtx), (ox, oy) if flags & 1 else None, col if flags & 2 else None)
    return errors


def first_error(errors, ctx) -> int:
    """The first error in call order: those the calls returned (index =
    rendering calls issued so far, the failing call included) merged with
    the one the device found during the flush (0-based call index)."""
    found = list(errors)
    d = ctx.deferred_error()
    if d is not None:
        found.append((d[0] + 1, d[1]))
    return min(found)[1] if found else 0


def render_scene(sc: "S.Scene", return_context: bool = False):
    """Renders one scene on the GPU; returns (pixels, first_error[, ctx])."""
    from . import BLContext, BLImage
    images = _images(BLContext, sc)
    img = BLImage(sc.width, sc.height, sc.format)
    img.fill_raw(sc.clear_value)
    ctx = BLContext(img)
    errs = replay_into(ctx, sc, images)
    ctx.flush()
    stats = ctx.last_stats()
    err = first_error(errs, ctx)
    ctx.end()
    out = img.download()
    if return_context:
        return out, err, stats
    return out, err


def render_scenes_batched(scenes):
    """Renders many scenes (canvases) with one batched flush."""
    from . import BLBatch, BLContext, BLImage
    batch = BLBatch()
    keep = []
    errors = []
    for sc in scenes:
        images = _images(BLContext, sc)
        img = BLImage(sc.width, sc.height, sc.format)
        img.fill_raw(sc.clear_value)
        ctx = BLContext(img)
        errs = replay_into(ctx, sc, images)
        errors.append(errs)
        batch.add(ctx)
        keep.append((img, ctx, images))
    batch.flush()
    errors = [first_error(errs, ctx) for errs, (_, ctx, _) in zip(errors, keep)]
    outs = [img.download() for img, _, _ in keep]
    for img, ctx, _ in keep:
        ctx.end()
    return outs, errors
