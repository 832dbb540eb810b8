"""Recorded BLContext command streams ("scenes") and the synthetic workloads.

A scene is a canvas description plus an ordered list of BLContext calls
(set_comp_op, set_fill_style, fill_path, ...) serialized to a little-endian
byte string.  The same bytes replay through the MI355X backend
(`blend2d_amd.replay`), and the test suite replays them through the CPU
oracle and the reference library for parity checks.

The generators reproduce the benchmark workloads: seeded random polygons,
cubic-Bezier paths with gradients, rotated rectangles/ellipses with patterns
and glyph-like quadratic contours.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = 0x43533242
VERSION = 1

FMT_PRGB32, FMT_XRGB32, FMT_A8 = 1, 2, 3

OP_SET_COMP_OP = 1
OP_SET_FILL_RULE = 2
OP_SET_TRANSFORM = 3
OP_SET_GLOBAL_ALPHA = 4
OP_SET_STYLE_SOLID = 5
OP_SET_STYLE_GRADIENT = 6
OP_SET_STYLE_PATTERN = 7
OP_SET_PATTERN_QUALITY = 8
OP_FILL_PATH = 9
OP_FILL_POLYGON = 10
OP_FILL_RECT_D = 11
OP_FILL_RECT_I = 12
OP_FILL_ALL = 13
OP_SET_FILL_ALPHA = 14
OP_SET_STYLE_RGBA64 = 15
OP_SET_STYLE_RGBA = 16
OP_CLEAR_ALL = 17
OP_CLEAR_RECT_I = 18
OP_CLEAR_RECT_D = 19
OP_FILL_GEOMETRY = 20
OP_FILL_ALL_RGBA32 = 21
OP_FILL_RECT_I_RGBA32 = 22
OP_FILL_RECT_D_RGBA32 = 23
OP_FILL_PATH_EX = 24
OP_FILL_EXPLICIT = 25
OP_BLIT_IMAGE = 26
BLIT_I, BLIT_D, BLIT_SCALED_I, BLIT_SCALED_D = 0, 1, 2, 3
OP_FILL_MASK = 27
MASK_I, MASK_D = 0, 1
OP_SET_GRADIENT_QUALITY = 28
GRADIENT_QUALITY_NEAREST, GRADIENT_QUALITY_SMOOTH, GRADIENT_QUALITY_DITHER = 0, 1, 2

# FILL_EXPLICIT targets and style kinds
TARGET_ALL, TARGET_RECT_I, TARGET_RECT_D, TARGET_PATH, TARGET_GEOMETRY, TARGET_MASK = range(6)
STYLE_RGBA32, STYLE_RGBA64, STYLE_GRADIENT, STYLE_PATTERN = range(4)

# BLGeometryType values (core/geometry.h).
(GEOMETRY_NONE, GEOMETRY_BOXI, GEOMETRY_BOXD, GEOMETRY_RECTI, GEOMETRY_RECTD, GEOMETRY_CIRCLE, GEOMETRY_ELLIPSE,
 GEOMETRY_ROUND_RECT, GEOMETRY_ARC, GEOMETRY_CHORD, GEOMETRY_PIE, GEOMETRY_LINE, GEOMETRY_TRIANGLE,
 GEOMETRY_POLYLINEI, GEOMETRY_POLYLINED, GEOMETRY_POLYGONI, GEOMETRY_POLYGOND, GEOMETRY_ARRAY_VIEW_BOXI,
 GEOMETRY_ARRAY_VIEW_BOXD, GEOMETRY_ARRAY_VIEW_RECTI, GEOMETRY_ARRAY_VIEW_RECTD) = range(21)

# struct layout of each fixed-size geometry: (struct format, number of fields)
GEOMETRY_STRUCT = {
    GEOMETRY_BOXI: "<4i", GEOMETRY_BOXD: "<4d", GEOMETRY_RECTI: "<4i", GEOMETRY_RECTD: "<4d",
    GEOMETRY_CIRCLE: "<3d", GEOMETRY_ELLIPSE: "<4d", GEOMETRY_ROUND_RECT: "<6d", GEOMETRY_ARC: "<6d",
    GEOMETRY_CHORD: "<6d", GEOMETRY_PIE: "<6d", GEOMETRY_LINE: "<4d", GEOMETRY_TRIANGLE: "<6d",
}
# element layout of array geometries
GEOMETRY_ARRAY = {
    GEOMETRY_POLYLINEI: "<2i", GEOMETRY_POLYLINED: "<2d", GEOMETRY_POLYGONI: "<2i", GEOMETRY_POLYGOND: "<2d",
    GEOMETRY_ARRAY_VIEW_BOXI: "<4i", GEOMETRY_ARRAY_VIEW_BOXD: "<4d",
    GEOMETRY_ARRAY_VIEW_RECTI: "<4i", GEOMETRY_ARRAY_VIEW_RECTD: "<4d",
}

# BLCompOp values (core/context.h).
COMP_OP_SRC_OVER, COMP_OP_SRC_COPY = 0, 1
COMP_OP_PLUS, COMP_OP_MULTIPLY, COMP_OP_SCREEN = 12, 15, 16

# BLPathCmd values (core/path.h).
CMD_MOVE, CMD_ON, CMD_QUAD, CMD_CUBIC, CMD_CLOSE = 0, 1, 2, 4, 5

GRADIENT_LINEAR, GRADIENT_RADIAL, GRADIENT_CONIC = 0, 1, 2
EXTEND_PAD, EXTEND_REPEAT, EXTEND_REFLECT = 0, 1, 2


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * ((-len(b)) % 8)


@dataclass
class Scene:
    """An in-memory BLContext command stream."""

    width: int
    height: int
    format: int = FMT_PRGB32
    clear_value: int = 0
    images: list = field(default_factory=list)
    cmds: list = field(default_factory=list)
    fill_count: int = 0

    # -- images ----------------------------------------------------------
    def add_image(self, pixels: np.ndarray, fmt: int = FMT_PRGB32) -> int:
        pixels = np.ascontiguousarray(pixels)
        if fmt == FMT_A8:
            assert pixels.dtype == np.uint8 and pixels.ndim == 2
        else:
            assert pixels.dtype == np.uint32 and pixels.ndim == 2
        self.images.append((fmt, pixels))
        return len(self.images) - 1

    # -- state -----------------------------------------------------------
    def _cmd(self, op: int, payload: bytes = b"") -> None:
        self.cmds.append((op, _pad8(payload)))

    def set_comp_op(self, op: int) -> None:
        self._cmd(OP_SET_COMP_OP, struct.pack("<II", op, 0))

    def set_fill_rule(self, rule: int) -> None:
        self._cmd(OP_SET_FILL_RULE, struct.pack("<II", rule, 0))

    def set_transform(self, m) -> None:
        self._cmd(OP_SET_TRANSFORM, struct.pack("<6d", *m))

    def set_global_alpha(self, a: float) -> None:
        self._cmd(OP_SET_GLOBAL_ALPHA, struct.pack("<d", a))

    def set_gradient_quality(self, q: int) -> None:
        self._cmd(OP_SET_GRADIENT_QUALITY, struct.pack("<II", q, 0))

    def set_pattern_quality(self, q: int) -> None:
        self._cmd(OP_SET_PATTERN_QUALITY, struct.pack("<II", q, 0))

    def set_fill_style_rgba32(self, rgba32: int) -> None:
        self._cmd(OP_SET_STYLE_SOLID, struct.pack("<II", rgba32 & 0xFFFFFFFF, 0))

    def set_fill_style_gradient(self, gtype: int, values, stops, extend: int = EXTEND_PAD,
                                transform=None) -> None:
        vals = list(values) + [0.0] * (6 - len(values))
        tr = list(transform) if transform is not None else [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
        payload = struct.pack("<IIII6d6d", gtype, extend, len(stops), 1 if transform is not None else 0, *vals, *tr)
        for off, rgba in stops:
            payload += struct.pack("<dII", off, rgba & 0xFFFFFFFF, 0)
        self._cmd(OP_SET_STYLE_GRADIENT, payload)

    def set_fill_style_pattern(self, image_index: int, extend: int = EXTEND_REPEAT, transform=None) -> None:
        tr = list(transform) if transform is not None else [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
        self._cmd(OP_SET_STYLE_PATTERN, struct.pack("<IIII6d", image_index, extend,
                                                     1 if transform is not None else 0, 0, *tr))

    # -- fills -----------------------------------------------------------
    def fill_path(self, cmds, vtx) -> None:
        cmds = np.asarray(cmds, dtype=np.uint8)
        vtx = np.asarray(vtx, dtype=np.float64).reshape(-1, 2)
        assert len(cmds) == len(vtx)
        self._cmd(OP_FILL_PATH, struct.pack("<II", len(cmds), 0) + _pad8(cmds.tobytes()) + vtx.tobytes())
        self.fill_count += 1

    def fill_polygon(self, vtx) -> None:
        vtx = np.asarray(vtx, dtype=np.float64).reshape(-1, 2)
        self._cmd(OP_FILL_POLYGON, struct.pack("<II", len(vtx), 0) + vtx.tobytes())
        self.fill_count += 1

    def fill_rect_d(self, x, y, w, h) -> None:
        self._cmd(OP_FILL_RECT_D, struct.pack("<4d", x, y, w, h))
        self.fill_count += 1

    def fill_rect_i(self, x, y, w, h) -> None:
        self._cmd(OP_FILL_RECT_I, struct.pack("<4i", x, y, w, h))
        self.fill_count += 1

    def fill_all(self) -> None:
        self._cmd(OP_FILL_ALL)
        self.fill_count += 1

    def blit_image(self, kind: int, image_index: int, dst, area=None) -> None:
        """blit_image_i / _d (dst: x, y) and blit_scaled_image_i / _d (dst: x, y,
        w, h) of image `image_index`, `area` (x, y, w, h) or the whole image
        (scene_format.h: SceneBlit)."""
        integer = kind in (BLIT_I, BLIT_SCALED_I)
        d = [float(int(v)) if integer else float(v) for v in dst] + [0.0] * (4 - len(dst))
        a = list(area) if area is not None else [0, 0, 0, 0]
        self._cmd(OP_BLIT_IMAGE, struct.pack("<4I4i4d", kind, image_index, 0 if area is None else 1, 0, *a, *d))
        self.fill_count += 1

    @staticmethod
    def _mask_payload(kind: int, image_index: int, dst, area) -> bytes:
        d = [float(int(v)) if kind == MASK_I else float(v) for v in dst]
        a = list(area) if area is not None else [0, 0, 0, 0]
        return struct.pack("<4I4i2d", kind, image_index, 0 if area is None else 1, 0, *a, *d)

    def fill_mask(self, kind: int, image_index: int, dst, area=None) -> None:
        """fill_mask_i / _d (kind MASK_I / MASK_D) with the fill style: image
        `image_index` (its `area` x, y, w, h, or all of it) as the coverage,
        placed at dst = (x, y) (scene_format.h: SceneMask)."""
        self._cmd(OP_FILL_MASK, self._mask_payload(kind, image_index, dst, area))
        self.fill_count += 1

    def set_fill_alpha(self, a: float) -> None:
        self._cmd(OP_SET_FILL_ALPHA, struct.pack("<d", a))

    def set_fill_style_rgba64(self, rgba64: int) -> None:
        self._cmd(OP_SET_STYLE_RGBA64, struct.pack("<Q", rgba64 & 0xFFFFFFFFFFFFFFFF))

    def set_fill_style_rgba(self, r: float, g: float, b: float, a: float) -> None:
        self._cmd(OP_SET_STYLE_RGBA, struct.pack("<4f", r, g, b, a))

    def clear_all(self) -> None:
        self._cmd(OP_CLEAR_ALL)
        self.fill_count += 1

    def clear_rect_i(self, x, y, w, h) -> None:
        self._cmd(OP_CLEAR_RECT_I, struct.pack("<4i", x, y, w, h))
        self.fill_count += 1

    def clear_rect_d(self, x, y, w, h) -> None:
        self._cmd(OP_CLEAR_RECT_D, struct.pack("<4d", x, y, w, h))
        self.fill_count += 1

    def fill_all_rgba32(self, rgba32: int) -> None:
        self._cmd(OP_FILL_ALL_RGBA32, struct.pack("<II", rgba32 & 0xFFFFFFFF, 0))
        self.fill_count += 1

    def fill_rect_i_rgba32(self, x, y, w, h, rgba32: int) -> None:
        self._cmd(OP_FILL_RECT_I_RGBA32, struct.pack("<4iII", x, y, w, h, rgba32 & 0xFFFFFFFF, 0))
        self.fill_count += 1

    def fill_rect_d_rgba32(self, x, y, w, h, rgba32: int) -> None:
        self._cmd(OP_FILL_RECT_D_RGBA32, struct.pack("<4dII", x, y, w, h, rgba32 & 0xFFFFFFFF, 0))
        self.fill_count += 1

    def fill_geometry(self, gtype: int, data, rgba32: int | None = None) -> None:
        """`data`: the geometry's fields (fixed-size types) or a sequence of
        elements, each a tuple of fields (array types)."""
        if gtype in GEOMETRY_ARRAY:
            fmt = GEOMETRY_ARRAY[gtype]
            items = list(data)
            body = b"".join(struct.pack(fmt, *it) for it in items)
            count = len(items)
        elif gtype in GEOMETRY_STRUCT:
            body = struct.pack(GEOMETRY_STRUCT[gtype], *data)
            count = 0
        else:
            body, count = b"", 0
        head = struct.pack("<4I", gtype, 0 if rgba32 is None else 1, 0 if rgba32 is None else rgba32 & 0xFFFFFFFF, count)
        self._cmd(OP_FILL_GEOMETRY, head + body)
        self.fill_count += 1

    def fill_path_ex(self, cmds, vtx, origin=None, rgba32: int | None = None) -> None:
        cmds = np.asarray(cmds, dtype=np.uint8)
        vtx = np.asarray(vtx, dtype=np.float64).reshape(-1, 2)
        assert len(cmds) == len(vtx)
        flags = (1 if origin is not None else 0) | (2 if rgba32 is not None else 0)
        ox, oy = origin if origin is not None else (0.0, 0.0)
        head = struct.pack("<4I2d", len(cmds), flags, 0 if rgba32 is None else rgba32 & 0xFFFFFFFF, 0, ox, oy)
        self._cmd(OP_FILL_PATH_EX, head + _pad8(cmds.tobytes()) + vtx.tobytes())
        self.fill_count += 1

    def fill_explicit(self, target: int, style, geometry=None) -> None:
        """A fill with a style for this call only: fill_*_rgba64 / fill_*_ext
        of the reference.  `style`: ("rgba32", v), ("rgba64", v),
        ("gradient", type, values, stops, extend, transform) or ("pattern",
        image_index, extend, transform).  `geometry`: None (fill_all),
        (x, y, w, h) for rects, (cmds, vtx) for paths, (type, data) for
        fill_geometry."""
        kind = {"rgba32": STYLE_RGBA32, "rgba64": STYLE_RGBA64, "gradient": STYLE_GRADIENT, "pattern": STYLE_PATTERN}[style[0]]
        rgba = style[1] if kind in (STYLE_RGBA32, STYLE_RGBA64) else 0
        body, gtype, count = b"", 0, 0
        if kind == STYLE_GRADIENT:
            _, gt, values, stops, extend, transform = style
            vals = list(values) + [0.0] * (6 - len(values))
            tr = list(transform) if transform is not None else [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
            body = struct.pack("<IIII6d6d", gt, extend, len(stops), 1 if transform is not None else 0, *vals, *tr)
            for off, c in stops:
                body += struct.pack("<dII", off, c & 0xFFFFFFFF, 0)
        elif kind == STYLE_PATTERN:
            _, idx, extend, transform = style
            tr = list(transform) if transform is not None else [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
            body = struct.pack("<IIII6d", idx, extend, 1 if transform is not None else 0, 0, *tr)
        if target == TARGET_RECT_I:
            body += struct.pack("<4i", *[int(v) for v in geometry])
        elif target == TARGET_RECT_D:
            body += struct.pack("<4d", *[float(v) for v in geometry])
        elif target == TARGET_PATH:
            cmds = np.asarray(geometry[0], dtype=np.uint8)
            vtx = np.asarray(geometry[1], dtype=np.float64).reshape(-1, 2)
            count = len(cmds)
            body += _pad8(cmds.tobytes()) + vtx.tobytes()
        elif target == TARGET_MASK:
            kind_m, index, dst, area = geometry
            body += self._mask_payload(kind_m, index, dst, area)
        elif target == TARGET_GEOMETRY:
            gtype, data = geometry
            if gtype in GEOMETRY_ARRAY:
                items = list(data)
                body += b"".join(struct.pack(GEOMETRY_ARRAY[gtype], *it) for it in items)
                count = len(items)
            else:
                body += struct.pack(GEOMETRY_STRUCT[gtype], *data)
        head = struct.pack("<IIIIQ", target, kind, gtype, count, rgba & 0xFFFFFFFFFFFFFFFF)
        self._cmd(OP_FILL_EXPLICIT, head + body)
        self.fill_count += 1

    # -- serialization ---------------------------------------------------
    def to_bytes(self) -> bytes:
        out = [struct.pack("<8I", MAGIC, VERSION, self.width, self.height, self.format,
                           len(self.images), len(self.cmds), self.clear_value & 0xFFFFFFFF)]
        for fmt, px in self.images:
            h, w = px.shape
            bpp = 1 if fmt == FMT_A8 else 4
            out.append(struct.pack("<4I", w, h, fmt, w * bpp))
            out.append(_pad8(px.tobytes()))
        for op, payload in self.cmds:
            out.append(struct.pack("<II", op, len(payload)))
            out.append(payload)
        return b"".join(out)

    def pixel_bytes(self) -> int:
        return self.width * self.height * (1 if self.format == FMT_A8 else 4)


# ---------------------------------------------------------------------------
# Geometry helpers
# ---------------------------------------------------------------------------

def ellipse_path(cx: float, cy: float, rx: float, ry: float):
    """Ellipse as a closed figure of four cubic arcs (one per quadrant, starting
    at angle 0), with the circle-to-cubic constant k = 4/3 (sqrt(2) - 1)."""
    k = 0.55228474983079339840
    kx, ky = rx * k, ry * k
    cmds = [CMD_MOVE] + [CMD_CUBIC, CMD_CUBIC, CMD_ON] * 4 + [CMD_CLOSE]
    v = [(cx + rx, cy),
         (cx + rx, cy + ky), (cx + kx, cy + ry), (cx, cy + ry),
         (cx - kx, cy + ry), (cx - rx, cy + ky), (cx - rx, cy),
         (cx - rx, cy - ky), (cx - kx, cy - ry), (cx, cy - ry),
         (cx + kx, cy - ry), (cx + rx, cy - ky), (cx + rx, cy),
         (cx + rx, cy)]
    return cmds, v


def rotation_about(angle: float, cx: float, cy: float):
    """Matrix rotating by `angle` around (cx, cy): [m00 m01 m10 m11 m20 m21]."""
    c, s = math.cos(angle), math.sin(angle)
    return [c, s, -s, c, cx - c * cx + s * cy, cy - s * cx - c * cy]


def premultiplied_noise(rng: np.random.Generator, w: int, h: int) -> np.ndarray:
    """Seeded PRGB32 noise image (channels <= alpha)."""
    a = rng.integers(0, 256, size=(h, w), dtype=np.uint32)
    r = (rng.integers(0, 256, size=(h, w), dtype=np.uint32) * a + 127) // 255
    g = (rng.integers(0, 256, size=(h, w), dtype=np.uint32) * a + 127) // 255
    b = (rng.integers(0, 256, size=(h, w), dtype=np.uint32) * a + 127) // 255
    return ((a << 24) | (r << 16) | (g << 8) | b).astype(np.uint32)


def _rand_rgba(rng) -> int:
    return int(rng.integers(0, 1 << 32, dtype=np.uint64))


def _rand_stops(rng, n_min=2, n_max=8):
    n = int(rng.integers(n_min, n_max + 1))
    offs = np.sort(rng.random(n))
    return [(float(o), _rand_rgba(rng)) for o in offs]


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def polygons_scene(seed: int = 1, width: int = 512, height: int = 512, count: int = 1000) -> Scene:
    """512x512 PRGB32 cleared to opaque white, `count` random simple polygons
    (3-12 vertices uniform in the canvas), non-zero rule, random straight-alpha
    solid colours, SRC_OVER, submission order."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, FMT_PRGB32, 0xFFFFFFFF)
    sc.set_fill_rule(0)
    for _ in range(count):
        n = int(rng.integers(3, 13))
        pts = rng.random((n, 2)) * np.array([width, height])
        sc.set_fill_style_rgba32(_rand_rgba(rng))
        sc.fill_polygon(pts)
    return sc


def cubic_gradient_scene(seed: int = 1, width: int = 1920, height: int = 1080, count: int = 10000) -> Scene:
    """1920x1080 PRGB32, `count` closed cubic-Bezier paths (4-16 segments, control
    points uniform in a random 16-600 px box), even-odd, 50% linear / 50% radial
    gradients with 2-8 stops, PAD extend, SRC_OVER."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, FMT_PRGB32, 0xFF000000)
    sc.set_fill_rule(1)
    for i in range(count):
        side = float(rng.uniform(16, 600))
        # boxes larger than the canvas start a quarter side above / left of it
        bx = float(rng.uniform(-side * 0.25, max(width - side * 0.75, -side * 0.25)))
        by = float(rng.uniform(-side * 0.25, max(height - side * 0.75, -side * 0.25)))
        k = int(rng.integers(4, 17))
        pts = rng.random((3 * k, 2)) * side + np.array([bx, by])
        cmds = [CMD_MOVE] + [CMD_CUBIC, CMD_CUBIC, CMD_ON] * k + [CMD_CLOSE]
        vtx = np.vstack([pts[-1:], pts, pts[-1:]])
        stops = _rand_stops(rng)
        if i % 2 == 0:
            p0 = rng.random(2) * side + np.array([bx, by])
            p1 = rng.random(2) * side + np.array([bx, by])
            sc.set_fill_style_gradient(GRADIENT_LINEAR, [p0[0], p0[1], p1[0], p1[1]], stops, EXTEND_PAD)
        else:
            c = rng.random(2) * side + np.array([bx, by])
            fpt = c + (rng.random(2) - 0.5) * side * 0.2
            r = float(rng.uniform(8, 300))
            sc.set_fill_style_gradient(GRADIENT_RADIAL, [c[0], c[1], fpt[0], fpt[1], r, 0.0], stops, EXTEND_PAD)
        sc.fill_path(cmds, vtx)
    return sc


def pattern_shapes_scene(seed: int = 1, width: int = 3840, height: int = 2160, count: int = 2000,
                         comp_ops=(COMP_OP_SRC_OVER, COMP_OP_MULTIPLY, COMP_OP_PLUS, COMP_OP_SCREEN),
                         pattern_size: int = 512) -> Scene:
    """3840x2160 PRGB32, `count` rotated rectangles and ellipses filled with an
    affine pattern of a seeded 512x512 PRGB32 noise image (rotation 0-2pi, scale
    log-uniform 0.5-2), bilinear, REPEAT, comp ops drawn from `comp_ops`."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, FMT_PRGB32, 0xFF808080)
    img = sc.add_image(premultiplied_noise(rng, pattern_size, pattern_size))
    sc.set_pattern_quality(1)
    for i in range(count):
        w = float(rng.uniform(16, 600))
        h = float(rng.uniform(16, 600))
        cx = float(rng.uniform(0, width))
        cy = float(rng.uniform(0, height))
        ang = float(rng.uniform(0, 2 * math.pi))
        pang = float(rng.uniform(0, 2 * math.pi))
        psc = float(2.0 ** rng.uniform(-1, 1))
        op = int(comp_ops[int(rng.integers(0, len(comp_ops)))])
        c, s = math.cos(pang) * psc, math.sin(pang) * psc
        sc.set_comp_op(op)
        if i % 2 == 0:
            sc.set_transform(rotation_about(ang, cx, cy))
            sc.set_fill_style_pattern(img, EXTEND_REPEAT, [c, s, -s, c, cx, cy])
            sc.fill_rect_d(cx - w / 2, cy - h / 2, w, h)
        else:
            sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
            sc.set_fill_style_pattern(img, EXTEND_REPEAT, [c, s, -s, c, cx, cy])
            cm, v = ellipse_path(cx, cy, w / 2, h / 2)
            sc.fill_path(cm, v)
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def glyph_path(rng, size_w: float, size_h: float):
    """Closed quadratic contour with 20-80 segments around a wobbly ring."""
    n = int(rng.integers(20, 81))
    cx, cy = size_w * 0.5, size_h * 0.5
    ang = np.sort(rng.random(n)) * 2 * math.pi
    rad = 0.25 + 0.25 * rng.random(n)
    pts = np.stack([cx + np.cos(ang) * rad * size_w, cy + np.sin(ang) * rad * size_h], 1)
    ctl_r = 0.2 + 0.3 * rng.random(n)
    mid = ang + math.pi / n
    ctl = np.stack([cx + np.cos(mid) * ctl_r * size_w, cy + np.sin(mid) * ctl_r * size_h], 1)
    cmds = np.empty(2 * n + 2, dtype=np.uint8)
    cmds[0] = CMD_MOVE
    cmds[1:-1:2] = CMD_QUAD
    cmds[2:-1:2] = CMD_ON
    cmds[-1] = CMD_CLOSE
    vtx = np.empty((2 * n + 2, 2))
    vtx[0] = pts[0]
    vtx[1:-1:2] = ctl
    vtx[2:-1:2] = np.roll(pts, -1, axis=0)
    vtx[-1] = pts[0]
    return cmds, vtx


def glyph_canvases(seed: int = 1, count: int = 4096, max_paths: int = 200):
    """`count` independent canvases, sides log-uniform 32-256 px, half A8 and
    half PRGB32, each with 1-`max_paths` glyph-like quadratic contours
    (20-80 segments around a wobbly ring, see glyph_path), non-zero, solid
    and linear-gradient fills, SRC_OVER.  Generated per canvas with array
    operations."""
    rng = np.random.default_rng(seed)
    scenes = []
    two_pi = 2 * math.pi
    for i in range(count):
        w = int(round(2.0 ** rng.uniform(5, 8)))
        h = int(round(2.0 ** rng.uniform(5, 8)))
        fmt = FMT_A8 if i % 2 == 0 else FMT_PRGB32
        sc = Scene(w, h, fmt, 0)
        sc.set_fill_rule(0)
        k = int(rng.integers(1, max_paths + 1))
        gw = rng.uniform(8, w, k)
        gh = rng.uniform(8, h, k)
        ox = rng.uniform(-gw * 0.2, w - gw * 0.8)
        oy = rng.uniform(-gh * 0.2, h - gh * 0.8)
        ns = rng.integers(20, 81, k)
        total = int(ns.sum())
        pid = np.repeat(np.arange(k), ns)
        ang = np.sort(rng.random(total) + pid) - pid          # sorted within each path
        ang *= two_pi
        rad = 0.25 + 0.25 * rng.random(total)
        ctl_r = 0.2 + 0.3 * rng.random(total)
        gwp, ghp = gw[pid], gh[pid]
        cx, cy = gwp * 0.5 + ox[pid], ghp * 0.5 + oy[pid]
        px = cx + np.cos(ang) * rad * gwp
        py = cy + np.sin(ang) * rad * ghp
        mid = ang + math.pi / ns[pid]
        qx = cx + np.cos(mid) * ctl_r * gwp
        qy = cy + np.sin(mid) * ctl_r * ghp
        solid = rng.random(k) < 0.5
        colors = rng.integers(0, 1 << 32, k, dtype=np.uint64)
        gp = rng.random((k, 4))
        starts = np.concatenate([[0], np.cumsum(ns)])
        for j in range(k):
            s0, s1 = starts[j], starts[j + 1]
            n = s1 - s0
            if solid[j]:
                sc.set_fill_style_rgba32(int(colors[j]))
            else:
                p0x, p0y = gp[j, 0] * gw[j] + ox[j], gp[j, 1] * gh[j] + oy[j]
                p1x, p1y = gp[j, 2] * gw[j] + ox[j], gp[j, 3] * gh[j] + oy[j]
                sc.set_fill_style_gradient(GRADIENT_LINEAR, [p0x, p0y, p1x, p1y], _rand_stops(rng, 2, 4))
            cmds = np.empty(2 * n + 2, dtype=np.uint8)
            cmds[0] = CMD_MOVE
            cmds[1:-1:2] = CMD_QUAD
            cmds[2:-1:2] = CMD_ON
            cmds[-1] = CMD_CLOSE
            vtx = np.empty((2 * n + 2, 2))
            vtx[0, 0], vtx[0, 1] = px[s0], py[s0]
            vtx[1:-1:2, 0], vtx[1:-1:2, 1] = qx[s0:s1], qy[s0:s1]
            vtx[2:-2:2, 0], vtx[2:-2:2, 1] = px[s0 + 1:s1], py[s0 + 1:s1]
            vtx[-2] = vtx[0]
            vtx[-1] = vtx[0]
            sc.fill_path(cmds, vtx)
        scenes.append(sc)
    return scenes


def srcover_layers_scene(seed: int = 1, width: int = 3840, height: int = 2160, layers: int = 1) -> Scene:
    """A canvas cleared to seeded opaque colour with `layers` full-canvas
    SRC_OVER fills of semi-transparent solid colours (fill_all): the pure
    compositing load, 8 bytes of destination traffic per pixel and layer."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24))
    sc = Scene(width, height, FMT_PRGB32, clear)
    for _ in range(layers):
        a = int(rng.integers(16, 240))
        sc.set_fill_style_rgba32((a << 24) | int(rng.integers(0, 1 << 24)))
        sc.fill_all()
    return sc


def blit_layer_scene(image: np.ndarray, seed: int = 1) -> Scene:
    """A canvas the size of `image` (premultiplied PRGB32) cleared to seeded
    opaque colour, with the image blitted over it once at the origin
    (blit_image_i, SRC_OVER): 12 bytes of traffic per pixel (destination
    read and write, source read)."""
    rng = np.random.default_rng(seed)
    h, w = image.shape
    sc = Scene(w, h, FMT_PRGB32, 0xFF000000 | int(rng.integers(0, 1 << 24)))
    k = sc.add_image(image)
    sc.blit_image(BLIT_I, k, (0, 0))
    return sc


def mask_layer_scene(mask: np.ndarray, seed: int = 1) -> Scene:
    """A PRGB32 canvas the size of the A8 `mask` cleared to seeded opaque
    colour, with one fill_mask of a semi-transparent solid colour through the
    mask at the origin (SRC_OVER): 9 bytes of traffic per pixel."""
    rng = np.random.default_rng(seed)
    h, w = mask.shape
    sc = Scene(w, h, FMT_PRGB32, 0xFF000000 | int(rng.integers(0, 1 << 24)))
    k = sc.add_image(mask, FMT_A8)
    sc.set_fill_style_rgba32((int(rng.integers(64, 240)) << 24) | int(rng.integers(0, 1 << 24)))
    sc.fill_mask(MASK_I, k, (0, 0))
    return sc


def mixed_scene(seed: int = 1, width: int = 513, height: int = 513, count: int = 120,
                fmt: int = FMT_PRGB32, comp_ops=(COMP_OP_SRC_OVER, COMP_OP_SRC_COPY),
                pattern_size: int = 64) -> Scene:
    """Every call of the fill path in one seeded stream: integer and fractional
    rects, polygons, cubic ellipses, quadratic contours and fill_all; solid,
    linear / radial / conic gradients (PAD / REPEAT / REFLECT, optional
    gradient transform) and nearest / bilinear affine patterns; both fill
    rules, global alpha, scale / rotation / skew user transforms and the
    given comp ops.  Pattern transforms always rotate, so the affine fetchers
    are used."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    img = sc.add_image(premultiplied_noise(rng, pattern_size, pattern_size))
    ident = [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    for i in range(count):
        # state
        sc.set_comp_op(int(comp_ops[int(rng.integers(0, len(comp_ops)))]))
        sc.set_fill_rule(int(rng.integers(0, 2)))
        sc.set_global_alpha(1.0 if rng.random() < 0.6 else float(rng.uniform(0.0, 1.0)))
        t = rng.random()
        cx, cy = float(rng.uniform(0, width)), float(rng.uniform(0, height))
        if t < 0.45:
            xf = ident
        elif t < 0.6:
            xf = [1.0, 0.0, 0.0, 1.0, float(rng.integers(-8, 9)), float(rng.integers(-8, 9))]
        elif t < 0.75:
            sx, sy = float(2.0 ** rng.uniform(-1, 1)), float(2.0 ** rng.uniform(-1, 1))
            xf = [sx, 0.0, 0.0, sy, cx - sx * cx, cy - sy * cy]
        elif t < 0.9:
            xf = rotation_about(float(rng.uniform(0, 2 * math.pi)), cx, cy)
        else:
            k = float(rng.uniform(-0.5, 0.5))
            xf = [1.0, 0.0, k, 1.0, -k * cy, 0.0]
        sc.set_transform(xf)
        # style
        st = rng.random()
        size = float(rng.uniform(8, max(width, height) * 0.6))
        if st < 0.3:
            sc.set_fill_style_rgba32(_rand_rgba(rng))
        elif st < 0.75:
            gtype = int(rng.integers(0, 3))
            extend = int(rng.integers(0, 3))
            stops = _rand_stops(rng, 1, 6)
            gtr = None
            if rng.random() < 0.3:
                a = float(rng.uniform(0, 2 * math.pi))
                gtr = rotation_about(a, cx, cy)
            if gtype == GRADIENT_LINEAR:
                p = rng.random(4) * size + np.array([cx, cy, cx, cy]) - size / 2
                vals = [float(v) for v in p]
            elif gtype == GRADIENT_RADIAL:
                r = float(rng.uniform(4, size))
                f = (rng.random(2) - 0.5) * r
                vals = [cx, cy, cx + float(f[0]), cy + float(f[1]), r, 0.0]
            else:
                vals = [cx, cy, float(rng.uniform(0, 2 * math.pi)), float(rng.choice([1.0, 1.0, 2.0, 0.5]))]
            sc.set_fill_style_gradient(gtype, vals, stops, extend, gtr)
        else:
            sc.set_pattern_quality(int(rng.integers(0, 2)))
            a = float(rng.uniform(0.1, 2 * math.pi - 0.1))
            s = float(2.0 ** rng.uniform(-1, 1))
            c_, s_ = math.cos(a) * s, math.sin(a) * s
            sc.set_fill_style_pattern(img, int(rng.integers(0, 3)), [c_, s_, -s_, c_, cx, cy])
        # geometry
        g = rng.random()
        x0, y0 = float(rng.uniform(-0.2, 1.0) * width), float(rng.uniform(-0.2, 1.0) * height)
        w, h = float(rng.uniform(1, size)), float(rng.uniform(1, size))
        if g < 0.15:
            sc.fill_rect_i(int(x0), int(y0), int(w), int(h))
        elif g < 0.3:
            q = 1.0 / 256.0
            sc.fill_rect_d(round(x0 / q) * q if rng.random() < 0.5 else x0, y0, w, h)
        elif g < 0.5:
            n = int(rng.integers(3, 13))
            sc.fill_polygon(rng.random((n, 2)) * np.array([w, h]) + np.array([x0, y0]))
        elif g < 0.7:
            cm, v = ellipse_path(x0 + w / 2, y0 + h / 2, w / 2, h / 2)
            sc.fill_path(cm, v)
        elif g < 0.98:
            cm, v = glyph_path(rng, w, h)
            sc.fill_path(cm, np.asarray(v) + np.array([x0, y0]))
        else:
            sc.set_global_alpha(float(rng.uniform(0.05, 0.3)))
            sc.fill_all()
    sc.set_transform(ident)
    return sc


def translated_patterns_scene(seed: int = 1, width: int = 301, height: int = 211, count: int = 120,
                              fmt: int = FMT_PRGB32,
                              comp_ops=(COMP_OP_SRC_OVER, COMP_OP_SRC_COPY, COMP_OP_PLUS, COMP_OP_MULTIPLY)) -> Scene:
    """Patterns whose final transform is a pure translation, the case that
    selects the aligned and fractional-translation fetchers: integer, 1/256
    and arbitrary fractional offsets (positive, negative and far outside the
    image), set on the pattern, on the user transform or split between them;
    PAD / REPEAT / REFLECT; nearest and bilinear; PRGB32, XRGB32 and A8
    images including 1-pixel-wide / -high and tiny (3x2) ones; aligned and
    fractional rects, ellipses and quadratic contours."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    imgs = [sc.add_image(premultiplied_noise(rng, 61, 47)),
            sc.add_image(premultiplied_noise(rng, 1, 37)),
            sc.add_image(premultiplied_noise(rng, 29, 1)),
            sc.add_image(premultiplied_noise(rng, 3, 2)),
            sc.add_image(rng.integers(0, 1 << 32, size=(23, 19), dtype=np.uint32), FMT_XRGB32),
            sc.add_image(rng.integers(0, 256, size=(17, 33), dtype=np.uint8), FMT_A8)]

    def offset():
        t = rng.random()
        if t < 0.3:
            return float(rng.integers(-400, 400))
        if t < 0.6:
            return float(rng.integers(-400 * 256, 400 * 256)) / 256.0
        if t < 0.9:
            return float(rng.uniform(-400, 400))
        return float(rng.choice([-1e5, 1e5])) + float(rng.uniform(-1, 1))

    for _ in range(count):
        sc.set_comp_op(int(comp_ops[int(rng.integers(0, len(comp_ops)))]))
        sc.set_global_alpha(1.0 if rng.random() < 0.7 else float(rng.uniform(0.0, 1.0)))
        sc.set_pattern_quality(int(rng.integers(0, 2)))
        u = rng.random()
        ux, uy = (offset(), offset()) if u < 0.3 else (float(rng.integers(-20, 20)), 0.0) if u < 0.5 else (0.0, 0.0)
        sc.set_transform([1.0, 0.0, 0.0, 1.0, ux, uy])
        img = imgs[int(rng.integers(0, len(imgs)))]
        ptr = [1.0, 0.0, 0.0, 1.0, offset(), offset()] if rng.random() < 0.8 else None
        sc.set_fill_style_pattern(img, int(rng.integers(0, 3)), ptr)
        x0, y0 = float(rng.uniform(-0.2, 1.0) * width) - ux, float(rng.uniform(-0.2, 1.0) * height) - uy
        w, h = float(rng.uniform(1, width * 0.7)), float(rng.uniform(1, height * 0.7))
        g = rng.random()
        if g < 0.35:
            sc.fill_rect_i(int(x0), int(y0), int(w), int(h))
        elif g < 0.6:
            sc.fill_rect_d(x0, y0, w, h)
        elif g < 0.8:
            cm, v = ellipse_path(x0 + w / 2, y0 + h / 2, w / 2, h / 2)
            sc.fill_path(cm, v)
        else:
            cm, v = glyph_path(rng, w, h)
            sc.fill_path(cm, np.asarray(v) + np.array([x0, y0]))
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def cancelling_pattern_transforms_scene(seed: int = 1, width: int = 200, height: int = 150, count: int = 40,
                                        fmt: int = FMT_PRGB32) -> Scene:
    """Patterns whose transform cancels the user transform (scale 2 / 0.5,
    3 / (1/3), rotation by +a / -a), so the final transform is a translation
    exactly or to within rounding (the near-identity test of the aligned /
    fractional fetchers), PAD / REPEAT / REFLECT, nearest and bilinear."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, fmt, 0xFF203040 if fmt != FMT_A8 else 0x40)
    img = sc.add_image(premultiplied_noise(rng, 37, 29))
    for i in range(count):
        sc.set_comp_op(COMP_OP_SRC_OVER if i % 3 else COMP_OP_SRC_COPY)
        sc.set_pattern_quality(int(rng.integers(0, 2)))
        k = i % 3
        cx, cy = float(rng.uniform(0, width)), float(rng.uniform(0, height))
        if k == 0:
            u, p = [2.0, 0.0, 0.0, 2.0, cx, cy], [0.5, 0.0, 0.0, 0.5, float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20))]
        elif k == 1:
            u, p = [3.0, 0.0, 0.0, 3.0, cx, cy], [1.0 / 3.0, 0.0, 0.0, 1.0 / 3.0, float(rng.uniform(-9, 9)), float(rng.uniform(-9, 9))]
        else:
            a = float(rng.uniform(0, 2 * math.pi))
            c_, s_ = math.cos(a), math.sin(a)
            u = [c_, s_, -s_, c_, cx, cy]
            p = [c_, -s_, s_, c_, float(rng.uniform(-30, 30)), float(rng.uniform(-30, 30))]
        sc.set_transform(u)
        sc.set_fill_style_pattern(img, int(rng.integers(0, 3)), p)
        w, h = float(rng.uniform(5, 60)), float(rng.uniform(5, 60))
        if i % 2:
            sc.fill_rect_d(-w / 2, -h / 2, w, h)
        else:
            cm, v = ellipse_path(0.0, 0.0, w / 2, h / 2)
            sc.fill_path(cm, v)
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def box_scene(seed: int = 1, width: int = 333, height: int = 77, count: int = 60, fmt: int = FMT_PRGB32,
              comp_ops=(COMP_OP_SRC_OVER, COMP_OP_SRC_COPY)) -> Scene:
    """Box fills only (integer and fractional rects, fill_all, clears) under
    identity, translate and scale transforms, so every band holds box fills
    only: solid, linear / radial / conic gradients (PAD / REPEAT / REFLECT)
    and nearest / bilinear affine patterns, global and fill alpha, the given
    comp ops.  Canvas sizes need not be multiples of the tile size."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    img = sc.add_image(premultiplied_noise(rng, 32, 32))
    for i in range(count):
        sc.set_comp_op(int(comp_ops[int(rng.integers(0, len(comp_ops)))]))
        sc.set_global_alpha(1.0 if rng.random() < 0.7 else float(rng.uniform(0.0, 1.0)))
        sc.set_fill_alpha(1.0 if rng.random() < 0.8 else float(rng.uniform(0.0, 1.0)))
        t = rng.random()
        if t < 0.5:
            sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
        elif t < 0.75:
            sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.uniform(-9, 9)), float(rng.uniform(-9, 9))])
        else:
            sc.set_transform([float(2.0 ** rng.uniform(-1, 1)), 0.0, 0.0, float(2.0 ** rng.uniform(-1, 1)),
                              float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20))])
        cx, cy = float(rng.uniform(0, width)), float(rng.uniform(0, height))
        size = float(rng.uniform(8, max(width, height, 9)))
        st = rng.random()
        if st < 0.3:
            sc.set_fill_style_rgba32(_rand_rgba(rng))
        elif st < 0.8:
            gtype = int(rng.integers(0, 3))
            extend = int(rng.integers(0, 3))
            stops = _rand_stops(rng, 1, 6)
            if gtype == GRADIENT_LINEAR:
                vals = [float(v) for v in rng.random(4) * size + np.array([cx, cy, cx, cy]) - size / 2]
            elif gtype == GRADIENT_RADIAL:
                r = float(rng.uniform(4, size))
                f = (rng.random(2) - 0.5) * r
                vals = [cx, cy, cx + float(f[0]), cy + float(f[1]), r, 0.0]
            else:
                vals = [cx, cy, float(rng.uniform(0, 2 * math.pi)), float(rng.choice([1.0, 2.0, 0.5]))]
            sc.set_fill_style_gradient(gtype, vals, stops, extend)
        else:
            sc.set_pattern_quality(int(rng.integers(0, 2)))
            a = float(rng.uniform(0.1, 2 * math.pi - 0.1))
            s = float(2.0 ** rng.uniform(-1, 1))
            c_, s_ = math.cos(a) * s, math.sin(a) * s
            sc.set_fill_style_pattern(img, int(rng.integers(0, 3)), [c_, s_, -s_, c_, cx, cy])
        g = rng.random()
        x0, y0 = float(rng.uniform(-0.2, 1.0) * width), float(rng.uniform(-0.2, 1.0) * height)
        w, h = float(rng.uniform(1, size)), float(rng.uniform(1, size * 0.5))
        if g < 0.4:
            sc.fill_rect_i(int(x0), int(y0), int(w), int(h))
        elif g < 0.85:
            sc.fill_rect_d(x0, y0, w, h)
        elif g < 0.9:
            sc.fill_all()
        elif g < 0.95:
            sc.clear_rect_d(x0, y0, w, h)
        else:
            sc.fill_rect_d_rgba32(x0, y0, w, h, _rand_rgba(rng))
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def edge_case_scene(width: int = 257, height: int = 131, fmt: int = FMT_PRGB32) -> Scene:
    """Degenerate and boundary inputs: empty and one-point paths, zero-area
    and zero-size rects, geometry exactly on and far outside the canvas
    edges, huge coordinates, horizontal-only polygons, self-intersecting
    figures under both fill rules and unclosed figures."""
    sc = Scene(width, height, fmt, 0xFF102030 if fmt != FMT_A8 else 0x10)
    W, H = float(width), float(height)
    sc.set_fill_style_rgba32(0xC0FF8040)
    sc.fill_path([], np.zeros((0, 2)))
    sc.fill_path([CMD_MOVE], [(5.0, 5.0)])
    sc.fill_polygon([(1.0, 1.0), (9.0, 1.0)])
    sc.fill_polygon([(0.0, 10.0), (W, 10.0), (W / 2, 10.0)])
    sc.fill_rect_i(3, 3, 0, 10)
    sc.fill_rect_i(-5, -5, 10, 10)
    sc.fill_rect_i(width - 4, height - 4, 100, 100)
    sc.fill_rect_d(10.25, 20.5, 0.0, 5.0)
    sc.fill_rect_d(10.25, 20.5, 0.5, 0.25)
    sc.fill_rect_d(-1e9, -1e9, 2e9, 12.75)
    sc.fill_polygon([(0.0, 0.0), (W, 0.0), (W, 0.5), (0.0, 0.5)])
    sc.fill_polygon([(W - 0.5, 0.0), (W, 0.0), (W, H), (W - 0.5, H)])
    sc.fill_polygon([(-1e7, 30.0), (1e7, 40.0), (1e7, 60.0), (-1e7, 50.0)])
    sc.fill_polygon([(W + 10, 10.0), (W + 50, 10.0), (W + 30, 90.0)])
    sc.fill_polygon([(-50.0, 10.0), (-10.0, 10.0), (-30.0, 90.0)])
    sc.fill_polygon([(20.0, H + 5), (60.0, H + 5), (40.0, H + 50)])
    sc.set_fill_style_rgba32(0x8020C0F0)
    for rule in (0, 1):
        sc.set_fill_rule(rule)
        ox = 60.0 + rule * 90.0
        star = [(ox + 40 * math.cos(2 * math.pi * k * 2 / 5), 70 + 40 * math.sin(2 * math.pi * k * 2 / 5)) for k in range(5)]
        sc.fill_polygon(star)
        sc.fill_path([CMD_MOVE, CMD_ON, CMD_ON, CMD_MOVE, CMD_ON, CMD_ON],
                     [(ox - 30, 100.0), (ox + 30, 100.0), (ox, 125.0), (ox - 20, 105.0), (ox + 20, 105.0), (ox, 120.0)])
    sc.set_fill_style_rgba32(0xFFFFFFFF)
    sc.fill_polygon([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0)])
    sc.fill_polygon([(W - 1.0, H - 1.0), (W, H - 1.0), (W, H)])
    sc.set_fill_style_rgba32(0x30FFFFFF)
    sc.fill_path([CMD_MOVE, CMD_CUBIC, CMD_CUBIC, CMD_ON],
                 [(0.0, H), (W * 3, -H * 2), (-W * 2, -H * 2), (W, H)])
    sc.fill_path([CMD_MOVE, CMD_QUAD, CMD_ON, CMD_QUAD, CMD_ON],
                 [(W * 0.5, -H), (W * 4, H * 0.5), (W * 0.5, H * 2), (-W * 3, H * 0.5), (W * 0.5, -H)])
    return sc


INT32_MIN, INT32_MAX = -(1 << 31), (1 << 31) - 1


def _clip_tables(cw: int, ch: int, sw: int, sh: int):
    """The user transforms, points and sizes of the reference's clipping
    robustness test (core/context_test.cpp:389-545) for a cw x ch canvas and
    an sw x sh sprite."""
    nan, inf = float("nan"), float("inf")
    matrices = [
        (1.0, 0.0, 0.0, 1.0, 0.0, 0.0), (1.0, 0.0, 0.0, 1.0, 11.3, 11.9),
        (100.0, 0.0, 0.0, 100.0, 0.0, 0.0), (-100.0, 0.0, 0.0, -100.0, 0.0, 0.0),
        (1.0, 0.0, 0.0, 0.000001, 0.0, 0.0), (0.000001, 0.0, 0.0, 1.0, 0.0, 0.0),
        (0.000001, 0.0, 0.0, 0.000001, 0.0, 0.0), (1e-20, 0.0, 0.0, 1e-20, 0.0, 0.0),
        (1e-100, 0.0, 0.0, 1e-100, 0.0, 0.0)]
    pts_i = [(0, 0), (0, -1), (-1, 0), (-1, -1), (0, ch - 1), (cw - 1, 0), (cw - 1, ch - 1), (0, -sh + 1),
             (-sw + 1, 0), (-sw + 1, -sh + 1), (INT32_MIN, 0), (INT32_MIN, -1), (0, INT32_MIN), (-1, INT32_MIN),
             (INT32_MIN, INT32_MIN), (INT32_MAX, 0), (INT32_MAX, -1), (0, INT32_MAX), (-1, INT32_MAX),
             (INT32_MAX, INT32_MAX)]
    pts_d = [(0.0, 0.0), (0.0, 0.3), (0.3, 0.0), (0.3, 0.3), (0.0, 100.0), (100.0, 0.0), (100.0, 100.0)]
    pts_d += [(0.0, -sh + 10.0 ** -k) for k in range(1, 8)] + [(0.0, -sh + 1.0)]
    pts_d += [(-sw + 10.0 ** -k, 0.0) for k in range(1, 8)] + [(-sw + 1.0, 0.0)]
    pts_d += [(0.0, ch - 10.0 ** -k) for k in range(1, 6)] + [(cw - 10.0 ** -k, 0.0) for k in range(1, 6)]
    pts_d += [(cw - 0.00001, ch - 0.00001), (-1000.0, 0.0), (-1000000.0, 0.0), (0.0, -1000.0), (0.0, -1000000.0),
              (-1000.0, -1000.0), (-1000000.0, -1000000.0), (-1e50, -1e50), (-1e100, -1e100), (-1e200, -1e200),
              (1e50, 1e50), (1e100, 1e100), (1e200, 1e200), (inf, 0.0), (0.0, inf), (inf, inf), (-inf, 0.0),
              (0.0, -inf), (-inf, -inf), (nan, 0.0), (0.0, nan), (nan, nan)]
    sizes_i = [(sw, sh), (sw // 2, sh // 2), (1, 1), (0, 0), (0, INT32_MIN), (INT32_MIN, 0), (INT32_MIN, INT32_MIN),
               (0, INT32_MAX), (INT32_MAX, 0), (INT32_MAX, INT32_MAX)]
    sizes_d = [(float(sw), float(sh)), (float(sw // 2), float(sh // 2)), (1.0, 1.0), (0.0, 0.0), (0.0, float(sh)),
               (float(sw), 0.0), (0.1, 0.1), (0.00001, float(sh)), (float(sw), 0.00001), (0.00001, 0.00001),
               (0.0000001, 0.0000001), (-0.00001, float(sh)), (float(sw), -0.00001), (-0.00001, -0.00001)]
    sizes_d += [(float(sw), 10.0 ** e) for e in (40, 80, 120, 160, 200)]
    sizes_d += [(10.0 ** e, float(sh)) for e in (40, 80, 120, 160, 200)]
    sizes_d += [(inf, float(sh)), (inf, float(sh)), (inf, inf), (float(sw), -inf), (-inf, float(sh)), (-inf, -inf),
                (float(sw), nan), (nan, float(sh)), (nan, nan)]
    return matrices, pts_i, pts_d, sizes_i, sizes_d


def clip_robustness_scenes(width: int = 97, height: int = 67, fmt: int = FMT_PRGB32, seed: int = 1) -> list:
    """The fill-clipping matrix of the reference's robustness test
    (core/context_test.cpp:389-586): under each of its nine user transforms,
    fill_rect(BLRectI) for every (point, size) pair of its integer tables and
    fill_rect(BLRect) for every pair of its double tables (NaN, +-Inf, INT32
    extremes, 1e-100 scales, sizes just below a pixel).  The reference fills
    opaque white; here each call gets its own seeded translucent colour so a
    coverage difference anywhere shows in the pixels.  One canvas per
    transform."""
    matrices, pts_i, pts_d, sizes_i, sizes_d = _clip_tables(width, height, 23, 23)
    rng = np.random.default_rng(seed)
    scenes = []
    for m in matrices:
        sc = Scene(width, height, fmt, 0xFF000000 if fmt != FMT_A8 else 0)
        sc.set_transform(m)
        for x, y in pts_i:
            for w, h in sizes_i:
                sc.set_fill_style_rgba32(int(rng.integers(0, 1 << 24)) | (int(rng.integers(8, 48)) << 24))
                sc.fill_rect_i(x, y, w, h)
        for x, y in pts_d:
            for w, h in sizes_d:
                sc.set_fill_style_rgba32(int(rng.integers(0, 1 << 24)) | (int(rng.integers(8, 48)) << 24))
                sc.fill_rect_d(x, y, w, h)
        scenes.append(sc)
    return scenes


def blit_clip_robustness_scenes(width: int = 97, height: int = 67, fmt: int = FMT_PRGB32, seed: int = 1) -> list:
    """The blit-clipping matrix of the reference's robustness test
    (core/context_test.cpp:567-586): under each of its nine user transforms,
    blit_image(BLPointI) and blit_image(BLRectI) for every point and (point,
    size) pair of its integer tables, blit_image(BLPoint) and
    blit_image(BLRect) for its double tables, of a 23 x 23 sprite.  The
    reference blits an opaque white sprite; here the sprite is seeded noise
    and each call gets its own global alpha, so a difference in coverage or
    sampling anywhere shows in the pixels.  One canvas per transform."""
    matrices, pts_i, pts_d, sizes_i, sizes_d = _clip_tables(width, height, 23, 23)
    rng = np.random.default_rng(seed)
    scenes = []
    for m in matrices:
        sc = Scene(width, height, fmt, 0xFF000000 if fmt != FMT_A8 else 0)
        sprite = sc.add_image(premultiplied_noise(rng, 23, 23))
        sc.set_transform(m)
        for x, y in pts_i:
            sc.set_global_alpha(float(rng.uniform(0.05, 0.3)))
            sc.blit_image(BLIT_I, sprite, (x, y))
            for w, h in sizes_i:
                sc.set_global_alpha(float(rng.uniform(0.05, 0.3)))
                sc.blit_image(BLIT_SCALED_I, sprite, (x, y, w, h))
        for x, y in pts_d:
            sc.set_global_alpha(float(rng.uniform(0.05, 0.3)))
            sc.blit_image(BLIT_D, sprite, (x, y))
            for w, h in sizes_d:
                sc.set_global_alpha(float(rng.uniform(0.05, 0.3)))
                sc.blit_image(BLIT_SCALED_D, sprite, (x, y, w, h))
        scenes.append(sc)
    return scenes


def invalid_geometry_scene(seed: int = 1, width: int = 160, height: int = 120, fmt: int = FMT_PRGB32,
                           count: int = 160) -> Scene:
    """Polygons and line paths, part of them carrying a NaN vertex,
    interleaved with valid polygons and cubic paths, under identity, scaling
    and rotating transforms.  A line from a vertex inside the canvas to a NaN
    vertex gets an invalid clip-flag combination in the reference: the whole
    fill fails with BL_ERROR_INVALID_GEOMETRY and its edges are reverted
    (raster/edgebuilder_p.h:1584-1586, raster/rastercontextops.cpp:38-45).
    The NaN vertex is therefore the second one, after a first vertex inside
    the canvas, so the failing line is the first line of the fill.  Two
    other placements are left out on purpose: where the line before the NaN
    vertex starts outside, the reference's clipper decides from clip flags
    it carries from line to line, partly recomputed
    (raster/edgebuilder_p.h:1300-1420); and where edges of the same fill were
    already stored, the reference's revert leaves its band links behind (its
    own assertions fire: support/zeroallocator.cpp:39,
    raster/analyticrasterizer_p.h:468).  Infinite or astronomically large vertices are covered by the
    rect matrix of clip_robustness_scenes only: in polygons and curves the
    reference itself writes out of bounds, trips assertions
    (raster/edgebuilder_p.h:2506, raster/analyticrasterizer_p.h:468) or
    subdivides curves to its recursion limit."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, fmt, 0xFF204060 if fmt != FMT_A8 else 0x20)
    W, H = float(width), float(height)
    nan = float("nan")
    transforms = [(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), (1.5, 0.0, 0.0, 0.75, 3.0, -2.0), (0.8, 0.6, -0.6, 0.8, 40.0, -10.0)]
    tr = transforms[0]
    for i in range(count):
        if i % 40 == 0:
            tr = transforms[(i // 40) % len(transforms)]
            sc.set_transform(tr)
            sc.set_fill_rule(int(rng.integers(0, 2)))
        sc.set_fill_style_rgba32(_rand_rgba(rng) | 0x40000000)
        kind = int(rng.integers(0, 3))
        n = int(rng.integers(3, 9))
        cx, cy = rng.uniform(-0.2 * W, 1.2 * W), rng.uniform(-0.2 * H, 1.2 * H)
        vtx = np.stack([cx + rng.uniform(-0.4, 0.4, n) * W, cy + rng.uniform(-0.4, 0.4, n) * H], axis=1)
        if kind == 2:
            # valid cubic path between the invalid fills
            m = (n - 1) // 3
            if m:
                sc.fill_path([CMD_MOVE] + [CMD_CUBIC] * (3 * m) + [CMD_CLOSE],
                             np.vstack([vtx[:1 + 3 * m], vtx[:1]]))
                continue
        if rng.random() < 0.6:
            # vertices (after the user transform) strictly inside the canvas
            ux = vtx[:, 0] * tr[0] + vtx[:, 1] * tr[2] + tr[4]
            uy = vtx[:, 0] * tr[1] + vtx[:, 1] * tr[3] + tr[5]
            inside = (ux > 0) & (ux < W) & (uy > 0) & (uy < H)
            if inside[0]:
                j = 1
                which = int(rng.integers(0, 3))
                if which != 1:
                    vtx[j, 0] = nan
                if which != 0:
                    vtx[j, 1] = nan
        if kind == 1:
            sc.fill_path([CMD_MOVE] + [CMD_ON] * (n - 1) + [CMD_CLOSE], np.vstack([vtx, vtx[:1]]))
        else:
            sc.fill_polygon(vtx)
    return sc


def comp_op_folds_scene(seed: int = 1, width: int = 211, height: int = 157, fmt: int = FMT_PRGB32,
                        count: int = 160) -> Scene:
    """PLUS / MULTIPLY / SCREEN where the comp-op simplification of the
    reference (core/compopsimplifyimpl_p.h) folds them into operators its
    portable pipeline renders: a fully transparent solid source (ZERO32)
    makes PLUS, MULTIPLY and SCREEN a NOP on every target; on an A8 target
    MULTIPLY becomes SRC_OVER (SRC_COPY for an opaque source) and SCREEN
    SRC_OVER (SRC_COPY); on XRGB32 an opaque MULTIPLY becomes MODULATE, which
    neither side renders (BL_ERROR_NOT_IMPLEMENTED).  Interleaved with
    SRC_OVER / SRC_COPY fills of translucent, opaque and transparent solids and
    linear gradients, global alpha and box / polygon / path geometry.
    Combinations the portable pipeline cannot render (PLUS, MULTIPLY, SCREEN
    of a non-zero source on PRGB32 / XRGB32, PLUS on A8) are left out."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    ops = (COMP_OP_SRC_OVER, COMP_OP_SRC_COPY, COMP_OP_PLUS, COMP_OP_MULTIPLY, COMP_OP_SCREEN)
    W, H = float(width), float(height)
    for i in range(count):
        op = int(ops[int(rng.integers(0, len(ops)))])
        kind = int(rng.integers(0, 4))   # 0 zero, 1 opaque solid, 2 translucent solid, 3 linear gradient
        folds_to_renderable = (
            op in (COMP_OP_SRC_OVER, COMP_OP_SRC_COPY) or kind == 0 or
            (fmt == FMT_A8 and op in (COMP_OP_MULTIPLY, COMP_OP_SCREEN)) or
            (fmt == FMT_XRGB32 and op == COMP_OP_MULTIPLY and kind == 1))
        if not folds_to_renderable:
            kind = 0
        sc.set_comp_op(op)
        sc.set_global_alpha(1.0 if rng.random() < 0.7 else float(rng.uniform(0.1, 1.0)))
        sc.set_fill_rule(int(rng.integers(0, 2)))
        if kind == 0:
            sc.set_fill_style_rgba32(0x00000000)
        elif kind == 1:
            sc.set_fill_style_rgba32(0xFF000000 | int(rng.integers(0, 1 << 24)))
        elif kind == 2:
            sc.set_fill_style_rgba32(int(rng.integers(1, 255)) << 24 | int(rng.integers(0, 1 << 24)))
        else:
            x0, y0 = float(rng.uniform(0, W)), float(rng.uniform(0, H))
            sc.set_fill_style_gradient(GRADIENT_LINEAR, [x0, y0, x0 + float(rng.uniform(-80, 80)), y0 + float(rng.uniform(-80, 80))],
                                       _rand_stops(rng, 2, 4), int(rng.integers(0, 3)))
        g = rng.random()
        x0, y0 = float(rng.uniform(-0.1, 0.9) * W), float(rng.uniform(-0.1, 0.9) * H)
        w, h = float(rng.uniform(4, W * 0.5)), float(rng.uniform(4, H * 0.5))
        if g < 0.3:
            sc.fill_rect_i(int(x0), int(y0), int(w), int(h))
        elif g < 0.5:
            sc.fill_rect_d(x0, y0, w, h)
        elif g < 0.8:
            n = int(rng.integers(3, 9))
            sc.fill_polygon(rng.random((n, 2)) * np.array([w, h]) + np.array([x0, y0]))
        else:
            cm, v = ellipse_path(x0 + w / 2, y0 + h / 2, w / 2, h / 2)
            sc.fill_path(cm, v)
    return sc


def explicit_styles_scene(seed: int = 1, width: int = 233, height: int = 171, fmt: int = FMT_PRGB32,
                          count: int = 90) -> Scene:
    """Fills with a style for one call only -- fill_all / fill_rect_i /
    fill_rect_d / fill_path_d / fill_geometry in their _rgba64 and _ext forms
    (core/context.h:735-758; rgba32, rgba64, linear / radial / conic gradient
    and pattern styles) -- interleaved with ordinary fills of the context's
    own style, under changing user transforms (an explicit style takes the
    transform of its call), global and fill alpha and SRC_OVER / SRC_COPY."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    img = sc.add_image(premultiplied_noise(rng, 37, 29))
    W, H = float(width), float(height)
    sc.set_fill_style_rgba32(0x80FF8040)
    for i in range(count):
        if i % 15 == 0:
            t = rng.random()
            cx, cy = float(rng.uniform(0, W)), float(rng.uniform(0, H))
            xf = [1.0, 0.0, 0.0, 1.0, 0.0, 0.0] if t < 0.4 else rotation_about(float(rng.uniform(0, 2 * math.pi)), cx, cy)
            sc.set_transform(xf)
            sc.set_comp_op(COMP_OP_SRC_OVER if rng.random() < 0.8 else COMP_OP_SRC_COPY)
            sc.set_global_alpha(1.0 if rng.random() < 0.6 else float(rng.uniform(0.2, 1.0)))
            sc.set_fill_alpha(1.0 if rng.random() < 0.7 else float(rng.uniform(0.2, 1.0)))
        k = int(rng.integers(0, 4))
        if k == 0:
            style = ("rgba32", _rand_rgba(rng))
        elif k == 1:
            style = ("rgba64", int(rng.integers(0, 1 << 63, dtype=np.uint64)) * 2 + int(rng.integers(0, 2)))
        elif k == 2:
            gt = int(rng.integers(0, 3))
            x0, y0 = float(rng.uniform(0, W)), float(rng.uniform(0, H))
            if gt == GRADIENT_LINEAR:
                vals = [x0, y0, x0 + float(rng.uniform(-90, 90)), y0 + float(rng.uniform(-90, 90))]
            elif gt == GRADIENT_RADIAL:
                r = float(rng.uniform(5, 90))
                vals = [x0, y0, x0 + float(rng.uniform(-r, r)) * 0.5, y0, r, 0.0]
            else:
                vals = [x0, y0, float(rng.uniform(0, 6.28)), 1.0]
            style = ("gradient", gt, vals, _rand_stops(rng, 2, 5), int(rng.integers(0, 3)), None)
        else:
            a = float(rng.uniform(0.2, 6.0))
            style = ("pattern", img, int(rng.integers(0, 3)), [math.cos(a), math.sin(a), -math.sin(a), math.cos(a), 11.0, 7.0])
        target = int(rng.integers(0, 5))
        x0, y0 = float(rng.uniform(-0.1, 0.9) * W), float(rng.uniform(-0.1, 0.9) * H)
        w, h = float(rng.uniform(4, W * 0.6)), float(rng.uniform(4, H * 0.6))
        if target == TARGET_ALL:
            if rng.random() < 0.7:
                target = TARGET_RECT_D
            geom = None if target == TARGET_ALL else (x0, y0, w, h)
        elif target == TARGET_RECT_I:
            geom = (int(x0), int(y0), int(w), int(h))
        elif target == TARGET_RECT_D:
            geom = (x0, y0, w, h)
        elif target == TARGET_PATH:
            geom = ellipse_path(x0 + w / 2, y0 + h / 2, w / 2, h / 2)
        else:
            n = int(rng.integers(3, 8))
            geom = (GEOMETRY_POLYGOND, [tuple(v) for v in (rng.random((n, 2)) * np.array([w, h]) + np.array([x0, y0]))])
        sc.fill_explicit(target, style, geom)
        if rng.random() < 0.3:
            # the context's own style is untouched by the explicit calls
            n = int(rng.integers(3, 7))
            sc.fill_polygon(rng.random((n, 2)) * np.array([W, H]))
    return sc


def blit_scene(seed: int = 1, width: int = 257, height: int = 181, fmt: int = FMT_PRGB32, count: int = 90,
               bad_areas: bool = False) -> Scene:
    """blit_image_i / _d and blit_scaled_image_i / _d (core/context.h:870-873)
    of PRGB32, XRGB32 and A8 images, whole or through an area (with
    `bad_areas`, now and then an area outside the image, which the reference
    rejects with BL_ERROR_INVALID_VALUE and renders nothing): integer, 1/256 and
    arbitrary fractional origins, partly or wholly off the canvas; scaled
    rectangles larger, smaller, mirrored, empty or of the area's own size;
    under identity, integral and fractional translations, scales, rotations
    and a singular user transform; SRC_OVER / SRC_COPY, global alpha (0 makes
    a blit a NOP; fill alpha and the fill style are ignored) and both
    pattern qualities, interleaved with ordinary fills."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    imgs = [sc.add_image(premultiplied_noise(rng, 61, 47)),
            sc.add_image(premultiplied_noise(rng, 1, 37)),
            sc.add_image(premultiplied_noise(rng, 29, 1)),
            sc.add_image(premultiplied_noise(rng, 3, 2)),
            sc.add_image(premultiplied_noise(rng, 130, 90)),
            sc.add_image(rng.integers(0, 1 << 32, size=(23, 19), dtype=np.uint32), FMT_XRGB32),
            sc.add_image(rng.integers(0, 256, size=(17, 33), dtype=np.uint8), FMT_A8)]
    sizes = [(61, 47), (1, 37), (29, 1), (3, 2), (130, 90), (19, 23), (33, 17)]
    W, H = float(width), float(height)

    def coord(extent):
        t = rng.random()
        if t < 0.35:
            return float(rng.integers(int(-0.3 * extent), int(extent)))
        if t < 0.6:
            return float(rng.integers(int(-0.3 * extent) * 256, int(extent) * 256)) / 256.0
        return float(rng.uniform(-0.3 * extent, extent))

    for i in range(count):
        r = rng.random()
        if r < 0.06:
            t = rng.random()
            if t < 0.35:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.integers(-20, 21)), float(rng.integers(-20, 21))])
            elif t < 0.55:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20))])
            elif t < 0.7:
                sc.set_transform([float(rng.uniform(0.5, 2)), 0.0, 0.0, float(rng.uniform(0.5, 2)),
                                  float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20))])
            elif t < 0.82:
                sc.set_transform(rotation_about(float(rng.uniform(0, 6.3)), W / 2, H / 2))
            elif t < 0.87:
                sc.set_transform([0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
            else:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
            continue
        if r < 0.1:
            sc.set_comp_op(int(rng.choice([COMP_OP_SRC_OVER, COMP_OP_SRC_COPY])))
            continue
        if r < 0.14:
            a = rng.random()
            sc.set_global_alpha(0.0 if a < 0.15 else 1.0 if a < 0.6 else float(rng.uniform(0.0, 1.0)))
            continue
        if r < 0.17:
            sc.set_fill_alpha(float(rng.uniform(0.0, 1.0)))
            sc.set_pattern_quality(int(rng.integers(0, 2)))
            continue
        if r < 0.22:
            sc.set_fill_style_rgba32(_rand_rgba(rng))
            sc.fill_rect_d(coord(W), coord(H), float(rng.uniform(1, W / 2)), float(rng.uniform(1, H / 2)))
            continue
        k = int(rng.integers(0, len(imgs)))
        iw, ih = sizes[k]
        area = None
        a = rng.random()
        if a < 0.4:
            ax, ay = int(rng.integers(0, iw)), int(rng.integers(0, ih))
            # zero widths are NOPs; zero heights are left out (the reference's
            # pattern fetch divides by the area height)
            area = (ax, ay, int(rng.integers(0, iw - ax + 1)), int(rng.integers(1, ih - ay + 1)))
        elif a < 0.45 and bad_areas:
            area = (int(rng.integers(-3, iw + 3)), int(rng.integers(-3, ih + 3)), iw, ih)
            if area[0] == 0 and area[1] == 0:
                area = (1, 0, iw, ih)
        aw, ah = (area[2], area[3]) if area is not None else (iw, ih)
        kind = int(rng.integers(0, 4))
        x, y = coord(W), coord(H)
        if kind == BLIT_I:
            sc.blit_image(kind, imgs[k], (int(x), int(y)), area)
        elif kind == BLIT_D:
            sc.blit_image(kind, imgs[k], (x, y), area)
        else:
            s = rng.random()
            if s < 0.15:
                w, h = float(aw), float(ah)
            elif s < 0.25:
                w, h = float(rng.uniform(-80, -1)), float(rng.uniform(1, 80))
            elif s < 0.3:
                w, h = 0.0, float(rng.uniform(1, 40))
            else:
                w, h = float(rng.uniform(1, W * 0.6)), float(rng.uniform(1, H * 0.6))
            if kind == BLIT_SCALED_I:
                sc.blit_image(kind, imgs[k], (int(x), int(y), int(w), int(h)), area)
            else:
                sc.blit_image(kind, imgs[k], (x, y, w, h), area)
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def mask_scene(seed: int = 1, width: int = 241, height: int = 173, fmt: int = FMT_PRGB32, count: int = 90,
               unaligned: bool = False) -> Scene:
    """fill_mask_i / _d (core/context.h:800-808) with the fill style (solid,
    linear / radial gradient, pattern) and with a style for the call alone
    (rgba32, rgba64, gradient, pattern): A8 masks (noise, smooth ramps, one
    pixel wide / high), and a PRGB32 and an XRGB32 image used as masks (the
    reference reads their bytes as an A8 row); whole masks and areas,
    partly or wholly off the canvas; identity and integral translations, and
    fractional origins whose translation makes them pixel-aligned; SRC_OVER
    / SRC_COPY, global and fill alpha (full alpha and below).  With
    `unaligned`, also the placements the reference reports as
    BL_ERROR_NOT_IMPLEMENTED (a fractional start, a rotation or scale), global
    alpha 0 over such placements (a NOP, no error), and areas outside the mask
    (BL_ERROR_INVALID_VALUE)."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    yy, xx = np.mgrid[0:40, 0:57]
    ramp = ((xx * 4 + yy * 3) % 256).astype(np.uint8)
    masks = [sc.add_image(rng.integers(0, 256, size=(45, 61), dtype=np.uint8), FMT_A8),
             sc.add_image(ramp, FMT_A8),
             sc.add_image(rng.integers(0, 256, size=(1, 30), dtype=np.uint8), FMT_A8),
             sc.add_image(rng.integers(0, 256, size=(25, 1), dtype=np.uint8), FMT_A8),
             sc.add_image(premultiplied_noise(rng, 33, 21)),
             sc.add_image(rng.integers(0, 1 << 32, size=(17, 29), dtype=np.uint32), FMT_XRGB32)]
    sizes = [(61, 45), (57, 40), (30, 1), (1, 25), (33, 21), (29, 17)]
    pattern = sc.add_image(premultiplied_noise(rng, 23, 19))
    W, H = float(width), float(height)

    def style():
        k = rng.random()
        if k < 0.4:
            return ("rgba32", _rand_rgba(rng))
        if k < 0.55:
            return ("rgba64", int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2)))
        if k < 0.8:
            x0, y0 = float(rng.uniform(0, W)), float(rng.uniform(0, H))
            if rng.random() < 0.5:
                return ("gradient", GRADIENT_LINEAR, [x0, y0, x0 + float(rng.uniform(-90, 90)), y0 + float(rng.uniform(-90, 90))],
                        _rand_stops(rng, 2, 4), int(rng.integers(0, 3)), None)
            return ("gradient", GRADIENT_RADIAL, [x0, y0, x0 + 5.0, y0 - 3.0, float(rng.uniform(20, 90)), 0.0],
                    _rand_stops(rng, 2, 4), int(rng.integers(0, 3)), None)
        return ("pattern", pattern, int(rng.integers(0, 3)),
                [1.0, 0.0, 0.0, 1.0, float(rng.integers(-30, 30)), float(rng.integers(-30, 30))] if rng.random() < 0.5 else None)

    translation = (0.0, 0.0)
    for i in range(count):
        r = rng.random()
        if r < 0.07:
            t = rng.random()
            if t < 0.4:
                translation = (float(rng.integers(-15, 16)), float(rng.integers(-15, 16)))
            elif t < 0.6:
                translation = (0.5, -0.25)   # fractional: aligned again by origins x.5 / y.25
            elif t < 0.8 or not unaligned:
                translation = (0.0, 0.0)
            if unaligned and t >= 0.8:
                sc.set_transform(rotation_about(float(rng.uniform(0, 6.3)), W / 2, H / 2) if t < 0.9 else
                                 [1.5, 0.0, 0.0, 1.5, 0.0, 0.0])
                translation = None
            else:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, translation[0], translation[1]])
            continue
        if r < 0.1:
            sc.set_comp_op(int(rng.choice([COMP_OP_SRC_OVER, COMP_OP_SRC_COPY])))
            continue
        if r < 0.14:
            a = rng.random()
            sc.set_global_alpha(0.0 if (a < 0.1 and unaligned) else 1.0 if a < 0.6 else float(rng.uniform(0.05, 1.0)))
            continue
        if r < 0.17:
            sc.set_fill_alpha(1.0 if rng.random() < 0.5 else float(rng.uniform(0.05, 1.0)))
            continue
        if r < 0.25:
            s_ = style()
            if s_[0] == "rgba32":
                sc.set_fill_style_rgba32(s_[1])
            elif s_[0] == "rgba64":
                sc.set_fill_style_rgba64(s_[1])
            elif s_[0] == "gradient":
                sc.set_fill_style_gradient(s_[1], s_[2], s_[3], s_[4], s_[5])
            else:
                sc.set_fill_style_pattern(s_[1], s_[2], s_[3])
            continue
        k = int(rng.integers(0, len(masks)))
        mw, mh = sizes[k]
        area = None
        a = rng.random()
        if a < 0.35:
            ax, ay = int(rng.integers(0, mw)), int(rng.integers(0, mh))
            area = (ax, ay, int(rng.integers(0, mw - ax + 1)), int(rng.integers(0, mh - ay + 1)))
        elif a < 0.4 and unaligned:
            area = (int(rng.integers(-2, mw + 2)), int(rng.integers(-2, mh + 2)), mw, mh)
            if area[0] == 0 and area[1] == 0:
                area = (0, 1, mw, mh)
        x, y = int(rng.integers(int(-0.2 * W), int(W))), int(rng.integers(int(-0.2 * H), int(H)))
        kind = MASK_I if rng.random() < 0.5 else MASK_D
        dst = (x, y)
        if kind == MASK_D:
            if translation == (0.5, -0.25):
                dst = (x + 0.5, y + 0.25)
            elif unaligned and rng.random() < 0.3:
                dst = (x + float(rng.uniform(0.01, 0.99)), float(y))
            else:
                dst = (float(x), float(y))
        elif translation == (0.5, -0.25) and not unaligned:
            kind, dst = MASK_D, (x + 0.5, y + 0.25)
        if rng.random() < 0.4:
            sc.fill_explicit(TARGET_MASK, style(), (kind, masks[k], dst, area))
        else:
            sc.fill_mask(kind, masks[k], dst, area)
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def dither_scene(seed: int = 1, width: int = 223, height: int = 167, fmt: int = FMT_PRGB32, count: int = 80,
                 bad_hint: bool = False) -> Scene:
    """Gradients under the gradient quality hint (BL_CONTEXT_HINT_GRADIENT_
    QUALITY): NEAREST, SMOOTH (rendered as nearest) and DITHER (a 64-bit table
    of twice the size plus the pixel's Bayer threshold); linear, radial and
    conic, PAD / REPEAT / REFLECT, 1-12 stops opaque and translucent, with
    gradient and user transforms, as the fill style (which keeps the hint of
    the moment it was set) and as a style for one call; rects, ellipses,
    glyph contours and fill_all; SRC_OVER / SRC_COPY and global alpha.  A8
    targets never dither.  With `bad_hint`, also a hint value out of range
    (BL_ERROR_INVALID_VALUE, the hint unchanged)."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    W, H = float(width), float(height)

    def gradient():
        gtype = int(rng.integers(0, 3))
        cx, cy = float(rng.uniform(0, W)), float(rng.uniform(0, H))
        size = float(rng.uniform(10, max(W, H)))
        if gtype == GRADIENT_LINEAR:
            p = rng.random(4) * size + np.array([cx, cy, cx, cy]) - size / 2
            vals = [float(v) for v in p]
        elif gtype == GRADIENT_RADIAL:
            r = float(rng.uniform(4, size))
            f = (rng.random(2) - 0.5) * r
            vals = [cx, cy, cx + float(f[0]), cy + float(f[1]), r, 0.0]
        else:
            vals = [cx, cy, float(rng.uniform(0, 2 * math.pi)), float(rng.choice([1.0, 2.0, 0.5]))]
        stops = _rand_stops(rng, 1, 12)
        if rng.random() < 0.4:   # opaque stops: the table keeps 16-bit channels between them
            stops = [(o, c | 0xFF000000) for o, c in stops]
        gtr = rotation_about(float(rng.uniform(0, 2 * math.pi)), cx, cy) if rng.random() < 0.25 else None
        return gtype, vals, stops, int(rng.integers(0, 3)), gtr

    for i in range(count):
        r = rng.random()
        if r < 0.15:
            q = int(rng.choice([0, 1, 2, 2, 2]))
            if bad_hint and rng.random() < 0.2:
                q = 3
            sc.set_gradient_quality(q)
            continue
        if r < 0.2:
            t = rng.random()
            cx, cy = W / 2, H / 2
            if t < 0.4:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.integers(-9, 10)), float(rng.integers(-9, 10))])
            elif t < 0.7:
                sc.set_transform(rotation_about(float(rng.uniform(0, 2 * math.pi)), cx, cy))
            else:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
            continue
        if r < 0.25:
            sc.set_comp_op(int(rng.choice([COMP_OP_SRC_OVER, COMP_OP_SRC_COPY])))
            sc.set_global_alpha(1.0 if rng.random() < 0.6 else float(rng.uniform(0.1, 1.0)))
            continue
        explicit = rng.random() < 0.3
        if not explicit:
            gtype, vals, stops, extend, gtr = gradient()
            sc.set_fill_style_gradient(gtype, vals, stops, extend, gtr)
        g = rng.random()
        x0, y0 = float(rng.uniform(-0.2, 0.9) * W), float(rng.uniform(-0.2, 0.9) * H)
        w, h = float(rng.uniform(4, W * 0.7)), float(rng.uniform(4, H * 0.7))
        if explicit:
            gtype, vals, stops, extend, gtr = gradient()
            style = ("gradient", gtype, vals, stops, extend, gtr)
            sc.fill_explicit(TARGET_RECT_D if g < 0.6 else TARGET_ALL, style, (x0, y0, w, h) if g < 0.6 else None)
        elif g < 0.25:
            sc.fill_rect_i(int(x0), int(y0), int(w), int(h))
        elif g < 0.5:
            sc.fill_rect_d(x0, y0, w, h)
        elif g < 0.75:
            cm, v = ellipse_path(x0 + w / 2, y0 + h / 2, w / 2, h / 2)
            sc.fill_path(cm, v)
        elif g < 0.95:
            cm, v = glyph_path(rng, w, h)
            sc.fill_path(cm, np.asarray(v) + np.array([x0, y0]))
        else:
            sc.fill_all()
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def frontend_scene(seed: int = 1, width: int = 300, height: int = 220, fmt: int = FMT_PRGB32, count: int = 120) -> Scene:
    """The rest of the fill frontend in one seeded stream: fill_geometry of
    every non-arc geometry type, explicit-colour fills, clear_all /
    clear_rect, fill alpha, 64-bit and float colours (NaN disables the
    style), fill_path with an origin, and non-invertible user transforms
    (geometry fills become no-ops, fill_all still renders)."""
    rng = np.random.default_rng(seed)
    clear = 0xFF000000 | int(rng.integers(0, 1 << 24)) if fmt != FMT_A8 else int(rng.integers(0, 256))
    sc = Scene(width, height, fmt, clear)
    W, H = float(width), float(height)

    def pt():
        return float(rng.uniform(-0.1, 1.1) * W), float(rng.uniform(-0.1, 1.1) * H)

    for i in range(count):
        r = rng.random()
        if r < 0.08:
            sc.set_fill_alpha(float(rng.uniform(0, 1)) if rng.random() < 0.8 else 1.0)
        elif r < 0.14:
            sc.set_fill_style_rgba64(int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2)))
        elif r < 0.2:
            v = [float(x) for x in rng.uniform(-0.2, 1.2, 4)]
            if rng.random() < 0.1:
                v[int(rng.integers(0, 4))] = float("nan")
            sc.set_fill_style_rgba(*v)
        elif r < 0.26:
            sc.set_fill_style_rgba32(_rand_rgba(rng))
        elif r < 0.3:
            t = rng.random()
            if t < 0.25:
                sc.set_transform([0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
            elif t < 0.5:
                sc.set_transform(rotation_about(float(rng.uniform(0, 6.3)), W / 2, H / 2))
            else:
                sc.set_transform([1.0, 0.0, 0.0, 1.0, float(rng.integers(-5, 6)), float(rng.integers(-5, 6))])
        elif r < 0.33:
            sc.set_comp_op(int(rng.choice([COMP_OP_SRC_OVER, COMP_OP_SRC_COPY])))
        elif r < 0.35:
            if rng.random() < 0.2:
                sc.clear_all()
            else:
                sc.fill_all_rgba32((int(rng.integers(0, 0x80)) << 24) | int(rng.integers(0, 1 << 24)))
        elif r < 0.4:
            x, y = pt()
            w, h = float(rng.uniform(-5, W / 2)), float(rng.uniform(-5, H / 2))
            k = int(rng.integers(0, 4))
            if k == 0:
                sc.clear_rect_i(int(x), int(y), int(w), int(h))
            elif k == 1:
                sc.clear_rect_d(x, y, w, h)
            elif k == 2:
                sc.fill_rect_i_rgba32(int(x), int(y), int(w), int(h), _rand_rgba(rng))
            else:
                sc.fill_rect_d_rgba32(x, y, w, h, _rand_rgba(rng))
        elif r < 0.46:
            cm, v = glyph_path(rng, float(rng.uniform(10, W / 2)), float(rng.uniform(10, H / 2)))
            org = pt() if rng.random() < 0.7 else None
            col = _rand_rgba(rng) if rng.random() < 0.5 else None
            sc.fill_path_ex(cm, v, org, col)
        else:
            g = int(rng.choice([GEOMETRY_BOXI, GEOMETRY_BOXD, GEOMETRY_RECTI, GEOMETRY_RECTD, GEOMETRY_CIRCLE,
                                GEOMETRY_ELLIPSE, GEOMETRY_ROUND_RECT, GEOMETRY_LINE, GEOMETRY_TRIANGLE,
                                GEOMETRY_POLYLINEI, GEOMETRY_POLYLINED, GEOMETRY_POLYGONI, GEOMETRY_POLYGOND,
                                GEOMETRY_ARRAY_VIEW_BOXI, GEOMETRY_ARRAY_VIEW_BOXD, GEOMETRY_ARRAY_VIEW_RECTI,
                                GEOMETRY_ARRAY_VIEW_RECTD, GEOMETRY_NONE]))
            x, y = pt()
            w, h = float(rng.uniform(-4, W / 2)), float(rng.uniform(-4, H / 2))
            if g in (GEOMETRY_BOXI, GEOMETRY_RECTI):
                data = (int(x), int(y), int(x + w) if g == GEOMETRY_BOXI else int(w), int(y + h) if g == GEOMETRY_BOXI else int(h))
            elif g == GEOMETRY_BOXD:
                data = (x, y, x + w, y + h)
            elif g == GEOMETRY_RECTD:
                data = (x, y, w, h)
            elif g == GEOMETRY_CIRCLE:
                data = (x, y, float(rng.uniform(-40, 80)))
            elif g == GEOMETRY_ELLIPSE:
                data = (x, y, float(rng.uniform(-40, 80)), float(rng.uniform(-40, 80)))
            elif g == GEOMETRY_ROUND_RECT:
                data = (x, y, abs(w), abs(h), float(rng.uniform(-10, 40)), float(rng.uniform(0, 40)) if rng.random() < 0.9 else 0.0)
            elif g == GEOMETRY_LINE:
                data = (*pt(), *pt())
            elif g == GEOMETRY_TRIANGLE:
                data = (*pt(), *pt(), *pt())
            elif g in (GEOMETRY_POLYLINEI, GEOMETRY_POLYGONI):
                data = [(int(a), int(b)) for a, b in (pt() for _ in range(int(rng.integers(0, 9))))]
            elif g in (GEOMETRY_POLYLINED, GEOMETRY_POLYGOND):
                data = [pt() for _ in range(int(rng.integers(0, 9)))]
            elif g in (GEOMETRY_ARRAY_VIEW_BOXI, GEOMETRY_ARRAY_VIEW_RECTI):
                data = [tuple(int(v) for v in (*pt(), float(rng.uniform(-20, 60)), float(rng.uniform(-20, 60))))
                        for _ in range(int(rng.integers(0, 5)))]
                if g == GEOMETRY_ARRAY_VIEW_BOXI:
                    data = [(a, b, a + c, b + d) for a, b, c, d in data]
            elif g in (GEOMETRY_ARRAY_VIEW_BOXD, GEOMETRY_ARRAY_VIEW_RECTD):
                data = [(*pt(), float(rng.uniform(-20, 60)), float(rng.uniform(-20, 60))) for _ in range(int(rng.integers(0, 5)))]
                if g == GEOMETRY_ARRAY_VIEW_BOXD:
                    data = [(a, b, a + c, b + d) for a, b, c, d in data]
            else:
                data = ()
            sc.fill_geometry(g, data, _rand_rgba(rng) if rng.random() < 0.3 else None)
    sc.set_transform([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    return sc


def cost_canvases(seed: int = 1, count: int = 256, side_lo: int = 1024, side_hi: int = 4096, paths: int = 64,
                  max_segments: int = 20000):
    """Cost-model workload: `count` PRGB32 canvases, sides log-uniform in
    [side_lo, side_hi]; per path a segment count clip(Pareto(1.5) * 8, 8,
    max_segments) of lines, quads, cubics or a mix, control points uniform in
    a box of random size; solid / linear / radial / pattern fills and comp
    ops SRC_OVER, MULTIPLY, PLUS, SCREEN; both fill rules."""
    rng = np.random.default_rng(seed)
    ops = (COMP_OP_SRC_OVER, COMP_OP_MULTIPLY, COMP_OP_PLUS, COMP_OP_SCREEN)
    scenes = []
    for _ in range(count):
        w = int(round(2.0 ** rng.uniform(math.log2(side_lo), math.log2(side_hi))))
        h = int(round(2.0 ** rng.uniform(math.log2(side_lo), math.log2(side_hi))))
        sc = Scene(w, h, FMT_PRGB32, 0xFF000000 | int(rng.integers(0, 1 << 24)))
        img = sc.add_image(premultiplied_noise(rng, 64, 64))
        k = int(rng.integers(1, paths + 1))
        segs = np.clip(((rng.pareto(1.5, k) + 1.0) * 8.0).astype(np.int64), 8, max_segments)
        for j in range(k):
            n = int(segs[j])
            side = float(2.0 ** rng.uniform(4, math.log2(max(w, h))))
            bx = float(rng.uniform(-side * 0.3, max(w - side * 0.7, 1.0 - side * 0.3)))
            by = float(rng.uniform(-side * 0.3, max(h - side * 0.7, 1.0 - side * 0.3)))
            mix = int(rng.integers(0, 4))          # 0 lines, 1 quads, 2 cubics, 3 mixed
            kinds = np.full(n, mix) if mix < 3 else rng.integers(0, 3, n)
            per = np.array([1, 2, 3])[kinds]
            nv = int(per.sum())
            pts = rng.random((nv, 2)) * side + np.array([bx, by])
            cmds = np.empty(nv + 2, dtype=np.uint8)
            cmds[0] = CMD_MOVE
            # quads: QUAD ON, cubics: CUBIC CUBIC ON, lines: ON
            pos = 1 + np.concatenate([[0], np.cumsum(per)[:-1]])
            body = np.full(nv, CMD_ON, dtype=np.uint8)
            q = pos[kinds == 1] - 1
            body[q] = CMD_QUAD
            c = pos[kinds == 2] - 1
            body[c] = CMD_CUBIC
            body[c + 1] = CMD_CUBIC
            cmds[1:-1] = body
            cmds[-1] = CMD_CLOSE
            vtx = np.vstack([pts[-1:], pts, pts[-1:]])
            sc.set_comp_op(int(ops[int(rng.integers(0, 4))]))
            sc.set_fill_rule(int(rng.integers(0, 2)))
            st = rng.random()
            if st < 0.3:
                sc.set_fill_style_rgba32(_rand_rgba(rng))
            elif st < 0.55:
                p = rng.random(4) * side + np.array([bx, by, bx, by])
                sc.set_fill_style_gradient(GRADIENT_LINEAR, [float(v) for v in p], _rand_stops(rng, 2, 4))
            elif st < 0.8:
                cx, cy = float(bx + side / 2), float(by + side / 2)
                sc.set_fill_style_gradient(GRADIENT_RADIAL, [cx, cy, cx, cy, float(side / 2), 0.0], _rand_stops(rng, 2, 4))
            else:
                a = float(rng.uniform(0.1, 6.1))
                sc.set_fill_style_pattern(img, EXTEND_REPEAT, [math.cos(a), math.sin(a), -math.sin(a), math.cos(a), bx, by])
            sc.fill_path(cmds, vtx)
        scenes.append(sc)
    return scenes
