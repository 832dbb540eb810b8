"""blend2d_amd -- MI355X backend for Blend2D's fill path.

Python mirror of the reference's BLContext / BLImage / BLPath / BLGradient /
BLPattern fill API (core/context.h, core/path.h, core/gradient.h,
core/pattern.h), calling the HIP implementation through the C ABI in
include/blend2d_amd.h (libblend2d_amd.so, built by blend2d_amd.build).
There is no CPU fallback: without the built library, importing the GPU entry
points raises.
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libblend2d_amd.so")

FORMAT_PRGB32, FORMAT_XRGB32, FORMAT_A8 = 1, 2, 3
COMP_OP_SRC_OVER, COMP_OP_SRC_COPY, COMP_OP_PLUS, COMP_OP_MULTIPLY, COMP_OP_SCREEN = 0, 1, 12, 15, 16
FILL_RULE_NON_ZERO, FILL_RULE_EVEN_ODD = 0, 1
EXTEND_PAD, EXTEND_REPEAT, EXTEND_REFLECT = 0, 1, 2
GRADIENT_LINEAR, GRADIENT_RADIAL, GRADIENT_CONIC = 0, 1, 2
PATTERN_QUALITY_NEAREST, PATTERN_QUALITY_BILINEAR = 0, 1
FLUSH_SYNC = 0x80000000

CMD_MOVE, CMD_ON, CMD_QUAD, CMD_CUBIC, CMD_CLOSE = 0, 1, 2, 4, 5

SYMBOLS = [
    "bl_amd_version", "bl_amd_device_count",
    "bl_amd_image_create", "bl_amd_image_destroy", "bl_amd_image_upload", "bl_amd_image_download",
    "bl_amd_image_fill_raw", "bl_amd_image_get_data",
    "bl_amd_context_create", "bl_amd_context_destroy", "bl_amd_context_begin", "bl_amd_context_end",
    "bl_amd_context_flush", "bl_amd_context_set_comp_op", "bl_amd_context_set_fill_rule",
    "bl_amd_context_set_global_alpha", "bl_amd_context_set_transform", "bl_amd_context_set_pattern_quality",
    "bl_amd_context_set_fill_style_rgba32", "bl_amd_context_set_fill_style_gradient",
    "bl_amd_context_set_fill_style_pattern", "bl_amd_context_fill_path_d", "bl_amd_context_fill_polygon_d",
    "bl_amd_context_fill_rect_d", "bl_amd_context_fill_rect_i", "bl_amd_context_fill_all",
    "bl_amd_context_get_stats", "bl_amd_context_replay", "bl_amd_context_replay_noalloc",
    "bl_amd_context_get_error_flags", "bl_amd_context_get_deferred_error", "bl_amd_context_reset",
    "bl_amd_context_fill_all_rgba64", "bl_amd_context_fill_rect_i_rgba64", "bl_amd_context_fill_rect_d_rgba64",
    "bl_amd_context_fill_path_d_rgba64", "bl_amd_context_fill_geometry_rgba64",
    "bl_amd_context_fill_all_ext", "bl_amd_context_fill_rect_i_ext", "bl_amd_context_fill_rect_d_ext",
    "bl_amd_context_fill_path_d_ext", "bl_amd_context_fill_geometry_ext",
    "bl_amd_context_fill_all_rgba32", "bl_amd_context_fill_rect_i_rgba32", "bl_amd_context_fill_rect_d_rgba32",
    "bl_amd_context_fill_path_d_rgba32", "bl_amd_context_fill_geometry", "bl_amd_context_fill_geometry_rgba32",
    "bl_amd_context_clear_all", "bl_amd_context_clear_rect_i", "bl_amd_context_clear_rect_d",
    "bl_amd_context_set_fill_style_rgba64", "bl_amd_context_set_fill_style_rgba", "bl_amd_context_set_fill_alpha",
    "bl_amd_batch_create", "bl_amd_batch_destroy", "bl_amd_batch_add", "bl_amd_batch_clear", "bl_amd_batch_flush",
    "bl_amd_batch_replay", "bl_amd_batch_get_stats",
    "bl_amd_set_sizing", "bl_amd_cost_model_set", "bl_amd_context_collect_costs",
    "bl_amd_cost_model_forward", "bl_amd_cost_model_grad", "bl_amd_cost_model_train",
    "bl_amd_set_render_schedule", "bl_amd_tile_model_set", "bl_amd_context_collect_tile_costs",
    "bl_amd_batch_collect_tile_costs", "bl_amd_context_autotune", "bl_amd_batch_autotune", "bl_amd_autotune_log",
    "bl_amd_autotune_reset", "bl_amd_set_graph_replay",
    "bl_amd_context_blit_image_i", "bl_amd_context_blit_image_d", "bl_amd_context_blit_scaled_image_i",
    "bl_amd_context_blit_scaled_image_d",
    "bl_amd_context_fill_mask_i", "bl_amd_context_fill_mask_d", "bl_amd_context_fill_mask_i_rgba32",
    "bl_amd_context_fill_mask_d_rgba32", "bl_amd_context_fill_mask_i_rgba64", "bl_amd_context_fill_mask_d_rgba64",
    "bl_amd_context_fill_mask_i_ext", "bl_amd_context_fill_mask_d_ext", "bl_amd_context_set_gradient_quality",
]

SIZING_EXACT, SIZING_LEARNED, SIZING_HISTORY = 0, 1, 2
SCHEDULE_GRID, SCHEDULE_QUEUE, SCHEDULE_QUEUE_HEURISTIC, SCHEDULE_QUEUE_LEARNED, SCHEDULE_AUTO = 0, 1, 2, 3, 4
SCHEDULE_NAMES = {0: "grid", 1: "queue", 2: "queue_heuristic", 3: "queue_learned", 4: "auto"}
TILE_FEATURES, TILE_OUTPUTS = 16, 2
COST_FEATURES, COST_OUTPUTS, COST_PARAMS = 16, 2, 5378


class BLAmdStats(ctypes.Structure):
    """Mirror of BLAmdStats (include/blend2d_amd.h)."""
    _fields_ = [("commands", ctypes.c_uint64), ("sources", ctypes.c_uint64), ("segments", ctypes.c_uint64),
                ("binned", ctypes.c_uint64), ("tiles", ctypes.c_uint64), ("pixels", ctypes.c_uint64),
                ("frame_bytes", ctypes.c_uint64), ("lut_bytes", ctypes.c_uint64), ("image_bytes", ctypes.c_uint64),
                ("pred_segs", ctypes.c_uint64), ("pred_bins", ctypes.c_uint64),
                ("sizing", ctypes.c_uint32), ("fallback", ctypes.c_uint32), ("syncs", ctypes.c_uint32),
                ("render_kernels", ctypes.c_uint32),
                ("render_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("merge_ms", ctypes.c_double), ("upload_ms", ctypes.c_double), ("pack_ms", ctypes.c_double),
                ("schedule", ctypes.c_uint32), ("queue_wgs", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BLAmdTune(ctypes.Structure):
    """Mirror of BLAmdTune (include/blend2d_amd.h)."""
    _fields_ = [("schedule", ctypes.c_uint32), ("wgs_per_cu", ctypes.c_uint32), ("candidates", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("best_ms", ctypes.c_double), ("grid_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BLRgba64(int):
    """A 64-bit colour (BLRgba64, core/rgba.h): 16 bits per channel, ARGB order."""


class BLRgba(tuple):
    """A float colour (BLRgba, core/rgba.h): (r, g, b, a) in [0, 1]."""

    def __new__(cls, r, g, b, a=1.0):
        return super().__new__(cls, (float(r), float(g), float(b), float(a)))


# BLGeometryType values (core/geometry.h) and the struct / element layouts
# fill_geometry packs (same memory layout as the reference's structs).
from .scene import (GEOMETRY_ARRAY as _GEOMETRY_ARRAY, GEOMETRY_STRUCT as _GEOMETRY_STRUCT,  # noqa: E402
                    GEOMETRY_NONE, GEOMETRY_BOXI, GEOMETRY_BOXD, GEOMETRY_RECTI, GEOMETRY_RECTD,
                    GEOMETRY_CIRCLE, GEOMETRY_ELLIPSE, GEOMETRY_ROUND_RECT, GEOMETRY_ARC, GEOMETRY_CHORD,
                    GEOMETRY_PIE, GEOMETRY_LINE, GEOMETRY_TRIANGLE, GEOMETRY_POLYLINEI, GEOMETRY_POLYLINED,
                    GEOMETRY_POLYGONI, GEOMETRY_POLYGOND, GEOMETRY_ARRAY_VIEW_BOXI, GEOMETRY_ARRAY_VIEW_BOXD,
                    GEOMETRY_ARRAY_VIEW_RECTI, GEOMETRY_ARRAY_VIEW_RECTD)


class _ArrayView(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("size", ctypes.c_size_t)]


def _geometry_arg(gtype: int, data):
    """ctypes buffer (kept alive by the caller) holding the geometry struct."""
    if gtype in _GEOMETRY_ARRAY:
        items = list(data)
        body = b"".join(struct.pack(_GEOMETRY_ARRAY[gtype], *it) for it in items)
        buf = ctypes.create_string_buffer(body, max(len(body), 1))
        view = _ArrayView(ctypes.cast(buf, ctypes.c_void_p), len(items))
        return view, (buf, view)
    if gtype in _GEOMETRY_STRUCT:
        buf = ctypes.create_string_buffer(struct.pack(_GEOMETRY_STRUCT[gtype], *data))
        return buf, (buf,)
    return None, ()


class BLAmdStyle(ctypes.Structure):
    """Mirror of BLAmdStyle (include/blend2d_amd.h): the style of one call."""
    _fields_ = [("kind", ctypes.c_uint32), ("rgba32", ctypes.c_uint32), ("rgba64", ctypes.c_uint64),
                ("gradient_type", ctypes.c_uint32), ("extend", ctypes.c_uint32),
                ("values", ctypes.c_void_p), ("stop_offsets", ctypes.c_void_p), ("stop_rgba32", ctypes.c_void_p),
                ("n_stops", ctypes.c_size_t), ("transform", ctypes.c_void_p), ("image", ctypes.c_void_p)]


STYLE_RGBA32, STYLE_RGBA64, STYLE_GRADIENT, STYLE_PATTERN = 0, 1, 2, 3


class BLError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with BLResult 0x{code:X}")
        self.code = code


_lib = None


def lib():
    """Loads libblend2d_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m blend2d_amd.build`")
    _lib = bind(ctypes.CDLL(LIB_PATH))
    return _lib


def bind(L):
    """Sets the argument and result types of every C-ABI entry point on the
    loaded library `L` (include/blend2d_amd.h); returns `L`."""
    P, D, U32, SZ = ctypes.c_void_p, ctypes.c_double, ctypes.c_uint32, ctypes.c_size_t
    I, IP = ctypes.c_int, ctypes.c_ssize_t
    sig = {
        "bl_amd_version": ([], U32),
        "bl_amd_device_count": ([], I),
        "bl_amd_image_create": ([ctypes.POINTER(P), I, I, U32], U32),
        "bl_amd_image_destroy": ([P], U32),
        "bl_amd_image_upload": ([P, P, IP], U32),
        "bl_amd_image_download": ([P, P, IP], U32),
        "bl_amd_image_fill_raw": ([P, U32], U32),
        "bl_amd_image_get_data": ([P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(U32), ctypes.POINTER(P), ctypes.POINTER(IP)], U32),
        "bl_amd_context_create": ([ctypes.POINTER(P)], U32),
        "bl_amd_context_destroy": ([P], U32),
        "bl_amd_context_begin": ([P, P], U32),
        "bl_amd_context_end": ([P], U32),
        "bl_amd_context_flush": ([P, U32], U32),
        "bl_amd_context_set_comp_op": ([P, U32], U32),
        "bl_amd_context_set_fill_rule": ([P, U32], U32),
        "bl_amd_context_set_global_alpha": ([P, D], U32),
        "bl_amd_context_set_transform": ([P, P], U32),
        "bl_amd_context_set_pattern_quality": ([P, U32], U32),
        "bl_amd_context_set_fill_style_rgba32": ([P, U32], U32),
        "bl_amd_context_set_fill_style_gradient": ([P, U32, P, U32, P, P, SZ, P], U32),
        "bl_amd_context_set_fill_style_pattern": ([P, P, U32, P], U32),
        "bl_amd_context_fill_path_d": ([P, P, P, P, SZ], U32),
        "bl_amd_context_fill_polygon_d": ([P, P, SZ], U32),
        "bl_amd_context_fill_rect_d": ([P, D, D, D, D], U32),
        "bl_amd_context_fill_rect_i": ([P, I, I, I, I], U32),
        "bl_amd_context_fill_all": ([P], U32),
        "bl_amd_context_get_stats": ([P, ctypes.POINTER(BLAmdStats)], U32),
        "bl_amd_context_replay": ([P], U32),
        "bl_amd_context_replay_noalloc": ([P], U32),
        "bl_amd_context_get_error_flags": ([P, ctypes.POINTER(U32)], U32),
        "bl_amd_context_reset": ([P], U32),
        "bl_amd_context_fill_all_rgba64": ([P, ctypes.c_uint64], U32),
        "bl_amd_context_fill_rect_i_rgba64": ([P, I, I, I, I, ctypes.c_uint64], U32),
        "bl_amd_context_fill_rect_d_rgba64": ([P, D, D, D, D, ctypes.c_uint64], U32),
        "bl_amd_context_fill_path_d_rgba64": ([P, P, P, P, SZ, ctypes.c_uint64], U32),
        "bl_amd_context_fill_geometry_rgba64": ([P, U32, P, ctypes.c_uint64], U32),
        "bl_amd_context_fill_all_ext": ([P, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_fill_rect_i_ext": ([P, I, I, I, I, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_fill_rect_d_ext": ([P, D, D, D, D, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_fill_path_d_ext": ([P, P, P, P, SZ, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_fill_geometry_ext": ([P, U32, P, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_get_deferred_error": ([P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(U32)], U32),
        "bl_amd_context_fill_all_rgba32": ([P, U32], U32),
        "bl_amd_context_fill_rect_i_rgba32": ([P, I, I, I, I, U32], U32),
        "bl_amd_context_fill_rect_d_rgba32": ([P, D, D, D, D, U32], U32),
        "bl_amd_context_fill_path_d_rgba32": ([P, P, P, P, SZ, U32], U32),
        "bl_amd_context_fill_geometry": ([P, U32, P], U32),
        "bl_amd_context_fill_geometry_rgba32": ([P, U32, P, U32], U32),
        "bl_amd_context_clear_all": ([P], U32),
        "bl_amd_context_clear_rect_i": ([P, I, I, I, I], U32),
        "bl_amd_context_clear_rect_d": ([P, D, D, D, D], U32),
        "bl_amd_context_set_fill_style_rgba64": ([P, ctypes.c_uint64], U32),
        "bl_amd_context_set_fill_style_rgba": ([P, P], U32),
        "bl_amd_context_set_fill_alpha": ([P, D], U32),
        "bl_amd_batch_create": ([ctypes.POINTER(P)], U32),
        "bl_amd_batch_destroy": ([P], U32),
        "bl_amd_batch_add": ([P, P], U32),
        "bl_amd_batch_clear": ([P], U32),
        "bl_amd_batch_flush": ([P], U32),
        "bl_amd_batch_replay": ([P], U32),
        "bl_amd_batch_get_stats": ([P, ctypes.POINTER(BLAmdStats)], U32),
        "bl_amd_set_sizing": ([U32], U32),
        "bl_amd_cost_model_set": ([P], U32),
        "bl_amd_context_collect_costs": ([P, P, P, SZ, ctypes.POINTER(SZ)], U32),
        "bl_amd_cost_model_forward": ([P, SZ, P, P], U32),
        "bl_amd_cost_model_grad": ([P, P, SZ, P, P, ctypes.POINTER(ctypes.c_float)], U32),
        "bl_amd_cost_model_train": ([P, P, SZ, P, U32, U32, ctypes.c_float, P], U32),
        "bl_amd_set_render_schedule": ([U32, U32], U32),
        "bl_amd_tile_model_set": ([P], U32),
        "bl_amd_context_collect_tile_costs": ([P, P, P, SZ, ctypes.POINTER(SZ)], U32),
        "bl_amd_batch_collect_tile_costs": ([P, P, P, SZ, ctypes.POINTER(SZ)], U32),
        "bl_amd_context_autotune": ([P, U32, ctypes.POINTER(BLAmdTune)], U32),
        "bl_amd_batch_autotune": ([P, U32, ctypes.POINTER(BLAmdTune)], U32),
        "bl_amd_autotune_log": ([ctypes.c_char_p, SZ], SZ),
        "bl_amd_autotune_reset": ([], U32),
        "bl_amd_set_graph_replay": ([U32], U32),
        "bl_amd_context_blit_image_i": ([P, I, I, P, P], U32),
        "bl_amd_context_blit_image_d": ([P, D, D, P, P], U32),
        "bl_amd_context_blit_scaled_image_i": ([P, I, I, I, I, P, P], U32),
        "bl_amd_context_blit_scaled_image_d": ([P, D, D, D, D, P, P], U32),
        "bl_amd_context_set_gradient_quality": ([P, U32], U32),
        "bl_amd_context_fill_mask_i": ([P, I, I, P, P], U32),
        "bl_amd_context_fill_mask_d": ([P, D, D, P, P], U32),
        "bl_amd_context_fill_mask_i_rgba32": ([P, I, I, P, P, U32], U32),
        "bl_amd_context_fill_mask_d_rgba32": ([P, D, D, P, P, U32], U32),
        "bl_amd_context_fill_mask_i_rgba64": ([P, I, I, P, P, ctypes.c_uint64], U32),
        "bl_amd_context_fill_mask_d_rgba64": ([P, D, D, P, P, ctypes.c_uint64], U32),
        "bl_amd_context_fill_mask_i_ext": ([P, I, I, P, P, ctypes.POINTER(BLAmdStyle)], U32),
        "bl_amd_context_fill_mask_d_ext": ([P, D, D, P, P, ctypes.POINTER(BLAmdStyle)], U32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def _check(code: int, what: str) -> None:
    if code:
        raise BLError(code, what)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mat(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(6))


class BLImage:
    """Device-resident image (BLImage)."""

    def __init__(self, w: int, h: int, fmt: int = FORMAT_PRGB32):
        self._h = ctypes.c_void_p()
        _check(lib().bl_amd_image_create(ctypes.byref(self._h), w, h, fmt), "bl_amd_image_create")
        self.width, self.height, self.format = w, h, fmt

    @property
    def handle(self):
        return self._h

    @property
    def bpp(self) -> int:
        return 1 if self.format == FORMAT_A8 else 4

    def upload(self, pixels: np.ndarray) -> None:
        pixels = np.ascontiguousarray(pixels)
        _check(lib().bl_amd_image_upload(self._h, _dptr(pixels), self.width * self.bpp), "bl_amd_image_upload")

    def download(self) -> np.ndarray:
        if self.format == FORMAT_A8:
            out = np.empty((self.height, self.width), dtype=np.uint8)
        else:
            out = np.empty((self.height, self.width), dtype=np.uint32)
        _check(lib().bl_amd_image_download(self._h, _dptr(out), self.width * self.bpp), "bl_amd_image_download")
        return out

    def fill_raw(self, value: int) -> None:
        _check(lib().bl_amd_image_fill_raw(self._h, value & 0xFFFFFFFF), "bl_amd_image_fill_raw")

    def close(self) -> None:
        if self._h:
            lib().bl_amd_image_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BLPath:
    """Command/vertex arrays of a path (BLPath, core/path.h)."""

    def __init__(self):
        self.cmds: list[int] = []
        self.vtx: list[tuple[float, float]] = []

    def move_to(self, x, y):
        self.cmds.append(CMD_MOVE); self.vtx.append((x, y)); return self

    def line_to(self, x, y):
        self.cmds.append(CMD_ON); self.vtx.append((x, y)); return self

    def quad_to(self, x1, y1, x2, y2):
        self.cmds += [CMD_QUAD, CMD_ON]; self.vtx += [(x1, y1), (x2, y2)]; return self

    def cubic_to(self, x1, y1, x2, y2, x3, y3):
        self.cmds += [CMD_CUBIC, CMD_CUBIC, CMD_ON]; self.vtx += [(x1, y1), (x2, y2), (x3, y3)]; return self

    def close(self):
        self.cmds.append(CMD_CLOSE); self.vtx.append((float("nan"), float("nan"))); return self

    def arrays(self):
        return np.asarray(self.cmds, dtype=np.uint8), np.asarray(self.vtx, dtype=np.float64).reshape(-1, 2)


class BLGradient:
    """Gradient style (BLGradient): type, values, extend mode, stops, transform."""

    def __init__(self, gtype: int, values, extend: int = EXTEND_PAD, stops=(), transform=None):
        self.type = gtype
        self.values = list(values) + [0.0] * (6 - len(values))
        self.extend = extend
        self.stops = list(stops)
        self.transform = transform

    def add_stop(self, offset: float, rgba32: int):
        self.stops.append((offset, rgba32))
        return self


class BLPattern:
    """Pattern style (BLPattern): image, extend mode, transform."""

    def __init__(self, image: BLImage, extend: int = EXTEND_REPEAT, transform=None):
        self.image = image
        self.extend = extend
        self.transform = transform


class BLContext:
    """Rendering context (BLContext) whose fills run on the MI355X."""

    def __init__(self, image: BLImage | None = None):
        self._h = ctypes.c_void_p()
        _check(lib().bl_amd_context_create(ctypes.byref(self._h)), "bl_amd_context_create")
        self.target = None
        self._keep = []
        if image is not None:
            self.begin(image)

    @property
    def handle(self):
        return self._h

    def begin(self, image: BLImage):
        _check(lib().bl_amd_context_begin(self._h, image.handle), "bl_amd_context_begin")
        self.target = image

    def end(self):
        r = lib().bl_amd_context_end(self._h)
        self.target = None
        self._keep.clear()
        _check(r, "bl_amd_context_end")

    def flush(self, flags: int = FLUSH_SYNC):
        r = lib().bl_amd_context_flush(self._h, flags)
        self._keep.clear()
        _check(r, "bl_amd_context_flush")

    def set_comp_op(self, op: int):
        _check(lib().bl_amd_context_set_comp_op(self._h, op), "set_comp_op")

    def set_fill_rule(self, rule: int):
        _check(lib().bl_amd_context_set_fill_rule(self._h, rule), "set_fill_rule")

    def set_global_alpha(self, a: float):
        _check(lib().bl_amd_context_set_global_alpha(self._h, a), "set_global_alpha")

    def set_transform(self, m):
        m = _mat(m)
        _check(lib().bl_amd_context_set_transform(self._h, _dptr(m)), "set_transform")

    def set_pattern_quality(self, q: int):
        _check(lib().bl_amd_context_set_pattern_quality(self._h, q), "set_pattern_quality")

    def set_gradient_quality(self, q: int):
        _check(lib().bl_amd_context_set_gradient_quality(self._h, q), "set_gradient_quality")

    def set_fill_style(self, style):
        L = lib()
        if isinstance(style, BLRgba64):
            return _check(L.bl_amd_context_set_fill_style_rgba64(self._h, int(style) & 0xFFFFFFFFFFFFFFFF), "set_fill_style")
        if isinstance(style, BLRgba):
            arr = (ctypes.c_float * 4)(*style)
            return _check(L.bl_amd_context_set_fill_style_rgba(self._h, arr), "set_fill_style")
        if isinstance(style, int):
            return _check(L.bl_amd_context_set_fill_style_rgba32(self._h, style & 0xFFFFFFFF), "set_fill_style")
        if isinstance(style, BLGradient):
            vals = np.ascontiguousarray(style.values, dtype=np.float64)
            offs = np.ascontiguousarray([s[0] for s in style.stops], dtype=np.float64)
            cols = np.ascontiguousarray([s[1] & 0xFFFFFFFF for s in style.stops], dtype=np.uint32)
            tr = _mat(style.transform) if style.transform is not None else None
            return _check(L.bl_amd_context_set_fill_style_gradient(
                self._h, style.type, _dptr(vals), style.extend, _dptr(offs), _dptr(cols), len(style.stops),
                _dptr(tr) if tr is not None else None), "set_fill_style")
        if isinstance(style, BLPattern):
            tr = _mat(style.transform) if style.transform is not None else None
            self._keep.append(style.image)
            return _check(L.bl_amd_context_set_fill_style_pattern(
                self._h, style.image.handle, style.extend, _dptr(tr) if tr is not None else None), "set_fill_style")
        raise TypeError("unsupported style")

    def set_fill_alpha(self, a: float):
        _check(lib().bl_amd_context_set_fill_alpha(self._h, a), "set_fill_alpha")

    def clear_all(self):
        _check(lib().bl_amd_context_clear_all(self._h), "clear_all")

    def blit_image(self, dst, image: "BLImage", area=None, integer: bool = False):
        """BLContext::blit_image (core/context.h:3860-3900): `dst` is an origin
        (x, y) or a rectangle (x, y, w, h) the area is scaled onto; `area`
        (x, y, w, h) selects part of the image; `integer` takes the _i form."""
        self._keep.append(image)
        a = None
        if area is not None:
            a = (ctypes.c_int32 * 4)(*[int(v) for v in area])
        L = lib()
        if len(dst) == 2:
            if integer:
                r = L.bl_amd_context_blit_image_i(self._h, int(dst[0]), int(dst[1]), image.handle, a)
            else:
                r = L.bl_amd_context_blit_image_d(self._h, float(dst[0]), float(dst[1]), image.handle, a)
        elif integer:
            r = L.bl_amd_context_blit_scaled_image_i(self._h, *[int(v) for v in dst], image.handle, a)
        else:
            r = L.bl_amd_context_blit_scaled_image_d(self._h, *[float(v) for v in dst], image.handle, a)
        _check(r, "blit_image")

    def fill_mask(self, dst, mask: "BLImage", area=None, integer: bool = False, rgba32: int | None = None, style=None):
        """BLContext::fill_mask (core/context.h:800-808): the fill style (or
        `rgba32` / `style` for this call) through the bytes of `mask`'s
        `area` (x, y, w, h; None: all of it) placed at dst = (x, y);
        `integer` takes the _i form."""
        self._keep.append(mask)
        a = None
        if area is not None:
            a = (ctypes.c_int32 * 4)(*[int(v) for v in area])
        L = lib()
        head = ((self._h, int(dst[0]), int(dst[1])) if integer else (self._h, float(dst[0]), float(dst[1])))
        sfx = "i" if integer else "d"
        if style is not None:
            kind, v, skeep = self._style_arg(style)
            if kind == "rgba64":
                r = getattr(L, f"bl_amd_context_fill_mask_{sfx}_rgba64")(*head, mask.handle, a, v)
            else:
                r = getattr(L, f"bl_amd_context_fill_mask_{sfx}_ext")(*head, mask.handle, a, ctypes.byref(v))
            del skeep
        elif rgba32 is not None:
            r = getattr(L, f"bl_amd_context_fill_mask_{sfx}_rgba32")(*head, mask.handle, a, rgba32 & 0xFFFFFFFF)
        else:
            r = getattr(L, f"bl_amd_context_fill_mask_{sfx}")(*head, mask.handle, a)
        _check(r, "fill_mask")

    def clear_rect(self, x, y, w, h):
        if all(isinstance(v, (int, np.integer)) for v in (x, y, w, h)):
            _check(lib().bl_amd_context_clear_rect_i(self._h, int(x), int(y), int(w), int(h)), "clear_rect")
        else:
            _check(lib().bl_amd_context_clear_rect_d(self._h, float(x), float(y), float(w), float(h)), "clear_rect")

    @staticmethod
    def _style_arg(style):
        """("rgba64", value) for a BLRgba64, else a BLAmdStyle and the
        buffers it points into (kept alive by the caller)."""
        st = BLAmdStyle()
        keep = []
        if isinstance(style, BLRgba64):
            return "rgba64", int(style) & 0xFFFFFFFFFFFFFFFF, keep
        if isinstance(style, int):
            st.kind, st.rgba32 = STYLE_RGBA32, style & 0xFFFFFFFF
        elif isinstance(style, BLGradient):
            vals = np.ascontiguousarray(style.values, dtype=np.float64)
            offs = np.ascontiguousarray([x[0] for x in style.stops] or [0.0], dtype=np.float64)
            cols = np.ascontiguousarray([x[1] & 0xFFFFFFFF for x in style.stops] or [0], dtype=np.uint32)
            tr = _mat(style.transform) if style.transform is not None else None
            keep += [vals, offs, cols, tr]
            st.kind, st.gradient_type, st.extend = STYLE_GRADIENT, style.type, style.extend
            st.values, st.stop_offsets, st.stop_rgba32 = vals.ctypes.data, offs.ctypes.data, cols.ctypes.data
            st.n_stops = len(style.stops)
            st.transform = tr.ctypes.data if tr is not None else None
        elif isinstance(style, BLPattern):
            tr = _mat(style.transform) if style.transform is not None else None
            keep.append(tr)
            st.kind, st.extend, st.image = STYLE_PATTERN, style.extend, style.image.handle.value
            st.transform = tr.ctypes.data if tr is not None else None
        else:
            raise TypeError("unsupported style")
        return "ext", st, keep

    def fill_geometry(self, gtype: int, data=(), rgba32: int | None = None, style=None):
        arg, keep = _geometry_arg(gtype, data)
        ptr = ctypes.byref(arg) if arg is not None else None
        L = lib()
        if style is not None:
            kind, v, skeep = self._style_arg(style)
            r = (L.bl_amd_context_fill_geometry_rgba64(self._h, gtype, ptr, v) if kind == "rgba64"
                 else L.bl_amd_context_fill_geometry_ext(self._h, gtype, ptr, ctypes.byref(v)))
            del skeep
        elif rgba32 is None:
            r = L.bl_amd_context_fill_geometry(self._h, gtype, ptr)
        else:
            r = L.bl_amd_context_fill_geometry_rgba32(self._h, gtype, ptr, rgba32 & 0xFFFFFFFF)
        del keep
        _check(r, "fill_geometry")

    def reset(self):
        """bl_context_reset: renders pending calls, detaches, default state."""
        r = lib().bl_amd_context_reset(self._h)
        self.target = None
        self._keep.clear()
        _check(r, "reset")

    def fill_path(self, path, origin=None, rgba32: int | None = None, style=None):
        if isinstance(path, BLPath):
            cmds, vtx = path.arrays()
        else:
            cmds, vtx = path
            cmds = np.ascontiguousarray(cmds, dtype=np.uint8)
            vtx = np.ascontiguousarray(vtx, dtype=np.float64).reshape(-1, 2)
        org = np.ascontiguousarray(origin, dtype=np.float64) if origin is not None else None
        if style is not None:
            kind, v, skeep = self._style_arg(style)
            a = (self._h, _dptr(org) if org is not None else None, _dptr(cmds), _dptr(vtx), len(cmds))
            r = (lib().bl_amd_context_fill_path_d_rgba64(*a, v) if kind == "rgba64"
                 else lib().bl_amd_context_fill_path_d_ext(*a, ctypes.byref(v)))
            del skeep
        elif rgba32 is None:
            r = lib().bl_amd_context_fill_path_d(self._h, _dptr(org) if org is not None else None,
                                                 _dptr(cmds), _dptr(vtx), len(cmds))
        else:
            r = lib().bl_amd_context_fill_path_d_rgba32(self._h, _dptr(org) if org is not None else None,
                                                        _dptr(cmds), _dptr(vtx), len(cmds), rgba32 & 0xFFFFFFFF)
        _check(r, "fill_path")

    def fill_polygon(self, vtx):
        vtx = np.ascontiguousarray(vtx, dtype=np.float64).reshape(-1, 2)
        _check(lib().bl_amd_context_fill_polygon_d(self._h, _dptr(vtx), len(vtx)), "fill_polygon")

    def fill_rect(self, x, y, w, h, rgba32: int | None = None, style=None):
        L = lib()
        if style is not None:
            kind, v, skeep = self._style_arg(style)
            if all(isinstance(q, (int, np.integer)) for q in (x, y, w, h)):
                a = (self._h, int(x), int(y), int(w), int(h))
                r = L.bl_amd_context_fill_rect_i_rgba64(*a, v) if kind == "rgba64" else L.bl_amd_context_fill_rect_i_ext(*a, ctypes.byref(v))
            else:
                a = (self._h, float(x), float(y), float(w), float(h))
                r = L.bl_amd_context_fill_rect_d_rgba64(*a, v) if kind == "rgba64" else L.bl_amd_context_fill_rect_d_ext(*a, ctypes.byref(v))
            del skeep
            return _check(r, "fill_rect")
        if all(isinstance(v, (int, np.integer)) for v in (x, y, w, h)):
            a = (self._h, int(x), int(y), int(w), int(h))
            r = L.bl_amd_context_fill_rect_i(*a) if rgba32 is None else L.bl_amd_context_fill_rect_i_rgba32(*a, rgba32 & 0xFFFFFFFF)
        else:
            a = (self._h, float(x), float(y), float(w), float(h))
            r = L.bl_amd_context_fill_rect_d(*a) if rgba32 is None else L.bl_amd_context_fill_rect_d_rgba32(*a, rgba32 & 0xFFFFFFFF)
        _check(r, "fill_rect")

    def fill_all(self, rgba32: int | None = None, style=None):
        if style is not None:
            kind, v, skeep = self._style_arg(style)
            r = (lib().bl_amd_context_fill_all_rgba64(self._h, v) if kind == "rgba64"
                 else lib().bl_amd_context_fill_all_ext(self._h, ctypes.byref(v)))
            del skeep
            return _check(r, "fill_all")
        if rgba32 is None:
            _check(lib().bl_amd_context_fill_all(self._h), "fill_all")
        else:
            _check(lib().bl_amd_context_fill_all_rgba32(self._h, rgba32 & 0xFFFFFFFF), "fill_all")

    def error_flags(self) -> int:
        """BLContextErrorFlags accumulated since begin() (accumulated_error_flags)."""
        v = ctypes.c_uint32()
        _check(lib().bl_amd_context_get_error_flags(self._h, ctypes.byref(v)), "get_error_flags")
        return v.value

    def deferred_error(self):
        """(call index, BLResult) of the first rendering call of the last
        flush that failed on the device, or None."""
        idx, code = ctypes.c_uint64(), ctypes.c_uint32()
        _check(lib().bl_amd_context_get_deferred_error(self._h, ctypes.byref(idx), ctypes.byref(code)), "get_deferred_error")
        return (idx.value, code.value) if code.value else None

    def last_stats(self) -> dict:
        """Counters of the last flush or replay (BLAmdStats)."""
        st = BLAmdStats()
        _check(lib().bl_amd_context_get_stats(self._h, ctypes.byref(st)), "get_stats")
        return st.as_dict()

    def replay(self):
        """Executes the last flushed frame again from its device-resident copy."""
        _check(lib().bl_amd_context_replay(self._h), "replay")

    def replay_noalloc(self):
        """replay() with exact sizing and every intermediate buffer in the
        library's fixed per-device scratch region (no device allocation)."""
        _check(lib().bl_amd_context_replay_noalloc(self._h), "replay_noalloc")

    def collect_costs(self):
        """Features and exact-count targets of the resident frame's fill
        commands (re-executes the frame with exact counting)."""
        n = ctypes.c_size_t()
        _check(lib().bl_amd_context_collect_costs(self._h, None, None, 0, ctypes.byref(n)), "collect_costs")
        X = np.zeros((n.value, COST_FEATURES), dtype=np.float32)
        T = np.zeros((n.value, COST_OUTPUTS), dtype=np.float32)
        if n.value:
            _check(lib().bl_amd_context_collect_costs(self._h, _dptr(X), _dptr(T), n.value, ctypes.byref(n)), "collect_costs")
        return X, T

    def close(self):
        if self._h:
            lib().bl_amd_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BLBatch:
    """Renders the pending calls of many contexts with one pass per kernel stage."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(lib().bl_amd_batch_create(ctypes.byref(self._h)), "bl_amd_batch_create")

    def add(self, ctx: BLContext):
        _check(lib().bl_amd_batch_add(self._h, ctx.handle), "bl_amd_batch_add")

    def clear(self):
        _check(lib().bl_amd_batch_clear(self._h), "bl_amd_batch_clear")

    def flush(self):
        _check(lib().bl_amd_batch_flush(self._h), "bl_amd_batch_flush")

    def replay(self):
        _check(lib().bl_amd_batch_replay(self._h), "bl_amd_batch_replay")

    def last_stats(self) -> dict:
        st = BLAmdStats()
        _check(lib().bl_amd_batch_get_stats(self._h, ctypes.byref(st)), "bl_amd_batch_get_stats")
        return st.as_dict()

    def close(self):
        if self._h:
            lib().bl_amd_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .replay import replay_into, render_scene, render_scenes_batched  # noqa: E402,F401


# ---------------------------------------------------------------------------
# Learned sizing (cost model)
# ---------------------------------------------------------------------------

def set_sizing(mode: int) -> None:
    """Process-wide sizing mode: SIZING_EXACT, SIZING_LEARNED or SIZING_HISTORY."""
    _check(lib().bl_amd_set_sizing(mode), "set_sizing")


def cost_model_set(params) -> None:
    """Installs cost-model parameters (COST_PARAMS float32) or removes them (None)."""
    if params is None:
        _check(lib().bl_amd_cost_model_set(None), "cost_model_set")
        return
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(COST_PARAMS)
    _check(lib().bl_amd_cost_model_set(_dptr(p)), "cost_model_set")


def cost_model_forward(X, params) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, COST_FEATURES)
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(COST_PARAMS)
    Y = np.zeros((len(X), COST_OUTPUTS), dtype=np.float32)
    _check(lib().bl_amd_cost_model_forward(_dptr(X), len(X), _dptr(p), _dptr(Y)), "cost_model_forward")
    return Y


def cost_model_grad(X, T, params):
    """(gradient of the mean squared error, loss) computed on the GPU."""
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, COST_FEATURES)
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, COST_OUTPUTS)
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(COST_PARAMS)
    g = np.zeros(COST_PARAMS, dtype=np.float32)
    loss = ctypes.c_float()
    _check(lib().bl_amd_cost_model_grad(_dptr(X), _dptr(T), len(X), _dptr(p), _dptr(g), ctypes.byref(loss)), "cost_model_grad")
    return g, loss.value


def cost_model_train(X, T, params, epochs: int = 10, batch: int = 4096, lr: float = 1e-3):
    """Adam training on the GPU; returns (trained params, loss per epoch)."""
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, COST_FEATURES)
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, COST_OUTPUTS)
    p = np.array(params, dtype=np.float32).reshape(COST_PARAMS)
    losses = np.zeros(epochs, dtype=np.float32)
    _check(lib().bl_amd_cost_model_train(_dptr(X), _dptr(T), len(X), _dptr(p), epochs, batch, lr, _dptr(losses)),
           "cost_model_train")
    return p, losses


# ---------------------------------------------------------------------------
# Tile scheduling, tile cost model and autotuner
# ---------------------------------------------------------------------------

def set_graph_replay(on: bool) -> None:
    """Process-wide: replays run as captured HIP graphs."""
    _check(lib().bl_amd_set_graph_replay(1 if on else 0), "set_graph_replay")


def set_render_schedule(schedule: int, wgs_per_cu: int = 0) -> None:
    """Process-wide schedule of the analytic tiles (SCHEDULE_*)."""
    _check(lib().bl_amd_set_render_schedule(schedule, wgs_per_cu), "set_render_schedule")


def tile_model_set(params) -> None:
    """Installs the tile cost model (COST_PARAMS float32) or removes it (None)."""
    if params is None:
        _check(lib().bl_amd_tile_model_set(None), "tile_model_set")
        return
    p = np.ascontiguousarray(params, dtype=np.float32).reshape(COST_PARAMS)
    _check(lib().bl_amd_tile_model_set(_dptr(p)), "tile_model_set")


def collect_tile_costs(target, cap: int = 1 << 22):
    """(features, targets) of every analytic tile of the resident frame of a
    BLContext or BLBatch (re-executed with tile timing)."""
    X = np.zeros((cap, TILE_FEATURES), dtype=np.float32)
    T = np.zeros((cap, TILE_OUTPUTS), dtype=np.float32)
    n = ctypes.c_size_t()
    fn = lib().bl_amd_batch_collect_tile_costs if isinstance(target, BLBatch) else lib().bl_amd_context_collect_tile_costs
    h = target._h if isinstance(target, BLBatch) else target.handle
    _check(fn(h, _dptr(X), _dptr(T), cap, ctypes.byref(n)), "collect_tile_costs")
    k = min(n.value, cap)
    return X[:k].copy(), T[:k].copy()


def autotune(target, reps: int = 3) -> dict:
    """Runs the autotuner on the resident frame of a BLContext or BLBatch."""
    out = BLAmdTune()
    fn = lib().bl_amd_batch_autotune if isinstance(target, BLBatch) else lib().bl_amd_context_autotune
    h = target._h if isinstance(target, BLBatch) else target.handle
    _check(fn(h, reps, ctypes.byref(out)), "autotune")
    return out.as_dict()


def autotune_log() -> str:
    n = lib().bl_amd_autotune_log(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().bl_amd_autotune_log(buf, n + 1)
    return buf.value.decode()


def autotune_reset() -> None:
    _check(lib().bl_amd_autotune_reset(), "autotune_reset")
