"""ctypes loaders for the checkers: the CPU oracle (oracle/_build/liborc.so)
and the reference driver (oracle/_ref/libref_render.so).  Test-side only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_SO = os.path.join(ROOT, "oracle", "_build", "liborc.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_render.so")

_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        lib = ctypes.CDLL(ORC_SO)
        lib.orc_render.restype = ctypes.c_uint32
        lib.orc_render.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        lib.orc_path_segments.restype = ctypes.c_size_t
        lib.orc_path_segments.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t]
        lib.orc_gradient_lut.restype = None
        lib.orc_gradient_lut.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_composite_pixel.restype = ctypes.c_uint32
        lib.orc_composite_pixel.argtypes = [ctypes.c_uint32] * 5
        lib.orc_composite_array.restype = None
        lib.orc_composite_array.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        lib.orc_simplify_op.restype = ctypes.c_uint32
        lib.orc_simplify_op.argtypes = [ctypes.c_uint32] * 3
        lib.orc_div255.restype = ctypes.c_uint32
        lib.orc_div255.argtypes = [ctypes.c_uint32]
        lib.orc_premultiply.restype = ctypes.c_uint32
        lib.orc_premultiply.argtypes = [ctypes.c_uint32]
        _orc = lib
    return _orc


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_render.restype = ctypes.c_uint32
        lib.ref_render.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
        _ref = lib
    return _ref


def _canvas(scene):
    bpp = 1 if scene.format == 3 else 4
    return np.zeros((scene.height, scene.width * bpp), dtype=np.uint8)


def as_pixels(scene, buf):
    if scene.format == 3:
        return buf.reshape(scene.height, scene.width)
    return buf.view(np.uint32).reshape(scene.height, scene.width)


def render_oracle(scene):
    data = scene.to_bytes()
    out = _canvas(scene)
    err = oracle().orc_render(data, len(data), out.ctypes.data, out.nbytes)
    return as_pixels(scene, out), err


def render_reference(scene, threads: int = 0):
    data = scene.to_bytes()
    out = _canvas(scene)
    ms = ctypes.c_double(0.0)
    err = reference().ref_render(data, len(data), out.ctypes.data, out.nbytes, threads, ctypes.byref(ms))
    return as_pixels(scene, out), err, ms.value


def diff_info(a, b):
    """Max and mean per-channel absolute difference (blend2d-testing/commons/imagediff.h)."""
    if a.dtype == np.uint32:
        a8 = a.view(np.uint8).astype(np.int16)
        b8 = b.view(np.uint8).astype(np.int16)
    else:
        a8 = a.astype(np.int16)
        b8 = b.astype(np.int16)
    d = np.abs(a8 - b8)
    return int(d.max()) if d.size else 0, float(d.mean()) if d.size else 0.0, int((d != 0).sum())
