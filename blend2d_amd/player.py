"""Native scene player (libb2amd_player.so, csrc/scene_player.cpp).

Plays recorded scenes (blend2d_amd.scene) through the C ABI from native code,
so a frame of thousands of canvases is issued without a Python call per fill,
and times it the way the reference's benchmark times a backend: wall clock
from the first begin() to the last end() of a synchronous flush
(blend2d-testing/bench/bl_bench_backend.cpp:50-89), every call object built
beforehand.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import BLAmdStats as BLStats, BLAmdTune, BLError, lib as _lib

HERE = os.path.dirname(os.path.abspath(__file__))
PLAYER_SO = os.path.join(HERE, "libb2amd_player.so")

PLAIN, BATCHED = 0, 1

_player = None


def player():
    global _player
    if _player is None:
        _lib()   # the rendering library first: the player binds to that same copy
        if not os.path.exists(PLAYER_SO):
            raise RuntimeError(f"{PLAYER_SO} is not built (python -m blend2d_amd.build)")
        _player = bind(ctypes.CDLL(PLAYER_SO))
    return _player


def bind(p):
    """Sets the argument and result types of the player's entry points on the
    loaded library `p`; returns `p`."""
    P, U32, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t
    p.b2p_frame_create.restype = U32
    p.b2p_frame_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(SZ), SZ]
    p.b2p_frame_destroy.restype = None
    p.b2p_frame_destroy.argtypes = [P]
    p.b2p_frame_clear.restype = U32
    p.b2p_frame_clear.argtypes = [P]
    p.b2p_frame_render.restype = U32
    p.b2p_frame_render.argtypes = [P, U32, U32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), P,
                                   ctypes.POINTER(BLStats)]
    p.b2p_frame_replay.restype = U32
    p.b2p_frame_replay.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(BLStats)]
    p.b2p_frame_download.restype = U32
    p.b2p_frame_download.argtypes = [P, SZ, P, SZ]
    p.b2p_frame_autotune.restype = U32
    p.b2p_frame_autotune.argtypes = [P, U32, ctypes.POINTER(BLAmdTune)]
    return p


class Frame:
    """Scenes with their canvases and source images on the device."""

    def __init__(self, scenes):
        self.scenes = list(scenes)
        blobs = [sc.to_bytes() for sc in self.scenes]
        n = len(blobs)
        arr = (ctypes.c_char_p * n)(*blobs)
        sizes = (ctypes.c_size_t * n)(*[len(b) for b in blobs])
        h = ctypes.c_void_p()
        r = player().b2p_frame_create(ctypes.byref(h), arr, sizes, n)
        if r:
            raise BLError(r, "b2p_frame_create")
        self._h = h

    def close(self):
        if self._h:
            player().b2p_frame_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps: int = 3) -> dict:
        """The autotuner on the batched frame resident on the device (after a
        BATCHED render); its replays render onto the canvases again."""
        out = BLAmdTune()
        r = player().b2p_frame_autotune(self._h, reps, ctypes.byref(out))
        if r:
            raise BLError(r, "b2p_frame_autotune")
        return out.as_dict()

    def clear(self):
        r = player().b2p_frame_clear(self._h)
        if r:
            raise BLError(r, "b2p_frame_clear")

    def render(self, mode: int, threads: int = 1):
        """Renders every scene onto its canvas (PLAIN: canvas by canvas
        through one context; BATCHED: one batch flush, the calls recorded on
        `threads` client threads).  Returns (elapsed ms from the first begin
        to the last end, per-canvas first errors, stats dict, result of the
        calls); the time until every call was issued is kept in
        `self.last_issue_ms`."""
        n = len(self.scenes)
        errors = np.zeros(n, dtype=np.uint32)
        ms = ctypes.c_double()
        issue = ctypes.c_double()
        st = BLStats()
        r = player().b2p_frame_render(self._h, mode, threads, ctypes.byref(ms), ctypes.byref(issue),
                                      errors.ctypes.data, ctypes.byref(st))
        self.last_issue_ms = issue.value
        return ms.value, [int(e) for e in errors], st.as_dict(), int(r)

    def replay(self):
        """Renders the batch's last frame again from its device-resident copy;
        returns (elapsed ms, stats dict)."""
        ms = ctypes.c_double()
        st = BLStats()
        r = player().b2p_frame_replay(self._h, ctypes.byref(ms), ctypes.byref(st))
        if r:
            raise BLError(r, "b2p_frame_replay")
        return ms.value, st.as_dict()

    def download(self, k: int) -> np.ndarray:
        sc = self.scenes[k]
        if sc.format == 3:
            out = np.empty((sc.height, sc.width), dtype=np.uint8)
        else:
            out = np.empty((sc.height, sc.width), dtype=np.uint32)
        r = player().b2p_frame_download(self._h, k, out.ctypes.data, out.nbytes)
        if r:
            raise BLError(r, "b2p_frame_download")
        return out
