"""The native scene player (libb2amd_player.so), which the bench drives
frames through: every golden case played in the batched and in the plain
form must give the reference's pixels and error codes, as the Python replay
does."""
import ctypes
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))
CASES = cases.cases()

PLAYER_SYMBOLS = ["b2p_frame_create", "b2p_frame_destroy", "b2p_frame_clear", "b2p_frame_render",
                  "b2p_frame_replay", "b2p_frame_download", "b2p_frame_autotune"]


def test_player_library_exports_its_entry_points():
    from blend2d_amd import player
    if not os.path.exists(player.PLAYER_SO):
        pytest.skip("player not built")
    lib = ctypes.CDLL(player.PLAYER_SO)
    for name in PLAYER_SYMBOLS:
        assert hasattr(lib, name), name


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_player_matches_golden(amd, name, mode):
    from blend2d_amd import player
    scenes = CASES[name]()
    frame = player.Frame(scenes)
    try:
        frame.clear()
        _, errors, _, r = frame.render(player.BATCHED if mode == "batched" else player.PLAIN)
        assert r == 0
        for k, want in enumerate(GOLDEN[name]):
            got = cases.digest(frame.download(k))
            assert errors[k] == want["error"], (name, k)
            assert got["sha256"] == want["sha256"], (name, k)
    finally:
        frame.close()


@pytest.mark.gpu
def test_player_frames_repeat_and_replay(amd):
    """A frame rendered again onto cleared canvases, and replayed from its
    device-resident copy onto cleared canvases, gives the same pixels."""
    from blend2d_amd import player
    from blend2d_amd import scene as S
    scenes = S.glyph_canvases(11, 40, 25)
    frame = player.Frame(scenes)
    try:
        frame.clear()
        _, e1, st1, r = frame.render(player.BATCHED)
        assert r == 0
        first = [frame.download(k) for k in range(len(scenes))]
        frame.clear()
        _, e2, st2, r = frame.render(player.BATCHED)
        assert r == 0 and e1 == e2 and st1["pixels"] == st2["pixels"] > 0
        for k in range(len(scenes)):
            assert (frame.download(k) == first[k]).all()
        frame.clear()
        frame.replay()
        for k in range(len(scenes)):
            assert (frame.download(k) == first[k]).all()
    finally:
        frame.close()


@pytest.mark.gpu
def test_player_uses_the_sizing_set_from_python(amd):
    """The player binds to the library copy Python loaded: a sizing mode set
    through the Python mirror is the one its flushes use."""
    from blend2d_amd import player
    from blend2d_amd import scene as S
    scenes = S.glyph_canvases(12, 8, 10)
    frame = player.Frame(scenes)
    try:
        amd.set_sizing(amd.SIZING_HISTORY)
        frame.clear()
        _, _, st, r = frame.render(player.BATCHED)
        assert r == 0 and st["sizing"] == amd.SIZING_HISTORY
        amd.set_sizing(amd.SIZING_EXACT)
        _, _, st, r = frame.render(player.BATCHED)
        assert r == 0 and st["sizing"] == amd.SIZING_EXACT
    finally:
        amd.set_sizing(amd.SIZING_EXACT)
        frame.close()
