"""The scene player's decoding, on the CPU: the native player
(blend2d_amd/csrc/scene_player.cpp) and the Python replay
(blend2d_amd/replay.py) drive the same scene through a recording stand-in
of the C ABI (tests/native/abi_stub.cpp), and must make the same calls with
the same arguments (pointed-to data hashed), for every golden case and for
generated scenes covering every scene op.  A player that mistakes a geometry
type's data layout (a struct for an array view, or the reverse) must fail."""
import ctypes
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUB = os.path.join(ROOT, "tests", "native", "abi_stub.cpp")
PLAYER = os.path.join(ROOT, "blend2d_amd", "csrc", "scene_player.cpp")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import cases  # noqa: E402

from blend2d_amd import scene as S  # noqa: E402


def _build(tmp_path, player_text=None):
    if not shutil.which("g++"):
        pytest.skip("g++ is missing")
    (tmp_path / "tests" / "native").mkdir(parents=True)
    (tmp_path / "blend2d_amd" / "csrc").mkdir(parents=True)
    (tmp_path / "include").mkdir(parents=True)
    shutil.copy(STUB, tmp_path / "tests" / "native" / "abi_stub.cpp")
    shutil.copy(os.path.join(ROOT, "include", "blend2d_amd.h"), tmp_path / "include" / "blend2d_amd.h")
    shutil.copy(os.path.join(ROOT, "blend2d_amd", "csrc", "scene_format.h"), tmp_path / "blend2d_amd" / "csrc")
    (tmp_path / "blend2d_amd" / "csrc" / "scene_player.cpp").write_text(
        player_text if player_text is not None else open(PLAYER).read())
    so = tmp_path / "abi_stub.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-o", str(so),
                    str(tmp_path / "tests" / "native" / "abi_stub.cpp")], check=True)
    stub = ctypes.CDLL(str(so))
    stub.stub_log.restype = ctypes.c_char_p
    stub.stub_reset.restype = None
    return stub


def _window(log: str):
    """The lines from the rendering context's begin to its first flush."""
    lines = log.splitlines()
    start = max(i for i, l in enumerate(lines) if l.startswith("begin"))
    end = next(i for i in range(start, len(lines)) if lines[i] == "flush")
    return lines[start:end + 1]


def _logs(stub, sc):
    import blend2d_amd as B
    from blend2d_amd import player, replay
    saved = (B._lib, player._player)
    try:
        B._lib = B.bind(stub)
        player._player = player.bind(stub)
        stub.stub_reset()
        images = replay._images(None, sc)
        img = B.BLImage(sc.width, sc.height, sc.format)
        ctx = B.BLContext(img)
        replay.replay_into(ctx, sc, images)
        ctx.flush()
        py = _window(stub.stub_log().decode())
        # the handles are the stub's: release them while it is bound
        ctx.close()
        for im in images + [img]:
            im.close()
        stub.stub_reset()
        frame = player.Frame([sc])
        frame.render(player.PLAIN)
        native = _window(stub.stub_log().decode())
        frame.close()
    finally:
        B._lib, player._player = saved
    return py, native


def _scenes():
    out = {}
    for name, make in cases.cases().items():
        for k, sc in enumerate(make()[:3]):
            out[f"{name}[{k}]"] = sc
    out["frontend_prgb32_s9"] = S.frontend_scene(9, 200, 150, S.FMT_PRGB32, 300)
    out["frontend_a8_s9"] = S.frontend_scene(9, 200, 150, S.FMT_A8, 300)
    out["explicit_styles_s9"] = S.explicit_styles_scene(9, 160, 120, S.FMT_PRGB32)
    return out


SCENES = _scenes()


@pytest.fixture(scope="module")
def stub(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("stub"))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_player_makes_the_calls_the_python_replay_makes(stub, name):
    py, native = _logs(stub, SCENES[name])
    assert len(py) > 2
    for k, (a, b) in enumerate(zip(py, native)):
        assert a == b, f"call {k}: python {a!r} player {b!r}"
    assert len(py) == len(native)


def test_every_scene_op_is_covered():
    ops = set()
    for sc in SCENES.values():
        ops.update(op for op, _ in sc.cmds)
    names = {v for k, v in vars(S).items() if k.startswith("OP_")}
    assert names <= ops, sorted(names - ops)


def test_a_mistaken_geometry_layout_is_caught(tmp_path):
    text = open(PLAYER).read()
    # a struct type taken for an array view (the stand-in then hashes the
    # view record instead of the triangle; the opposite mistake makes the
    # callee dereference vertex data as a pointer)
    wrong = text.replace("return type >= BL_AMD_GEOMETRY_POLYLINEI && type <= BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD;",
                         "return type >= BL_AMD_GEOMETRY_TRIANGLE && type <= BL_AMD_GEOMETRY_ARRAY_VIEW_RECTD;")
    assert wrong != text
    stub = _build(tmp_path, wrong)
    sc = S.Scene(64, 64, S.FMT_PRGB32, 0)
    sc.fill_geometry(S.GEOMETRY_TRIANGLE, (1.0, 2.0, 30.0, 4.0, 12.0, 40.0))
    py, native = _logs(stub, sc)
    assert py != native
