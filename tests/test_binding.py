"""The reference-side binding (blend2d_amd/binding/amdcontext.cpp): a
BLContext implementation of the reference whose fill entries forward to the
C ABI.  On the CPU: it compiles against the reference headers and sets every
slot of its BLContextVirt table.  On the GPU: the reference's own scene
driver (oracle/ref_render.cpp built with REF_RENDER_AMD into
oracle/_ref/libamd_ref_render.so) renders every golden scene through the
reference's C++ API, its virtual table and this binding into the GPU
backend, and gets the reference's pixels."""
import ctypes
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
BINDING = os.path.join(ROOT, "blend2d_amd", "binding", "amdcontext.cpp")
AMD_DRIVER = os.path.join(ROOT, "oracle", "_ref", "libamd_ref_render.so")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import cases  # noqa: E402


def _need_reference():
    if not os.path.isdir(os.path.join(REFERENCE, "blend2d")) or not shutil.which("g++"):
        pytest.skip("the reference sources (headers) are not here")


def test_binding_compiles_against_the_reference_headers():
    _need_reference()
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-DBL_BUILD_NO_JIT",
                          "-I" + REFERENCE, "-I" + os.path.join(ROOT, "include"), BINDING],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]


def test_every_virtual_table_slot_is_set(tmp_path):
    _need_reference()
    ref_lib = os.path.join(ROOT, "oracle", "_ref")
    amd_lib = os.path.join(ROOT, "blend2d_amd")
    if not os.path.exists(os.path.join(ref_lib, "libblend2d_ref.so")) or not os.path.exists(os.path.join(amd_lib, "libblend2d_amd.so")):
        pytest.skip("libraries not built")
    exe = tmp_path / "virt_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-DBL_BUILD_NO_JIT", "-I" + REFERENCE, "-I" + os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "native", "binding_virt_check.cpp"),
                    "-L" + ref_lib, "-lblend2d_ref", "-L" + amd_lib, "-lblend2d_amd",
                    "-Wl,-rpath," + ref_lib, "-Wl,-rpath," + amd_lib], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "null 0" in res.stdout


def _bound_driver_on_stub(tmp_path):
    """The bound scene driver linked against the recording stand-in of the C
    ABI (tests/native/abi_stub.cpp) instead of the GPU library."""
    ref_lib = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref_lib, "libblend2d_ref.so")):
        pytest.skip("the reference library is not built")
    so = tmp_path / "bound_stub.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-DREF_RENDER_AMD", "-DBL_BUILD_NO_JIT",
                    "-I" + REFERENCE, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "blend2d_amd", "csrc"),
                    "-o", str(so), os.path.join(ROOT, "oracle", "ref_render.cpp"), BINDING,
                    os.path.join(ROOT, "tests", "native", "abi_stub.cpp"),
                    "-L" + ref_lib, "-lblend2d_ref", "-Wl,-rpath," + ref_lib], check=True)
    lib = ctypes.CDLL(str(so))
    lib.ref_render_amd.restype = ctypes.c_uint32
    lib.ref_render_amd.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    lib.stub_log.restype = ctypes.c_char_p
    lib.stub_reset.restype = None
    return lib


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_write_back_reaches_the_callers_image(tmp_path, fmt):
    """The device pixels are written into the image the caller handed to
    the context (the stand-in's download writes a marker), not into a copy
    made because the context holds a second reference to it."""
    _need_reference()
    from blend2d_amd import scene as S
    lib = _bound_driver_on_stub(tmp_path)
    sc = S.Scene(37, 11, fmt, 0x80402010)
    sc.set_fill_style_rgba32(0xFF00FF00)
    sc.fill_rect_i(1, 2, 5, 6)
    bpp = 1 if fmt == 3 else 4
    out = np.zeros((sc.height, sc.width * bpp), dtype=np.uint8)
    data = sc.to_bytes()
    lib.stub_reset()
    err = lib.ref_render_amd(data, len(data), out.ctypes.data, out.nbytes, 0, ctypes.byref(ctypes.c_double()))
    assert err == 0
    want = ((3 * np.arange(sc.height)[:, None] + np.arange(sc.width * bpp)[None, :]) & 0xFF).astype(np.uint8)
    assert np.array_equal(out, want)
    log = lib.stub_log().decode().splitlines()
    assert any(l.startswith("fill_rect_i") for l in log), log
    assert any(l.startswith("image_download") for l in log), log


GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
CASES = cases.cases()


@pytest.fixture(scope="module")
def driver():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    if not os.path.exists(AMD_DRIVER):
        pytest.skip("the bound scene driver is not built")
    lib = ctypes.CDLL(AMD_DRIVER)
    lib.ref_render_amd.restype = ctypes.c_uint32
    lib.ref_render_amd.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    return lib


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_scenes_through_the_reference_api(driver, name):
    for k, (sc, want) in enumerate(zip(CASES[name](), GOLDEN[name])):
        data = sc.to_bytes()
        bpp = 1 if sc.format == 3 else 4
        out = np.zeros((sc.height, sc.width * bpp), dtype=np.uint8)
        ms = ctypes.c_double()
        err = driver.ref_render_amd(data, len(data), out.ctypes.data, out.nbytes, 0, ctypes.byref(ms))
        px = out.reshape(sc.height, sc.width) if bpp == 1 else out.view(np.uint32).reshape(sc.height, sc.width)
        got = cases.digest(px)
        assert err == want["error"], (name, k, hex(err))
        assert got["sha256"] == want["sha256"], (name, k)


def _normalized(lines):
    """Call-log lines with the differences that do not change what renders
    removed: the reference's fill_polygon arrives as fill_geometry(POLYGOND);
    a path without origin as origin (0, 0); pattern images are uploaded when
    a style first uses them (other image numbers)."""
    import re
    out = []
    for line in lines:
        line = re.sub(r"^begin \d+", "begin", line)
        line = re.sub(r"^fill_polygon n=(\d+) (\w+)", r"fill_geometry type=16 view n=\1 h=\2", line)
        line = line.replace("origin=0,0 ", "origin=none ")
        line = re.sub(r"image=\d+", "image=N", line)
        if line.startswith("image_create") or line.startswith("image_upload"):
            continue
        out.append(line)
    return out


def test_bound_driver_makes_the_calls_of_the_python_replay(tmp_path):
    """Every golden scene driven through the reference's C++ API, its
    virtual table and the binding reaches the C ABI as the same calls with
    the same arguments (pointed-to data hashed) as the Python mirror's
    replay of the scene: the binding translates styles (colours, gradients,
    patterns by object type), transforms, comp ops and geometry faithfully."""
    _need_reference()
    import test_player_calls as P
    lib = _bound_driver_on_stub(tmp_path)
    (tmp_path / "py").mkdir()
    stub = P._build(tmp_path / "py")
    checked = 0
    for name in sorted(CASES):
        sc = CASES[name]()[0]
        py, _ = P._logs(stub, sc)
        data = sc.to_bytes()
        bpp = 1 if sc.format == 3 else 4
        out = np.zeros((sc.height, sc.width * bpp), dtype=np.uint8)
        lib.stub_reset()
        lib.ref_render_amd(data, len(data), out.ctypes.data, out.nbytes, 0, ctypes.byref(ctypes.c_double()))
        log = lib.stub_log().decode().splitlines()
        start = max(i for i, l in enumerate(log) if l.startswith("begin"))
        end = next(i for i in range(start, len(log)) if log[i].startswith("flush"))
        a, b = _normalized(py), _normalized(log[start:end + 1])
        assert a == b, name
        checked += 1
    assert checked == len(CASES)
