"""The C ABI (include/blend2d_amd.h): the library loads, exports every symbol
the header declares, and its host-side argument / state checks behave as the
reference's BLContext does (invalid arguments and missing targets fail with
an error code, nothing crashes).  No device work is done here."""
import ctypes
import os
import re

import pytest

import blend2d_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "blend2d_amd.h")

E_INVALID_VALUE = 0x10001
E_INVALID_STATE = 0x10002


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:BLAmdResult|uint32_t|int|size_t)\s+(bl_amd_\w+)\s*\(", text, re.M)))


def test_header_declares_what_the_mirror_binds():
    declared = _declared()
    assert len(declared) >= 30
    assert sorted(B.SYMBOLS) == declared


def test_library_exports_every_declared_symbol():
    lib = B.lib()
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing


def test_version_and_device_count_are_callable():
    lib = B.lib()
    v = lib.bl_amd_version()
    assert v >> 16 == 0 and (v >> 8) & 0xFF >= 1
    assert lib.bl_amd_device_count() >= 0


def test_context_state_checks_without_target():
    lib = B.lib()
    h = ctypes.c_void_p()
    assert lib.bl_amd_context_create(ctypes.byref(h)) == 0
    try:
        assert lib.bl_amd_context_set_comp_op(h, 0) == E_INVALID_STATE
        assert lib.bl_amd_context_set_fill_rule(h, 1) == E_INVALID_STATE
        assert lib.bl_amd_context_fill_all(h) == E_INVALID_STATE
        assert lib.bl_amd_context_flush(h, 0) == E_INVALID_STATE
        assert lib.bl_amd_context_end(h) == E_INVALID_STATE
        assert lib.bl_amd_context_replay(h) == E_INVALID_STATE
        assert lib.bl_amd_context_begin(h, None) == E_INVALID_VALUE
        st = B.BLAmdStats()
        assert lib.bl_amd_context_get_stats(h, ctypes.byref(st)) == 0
        assert st.commands == 0 and st.pixels == 0
    finally:
        assert lib.bl_amd_context_destroy(h) == 0


def test_null_arguments_are_rejected():
    lib = B.lib()
    assert lib.bl_amd_context_create(None) == E_INVALID_VALUE
    assert lib.bl_amd_image_create(None, 4, 4, 1) == E_INVALID_VALUE
    assert lib.bl_amd_batch_create(None) == E_INVALID_VALUE
    assert lib.bl_amd_batch_flush(None) == E_INVALID_VALUE
    assert lib.bl_amd_batch_replay(None) == E_INVALID_VALUE
    assert lib.bl_amd_context_get_stats(None, None) == E_INVALID_VALUE


@pytest.mark.parametrize("w,h,fmt", [(0, 4, 1), (4, -1, 1), (70000, 4, 1), (4, 4, 0), (4, 4, 9)])
def test_image_create_rejects_bad_sizes_and_formats(w, h, fmt):
    out = ctypes.c_void_p()
    assert B.lib().bl_amd_image_create(ctypes.byref(out), w, h, fmt) == E_INVALID_VALUE
    assert not out.value


def test_empty_batch_flush_is_a_no_op():
    lib = B.lib()
    b = ctypes.c_void_p()
    assert lib.bl_amd_batch_create(ctypes.byref(b)) == 0
    try:
        assert lib.bl_amd_batch_flush(b) == 0
        assert lib.bl_amd_batch_replay(b) == 0
        assert lib.bl_amd_batch_add(b, None) == E_INVALID_VALUE
    finally:
        lib.bl_amd_batch_destroy(b)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(B, "_lib", None)
    monkeypatch.setattr(B, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        B.lib()


def test_python_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of the header's structs have the C sizes and field
    offsets (a mirror shorter than the struct would let the library write
    past the Python object)."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ is missing")
    structs = {"BLAmdStats": B.BLAmdStats, "BLAmdTune": B.BLAmdTune}
    lines = ["#include <cstdio>", "#include <cstddef>", '#include "blend2d_amd.h"', "int main() {"]
    for name, cls in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(cls, field).offset, (name, field)
