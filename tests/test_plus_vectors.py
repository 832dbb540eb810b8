"""PLUS against the reference's own operator: tests/golden/plus_vectors.json
holds the sha256 of CompOp_Plus_Op (pipeline/reference/compopgeneric_p.h:
66-81, compiled from the reference header by make_plus_vectors.py) over every
(destination, source, mask) channel triple, plus a sample of raw vectors.
The oracle's PLUS and the GPU's PLUS (an aligned pattern of the source grid
filled over the destination grid with global alpha m, one canvas per m) must
reproduce that hash bit for bit."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import checkers

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_plus_vectors as M  # noqa: E402

FIX = json.load(open(os.path.join(HERE, "golden", "plus_vectors.json")))
OP_PLUS, FMT_PRGB32 = 12, 1


def _digest(out):
    return hashlib.sha256(np.ascontiguousarray(out, dtype=np.uint32).tobytes()).hexdigest()


def _oracle_grid():
    d, s = M.grid()
    out = np.empty((256, M.PIXELS), dtype=np.uint32)
    lib = checkers.oracle()
    for m in range(256):
        mm = np.full(M.PIXELS, m, dtype=np.uint32)
        lib.orc_composite_array(OP_PLUS, FMT_PRGB32, d.ctypes.data, s.ctypes.data, mm.ctypes.data, M.PIXELS,
                                out[m].ctypes.data)
    return out


def test_oracle_plus_matches_reference_operator():
    out = _oracle_grid()
    assert _digest(out) == FIX["sha256"]
    d, s = M.grid()
    for i, m, dv, sv, r in FIX["sample"]:
        assert d[i] == dv and s[i] == sv
        assert out[m, i] == r


def test_a_plausible_rounding_slip_changes_the_digest():
    # addus8(d, (s * m) >> 8) instead of the reference's div255: the digest
    # must tell them apart
    d, s = M.grid()
    out = np.empty((256, M.PIXELS), dtype=np.uint32)
    for m in range(256):
        res = np.zeros(M.PIXELS, dtype=np.uint32)
        for c in range(4):
            dc = (d >> (8 * c)) & 255
            sc = (s >> (8 * c)) & 255
            res |= np.minimum(dc + ((sc * m) >> 8), 255).astype(np.uint32) << (8 * c)
        out[m] = res
    assert _digest(out) != FIX["sha256"]


@pytest.mark.gpu
def test_gpu_plus_matches_reference_operator():
    import blend2d_amd as amd
    if amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    from blend2d_amd import scene as S
    d, s = M.grid()
    side = 128
    pat = amd.BLImage(side, side, S.FMT_PRGB32)
    pat.upload(s.reshape(side, side))
    batch = amd.BLBatch()
    keep = []
    for m in range(256):
        img = amd.BLImage(side, side, S.FMT_PRGB32)
        img.upload(d.reshape(side, side))
        ctx = amd.BLContext(img)
        ctx.set_comp_op(S.COMP_OP_PLUS)
        ctx.set_global_alpha(m / 255.0)
        ctx.set_fill_style(amd.BLPattern(pat, S.EXTEND_PAD))
        ctx.fill_rect(0, 0, side, side)
        batch.add(ctx)
        keep.append((img, ctx))
    batch.flush()
    out = np.stack([img.download().reshape(-1) for img, _ in keep])
    for _, ctx in keep:
        ctx.end()
    assert _digest(out) == FIX["sha256"]
