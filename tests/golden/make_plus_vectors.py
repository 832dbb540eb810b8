"""Writes plus_vectors.json: the sha256 of the reference's PLUS results over
the grid of plus_vectors.cpp (every destination / source / mask channel
triple) and a sample of raw vectors.

    python tests/golden/make_plus_vectors.py

Needs the reference sources (/root/reference) and g++; the operator is
compiled from the reference's own header (pipeline/reference/
compopgeneric_p.h), nothing of it is copied."""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"
PIXELS = 16384


def grid():
    """Destination and source pixels of the grid (see plus_vectors.cpp)."""
    p = np.arange(PIXELS * 4, dtype=np.uint32).reshape(PIXELS, 4)
    shifts = np.array([0, 8, 16, 24], dtype=np.uint32)
    d = ((p >> 8) << shifts).sum(axis=1).astype(np.uint32)
    s = ((p & 255) << shifts).sum(axis=1).astype(np.uint32)
    return d, s


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "plus_vectors")
        subprocess.run(["g++", "-std=c++17", "-O2", "-w", "-I" + REFERENCE, "-DBL_STATIC", "-DBL_BUILD_NO_JIT",
                        os.path.join(HERE, "plus_vectors.cpp"), "-o", exe], check=True)
        out_path = os.path.join(tmp, "out.bin")
        subprocess.run([exe, out_path], check=True)
        out = np.fromfile(out_path, dtype=np.uint32).reshape(256, PIXELS)
    d, s = grid()
    rng = np.random.default_rng(1)
    sample = []
    for m in [0, 1, 2, 127, 128, 254, 255] + [int(v) for v in rng.integers(0, 256, 9)]:
        for i in [0, 1, PIXELS - 1] + [int(v) for v in rng.integers(0, PIXELS, 13)]:
            sample.append([i, m, int(d[i]), int(s[i]), int(out[m, i])])
    doc = {"operator": "CompOp_Plus_Op<P32_A8R8G8B8>::op_prgb32_prgb32(d, s, m), pipeline/reference/compopgeneric_p.h:78-80",
           "grid": "m in 0..255 (rows), pixel i in 0..16383: channel c of d = (4i+c) >> 8, of s = (4i+c) & 255",
           "sha256": hashlib.sha256(out.tobytes()).hexdigest(),
           "sample_fields": ["pixel", "m", "d", "s", "result"], "sample": sample}
    with open(os.path.join(HERE, "plus_vectors.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(doc["sha256"], len(sample))


if __name__ == "__main__":
    main()
