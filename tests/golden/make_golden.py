"""Writes golden.json: reference renderings of the scenes in cases.py.

    python tests/golden/make_golden.py

Needs the reference library built from the reference sources
(`make -f oracle/ref.mk`, output oracle/_ref/).  Each entry records, per
canvas, the first BLResult error of the replay and the digest (sha256 +
64-row band crc32) of the pixels the reference produced with its non-JIT
pipeline and thread_count = 0."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import checkers  # noqa: E402
import cases  # noqa: E402


def main():
    if not checkers.have_ref():
        raise SystemExit("oracle/_ref/libref_render.so is missing: make -f oracle/ref.mk")
    out = {}
    for name, make in cases.cases().items():
        entries = []
        for sc in make():
            px, err, _ = checkers.render_reference(sc)
            d = cases.digest(px)
            d.update({"width": sc.width, "height": sc.height, "format": sc.format, "error": err})
            entries.append(d)
        out[name] = entries
        print(name, len(entries), flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
