"""Trains the tile cost model on the GPU and writes its parameters.

    python -m blend2d_amd.train_tile_model [--out blend2d_amd/tile_model.f32]

Samples: every analytic tile (a 64x16 tile of a band with path fills) of
seeded frames of four kinds -- batched glyph canvases (32-256 px, half A8),
batched Pareto cost canvases (1024-2048 px, every style and comp op),
1920x1080 cubic-gradient scenes and 1920x1080 pattern-shape scenes -- each
frame executed once more on the work-queue schedule with tile timing
(bl_amd_*_collect_tile_costs): per tile the 16 features and the targets
log2(1 + wall-clock ticks at 100 MHz), log2(1 + composited pixels).
Training is Adam on mini-batches with the library's MLP kernels (the
network of the sizing model); one held-out frame of each kind reports the
prediction error and how well the predicted order matches the measured
one (Spearman rank correlation).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import blend2d_amd as B  # noqa: E402
from blend2d_amd import scene as S  # noqa: E402
from blend2d_amd.train_cost_model import init_params  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "tile_model.f32")


def load(path: str = DEFAULT_OUT) -> np.ndarray | None:
    if not os.path.exists(path):
        return None
    p = np.fromfile(path, dtype="<f4")
    return p if p.size == B.COST_PARAMS else None


def frame_kinds(seed: int, scale: int):
    """(name, scenes) of the four frame kinds for one seed."""
    yield "glyph", S.glyph_canvases(seed, 256 * scale, 200)
    yield "cost", S.cost_canvases(seed, 4 * scale, 1024, 2048, 64)
    yield "cubic", [S.cubic_gradient_scene(seed, 1920, 1080, 2500 * scale)]
    yield "pattern", [S.pattern_shapes_scene(seed, 1920, 1080, 500 * scale)]


def collect_frame(scenes):
    """(features, targets) of every analytic tile of one frame (one context,
    or one batch of contexts)."""
    keep = []
    batch = B.BLBatch() if len(scenes) > 1 else None
    ctx = None
    for sc in scenes:
        images = []
        for fmt, px in sc.images:
            im = B.BLImage(px.shape[1], px.shape[0], fmt)
            im.upload(px)
            images.append(im)
        img = B.BLImage(sc.width, sc.height, sc.format)
        img.fill_raw(sc.clear_value)
        ctx = B.BLContext(img)
        B.replay_into(ctx, sc, images)
        if batch is not None:
            batch.add(ctx)
        keep.append((img, ctx, images))
    if batch is not None:
        batch.flush()
        X, T = B.collect_tile_costs(batch)
        batch.close()
    else:
        ctx.flush()
        X, T = B.collect_tile_costs(ctx)
    for img, c, _ in keep:
        c.end()
    return X, T


def spearman(a, b) -> float:
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--scale", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args()

    t0 = time.time()
    xs, ts = [], []
    for seed in range(300, 300 + args.seeds):
        for name, scenes in frame_kinds(seed, args.scale):
            X, T = collect_frame(scenes)
            xs.append(X)
            ts.append(T)
            print(f"seed {seed} {name}: {len(X)} tiles", flush=True)
    X, T = np.concatenate(xs), np.concatenate(ts)
    held = {name: collect_frame(scenes) for name, scenes in frame_kinds(999, args.scale)}
    print(f"samples: {len(X)} train, {sum(len(v[0]) for v in held.values())} held out ({time.time() - t0:.1f} s)",
          flush=True)
    t0 = time.time()
    rng = np.random.default_rng(0)
    order = rng.permutation(len(X))
    params, losses = B.cost_model_train(X[order], T[order], init_params(1), epochs=args.epochs, batch=args.batch,
                                        lr=args.lr)
    print(f"trained {args.epochs} epochs in {time.time() - t0:.1f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}",
          flush=True)
    for name, (Xv, Tv) in held.items():
        Y = B.cost_model_forward(Xv, params)
        err = np.abs(Y - Tv)
        print(f"held-out {name}: {len(Xv)} tiles, |log2 error| median {np.median(err, 0)}, "
              f"rank correlation of ticks {spearman(Y[:, 0], Tv[:, 0]):.3f}", flush=True)
    params.astype("<f4").tofile(args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
