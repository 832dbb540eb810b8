"""Trains the sizing MLP on the GPU and writes its parameters.

    python -m blend2d_amd.train_cost_model [--out blend2d_amd/cost_model.f32]

Samples: every fill command of seeded cost-model canvases (1 to 128 paths
of Pareto-distributed segment counts, lines / quads / cubics, on canvases of
256-2048 px) and of glyph canvases (1-200 quadratic contours of 20-80
segments on canvases of 32-256 px), rendered with exact counting; per command the 16 features and
the exact targets log2(1 + segments), log2(1 + (segment, band) copies) come
from bl_amd_context_collect_costs.  Training is Adam on mini-batches with
the library's own forward / backward kernels.  A held-out set reports the
prediction error.
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

DEFAULT_OUT = os.path.join(HERE, "cost_model.f32")


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases (layout of BL_AMD_COST_PARAMS)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 16), (16, 64)), np.zeros(64),
             rng.normal(0, np.sqrt(2 / 64), (64, 64)), np.zeros(64),
             rng.normal(0, np.sqrt(1 / 64), (64, 2)), np.zeros(2)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def collect(seeds, count: int, side_lo: int, side_hi: int, paths: int, max_segments: int):
    """(features, targets) of every fill command of the cost canvases."""
    return collect_scenes(sc for seed in seeds
                          for sc in S.cost_canvases(seed, count, side_lo, side_hi, paths, max_segments))


def collect_scenes(scenes):
    """(features, targets) of every fill command of `scenes`, one context each."""
    xs, ts = [], []
    for sc in scenes:
        images = []
        for fmt, px in sc.images:
            im = B.BLImage(px.shape[1], px.shape[0], fmt)
            im.upload(px)
            images.append(im)
        img = B.BLImage(sc.width, sc.height, sc.format)
        ctx = B.BLContext(img)
        B.replay_into(ctx, sc, images)
        ctx.flush()
        X, T = ctx.collect_costs()
        ctx.end()
        xs.append(X)
        ts.append(T)
    return np.concatenate(xs), np.concatenate(ts)


def load(path: str = DEFAULT_OUT) -> np.ndarray | None:
    if not os.path.exists(path):
        return None
    p = np.fromfile(path, dtype="<f4")
    return p if p.size == B.COST_PARAMS else None


def train(X, T, epochs: int, batch: int, lr: float, seed: int = 0):
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(X))
    return B.cost_model_train(X[order], T[order], init_params(seed), epochs=epochs, batch=batch, lr=lr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--seeds", type=int, default=6)
    ap.add_argument("--canvases", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args()

    t0 = time.time()
    # cost canvases (Pareto segment counts, every source kind) and glyph
    # canvases (many small quadratic contours on small canvases)
    X, T = collect(range(100, 100 + args.seeds), args.canvases, 256, 2048, 128, 20000)
    Xg, Tg = collect_scenes(sc for seed in range(200, 200 + args.seeds) for sc in S.glyph_canvases(seed, args.canvases, 200))
    X, T = np.concatenate([X, Xg]), np.concatenate([T, Tg])
    Xv, Tv = collect([999], 32, 256, 2048, 128, 20000)
    Xgv, Tgv = collect_scenes(S.glyph_canvases(998, 32, 200))
    Xv, Tv = np.concatenate([Xv, Xgv]), np.concatenate([Tv, Tgv])
    print(f"samples: {len(X)} train, {len(Xv)} held out ({time.time() - t0:.1f} s)", flush=True)
    t0 = time.time()
    params, losses = train(X, T, args.epochs, args.batch, args.lr)
    print(f"trained {args.epochs} epochs in {time.time() - t0:.1f} s; loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    Y = B.cost_model_forward(Xv, params)
    err = np.abs(Y - Tv)
    print(f"held-out |log2 error|: median {np.median(err, 0)}, p99 {np.percentile(err, 99, 0)}", flush=True)
    params.astype("<f4").tofile(args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
