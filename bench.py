"""Fill-path throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3|c4|c5|srcover4k]

A step renders one whole synthetic frame through the device pipeline (edge
building, binning, tile rasterization + fetch + composite) from a frame
already resident in HBM (bl_amd_context_replay / bl_amd_batch_replay): the
commands, geometry, styles and LUTs were recorded and uploaded once before
the timed region.

value      Mpix/s = pixels composited with a non-zero mask per frame x K /
           wall time of the K steps (host clock, device synchronized on both
           sides).  paths/s and canvas Mpix/s are reported beside it.
roofline   the render stage (k_render for bands with analytic fills,
           k_render_box_solid / k_render_box for box-only bands), timed with HIP
           events on the stream it runs on, against the HBM roofline: algorithmic
           bytes per launch
           = 8 B per composited PRGB32 pixel (2 B for A8: destination read +
           write) + gradient LUT bytes + pattern image bytes + 16 B per binned
           edge segment it reads.
cpu_baseline  the reference library (oracle/_ref, compiled from the reference
           sources without its JIT) rendering the same frame with worker
           threads on this host's cores; its first frame is also the
           bit-exactness check of the GPU output.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpix/s + paths/s at bit-exact output, batched+learned vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0
HOST_CORES_MAX = 16          # CPU share of a one-GPU box


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def _workload(name: str, seed: int):
    from blend2d_amd import scene as S
    if name == "c1":
        return [S.polygons_scene(seed, 512, 512, 1000)], {
            "workload": "512x512 PRGB32, 1000 random polygons (3-12 vertices), non-zero, solid, SRC_OVER",
            "canvases": 1, "paths": 1000, "seed": seed}
    if name == "c2":
        return [S.cubic_gradient_scene(seed, 1920, 1080, 10000)], {
            "workload": "1920x1080 PRGB32, 10000 closed cubic paths (4-16 segments), even-odd, "
                        "50% linear / 50% radial gradients (2-8 stops, PAD), SRC_OVER",
            "canvases": 1, "paths": 10000, "seed": seed}
    if name == "c3":
        return [S.pattern_shapes_scene(seed, 3840, 2160, 2000)], {
            "workload": "3840x2160 PRGB32, 2000 rotated rects / ellipses, 512x512 bilinear REPEAT pattern, "
                        "SRC_OVER/MULTIPLY/PLUS/SCREEN",
            "canvases": 1, "paths": 2000, "seed": seed}
    if name == "c4":
        scenes = S.glyph_canvases(seed, 4096, 200)
        return scenes, {
            "workload": "4096 canvases 32-256 px (half A8, half PRGB32), 1-200 quadratic glyph contours each, "
                        "non-zero, solid / linear gradient, SRC_OVER, one batched frame",
            "canvases": len(scenes), "paths": int(sum(sc.fill_count for sc in scenes)), "seed": seed}
    if name == "c5":
        scenes = S.cost_canvases(seed, 256, 1024, 4096, 64)
        return scenes, {
            "workload": "256 PRGB32 canvases 1024-4096 px, 1-64 paths each with clip(Pareto(1.5) * 8, 8, 20000) "
                        "line / quad / cubic segments, solid / linear / radial / pattern, SRC_OVER / MULTIPLY / "
                        "PLUS / SCREEN, both fill rules, one batched frame",
            "canvases": len(scenes), "paths": int(sum(sc.fill_count for sc in scenes)), "seed": seed}
    if name == "srcover4k":
        scenes = [S.srcover_layers_scene(seed + i, 3840, 2160, 1) for i in range(10)]
        return scenes, {
            "workload": "10 distinct 3840x2160 PRGB32 canvases, one full-canvas semi-transparent SRC_OVER "
                        "layer each, one batched frame",
            "canvases": 10, "paths": 10, "seed": seed}
    raise SystemExit(f"unknown workload {name}")


class _Frame:
    """One recorded frame on the GPU: a single context, or a batch of
    contexts (one set of kernel launches for all canvases)."""

    def __init__(self, scenes):
        import blend2d_amd as B
        self.B = B
        self.scenes = scenes
        self.keep = []
        self.targets = []
        self.batch = B.BLBatch() if len(scenes) > 1 else None
        self.errors = []
        for sc in scenes:
            images = []
            for fmt, px in sc.images:
                h, w = px.shape
                im = B.BLImage(w, h, fmt)
                im.upload(px)
                images.append(im)
            img = B.BLImage(sc.width, sc.height, sc.format)
            img.fill_raw(sc.clear_value)
            ctx = B.BLContext(img)
            errs = B.replay_into(ctx, sc, images)
            self.errors.append(errs[0][1] if errs else 0)
            if self.batch is not None:
                self.batch.add(ctx)
            self.keep.append((ctx, images))
            self.targets.append(img)
        # first execution: upload + render onto the cleared canvases
        if self.batch is not None:
            self.batch.flush()
        else:
            self.keep[0][0].flush()

    def replay(self):
        if self.batch is not None:
            self.batch.replay()
        else:
            self.keep[0][0].replay()

    def stats(self):
        return self.batch.last_stats() if self.batch is not None else self.keep[0][0].last_stats()

    def download(self):
        return [img.download() for img in self.targets]


def _time_frame(frame, steps, warmup):
    for _ in range(warmup):
        frame.replay()
    _sync()
    render_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        frame.replay()
        render_ms.append(frame.stats()["render_ms"])
    _sync()
    return time.perf_counter() - t0, render_ms


def _batched_vs_plain(seed: int, repeats: int, learned: bool):
    """The 4096-canvas glyph workload, device pipeline time only (recording
    of the calls excluded): plain = one flush per canvas through one
    context (exact counting); batched = the same calls of all canvases in
    one batch flush (learned sizing when a trained model is installed), and
    its replay from device-resident inputs."""
    import blend2d_amd as B
    from blend2d_amd import scene as S
    scenes = S.glyph_canvases(seed, 4096, 200)
    out = {"workload": "4096 canvases 32-256 px (half A8, half PRGB32), 1-200 glyph contours each",
           "paths": int(sum(sc.fill_count for sc in scenes)),
           "batched_sizing": "learned" if learned else "exact"}
    targets = [B.BLImage(sc.width, sc.height, sc.format) for sc in scenes]

    B.set_sizing(B.SIZING_EXACT)
    best = None
    for _ in range(repeats):
        ctx = B.BLContext()
        el = 0.0
        px = 0
        up = dev = 0.0
        for sc, img in zip(scenes, targets):
            img.fill_raw(sc.clear_value)
            ctx.begin(img)
            B.replay_into(ctx, sc, [])
            _sync()
            t0 = time.perf_counter()
            ctx.flush()
            el += time.perf_counter() - t0
            st = ctx.last_stats()
            px += st["pixels"]
            up += st.get("upload_ms", 0.0)
            dev += st.get("total_ms", 0.0)
            ctx.end()
        if best is None or el < best:
            out["plain_flush_breakdown_ms"] = {"upload": round(up, 3), "device_stages": round(dev, 3)}
        best = el if best is None else min(best, el)
    out["plain_ms_per_frame"] = round(best * 1e3, 3)
    out["plain_mpix_per_s"] = round(px / best / 1e6, 3)
    out["plain_paths_per_s"] = round(out["paths"] / best, 1)

    B.set_sizing(B.SIZING_LEARNED if learned else B.SIZING_EXACT)
    best = None
    for _ in range(repeats):
        batch = B.BLBatch()
        ctxs = []
        for sc, img in zip(scenes, targets):
            img.fill_raw(sc.clear_value)
            ctx = B.BLContext(img)
            B.replay_into(ctx, sc, [])
            batch.add(ctx)
            ctxs.append(ctx)
        _sync()
        t0 = time.perf_counter()
        batch.flush()
        el = time.perf_counter() - t0
        st = batch.last_stats()
        best = el if best is None else min(best, el)
    out["batched_ms_per_frame"] = round(best * 1e3, 3)
    out["batched_mpix_per_s"] = round(st["pixels"] / best / 1e6, 3)
    out["batched_paths_per_s"] = round(out["paths"] / best, 1)
    out["batched_fallback"] = st.get("fallback")
    out["batched_flush_breakdown_ms"] = {"merge": round(st.get("merge_ms", 0.0), 3), "upload": round(st.get("upload_ms", 0.0), 3),
                                         "device_stages": round(st.get("total_ms", 0.0), 3),
                                         "render": round(st.get("render_ms", 0.0), 3)}
    el, _ = _time_frame(_BatchHandle(batch), 3, 1)
    out["batched_replay_ms_per_frame"] = round(el * 1e3 / 3, 3)
    for c in ctxs:
        c.end()
    B.set_sizing(B.SIZING_EXACT)
    out["batched_over_plain"] = round(out["batched_mpix_per_s"] / out["plain_mpix_per_s"], 3)
    return out


class _BatchHandle:
    def __init__(self, batch):
        self.batch = batch

    def replay(self):
        self.batch.replay()

    def stats(self):
        return self.batch.last_stats()


def _render_bytes(st, scenes):
    a8 = all(sc.format == 3 for sc in scenes)
    per_px = 2 if a8 else 8
    if not a8 and any(sc.format == 3 for sc in scenes):
        # mixed batch: weight by canvas area
        area32 = sum(sc.width * sc.height for sc in scenes if sc.format != 3)
        area8 = sum(sc.width * sc.height for sc in scenes if sc.format == 3)
        per_px = (8 * area32 + 2 * area8) / float(area32 + area8)
    return per_px * st["pixels"] + st["lut_bytes"] + st["image_bytes"] + 16 * st["binned"]


def _render_kernels(mask: int) -> str:
    names = [n for bit, n in ((1, "k_render"), (2, "k_render_box_solid"), (4, "k_render_box")) if mask & bit]
    return "+".join(names) if names else "none"


def _traffic(workload: str):
    """HBM bytes per render launch of this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic.json), each
    counter scaled by what it reads for a plain copy of known size; None when
    the workload was not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as fh:
            rec = json.load(fh).get(workload)
    except (OSError, ValueError):
        return None
    return int(rec["traffic_bytes"]) if rec else None


def _cpu_baseline(scenes, first_gpu, pixels_per_frame, budget_s: float):
    """Reference library on this host's cores (bounded sample), plus the
    bit-exactness check of the GPU's first frame against its output."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    threads = max(1, min(HOST_CORES_MAX, os.cpu_count() or 1))
    if checkers.have_ref():
        kind = "reference"

        def render(sc):
            out, err, _ = checkers.render_reference(sc, threads=threads)
            return out, err
    else:
        kind, threads = "port", 1

        def render(sc):
            return checkers.render_oracle(sc)

    parity = {"against": kind, "max_diff": 0, "mean_diff": 0.0, "channels_differing": 0, "errors_match": True}
    t0 = time.perf_counter()
    frames = 0
    done_check = False
    while True:
        for i, sc in enumerate(scenes):
            out, err = render(sc)
            if not done_check:
                mx, mean, n = checkers.diff_info(first_gpu[i], out)
                parity["max_diff"] = max(parity["max_diff"], int(mx))
                parity["channels_differing"] += int(n)
                parity["mean_diff"] = max(parity["mean_diff"], float(mean))
        done_check = True
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or frames >= 50:
            break
    mpix = pixels_per_frame * frames / el / 1e6
    sample = (f"{frames} whole frame(s) of the same workload, {el:.1f} s, "
              f"{'reference library (no JIT) with ' + str(threads) + ' worker thread(s)' if kind == 'reference' else 'scalar oracle port'}")
    return {"value": round(mpix, 3), "unit": "Mpix/s", "cores": threads, "kind": kind, "sample": sample}, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "srcover4k"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity check")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-batching", action="store_true", help="skip the 4096-canvas batched-vs-plain measurement")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy three times (scale for FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a frame's tiles are rendered by one device)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = 64 * 1024 * 1024
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        for i in range(3):
            torch.mul(x, 1.0, out=y)
            ev[i + 1].record()
        torch.cuda.synchronize()
        ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(3)]
        print(f"copy probe: 3 launches, {4 * n} bytes read and {4 * n} bytes written each; "
              f"{', '.join(f'{m:.3f}' for m in ms)} ms = {', '.join(f'{8 * n / m / 1e6:.0f}' for m in ms)} GB/s",
              flush=True)
        return

    import blend2d_amd as B
    from blend2d_amd import train_cost_model
    if B.lib().bl_amd_device_count() < 1:
        raise SystemExit("no HIP device visible")

    scenes, config = _workload(args.workload, args.seed)
    frame = _Frame(scenes)
    first = frame.download()
    st0 = frame.stats()
    print(f"# workload {args.workload}: {st0}", file=sys.stderr, flush=True)

    # Sizing: exact counting, and learned sizing when trained parameters are
    # present; the faster one is the headline (pixels are identical).
    sizing = {}
    params = train_cost_model.load()
    try:
        B.set_sizing(B.SIZING_LEARNED)
        B.set_sizing(B.SIZING_EXACT)
    except B.BLError:
        params = None          # learned sizing not available in this build
    modes = [("exact", B.SIZING_EXACT)] + ([("learned", B.SIZING_LEARNED)] if params is not None else [])
    if params is not None:
        B.cost_model_set(params)
    best = None
    runs = {}
    for name, mode in modes:
        B.set_sizing(mode)
        el_m, render_m = _time_frame(frame, args.steps, args.warmup)
        st_m = frame.stats()
        runs[name] = (render_m, st_m)
        sizing[name] = {"ms_per_step": round(el_m * 1e3 / args.steps, 4), "syncs": st_m.get("syncs"),
                        "fallback": st_m.get("fallback"), "pred_segs": st_m.get("pred_segs")}
        if best is None or el_m < best[1]:
            best = (name, el_m, render_m, st_m)
    B.set_sizing(B.SIZING_EXACT)
    config["sizing"] = best[0]
    _, el, _, st = best
    # The render kernels do the same work in every sizing mode; pool mode
    # launches them over upper-bound grids.  The roofline is taken from the
    # exact-sizing run, whose launches are the ones the kernel traces show.
    render_ms, st_r = runs["exact"]

    pixels = st["pixels"]
    ms_per_step = el * 1e3 / args.steps
    value = pixels * args.steps / el / 1e6
    canvas_px = sum(sc.width * sc.height for sc in scenes)
    avg_render_ms = float(np.mean(render_ms))
    bytes_render = _render_bytes(st_r, scenes)
    achieved = bytes_render / (avg_render_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": config,
        "paths_per_s": round(config["paths"] * args.steps / el, 1),
        "canvas_mpix_per_s": round(canvas_px * args.steps / el / 1e6, 3),
        "pixels_per_frame": int(pixels),
        "render_ms": round(avg_render_ms, 4),
        "sizing": sizing,
        "roofline": {
            "bound": "hbm",
            "kernel": _render_kernels(st_r.get("render_kernels", 1)),
            "sizing": "exact",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(args.workload),
            "bytes_per_launch": int(bytes_render),
        },
        "cpu_baseline": None,
    }
    if any(frame.errors):
        line["errors"] = frame.errors
    if not args.no_batching:
        print("# batched vs plain (4096 canvases)", file=sys.stderr, flush=True)
        line["batched_vs_plain"] = _batched_vs_plain(args.seed, 2, params is not None)
    if not args.no_cpu:
        cpu, parity = _cpu_baseline(scenes, first, pixels, args.cpu_seconds)
        line["cpu_baseline"] = cpu
        line["parity"] = parity
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
