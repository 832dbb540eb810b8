"""Fill-path throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c1|c2|c3|c5|srcover4k]

A step is one whole frame, end to end: every canvas's context begins, takes
its recorded calls (issued from native code through the C ABI by the scene
player, libb2amd_player.so), one synchronous batch flush packs, uploads and
renders them (edge building, binning, tile rasterization + fetch +
composite), and every context ends -- the reference benchmark's
begin -> calls -> flush(SYNC) -> end timing (blend2d-testing/bench/
bl_bench_backend.cpp:50-89).  The default workload is the 4096-canvas glyph
frame the metric is quoted on, rendered in the batched form with learned
sizing (the trained cost model, when present).

value      Mpix/s = pixels composited with a non-zero mask per frame x K /
           wall time of the K steps.  paths/s and canvas Mpix/s beside it.
batched_vs_plain  the same frame in the plain form (one context, one
           synchronous flush per canvas), and the batched form with exact
           sizing (the learned part's ablation); outputs of both are compared
           with the batched frame's pixels.
device_only  the batched frame replayed from its device-resident copy
           (no recording, packing or upload): the device stages alone.
roofline   the render stage of the headline frame (k_render, plus the box
           kernels for box-only bands), timed with HIP events on the stream
           it runs on, against the HBM roofline: algorithmic bytes per launch
           = 8 B per composited PRGB32 pixel (2 B for A8: destination read +
           write) + gradient LUT bytes + pattern image bytes + 16 B per binned
           edge segment it reads.  roofline_srcover4k: the same for ten
           distinct 3840x2160 canvases with one full-canvas SRC_OVER layer
           each (device-resident, larger than the last-level cache).
           roofline_blit4k / roofline_mask4k: ten 3840x2160 canvases with one
           full-canvas blit_image (12 B per pixel: destination read + write,
           source read) / fill_mask of an A8 mask (9 B per pixel) each, with
           the reference's single-core time for the same frame beside them
           (blit4k.cpu_reference, mask4k.cpu_reference).
cpu_baseline  the reference library (oracle/_ref, compiled from the reference
           sources without its JIT) rendering the same frame on this host's
           cores, timed by its own begin -> end clock, best of R, with
           thread_count 0 and with worker threads; its first pass is also the
           bit-exactness check of every canvas of the GPU's first frame.
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
    if name == "blit4k":
        img = S.premultiplied_noise(np.random.default_rng(seed), 3840, 2160)
        scenes = [S.blit_layer_scene(img, seed + i) for i in range(10)]
        return scenes, {
            "workload": "10 distinct 3840x2160 PRGB32 canvases, one full-canvas blit_image (integer origin, "
                        "SRC_OVER) of a 3840x2160 translucent PRGB32 image each, one batched frame",
            "canvases": 10, "paths": 10, "seed": seed}
    if name == "mask4k":
        mask = np.random.default_rng(seed).integers(0, 256, size=(2160, 3840), dtype=np.uint8)
        scenes = [S.mask_layer_scene(mask, seed + i) for i in range(10)]
        return scenes, {
            "workload": "10 distinct 3840x2160 PRGB32 canvases, one full-canvas fill_mask (A8 mask, integer "
                        "origin, translucent solid, SRC_OVER) each, one batched frame",
            "canvases": 10, "paths": 10, "seed": seed}
    raise SystemExit(f"unknown workload {name}")


def _render_bytes(st, scenes):
    a8 = all(sc.format == 3 for sc in scenes)
    per_px = 2 if a8 else 8
    if not a8 and any(sc.format == 3 for sc in scenes):
        # mixed frame: weight by canvas area
        area32 = sum(sc.width * sc.height for sc in scenes if sc.format != 3)
        area8 = sum(sc.width * sc.height for sc in scenes if sc.format == 3)
        per_px = (8 * area32 + 2 * area8) / float(area32 + area8)
    return per_px * st["pixels"] + st["lut_bytes"] + st["image_bytes"] + 16 * st["binned"]


def _render_kernels(mask: int, schedule: int = 0) -> str:
    # queue schedules run the analytic tiles through k_render_queue, after the
    # tile-ordering kernels (features, cost model, counting sort), all timed
    analytic = "k_render" if schedule == 0 else "k_render_queue"
    names = [n for bit, n in ((1, analytic), (2, "k_render_box_solid"), (4, "k_render_box"), (8, "k_render_box_fast"))
             if mask & bit]
    return "+".join(names) if names else "none"


def _tune_times() -> dict:
    """Render-stage ms of every candidate of the autotuner's latest run (the
    last line of its log)."""
    import blend2d_amd as B
    lines = [l for l in B.autotune_log().splitlines() if l.strip()]
    if not lines:
        return {}
    body = lines[-1].split(":", 1)[1].split(" ms ->")[0].split()
    return {body[i]: float(body[i + 1]) for i in range(0, len(body) - 1, 2)}


def _schedule(st) -> str:
    import blend2d_amd as B
    name = B.SCHEDULE_NAMES.get(int(st.get("schedule", 0)), "?")
    return name if not st.get("queue_wgs") else f"{name}/{int(st['queue_wgs'])}"


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


def _roofline(scenes, render_ms, st, workload):
    avg = float(np.mean(render_ms))
    nbytes = _render_bytes(st, scenes)
    achieved = nbytes / (avg * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": _render_kernels(st.get("render_kernels", 1), st.get("schedule", 0)),
        "schedule": _schedule(st),
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": _traffic(workload),
        "bytes_per_launch": int(nbytes),
        "avg_launch_ms": round(avg, 4),
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _compare(a_list, b_list):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    mx = n = 0
    mean = 0.0
    for a, b in zip(a_list, b_list):
        m, mn, c = checkers.diff_info(a, b)
        mx, n, mean = max(mx, m), n + c, max(mean, mn)
    return {"max_diff": int(mx), "mean_diff": float(mean), "channels_differing": int(n)}


def _reference_passes(scenes, threads, budget_s, max_passes=5):
    """Best whole-frame time (ms) of the reference library over the passes
    that fit the budget (at least one), and the number of passes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers
    best, passes, t0 = float("inf"), 0, time.perf_counter()
    while True:
        total = 0.0
        for sc in scenes:
            total += checkers.render_reference(sc, threads=threads)[2]
        best, passes = min(best, total), passes + 1
        if passes >= max_passes or time.perf_counter() - t0 >= budget_s:
            return best, passes


def _reference_passes_child(queue, scenes, threads, budget_s):
    queue.put(_reference_passes(scenes, threads, budget_s))


def _cpu_baseline(scenes, first_gpu, gpu_errors, pixels_per_frame, budget_s: float):
    """Reference library on this host's cores: each canvas through ref_render
    (its own begin -> end clock), summed over the frame, best of the passes
    that fit the budget, with thread_count 0 and with worker threads; the
    first pass also checks every canvas of the GPU's first frame.  The
    worker-thread passes run in a forked child process (it uses no GPU and
    leaves with os._exit): the reference library has been seen to crash with
    worker threads on some scenes, which then costs only that figure."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import multiprocessing
    import checkers
    workers = max(1, min(HOST_CORES_MAX, os.cpu_count() or 1))
    parity = {"canvases": len(scenes), "errors_match": True}
    if not checkers.have_ref():
        # the scalar port, one thread
        t0 = time.perf_counter()
        outs = []
        for i, sc in enumerate(scenes):
            out, err = checkers.render_oracle(sc)
            outs.append(out)
            parity["errors_match"] &= int(err) == int(gpu_errors[i])
        ms = (time.perf_counter() - t0) * 1e3
        parity.update(_compare(first_gpu, outs), against="port")
        return {"value": round(pixels_per_frame / (ms * 1e-3) / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
                "sample": f"one pass of the whole frame by the scalar oracle port; {_cpu_model()}",
                "ms_per_frame": round(ms, 3)}, parity

    # parity pass, thread_count 0 (also the first timing pass)
    outs, first_ms = [], 0.0
    for i, sc in enumerate(scenes):
        out, err, ms = checkers.render_reference(sc, threads=0)
        outs.append(out)
        first_ms += ms
        parity["errors_match"] &= int(err) == int(gpu_errors[i])
    parity.update(_compare(first_gpu, outs), against="reference")
    if not parity["errors_match"]:
        # The reference's non-JIT pipeline reports NOT_IMPLEMENTED for
        # MULTIPLY / SCREEN on PRGB32 and skips those fills (they are
        # JIT-only there), so such a frame is checked against the port too.
        port = {"errors_match": True}
        port_outs = []
        for i, sc in enumerate(scenes):
            out, err = checkers.render_oracle(sc)
            port_outs.append(out)
            port["errors_match"] &= int(err) == int(gpu_errors[i])
            if (i + 1) % 32 == 0:
                print(f"port check: {i + 1}/{len(scenes)} canvases", file=sys.stderr, flush=True)
        port.update(_compare(first_gpu, port_outs))
        parity["port"] = port
        parity["reference_note"] = ("the reference library returned an error the GPU did not (its non-JIT "
                                    "pipeline lacks MULTIPLY/SCREEN) and left those fills out")
    results = {}
    ms0, n0 = _reference_passes(scenes, 0, budget_s / 2, max_passes=4)
    results[0] = (min(ms0, first_ms), n0 + 1)
    ctx = multiprocessing.get_context("fork")
    queue = ctx.Queue()
    child = ctx.Process(target=_reference_passes_child, args=(queue, scenes, workers, budget_s / 2))
    child.start()
    results[workers] = None
    t_start = time.perf_counter()
    deadline = t_start + budget_s * 4 + 120
    next_note = t_start + 30
    while time.perf_counter() < deadline:
        try:
            results[workers] = queue.get(timeout=1.0)
            break
        except Exception:
            if not child.is_alive():   # ended without a result: it crashed
                break
        if time.perf_counter() >= next_note:
            print(f"# cpu baseline: {workers}-thread passes running "
                  f"({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
            next_note += 30
    child.join(timeout=10)
    if child.is_alive():
        child.kill()
        child.join(timeout=10)
    crashed = results[workers] is None or child.exitcode != 0
    if crashed:
        results.pop(workers)
    threads, (ms, passes) = min(results.items(), key=lambda kv: kv[1][0])
    cores = 1 if threads == 0 else threads
    others = {str(t): round(pixels_per_frame / (v[0] * 1e-3) / 1e6, 3) for t, v in results.items()}
    note = f"; with {workers} worker threads the reference library crashed (exit {child.exitcode})" if crashed else ""
    if not parity["errors_match"]:
        note += "; it skipped the fills it reports NOT_IMPLEMENTED for (MULTIPLY/SCREEN), so it did less work"
    sample = (f"the whole frame ({len(scenes)} canvas(es)), best of {passes} pass(es) by the reference library "
              f"(no JIT, asmjit is not available offline)'s own begin->end clock, thread_count {threads}; "
              f"{_cpu_model()}; Mpix/s by thread_count: {others}{note}")
    return {"value": round(pixels_per_frame / (ms * 1e-3) / 1e6, 3), "unit": "Mpix/s", "cores": cores,
            "kind": "reference", "sample": sample, "ms_per_frame": round(ms, 3)}, parity


def _steps(frame, mode, steps, threads=1):
    """K frames end to end (each ends with a synchronous flush); returns the
    wall time, the per-step player times, every step's stats and the time
    each step took to issue its calls."""
    sts, per, issue = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ms, errs, st, r = frame.render(mode, threads)
        if r:
            raise RuntimeError(f"render failed with BLResult 0x{r:X}")
        per.append(ms)
        sts.append(st)
        issue.append(frame.last_issue_ms)
    return time.perf_counter() - t0, per, sts, issue


def _side_workload(name, seed, steps, learned, threads):
    """End-to-end batched frames and device-only replays of another workload
    (one line per workload in the output)."""
    import blend2d_amd as B
    from blend2d_amd import player
    scenes, config = _workload(name, seed)
    frame = player.Frame(scenes)
    frame.clear()
    B.set_sizing(B.SIZING_LEARNED if learned else B.SIZING_EXACT)
    frame.render(player.BATCHED, threads)
    tune = frame.autotune(3)
    tune_times = _tune_times()
    print(f"# {name} autotune: {tune} {tune_times}", file=sys.stderr, flush=True)
    frame.clear()
    frame.render(player.BATCHED, threads)
    el, per, sts, issue = _steps(frame, player.BATCHED, steps, threads)
    st = sts[-1]
    # device-only replays with exact sizing: the render stage then launches
    # only the kernels and workgroups the frame uses (learned sizing launches
    # every render kernel over upper-bound grids, idle workgroups included)
    B.set_sizing(B.SIZING_EXACT)
    frame.replay()
    rep_ms, rep_render = [], []
    for _ in range(steps):
        ms, rst = frame.replay()
        rep_ms.append(ms)
        rep_render.append(rst["render_ms"])
    out = {
        "workload": config["workload"],
        "mpix_per_s": round(st["pixels"] * steps / el / 1e6, 3),
        "paths_per_s": round(config["paths"] * steps / el, 1),
        "ms_per_frame": round(el * 1e3 / steps, 3),
        "frame_breakdown_ms": {"record": round(float(np.median(issue)), 3), "merge": round(st["merge_ms"], 3),
                               "upload": round(st["upload_ms"], 3), "of_which_pack": round(st["pack_ms"], 3),
                               "device_stages": round(st["total_ms"], 3),
                               "render": round(st["render_ms"], 3)},
        "device_only_ms_per_frame": round(float(np.mean(rep_ms)), 3),
        "device_only_mpix_per_s": round(st["pixels"] / (float(np.mean(rep_ms)) * 1e-3) / 1e6, 3),
        "pixels_per_frame": int(st["pixels"]),
        "sizing": "learned" if learned else "exact",
        "fallback": int(st["fallback"]),
        "schedule": _schedule(st),
        "autotune": {"picked": B.SCHEDULE_NAMES[tune["schedule"]], "wgs_per_cu": tune["wgs_per_cu"],
                     "best_render_ms": round(tune["best_ms"], 4), "grid_render_ms": round(tune["grid_ms"], 4),
                     "render_ms_by_schedule": tune_times},
    }
    roof = _roofline(scenes, rep_render, rst, name)
    roof["measured_on"] = "device-resident replays, exact sizing"
    frame.close()
    if name in ("blit4k", "mask4k"):
        # the reference library on one host core, the same frame (its own
        # begin -> end clock, best of the passes that fit 10 s)
        ref_ms, passes = _reference_passes(scenes, 0, 10.0, 3)
        out["cpu_reference"] = {"ms_per_frame": round(ref_ms, 3),
                                "mpix_per_s": round(st["pixels"] / (ref_ms * 1e-3) / 1e6, 3),
                                "threads": 0, "passes": passes}
    return out, roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c5", "srcover4k", "blit4k", "mask4k"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the reference parity check")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain form")
    ap.add_argument("--no-side", action="store_true", help="skip the srcover4k, c2, blit4k and mask4k side measurements")
    ap.add_argument("--record-threads", type=int, default=0,
                    help="client threads recording the canvases' calls (default: host cores, at most 16)")
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
    from blend2d_amd import player, train_cost_model
    if B.lib().bl_amd_device_count() < 1:
        raise SystemExit("no HIP device visible")

    params = train_cost_model.load()
    if params is not None:
        B.cost_model_set(params)
    learned = params is not None
    from blend2d_amd import train_tile_model
    tile_params = train_tile_model.load()
    if tile_params is not None:
        B.tile_model_set(tile_params)

    threads = args.record_threads or max(1, min(HOST_CORES_MAX, os.cpu_count() or 1))
    print(f"# workload {args.workload}: generating", file=sys.stderr, flush=True)
    scenes, config = _workload(args.workload, args.seed)
    frame = player.Frame(scenes)

    # first frame onto cleared canvases: the parity sample (and the warm-up
    # that sizes the device buffers)
    B.set_sizing(B.SIZING_LEARNED if learned else B.SIZING_EXACT)
    frame.clear()
    _, gpu_errors, st0, r = frame.render(player.BATCHED, threads)
    if r:
        raise RuntimeError(f"first frame failed with BLResult 0x{r:X}")
    first = [frame.download(k) for k in range(len(scenes))]
    print(f"# first frame: {st0}", file=sys.stderr, flush=True)
    # the autotuner times the render stage of this resident frame under every
    # tile schedule and records the fastest for frames of its shape, which
    # every later frame of the run then uses (AUTO)
    tune = frame.autotune(3)
    tune_times = _tune_times()
    print(f"# autotune: {tune}\n# {B.autotune_log().strip()}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        frame.render(player.BATCHED, threads)
    el, per, sts, issue = _steps(frame, player.BATCHED, args.steps, threads)
    st = sts[-1]
    pixels = st["pixels"]
    value = pixels * args.steps / el / 1e6
    canvas_px = sum(sc.width * sc.height for sc in scenes)
    config["sizing"] = "learned" if learned else "exact"
    config["record_threads"] = threads
    print(f"# batched: {el * 1e3 / args.steps:.3f} ms/frame", file=sys.stderr, flush=True)

    # device-only: the same frame replayed from its resident copy, eagerly
    # and as a captured HIP graph (the first graph replay sizes exactly and
    # records the counts, the second captures; both untimed)
    rep_ms, rep_render = [], []
    for _ in range(args.steps):
        ms, rst = frame.replay()
        rep_ms.append(ms)
        rep_render.append(rst["render_ms"])
    B.set_graph_replay(True)
    frame.replay()
    frame.replay()
    graph_ms = []
    for _ in range(args.steps):
        ms, gst = frame.replay()
        graph_ms.append(ms)
    B.set_graph_replay(False)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": config,
        "paths_per_s": round(config["paths"] * args.steps / el, 1),
        "canvas_mpix_per_s": round(canvas_px * args.steps / el / 1e6, 3),
        "pixels_per_frame": int(pixels),
        "frame_ms": {"min": round(min(per), 3), "median": round(float(np.median(per)), 3)},
        "frame_breakdown_ms": {"record": round(float(np.median(issue)), 3), "merge": round(st["merge_ms"], 3),
                               "upload": round(st["upload_ms"], 3), "of_which_pack": round(st["pack_ms"], 3),
                               "device_stages": round(st["total_ms"], 3),
                               "render": round(st["render_ms"], 3)},
        "fallback": int(st["fallback"]),
        "syncs": int(st["syncs"]),
        "schedule": _schedule(st),
        "autotune": {"picked": B.SCHEDULE_NAMES[tune["schedule"]], "wgs_per_cu": tune["wgs_per_cu"],
                     "candidates": tune["candidates"], "best_render_ms": round(tune["best_ms"], 4),
                     "grid_render_ms": round(tune["grid_ms"], 4),
                     "render_ms_by_schedule": tune_times,
                     "tile_model": tile_params is not None},
        "device_only": {"ms_per_frame": round(float(np.mean(rep_ms)), 3),
                        "mpix_per_s": round(pixels / (float(np.mean(rep_ms)) * 1e-3) / 1e6, 3),
                        "graph_ms_per_frame": round(float(np.mean(graph_ms)), 3),
                        "graph_launched": bool(gst["graph"]),
                        "graph_mpix_per_s": round(gst["pixels"] / (float(np.mean(graph_ms)) * 1e-3) / 1e6, 3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    line["roofline"] = _roofline(scenes, [s["render_ms"] for s in sts], st, args.workload)
    line["roofline"]["measured_on"] = "the timed end-to-end frames"
    if any(gpu_errors):
        line["errors"] = sorted({int(e) for e in gpu_errors if e})

    # ablation: the same batched frame with exact sizing, and with every call
    # recorded on one client thread
    B.set_sizing(B.SIZING_EXACT)
    frame.render(player.BATCHED, threads)
    xel, xper, xsts, xissue = _steps(frame, player.BATCHED, args.steps, threads)
    B.set_sizing(B.SIZING_LEARNED if learned else B.SIZING_EXACT)
    oel, oper, osts, oissue = _steps(frame, player.BATCHED, args.steps, 1)
    bvp = {
        "batched_ms_per_frame": round(el * 1e3 / args.steps, 3),
        "batched_mpix_per_s": round(value, 3),
        "batched_sizing": config["sizing"],
        "batched_exact_ms_per_frame": round(xel * 1e3 / args.steps, 3),
        "batched_exact_mpix_per_s": round(xsts[-1]["pixels"] * args.steps / xel / 1e6, 3),
        "batched_exact_frame_breakdown_ms": {"record": round(float(np.median(xissue)), 3),
                                             "upload": round(xsts[-1]["upload_ms"], 3),
                                             "device_stages": round(xsts[-1]["total_ms"], 3)},
        "learned_over_exact": round((xel / args.steps) / (el / args.steps), 4),
        "batched_one_record_thread_ms_per_frame": round(oel * 1e3 / args.steps, 3),
        "batched_one_record_thread_mpix_per_s": round(osts[-1]["pixels"] * args.steps / oel / 1e6, 3),
        "batched_one_record_thread_record_ms": round(float(np.median(oissue)), 3),
    }
    if not args.no_plain:
        # plain form: one context, one synchronous flush per canvas, one
        # workgroup per tile (no work queue, no autotuning)
        B.set_sizing(B.SIZING_EXACT)
        B.set_render_schedule(B.SCHEDULE_GRID)
        frame.clear()
        frame.render(player.PLAIN)
        plain_out = [frame.download(k) for k in range(len(scenes))]
        pel, pper, psts, pissue = _steps(frame, player.PLAIN, 1)
        B.set_render_schedule(B.SCHEDULE_AUTO)
        bvp.update({
            "plain_ms_per_frame": round(pel * 1e3, 3),
            "plain_mpix_per_s": round(psts[-1]["pixels"] / pel / 1e6, 3),
            "plain_paths_per_s": round(config["paths"] / pel, 1),
            "plain_frame_breakdown_ms": {"record": round(pissue[-1], 3), "upload": round(psts[-1]["upload_ms"], 3),
                                         "device_stages": round(psts[-1]["total_ms"], 3)},
            "batched_over_plain": round(value / (psts[-1]["pixels"] / pel / 1e6), 3),
            "plain_vs_batched_pixels": _compare(plain_out, first),
        })
        print(f"# plain: {pel * 1e3:.1f} ms/frame", file=sys.stderr, flush=True)
    line["batched_vs_plain"] = bvp
    B.set_sizing(B.SIZING_LEARNED if learned else B.SIZING_EXACT)
    frame.close()

    if not args.no_side:
        for name in ("srcover4k", "c2", "blit4k", "mask4k"):
            if name == args.workload:
                continue
            print(f"# side workload {name}", file=sys.stderr, flush=True)
            side, roof = _side_workload(name, args.seed, args.steps, learned, threads)
            line[f"roofline_{name}"] = roof
            line[name] = side

    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu, parity = _cpu_baseline(scenes, first, gpu_errors, pixels, args.cpu_seconds)
        line["cpu_baseline"] = cpu
        line["parity"] = parity
    B.set_sizing(B.SIZING_EXACT)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
