"""Throughput of the ochre rasterisation path on one MI355X.

A step is one pass of the hot path (fill -> finish for every path) over one
batch of synthetic paths already resident in HBM: by default the glyph run
(10,000 paths, 1-4 closed contours of 8-40 segments, 40% line / 40%
quadratic / 20% cubic, 12-48 px, 4096x4096 canvas).  Both pipelines are
timed: `plain` (the five-kernel pipeline) and `batched` (the persistent
work-queue kernel replayed from a HIP graph); `value` is the batched rate.

Prints one JSON line (the bench contract) to stdout; details go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "glyph paths/s + Mtiles/s on 1 MI355X (batched vs plain), % of HBM peak"


WORKLOAD_INFO = {
    "glyph_run": {"canvas": 4096, "segments": "1-4 contours x 8-40 (40% line, 40% quad, 20% cubic)",
                  "size_px": "12-48"},
    "star_blobs": {"canvas": 4096, "segments": "one contour of log-uniform 1k-100k cubics",
                   "radius_px": "200-2000, noise 15%"},
    "scenes": {"scenes": 256, "canvas": 2048, "paths_per_scene": "log-uniform 1-5000, 5% blobs of 1k-20k cubics",
               "transform": "affine: scale 0.5-4, rotation, shear <= 0.2"},
    "adjoint_glyphs": {"canvas": 1024, "segments": "16 cubics", "size_px": "32-64"},
    "stroked_glyphs": {"canvas": 4096, "segments": "glyph geometry in px, stroked 0.75-3 px"},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(batch, n_tiles, n_spans):
    """Bytes the path's interface must move: the input path (verbs 1 B,
    points 8 B, weights 4 B, transform 24 B per path) and the output (64 B
    alpha + 4 B position per tile, 6 B per span, 24 B of per-path ranges)."""
    b_in = batch.algorithmic_input_bytes()
    b_out = 68 * n_tiles + 6 * n_spans + 24 * batch.n_paths
    return b_in, b_out


# Above this many verbs the CPU baseline is a sample and has no full result.
CPU_FULL_MAX_VERBS = 2_000_000


def cpu_baseline(batch, seconds):
    """The oracle (a sequential C port of the reference Rasterizer) on one
    host core, repeated over the same batch until `seconds` have elapsed.  A
    batch too large for that budget (more than 2M verbs) is sampled instead:
    consecutive paths from the start, one at a time, until the budget is
    spent (no oracle result is returned then)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    oracle_ref.load()
    if batch.verbs.size > CPU_FULL_MAX_VERBS:
        t0 = time.perf_counter()
        done = 0
        while done < batch.n_paths and (time.perf_counter() - t0 < seconds or done == 0):
            oracle_ref.run_batch(_range_batch(batch, done, done + 1), stroke=batch.widths is not None)
            done += 1
        el = time.perf_counter() - t0
        return {"value": done / el, "unit": "paths/s", "cores": 1, "kind": "port",
                "sample": f"{batch.name}: the first {done} of {batch.n_paths} paths in {el:.1f} s (single thread)"}, None
    reps, t0 = 0, time.perf_counter()
    times = []
    ref = None
    while True:
        a = time.perf_counter()
        ref = oracle_ref.run_batch(batch, stroke=batch.widths is not None)
        times.append(time.perf_counter() - a)
        reps += 1
        if time.perf_counter() - t0 >= seconds and reps >= 2:
            break
    times.sort()
    med = times[len(times) // 2]
    return {"value": batch.n_paths / med, "unit": "paths/s", "cores": 1, "kind": "port",
            "sample": f"{batch.name}: all {batch.n_paths} paths, median of {reps} passes (single thread)"}, ref


_ALL_CORES_BATCH = None


def _range_batch(batch, a, b):
    import numpy as np
    from ochre_amd import workloads as W
    vo, po, wo = batch.verb_off, batch.point_off, batch.weight_off
    sub = W.PathBatch(batch.verbs[vo[a]:vo[b]], vo[a:b + 1] - vo[a], batch.points[po[a]:po[b]],
                      po[a:b + 1] - po[a], batch.weights[wo[a]:wo[b]], wo[a:b + 1] - wo[a],
                      np.ascontiguousarray(batch.xforms[a:b]))
    if batch.widths is not None:
        sub.widths = np.ascontiguousarray(batch.widths[a:b])
    return sub


def _all_cores_worker(ab):
    a, b = ab
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref
    sub = _range_batch(_ALL_CORES_BATCH, a, b)
    t = time.perf_counter()
    oracle_ref.run_batch(sub, stroke=sub.widths is not None)
    return time.perf_counter() - t


def cpu_all_cores(batch, passes=3):
    """The oracle on every host core (one process per core, a contiguous
    range of paths each, forked before any core pinning); the pass time is the
    slowest worker's."""
    import multiprocessing as mp
    global _ALL_CORES_BATCH
    n = max(1, min(len(os.sched_getaffinity(0)), 64))
    _ALL_CORES_BATCH = batch
    cuts = [round(batch.n_paths * k / n) for k in range(n + 1)]
    ranges = [(cuts[k], cuts[k + 1]) for k in range(n) if cuts[k + 1] > cuts[k]]
    best = None
    with mp.get_context("fork").Pool(len(ranges)) as pool:
        for _ in range(passes):
            t = max(pool.map(_all_cores_worker, ranges))
            best = t if best is None else min(best, t)
    return {"value": batch.n_paths / best, "unit": "paths/s", "cores": len(ranges), "kind": "port",
            "sample": f"{batch.name}: all {batch.n_paths} paths split over {len(ranges)} processes, best of {passes}"}


def api_latency(reps=20):
    """The drop-in Rasterizer API end to end on the README path: fill, finish
    (upload, device pipeline, download) and the TileBuilder callbacks."""
    from ochre_amd import EventRecorder, Rasterizer, workloads as W
    b = W.readme_path()
    v, p, w, xf = b.path(0)
    times = []
    for _ in range(reps + 3):
        t = time.perf_counter()
        r = Rasterizer()
        r.fill_arrays(v, p, w, xf)
        rec = EventRecorder()
        r.finish(rec)
        times.append(time.perf_counter() - t)
    times = sorted(times[3:])
    return {"ms": times[len(times) // 2] * 1e3, "workload": "readme_path (1 path, 100 tiles, 23 spans)",
            "what": "Rasterizer.fill + finish through the C ABI, TileBuilder callbacks included, median"}


def plain_per_path(ctx, batch, n=200):
    """The plain pipeline launched once per path (the drop-in use: one
    Rasterizer per path, as examples/svg.rs:143-145 runs it), device-resident
    input, on the first n paths."""
    import torch
    from ochre_amd.device import DeviceOutput, rasterize_plain, upload
    n = min(n, batch.n_paths)
    subs = [upload(_range_batch(batch, i, i + 1)) for i in range(n)]
    out = DeviceOutput(1, 1 << 20, 1 << 16)
    rasterize_plain(ctx, subs[0], out)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for sb in subs:
        rasterize_plain(ctx, sb, out)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    return {"paths_per_s": n / el, "ms_per_path": el / n * 1e3,
            "sample": f"the first {n} paths of {batch.name}, one ochre_rasterize_plain call each"}


def traffic_bytes(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or
    None when no summary covers them."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload.split('(')[0]}_traffic.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":", 1)[1])
    return None


def adjoint_bench(args):
    """The differentiable configuration: loss and gradients (w.r.t. every
    control point and transform) of 4,096 16-cubic glyphs against a target
    rendered from control points perturbed by N(0, 0.5 px), through
    ochre_coverage_adjoint (the forward runs inside it).  The target render is
    made once, before the timed steps.  CPU baseline: the f64 oracle adjoint
    of the same formula on one core, on a sample of paths."""
    import numpy as np
    import torch
    from ochre_amd import workloads as W
    from ochre_amd.device import Context, DeviceOutput, coverage_adjoint, rasterize_batched, upload
    batch = W.adjoint_glyphs()
    rng = np.random.default_rng(900)
    tb = W.PathBatch(batch.verbs, batch.verb_off, batch.points.copy(), batch.point_off, batch.weights,
                     batch.weight_off, batch.xforms, name=batch.name + "[target]")
    scale = np.repeat(batch.xforms[:, 0], np.diff(batch.point_off))[:, None]
    tb.points += (rng.normal(0, 0.5, tb.points.shape) / scale).astype(np.float32)
    ctx = Context(0)
    db = upload(batch)
    tgt = DeviceOutput(tb.n_paths, tb.n_paths * 128, tb.n_paths * 16, coverage=True)
    rasterize_batched(ctx, upload(tb), tgt)
    torch.cuda.synchronize()
    fn = lambda: coverage_adjoint(ctx, db, tgt)  # noqa: E731
    ctx.set_timing(True)
    ms, wall = time_steps(fn, args.steps, args.warmup)
    kt, kn = ctx.timing("k_adjoint")
    ctx.set_timing(False)
    kms = kt / max(kn, 1)
    loss, dp, dx = fn()
    torch.cuda.synchronize()
    # interface bytes: the input paths, the target's f32 coverage, the
    # gradients (8 B per point, 24 B per transform) and the loss per path
    b_in = batch.algorithmic_input_bytes() + 4 * 64 * tgt.n_tiles + 6 * tgt.n_spans
    b_out = 8 * batch.points.shape[0] + 24 * batch.n_paths + 4 * batch.n_paths
    achieved = (b_in + b_out) / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    log(f"adjoint: {ms:.3f} ms/step, {batch.n_paths / ms * 1e3:,.0f} paths/s, k_adjoint {kms:.3f} ms, "
        f"loss {float(loss.sum()):.4f}")
    line = {"metric": "adjoint paths/s (f32 coverage loss and gradients w.r.t. control points and transforms)",
            "value": batch.n_paths / ms * 1e3, "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": dict({"workload": batch.name, "paths": batch.n_paths, "target": "control points + N(0, 0.5 px)",
                            "target_tiles": tgt.n_tiles}, **WORKLOAD_INFO["adjoint_glyphs"]),
            "kernels_ms": {"k_adjoint": kms},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_adjoint", "kernel_ms": kms,
                         "algorithmic_bytes": b_in + b_out}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ref as O
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except Exception:
            pass
        txy, tcov = tgt.tile_xy.cpu().numpy(), tgt.tile_cov.cpu().numpy()
        sxy, sw = tgt.span_xy.cpu().numpy(), tgt.span_w.cpu().numpy().view(np.uint16)
        ts, tc = tgt.path_tile_start.cpu().numpy(), tgt.path_tile_count.cpu().numpy()
        ss, sc = tgt.path_span_start.cpu().numpy(), tgt.path_span_count.cpu().numpy()
        t0 = time.perf_counter()
        done = 0
        while done < batch.n_paths and (time.perf_counter() - t0 < args.cpu_seconds or done == 0):
            v, p, w, xf = batch.path(done)
            t = O.Target(txy[ts[done]:ts[done] + tc[done]], tcov[ts[done]:ts[done] + tc[done]],
                         sxy[ss[done]:ss[done] + sc[done]], sw[ss[done]:ss[done] + sc[done]])
            O.path_loss64(v, p, w, xf, t, grad=True)
            done += 1
        el = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": done / el, "unit": "paths/s", "cores": 1, "kind": "port",
                                "sample": f"{batch.name}: the first {done} paths, f64 oracle adjoint (single thread)"}
        line["vs_cpu_1core"] = line["value"] / line["cpu_baseline"]["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps * 1e3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="glyph_run")
    ap.add_argument("--mode", default="auto", choices=["auto", "plain", "batched"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--params", default=None, metavar="W,G,TC",
                    help="batched pipeline parameters (waves/CU, paths per pop, tiles per chunk) instead of autotuning")
    ap.add_argument("--dense", default=None, metavar="G,W",
                    help="dense kernel launch shape (workgroups per CU, waves per workgroup) instead of autotuning")
    ap.add_argument("--dense-small", type=int, default=None, metavar="G",
                    help="workgroups per CU of the small-grid dense instance (0: off) instead of autotuning")
    ap.add_argument("--dense-large", type=int, default=None, choices=[0, 1],
                    help="the 256-tile dense instance off (0) or on (1) instead of autotuning")
    ap.add_argument("--no-dense", action="store_true", help="the ordered batched kernel alone (bit-exact sums)")
    ap.add_argument("--no-graph", action="store_true", help="launch the batched kernel directly, not from a graph")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the coverage adjoint on the differentiable configuration instead")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty(n, dtype=torch.uint8, device="cuda")
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "copies": 5}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build's path runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        adjoint_bench(args)
        return

    import torch

    from ochre_amd import workloads as W
    from ochre_amd.device import (Context, DeviceOutput, batched_params, rasterize_batched, rasterize_plain,
                                  set_batched_params, set_dense_params, upload)

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    batch = W.WORKLOADS[args.workload]()
    log(f"workload {batch.name}: paths={batch.n_paths} verbs={batch.verbs.size} points={batch.points.shape[0]}")
    ctx = Context(0)
    db = upload(batch)
    torch.cuda.synchronize()
    stroking = batch.widths is not None
    if stroking:  # Rasterizer::stroke for every path: expansion + fill on the device
        from ochre_amd.device import stroke
        wd = torch.from_numpy(batch.widths).to("cuda")
        run_plain = lambda o: stroke(ctx, db, wd, o, batched=False)  # noqa: E731
        run_batched = lambda o: stroke(ctx, db, wd, o, batched=True)  # noqa: E731
    else:
        run_plain = lambda o: rasterize_plain(ctx, db, o)  # noqa: E731
        run_batched = lambda o: rasterize_batched(ctx, db, o)  # noqa: E731

    # size the outputs with one call
    out = DeviceOutput(batch.n_paths, batch.n_paths * 64, batch.n_paths * 8)
    try:
        run_plain(out)
    except Exception as e:  # capacity: grow once
        if getattr(e, "status", None) != -4:
            raise
        out = DeviceOutput(batch.n_paths, out.n_tiles, out.n_spans)
        run_plain(out)
    torch.cuda.synchronize()
    n_tiles, n_spans = out.n_tiles, out.n_spans
    b_in, b_out = algorithmic_bytes(batch, n_tiles, n_spans)
    log(f"tiles={n_tiles} spans={n_spans} algorithmic bytes in={b_in} out={b_out}")

    have_batched = hasattr(ctx.lib, "ochre_rasterize_batched")
    tuned = None
    if have_batched and args.mode != "plain":
        from ochre_amd.device import autotune
        dg, dw = (int(x) for x in args.dense.split(",")) if args.dense else (4, 4)
        set_dense_params(ctx, not args.no_dense, dg, dw)
        if args.dense_small is not None:
            from ochre_amd.device import set_dense_small
            set_dense_small(ctx, args.dense_small)
        if args.dense_large is not None:
            from ochre_amd.device import set_dense_large
            set_dense_large(ctx, bool(args.dense_large))
        if args.params or (args.dense and not args.no_dense):
            w, g, tc = (int(x) for x in (args.params or "8,1,32").split(","))
            set_batched_params(ctx, w, g, tc, graph=not args.no_graph)
            tuned = batched_params(ctx)
            log(f"batched parameters set: {tuned}")
        elif stroking:
            tuned = batched_params(ctx)  # the fill of the outlines runs with the defaults
        else:
            best, allm = autotune(ctx, db, out, reps=3)
            if args.no_graph:
                set_batched_params(ctx, best["waves_per_cu"], best["paths_per_pop"], best["tiles_per_chunk"],
                                   graph=False)
            tuned = batched_params(ctx)
            tuned["ms"] = best["ms"]
            log(f"autotune: best {best} over {len(allm)} configurations")
            for m in sorted(allm, key=lambda m: m["ms"])[:5]:
                log(f"  {m}")
    res = {}
    modes = ["plain"] + (["batched"] if have_batched and args.mode != "plain" else [])
    if args.mode == "batched" and not have_batched:
        raise SystemExit("batched pipeline not built")
    outputs = {}
    for mode in modes:
        if mode == "plain":
            fn = lambda: run_plain(out)  # noqa: E731
            ctx.set_timing(True)
            ev_ms, wall_ms = time_steps(fn, args.steps, args.warmup)
            kms, kn = ctx.timing("k_tile_rows")
            ctx.set_timing(False)
            kernels = {"k_tile_rows": kms / max(kn, 1)}
            kname = "k_tile_rows"
        else:
            fn = lambda: run_batched(out)  # noqa: E731
            # the production launch (graph replay unless --no-graph)
            ev_ms, wall_ms = time_steps(fn, args.steps, args.warmup)
            # kernel durations: the same launches made directly, with HIP
            # events around each kernel on the stream they run on
            bp = batched_params(ctx)
            set_batched_params(ctx, bp["waves_per_cu"], bp["paths_per_pop"], bp["tiles_per_chunk"], graph=False)
            ctx.set_timing(True)
            time_steps(fn, args.steps, 1)
            kernels = {}
            for k in ("k_dense", "k_fused", "k_row_tiles", "k_tile_rows"):
                t, n = ctx.timing(k)
                if n:
                    kernels[k] = t / n  # the last two: the plain pipeline over the handed-on paths
            stats = {}
            for k in ("lines", "entries", "runs", "rows", "windows", "windows_used"):
                t, n = ctx.timing("pass." + k)
                if n:
                    stats[k] = int(t) // (args.steps + 1)
            if stats:
                stats["passes"] = n // (args.steps + 1)
                log(f"handed-on paths, per call: {stats}")
            ctx.set_timing(False)
            set_batched_params(ctx, bp["waves_per_cu"], bp["paths_per_pop"], bp["tiles_per_chunk"],
                               graph=bool(bp["graph"]))
            kname = "k_dense" if bp.get("dense") else "k_fused"
        torch.cuda.synchronize()
        if not args.no_cpu and batch.verbs.size <= CPU_FULL_MAX_VERBS:  # checked against the oracle below
            outputs[mode] = out.to_host()
        res[mode] = {"ms": ev_ms, "wall_ms": wall_ms, "kernel": kname, "kernel_ms": kernels.get(kname, 0.0),
                     "kernels_ms": kernels}
        log(f"{mode}: {ev_ms:.3f} ms/step (wall {wall_ms:.3f}), {batch.n_paths / ev_ms * 1e3:,.0f} paths/s, "
            f"kernels (ms/launch) {kernels}")

    head = "batched" if "batched" in res else "plain"
    ms = res[head]["ms"]
    paths_s = batch.n_paths / ms * 1e3
    kms = res[head]["kernel_ms"]
    # roofline of the dominant kernel: the batched kernels move the whole
    # interface (input paths + output tiles/spans); the plain pipeline's tile
    # kernel writes the output and reads nothing of the interface.
    alg = (b_in + b_out) if head == "batched" else b_out
    kname = res[head]["kernel"]
    traffic = traffic_bytes(kname, batch.name)
    if kms < 0.5 * ms:
        # no single kernel dominates (e.g. the batched scenes, whose large
        # paths go through the plain pipeline): the whole step is the unit
        kname, kms, alg, traffic = "step", ms, b_in + b_out, None
    achieved = alg / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    line = {
        "metric": METRIC,
        "value": paths_s,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict({"workload": batch.name, "paths": batch.n_paths, "pipeline": head},
                       **WORKLOAD_INFO.get(args.workload, {})),
        "mtiles_per_s": n_tiles / ms * 1e-3,
        "tiles": n_tiles,
        "spans": n_spans,
        "plain_paths_per_s": batch.n_paths / res["plain"]["ms"] * 1e3,
        "batched_paths_per_s": (batch.n_paths / res["batched"]["ms"] * 1e3) if "batched" in res else None,
        "batched_over_plain": (res["plain"]["ms"] / res["batched"]["ms"]) if "batched" in res else None,
        "hbm_frac_end_to_end": (b_in + b_out) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "batched_params": tuned,
        "kernels_ms": {m: r["kernels_ms"] for m, r in res.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kms, "algorithmic_bytes": alg},
    }
    line["rasterizer_api_latency"] = api_latency()
    if not stroking:
        try:
            pp = plain_per_path(ctx, batch)
            line["plain_per_path_launch"] = pp
            line["batched_over_plain_per_path"] = paths_s / pp["paths_per_s"]
        except Exception as e:  # a path too large for the probe's output: report without it
            log(f"per-path plain probe skipped: {e}")
    if not args.no_cpu:
        try:
            line["cpu_baseline_all_cores"] = cpu_all_cores(batch)
        except Exception as e:  # a restricted host: report the single-core figure only
            log(f"all-core CPU baseline skipped: {e}")
        cb, ref = cpu_baseline(batch, args.cpu_seconds)
        cb["cpu"] = platform.processor() or platform.machine()
        line["cpu_baseline"] = cb
        line["vs_cpu_1core"] = paths_s / cb["value"]
        # the oracle's result from the baseline leg checks every timed output
        if ref is not None:
            from parity import compare
            line["parity"] = {m: compare(ref, o, max_lsb=1) for m, o in outputs.items()}
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
