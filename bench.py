"""Throughput of the ochre rasterisation path on one MI355X.

A step is one pass of the hot path (fill -> finish for every path) over one
batch of synthetic paths already resident in HBM: by default the glyph run
(10,000 paths, 1-4 closed contours of 8-40 segments, 40% line / 40%
quadratic / 20% cubic, 12-48 px, 4096x4096 canvas).  Both pipelines are
timed: `plain` (the five-kernel pipeline) and `batched` (the persistent
work-queue kernels replayed from a HIP graph); `value` is the batched rate.

The same run also times the other workloads ("secondary": the 200 star blobs,
the 256 batched scenes, the stroked glyphs and the coverage adjoint), each
with its own kernel split, roofline and single-core CPU sample.  CPU
baselines run first, before anything touches the GPU, so that the forked
all-core workers never inherit a device context.

Prints one JSON line (the bench contract) to stdout; details go to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "glyph paths/s + Mtiles/s on 1 MI355X (batched vs plain), % of HBM peak"
CPU_POOL_MAX = 15  # forked oracle workers (a 16-core share of the host, one core left to the parent)


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

# The secondary workloads of a default run: (name, steps cap, CPU seconds).
SECONDARY = [("star_blobs", 5, 3.0), ("scenes", 3, 3.0), ("stroked_glyphs", 10, 2.0), ("adjoint_glyphs", 10, 2.0)]


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


def _pin_one_core():
    """Pins this process to its first allowed core; returns the previous set."""
    prev = os.sched_getaffinity(0)
    try:
        os.sched_setaffinity(0, {sorted(prev)[0]})
    except OSError:
        pass
    return prev


def cpu_baseline(batch, seconds):
    """The oracle (a sequential C port of the reference Rasterizer) on one
    host core, repeated over the same batch until `seconds` have elapsed.  A
    batch too large for that budget (more than 2M verbs) is sampled instead:
    consecutive paths from the start, one at a time, until the budget is
    spent (no oracle result is returned then)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref
    oracle_ref.load()
    prev = _pin_one_core()
    try:
        stroke = batch.widths is not None
        if batch.verbs.size > CPU_FULL_MAX_VERBS:
            t0 = time.perf_counter()
            done = 0
            while done < batch.n_paths and (time.perf_counter() - t0 < seconds or done == 0):
                oracle_ref.run_batch(_range_batch(batch, done, done + 1), stroke=stroke)
                done += 1
            el = time.perf_counter() - t0
            return {"value": done / el, "unit": "paths/s", "cores": 1, "kind": "port",
                    "sample": f"{batch.name}: the first {done} of {batch.n_paths} paths in {el:.1f} s "
                              f"(single thread)"}, None
        reps, t0 = 0, time.perf_counter()
        times = []
        ref = None
        while True:
            a = time.perf_counter()
            ref = oracle_ref.run_batch(batch, stroke=stroke)
            times.append(time.perf_counter() - a)
            reps += 1
            if time.perf_counter() - t0 >= seconds and reps >= 2:
                break
        times.sort()
        med = times[len(times) // 2]
        return {"value": batch.n_paths / med, "unit": "paths/s", "cores": 1, "kind": "port",
                "sample": f"{batch.name}: all {batch.n_paths} paths, median of {reps} passes (single thread)"}, ref
    finally:
        os.sched_setaffinity(0, prev)


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
    """The oracle on the host's cores (one forked process per core, at most
    CPU_POOL_MAX, a contiguous range of paths each); the pass time is the
    slowest worker's.  Called before the process touches the GPU."""
    import multiprocessing as mp
    global _ALL_CORES_BATCH
    n = max(1, min(len(os.sched_getaffinity(0)), CPU_POOL_MAX))
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
    """The drop-in Rasterizer API on the README path: (1) fill + finish with
    the TileBuilder callbacks in Python (what a Python caller pays), and (2)
    ochre_rasterizer_finish alone with a NULL builder, timed around the C
    call (upload, device pipeline, download: the library's own cost), on the
    README path and on three glyphs of the glyph run."""
    import numpy as np
    from ochre_amd import EventRecorder, Rasterizer, _lib
    from ochre_amd import workloads as W
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
    lib = _lib.load()

    def finish_c_us(v, p, xf):
        v = np.ascontiguousarray(v)
        p = np.ascontiguousarray(p)
        vv = v.ctypes.data_as(ctypes.c_void_p)
        pp = p.ctypes.data_as(ctypes.c_void_p)
        t = _lib.Transform((ctypes.c_float * 4)(*xf[:4]), (ctypes.c_float * 2)(*xf[4:]))
        fin = []
        for _ in range(reps + 3):
            h = lib.ochre_rasterizer_new()
            _lib.check(lib.ochre_rasterizer_fill(h, vv, v.size, pp, p.shape[0], None, 0, ctypes.byref(t)), "fill")
            a = time.perf_counter()
            _lib.check(lib.ochre_rasterizer_finish(h, None), "finish")
            fin.append(time.perf_counter() - a)
        fin = sorted(fin[3:])
        return fin[len(fin) // 2] * 1e6

    g = W.glyph_run(64, seed=5)
    glyphs = {f"glyph_run(64, seed=5)[{i}] ({int(g.verb_off[i + 1] - g.verb_off[i])} verbs)":
              finish_c_us(*g.path(i)[:2], g.path(i)[3]) for i in range(3)}
    return {"ms": times[len(times) // 2] * 1e3, "finish_c_us": finish_c_us(v, p, xf),
            "finish_c_us_glyphs": glyphs,
            "workload": "readme_path (1 path, 100 tiles, 23 spans); single glyphs",
            "what": "ms: Rasterizer.fill + finish with Python TileBuilder callbacks; finish_c_us: "
                    "ochre_rasterizer_finish with a NULL builder (C ABI, no callbacks); medians"}


def svg_demo_latency(reps=3):
    """The demo's loop on the Ghostscript tiger (examples/svg.rs:140-156): a
    Rasterizer per fill or stroke, each finished into the builder, through the
    C ABI with a NULL builder (the library's own cost: upload, device work,
    read-back): one ochre_rasterizer_finish per draw, then all of them in one
    ochre_rasterizer_finish_many.  Medians of `reps` passes."""
    import gzip
    from ochre_amd import _lib, svg
    path = os.path.join(ROOT, "tests", "golden", "Ghostscript_Tiger.svg.gz")
    if not os.path.exists(path):
        return None
    b = svg.draw_batch(svg.parse(gzip.open(path).read()))
    lib = _lib.load()
    P = ctypes.c_void_p

    def make():
        hs = []
        for i in range(b.n_paths):
            v, p, w, xf = b.path(i)
            h = lib.ochre_rasterizer_new()
            t = _lib.Transform((ctypes.c_float * 4)(*xf[:4]), (ctypes.c_float * 2)(*xf[4:]))
            vv, pp = v.ctypes.data_as(P), p.ctypes.data_as(P)
            if b.strokes[i]:
                st = lib.ochre_rasterizer_stroke(h, vv, v.size, pp, p.shape[0], None, 0, float(b.widths[i]),
                                                 ctypes.byref(t))
            else:
                st = lib.ochre_rasterizer_fill(h, vv, v.size, pp, p.shape[0], None, 0, ctypes.byref(t))
            _lib.check(st, "fill/stroke")
            hs.append(h)
        return hs

    one, many = [], []
    for _ in range(reps + 1):
        hs = make()
        t = time.perf_counter()
        for h in hs:
            _lib.check(lib.ochre_rasterizer_finish(h, None), "finish")
        one.append(time.perf_counter() - t)
        hs = make()
        arr = (P * len(hs))(*hs)
        t = time.perf_counter()
        _lib.check(lib.ochre_rasterizer_finish_many(arr, len(hs), None), "finish_many")
        many.append(time.perf_counter() - t)
    one, many = sorted(one[1:]), sorted(many[1:])
    n = b.n_paths
    return {"draws": n, "strokes": int(b.strokes.sum()),
            "finish_each_ms": one[len(one) // 2] * 1e3, "finish_each_us_per_draw": one[len(one) // 2] / n * 1e6,
            "finish_many_ms": many[len(many) // 2] * 1e3,
            "finish_many_us_per_draw": many[len(many) // 2] / n * 1e6,
            "what": "the tiger's draw list (examples/res/Ghostscript_Tiger.svg as a fixture), one Rasterizer per "
                    "draw, NULL TileBuilder; medians"}


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


def time_steps(fn, steps, warmup):
    """HIP-event time per step on the current stream (every launch of the
    library goes to that stream), and host wall time per step."""
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


def roofline(kname, kms, step_ms, alg_kernel, alg_step, workload):
    """Roofline of the dominant kernel, or of the whole step when no single
    kernel takes half of it (the batched scenes, whose large paths go through
    the plain pipeline inside the call)."""
    traffic = traffic_bytes(kname, workload)
    if kms < 0.5 * step_ms:
        kname, kms, alg, traffic = "step", step_ms, alg_step, None
    else:
        alg = alg_kernel
    achieved = alg / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname, "kernel_ms": kms,
            "algorithmic_bytes": alg}


def time_modes(ctx, batch, db, out, modes, steps, warmup, keep_outputs):
    """Times the plain and/or batched call over `batch` with the context's
    current batched parameters.  Per mode: ms per step (HIP events over the
    production launch), wall ms, and per-kernel ms per launch (HIP events
    around each kernel on its stream, in a direct-launch pass)."""
    import torch
    from ochre_amd.device import batched_params, rasterize_batched, rasterize_plain, set_batched_params, stroke
    stroking = batch.widths is not None
    if stroking:
        wd = torch.from_numpy(batch.widths).to("cuda")
        runs = {"plain": lambda o: stroke(ctx, db, wd, o, batched=False),
                "batched": lambda o: stroke(ctx, db, wd, o, batched=True)}
    else:
        runs = {"plain": lambda o: rasterize_plain(ctx, db, o), "batched": lambda o: rasterize_batched(ctx, db, o)}
    res, outputs = {}, {}
    for mode in modes:
        fn = lambda: runs[mode](out)  # noqa: E731
        if mode == "plain":
            ctx.set_timing(True)
            ev_ms, wall_ms = time_steps(fn, steps, warmup)
            kernels = {}
            for k in ("k_tile_rows", "k_row_tiles"):
                t, n = ctx.timing(k)
                if n:
                    kernels[k] = t / n
            ctx.set_timing(False)
            kname = "k_tile_rows"
        else:
            ev_ms, wall_ms = time_steps(fn, steps, warmup)
            bp = batched_params(ctx)
            set_batched_params(ctx, bp["waves_per_cu"], bp["paths_per_pop"], bp["tiles_per_chunk"], graph=False)
            ctx.set_timing(True)
            time_steps(fn, steps, 1)
            kernels = {}
            for k in ("k_dense", "k_fused", "k_row_tiles", "k_tile_rows"):
                t, n = ctx.timing(k)
                if n:
                    kernels[k] = t / n  # the last two: the plain pipeline over the handed-on paths
            stats = {}
            for k in ("lines", "entries", "runs", "rows", "windows", "windows_used"):
                t, n = ctx.timing("pass." + k)
                if n:
                    stats[k] = int(t) // (steps + 1)
            if stats:
                stats["passes"] = n // (steps + 1)
                log(f"  handed-on paths, per call: {stats}")
            ctx.set_timing(False)
            set_batched_params(ctx, bp["waves_per_cu"], bp["paths_per_pop"], bp["tiles_per_chunk"],
                               graph=bool(bp["graph"]))
            kname = "k_dense" if bp.get("dense") else "k_fused"
            if kernels.get("k_row_tiles", 0.0) > kernels.get(kname, 0.0):
                kname = "k_row_tiles"
        torch.cuda.synchronize()
        if keep_outputs:
            outputs[mode] = out.to_host()
        res[mode] = {"ms": ev_ms, "wall_ms": wall_ms, "kernel": kname, "kernel_ms": kernels.get(kname, 0.0),
                     "kernels_ms": kernels}
        log(f"  {batch.name} {mode}: {ev_ms:.3f} ms/step (wall {wall_ms:.3f}), "
            f"{batch.n_paths / ev_ms * 1e3:,.0f} paths/s, kernels (ms/launch) {kernels}")
    return res, outputs


def sized_output(ctx, batch, db, run_plain):
    """A DeviceOutput sized by one plain call (grown once on -4)."""
    import torch
    from ochre_amd.device import DeviceOutput
    out = DeviceOutput(batch.n_paths, batch.n_paths * 64, batch.n_paths * 8)
    try:
        run_plain(out)
    except Exception as e:  # capacity: grow once
        if getattr(e, "status", None) != -4:
            raise
        out = DeviceOutput(batch.n_paths, out.n_tiles, out.n_spans)
        run_plain(out)
    torch.cuda.synchronize()
    return out


# ------------------------------------------------------------------ adjoint

def adjoint_target(batch):
    import numpy as np
    from ochre_amd import workloads as W
    rng = np.random.default_rng(900)
    tb = W.PathBatch(batch.verbs, batch.verb_off, batch.points.copy(), batch.point_off, batch.weights,
                     batch.weight_off, batch.xforms, name=batch.name + "[target]")
    scale = np.repeat(batch.xforms[:, 0], np.diff(batch.point_off))[:, None]
    tb.points += (rng.normal(0, 0.5, tb.points.shape) / scale).astype(np.float32)
    return tb


def adjoint_measure(ctx, steps, warmup, cpu_seconds):
    """The differentiable configuration: loss and gradients (w.r.t. every
    control point and transform) of 4,096 16-cubic glyphs against a target
    rendered from control points perturbed by N(0, 0.5 px), through
    ochre_coverage_adjoint (the forward runs inside it).  The target render is
    made once, before the timed steps.  CPU baseline (cpu_seconds > 0): the
    f64 oracle adjoint of the same formula on one core, on the first paths;
    its losses also check the device's on those paths."""
    import numpy as np
    import torch
    from ochre_amd import workloads as W
    from ochre_amd.device import DeviceOutput, coverage_adjoint, rasterize_batched, upload
    batch = W.adjoint_glyphs()
    tb = adjoint_target(batch)
    db = upload(batch)
    tgt = DeviceOutput(tb.n_paths, tb.n_paths * 128, tb.n_paths * 16, coverage=True)
    rasterize_batched(ctx, upload(tb), tgt)
    torch.cuda.synchronize()
    fn = lambda: coverage_adjoint(ctx, db, tgt)  # noqa: E731
    ctx.set_timing(True)
    ms, wall = time_steps(fn, steps, warmup)
    kt, kn = ctx.timing("k_adjoint")
    ctx.set_timing(False)
    kms = kt / max(kn, 1)
    loss, dp, dx = fn()
    torch.cuda.synchronize()
    loss = loss.cpu().numpy()
    # interface bytes: the input paths, the target's f32 coverage, the
    # gradients (8 B per point, 24 B per transform) and the loss per path
    b_in = batch.algorithmic_input_bytes() + 4 * 64 * tgt.n_tiles + 6 * tgt.n_spans
    b_out = 8 * batch.points.shape[0] + 24 * batch.n_paths + 4 * batch.n_paths
    log(f"  adjoint: {ms:.3f} ms/step, {batch.n_paths / ms * 1e3:,.0f} paths/s, k_adjoint {kms:.3f} ms, "
        f"loss {float(loss.sum()):.4f}")
    r = {"workload": batch.name, "paths": batch.n_paths, "ms_per_step": ms, "wall_ms_per_step": wall,
         "paths_per_s": batch.n_paths / ms * 1e3, "steps": steps, "target_tiles": tgt.n_tiles,
         "kernels_ms": {"k_adjoint": kms},
         "roofline": roofline("k_adjoint", kms, ms, b_in + b_out, b_in + b_out, batch.name)}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ref as O
        txy, tcov = tgt.tile_xy.cpu().numpy(), tgt.tile_cov.cpu().numpy()
        sxy, sw = tgt.span_xy.cpu().numpy(), tgt.span_w.cpu().numpy().view(np.uint16)
        ts, tc = tgt.path_tile_start.cpu().numpy(), tgt.path_tile_count.cpu().numpy()
        ss, sc = tgt.path_span_start.cpu().numpy(), tgt.path_span_count.cpu().numpy()
        prev = _pin_one_core()
        worst = 0.0
        try:
            t0 = time.perf_counter()
            done = 0
            while done < batch.n_paths and (time.perf_counter() - t0 < cpu_seconds or done == 0):
                v, p, w, xf = batch.path(done)
                t = O.Target(txy[ts[done]:ts[done] + tc[done]], tcov[ts[done]:ts[done] + tc[done]],
                             sxy[ss[done]:ss[done] + sc[done]], sw[ss[done]:ss[done] + sc[done]])
                L = O.path_loss64(v, p, w, xf, t, grad=True)[1]
                worst = max(worst, abs(float(loss[done]) - L) / max(L, 1.0))
                done += 1
            el = time.perf_counter() - t0
        finally:
            os.sched_setaffinity(0, prev)
        r["cpu_baseline"] = {"value": done / el, "unit": "paths/s", "cores": 1, "kind": "port",
                             "sample": f"{batch.name}: the first {done} paths, f64 oracle adjoint (single thread)"}
        r["vs_cpu_1core"] = r["paths_per_s"] / r["cpu_baseline"]["value"]
        r["parity"] = {"loss_rel_err_max_vs_f64_oracle": worst, "paths": done}
    return r


def adjoint_bench(args):
    """--adjoint: the adjoint alone, as its own bench line."""
    from ochre_amd.device import Context
    ctx = Context(0)
    r = adjoint_measure(ctx, args.steps, args.warmup, 0.0 if args.no_cpu else args.cpu_seconds)
    line = {"metric": "adjoint paths/s (f32 coverage loss and gradients w.r.t. control points and transforms)",
            "value": r["paths_per_s"], "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": dict({"workload": r["workload"], "paths": r["paths"],
                            "target": "control points + N(0, 0.5 px)", "target_tiles": r["target_tiles"]},
                           **WORKLOAD_INFO["adjoint_glyphs"]),
            "kernels_ms": r["kernels_ms"], "roofline": r["roofline"],
            "cpu_baseline": r.get("cpu_baseline"), "vs_cpu_1core": r.get("vs_cpu_1core"), "parity": r.get("parity")}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- secondary

def star_blob_parity(outputs):
    """Plain (and the ordered batched call) against the oracle's committed
    per-path digests; the dense batched call's tile/span sets against them and
    its alpha against the plain output."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import path_digests
    path = os.path.join(ROOT, "tests", "golden", "star_blobs_digests.json")
    if not os.path.exists(path):
        return None
    doc = json.load(open(path))
    res = {}
    for mode, o in outputs.items():
        dig = path_digests(*o)
        sets = sum(d[:3] == (p["tiles"], p["spans"], p["sets"]) for d, p in zip(dig, doc["paths"]))
        full = sum(d == (p["tiles"], p["spans"], p["sets"], p["alpha"]) for d, p in zip(dig, doc["paths"]))
        res[mode] = {"paths": len(dig), "sets_match_oracle": sets, "alpha_match_oracle": full}
    if "plain" in outputs and "batched" in outputs and outputs["plain"][1].shape == outputs["batched"][1].shape:
        d = np.abs(outputs["batched"][1].astype(np.int16) - outputs["plain"][1].astype(np.int16))
        res["batched"]["alpha_vs_plain_max_lsb"] = int(d.max()) if d.size else 0
        res["batched"]["alpha_vs_plain_bit_exact"] = float((d == 0).mean()) if d.size else 1.0
    return res


def secondary_run(ctx, name, steps, warmup, cpu):
    """One secondary workload through plain and batched (dense kernels on,
    default launch shapes), or the adjoint."""
    import torch
    from ochre_amd import workloads as W
    from ochre_amd.device import rasterize_plain, stroke, upload
    if name == "adjoint_glyphs":
        return adjoint_measure(ctx, steps, warmup, cpu["seconds"] if cpu else 0.0)
    t0 = time.perf_counter()
    batch = W.WORKLOADS[name]()
    db = upload(batch)
    torch.cuda.synchronize()
    if batch.widths is not None:
        wd = torch.from_numpy(batch.widths).to("cuda")
        run_plain = lambda o: stroke(ctx, db, wd, o, batched=False)  # noqa: E731
    else:
        run_plain = lambda o: rasterize_plain(ctx, db, o)  # noqa: E731
    out = sized_output(ctx, batch, db, run_plain)
    keep = name == "star_blobs" or (cpu is not None and cpu.get("ref") is not None)
    res, outputs = time_modes(ctx, batch, db, out, ["plain", "batched"], steps, warmup, keep)
    b_in, b_out = algorithmic_bytes(batch, out.n_tiles, out.n_spans)
    r = {"workload": batch.name, "paths": batch.n_paths, "tiles": out.n_tiles, "spans": out.n_spans,
         "steps": steps, "algorithmic_bytes": b_in + b_out}
    for mode, m in res.items():
        alg_k = (b_in + b_out) if (mode == "batched" and m["kernel"] != "k_row_tiles") else b_out
        r[mode] = {"ms_per_step": m["ms"], "wall_ms_per_step": m["wall_ms"],
                   "paths_per_s": batch.n_paths / m["ms"] * 1e3, "mtiles_per_s": out.n_tiles / m["ms"] * 1e-3,
                   "kernels_ms": m["kernels_ms"],
                   "roofline": roofline(m["kernel"], m["kernel_ms"], m["ms"], alg_k, b_in + b_out, batch.name)}
    r["batched_over_plain"] = res["plain"]["ms"] / res["batched"]["ms"]
    if cpu is not None:
        r["cpu_baseline"] = cpu["baseline"]
        r["vs_cpu_1core"] = r["batched"]["paths_per_s"] / cpu["baseline"]["value"]
        if cpu.get("ref") is not None:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from parity import compare
            r["parity"] = {m: compare(cpu["ref"], o, max_lsb=1) for m, o in outputs.items()}
    if name == "star_blobs":
        r["parity"] = star_blob_parity(outputs)
    del out, outputs
    torch.cuda.empty_cache()
    log(f"  {name}: done in {time.perf_counter() - t0:.1f} s")
    return r


# --------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="glyph_run")
    ap.add_argument("--mode", default="auto", choices=["auto", "plain", "batched"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="time only the headline workload (no star blobs, scenes, strokes, adjoint)")
    ap.add_argument("--no-probes", action="store_true",
                    help="skip the latency probes (single fills, the SVG draw list, per-path plain launches): "
                         "a profiled run then traces the timed kernels only")
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

    from ochre_amd import workloads as W
    batch = W.WORKLOADS[args.workload]()
    log(f"workload {batch.name}: paths={batch.n_paths} verbs={batch.verbs.size} points={batch.points.shape[0]}")
    secondaries = [] if (args.no_secondary or args.workload != "glyph_run") else SECONDARY

    # ---- CPU baselines, before the first GPU call
    cb, cb_all, ref, sec_cpu = None, None, None, {}
    if not args.no_cpu:
        t0 = time.perf_counter()
        try:
            cb_all = cpu_all_cores(batch)
        except Exception as e:  # a restricted host: report the single-core figure only
            log(f"all-core CPU baseline skipped: {e}")
        cb, ref = cpu_baseline(batch, args.cpu_seconds)
        cb["cpu"] = platform.processor() or platform.machine()
        for name, _, secs in secondaries:
            if name == "adjoint_glyphs":  # needs the device's target render: timed after the GPU run
                sec_cpu[name] = {"seconds": secs}
                continue
            sb = W.WORKLOADS[name]()
            base, sref = cpu_baseline(sb, secs)
            sec_cpu[name] = {"baseline": base, "ref": sref}
            log(f"cpu {name}: {base['value']:,.1f} paths/s ({base['sample']})")
        log(f"CPU baselines: {time.perf_counter() - t0:.1f} s")

    import torch

    from ochre_amd.device import Context, batched_params, set_batched_params, set_dense_params, upload

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    ctx = Context(0)
    db = upload(batch)
    torch.cuda.synchronize()
    stroking = batch.widths is not None
    if stroking:  # Rasterizer::stroke for every path: expansion + fill on the device
        from ochre_amd.device import stroke
        wd = torch.from_numpy(batch.widths).to("cuda")
        run_plain = lambda o: stroke(ctx, db, wd, o, batched=False)  # noqa: E731
    else:
        from ochre_amd.device import rasterize_plain
        run_plain = lambda o: rasterize_plain(ctx, db, o)  # noqa: E731

    out = sized_output(ctx, batch, db, run_plain)
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
    if args.mode == "batched" and not have_batched:
        raise SystemExit("batched pipeline not built")
    modes = ["plain"] + (["batched"] if have_batched and args.mode != "plain" else [])
    res, outputs = time_modes(ctx, batch, db, out, modes, args.steps, args.warmup, keep_outputs=ref is not None)

    head = "batched" if "batched" in res else "plain"
    ms = res[head]["ms"]
    paths_s = batch.n_paths / ms * 1e3
    # roofline of the dominant kernel: the batched kernels move the whole
    # interface (input paths + output tiles/spans); the plain pipeline's tile
    # kernel writes the output and reads nothing of the interface.
    kname = res[head]["kernel"]
    alg_k = (b_in + b_out) if (head == "batched" and kname != "k_row_tiles") else b_out
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
        "roofline": roofline(kname, res[head]["kernel_ms"], ms, alg_k, b_in + b_out, batch.name),
    }
    if not args.no_probes:
        line["rasterizer_api_latency"] = api_latency()
        try:
            line["svg_demo_latency"] = svg_demo_latency()
        except Exception as e:  # report the bench line without it
            log(f"svg demo latency probe skipped: {e}")
    if not stroking and not args.no_probes:
        try:
            pp = plain_per_path(ctx, batch)
            line["plain_per_path_launch"] = pp
            line["batched_over_plain_per_path"] = paths_s / pp["paths_per_s"]
        except Exception as e:  # a path too large for the probe's output: report without it
            log(f"per-path plain probe skipped: {e}")
    if cb is not None:
        line["cpu_baseline"] = cb
        line["vs_cpu_1core"] = paths_s / cb["value"]
        if cb_all is not None:
            line["cpu_baseline_all_cores"] = cb_all
        # the oracle's result from the baseline leg checks every timed output
        if ref is not None:
            from parity import compare
            line["parity"] = {m: compare(ref, o, max_lsb=1) for m, o in outputs.items()}
    else:
        line["cpu_baseline"] = None
    del out, outputs
    torch.cuda.empty_cache()
    if secondaries:
        # the secondaries run with the default launch shapes (no autotuning)
        set_batched_params(ctx, 6, 2, 32, graph=True)
        set_dense_params(ctx, True, 4, 4)
        sec = {}
        for name, steps, _ in secondaries:
            log(f"secondary {name}")
            sec[name] = secondary_run(ctx, name, min(args.steps, steps), 1, sec_cpu.get(name))
        line["secondary"] = sec
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
