"""Device-resident batch interface (torch is used only to hold device memory
and to name the stream; all computation is in libochre_hip.so)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class Context:
    """An ochre_context on one device."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.ochre_context_create(device, ctypes.byref(h)), "ochre_context_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.ochre_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on: bool):
        _lib.check(self.lib.ochre_context_set_timing(self.h, 1 if on else 0), "set_timing")

    def timing(self, kernel: str):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check(self.lib.ochre_context_timing(self.h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)),
                   "timing")
        return ms.value, n.value


@dataclass
class DeviceBatch:
    """A PathBatch uploaded to device memory."""
    n_paths: int
    verbs: object
    verb_off: object
    points: object
    point_off: object
    weights: object
    weight_off: object
    xforms: object
    n_verbs: int
    n_points: int
    n_weights: int
    host: object = None

    def c(self) -> _lib.PathBatchC:
        def p(t):
            return t.data_ptr() if t.numel() else None
        return _lib.PathBatchC(self.n_paths, p(self.verbs), p(self.verb_off), p(self.points), p(self.point_off),
                               p(self.weights), p(self.weight_off), p(self.xforms), self.n_verbs, self.n_points,
                               self.n_weights)


def upload(batch, device="cuda") -> DeviceBatch:
    torch = _torch()

    def t(a, dt):
        a = np.ascontiguousarray(a)
        if a.size == 0:
            return torch.zeros(1, dtype=dt, device=device)[:0]
        return torch.from_numpy(a).to(device)

    return DeviceBatch(
        n_paths=batch.n_paths,
        verbs=t(batch.verbs.astype(np.uint8), torch.uint8),
        verb_off=t(batch.verb_off.astype(np.int64), torch.int64),
        points=t(batch.points.astype(np.float32).reshape(-1), torch.float32),
        point_off=t(batch.point_off.astype(np.int64), torch.int64),
        weights=t(batch.weights.astype(np.float32), torch.float32),
        weight_off=t(batch.weight_off.astype(np.int64), torch.int64),
        xforms=t(batch.xforms.astype(np.float32).reshape(-1), torch.float32),
        n_verbs=int(batch.verbs.size),
        n_points=int(batch.points.shape[0]),
        n_weights=int(batch.weights.size),
        host=batch,
    )


class DeviceOutput:
    """Caller-owned output arrays (see ochre_batch_output)."""

    def __init__(self, n_paths: int, tile_capacity: int, span_capacity: int, device="cuda", coverage=False):
        torch = _torch()
        self.n_paths = n_paths
        self.tile_capacity = max(int(tile_capacity), 1)
        self.span_capacity = max(int(span_capacity), 1)
        self.tile_xy = torch.empty((self.tile_capacity, 2), dtype=torch.int16, device=device)
        self.tile_alpha = torch.empty((self.tile_capacity, 64), dtype=torch.uint8, device=device)
        self.span_xy = torch.empty((self.span_capacity, 2), dtype=torch.int16, device=device)
        self.span_w = torch.empty((self.span_capacity,), dtype=torch.int16, device=device)
        n = max(n_paths, 1)
        self.path_tile_start = torch.empty((n,), dtype=torch.int64, device=device)
        self.path_tile_count = torch.empty((n,), dtype=torch.int32, device=device)
        self.path_span_start = torch.empty((n,), dtype=torch.int64, device=device)
        self.path_span_count = torch.empty((n,), dtype=torch.int32, device=device)
        self.tile_cov = (torch.empty((self.tile_capacity, 64), dtype=torch.float32, device=device)
                         if coverage else None)
        self.n_tiles = 0
        self.n_spans = 0

    def c(self) -> _lib.BatchOutputC:
        return _lib.BatchOutputC(self.tile_xy.data_ptr(), self.tile_alpha.data_ptr(), self.tile_capacity,
                                 self.span_xy.data_ptr(), self.span_w.data_ptr(), self.span_capacity,
                                 self.path_tile_start.data_ptr(), self.path_tile_count.data_ptr(),
                                 self.path_span_start.data_ptr(), self.path_span_count.data_ptr(), 0, 0,
                                 self.tile_cov.data_ptr() if self.tile_cov is not None else None)

    def target(self) -> _lib.CoverageTargetC:
        """This render (with coverage) as an adjoint target."""
        assert self.tile_cov is not None
        return _lib.CoverageTargetC(self.path_tile_start.data_ptr(), self.path_tile_count.data_ptr(),
                                    self.tile_xy.data_ptr(), self.tile_cov.data_ptr(),
                                    self.path_span_start.data_ptr(), self.path_span_count.data_ptr(),
                                    self.span_xy.data_ptr(), self.span_w.data_ptr())

    def to_host(self):
        """Per-path host arrays in path order: (tile_xy, tile_alpha, span_xy,
        span_w, path_tile_off, path_span_off) with contiguous path ranges."""
        ts = self.path_tile_start[:self.n_paths].cpu().numpy()
        tc = self.path_tile_count[:self.n_paths].cpu().numpy().astype(np.int64)
        ss = self.path_span_start[:self.n_paths].cpu().numpy()
        sc = self.path_span_count[:self.n_paths].cpu().numpy().astype(np.int64)
        txy = self.tile_xy[:self.n_tiles].cpu().numpy()
        ta = self.tile_alpha[:self.n_tiles].cpu().numpy()
        sxy = self.span_xy[:self.n_spans].cpu().numpy()
        sw = self.span_w[:self.n_spans].cpu().numpy().view(np.uint16)
        tidx = np.concatenate([np.arange(s, s + c) for s, c in zip(ts, tc)]) if self.n_paths else np.zeros(0, int)
        sidx = np.concatenate([np.arange(s, s + c) for s, c in zip(ss, sc)]) if self.n_paths else np.zeros(0, int)
        tidx = tidx.astype(np.int64)
        sidx = sidx.astype(np.int64)
        pto = np.zeros(self.n_paths + 1, np.int64)
        pso = np.zeros(self.n_paths + 1, np.int64)
        np.cumsum(tc, out=pto[1:])
        np.cumsum(sc, out=pso[1:])
        return txy[tidx], ta[tidx], sxy[sidx], sw[sidx], pto, pso


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def rasterize_plain(ctx: Context, db: DeviceBatch, out: DeviceOutput, stream=None):
    """Runs the five-kernel pipeline.  Returns (n_tiles, n_spans).  Raises
    OchreError(OCHRE_ERR_CAPACITY) with out.n_tiles/n_spans set when the
    output arrays are too small."""
    cb = db.c()
    co = out.c()
    st = ctx.lib.ochre_rasterize_plain(ctx.h, ctypes.byref(cb), ctypes.byref(co), _stream_ptr(stream))
    out.n_tiles, out.n_spans = co.n_tiles, co.n_spans
    _lib.check(st, "ochre_rasterize_plain")
    return co.n_tiles, co.n_spans


def rasterize_batched(ctx: Context, db: DeviceBatch, out: DeviceOutput, stream=None):
    """Runs the persistent work-queue pipeline.  Returns (n_tiles, n_spans)."""
    if not hasattr(ctx.lib, "ochre_rasterize_batched"):
        raise RuntimeError("libochre_hip.so has no ochre_rasterize_batched")
    cb = db.c()
    co = out.c()
    st = ctx.lib.ochre_rasterize_batched(ctx.h, ctypes.byref(cb), ctypes.byref(co), _stream_ptr(stream))
    out.n_tiles, out.n_spans = co.n_tiles, co.n_spans
    _lib.check(st, "ochre_rasterize_batched")
    return co.n_tiles, co.n_spans


def stroke(ctx: Context, db: DeviceBatch, widths, out: DeviceOutput, batched=True, stream=None):
    """Rasterizer::stroke for every path (widths: a device float32 tensor of
    n_paths): device flattening and stroke expansion, then the batched (or
    plain) fill.  Returns (n_tiles, n_spans)."""
    cb = db.c()
    co = out.c()
    fn = ctx.lib.ochre_stroke_batched if batched else ctx.lib.ochre_stroke_plain
    wp = widths.data_ptr() if widths.numel() else None
    st = fn(ctx.h, ctypes.byref(cb), wp, ctypes.byref(co), _stream_ptr(stream))
    out.n_tiles, out.n_spans = co.n_tiles, co.n_spans
    _lib.check(st, "ochre_stroke_batched" if batched else "ochre_stroke_plain")
    return co.n_tiles, co.n_spans


def draw(ctx: Context, db: DeviceBatch, stroke_flags, widths, out: DeviceOutput, batched=True, stream=None):
    """A draw list: path p filled (stroke_flags[p] == 0) or stroked with
    widths[p] (stroke_flags[p] == 1), each its own Rasterizer (device uint8
    and float32 tensors of n_paths).  Returns (n_tiles, n_spans)."""
    cb = db.c()
    co = out.c()
    fn = ctx.lib.ochre_draw_batched if batched else ctx.lib.ochre_draw_plain
    fp = stroke_flags.data_ptr() if stroke_flags.numel() else None
    wp = widths.data_ptr() if widths.numel() else None
    st = fn(ctx.h, ctypes.byref(cb), fp, wp, ctypes.byref(co), _stream_ptr(stream))
    out.n_tiles, out.n_spans = co.n_tiles, co.n_spans
    _lib.check(st, "ochre_draw_batched" if batched else "ochre_draw_plain")
    return co.n_tiles, co.n_spans


ATLAS_VERTEX_BYTES = 12  # i16 pos[2], u16 uv[2], u8 col[4] (examples/svg.rs:15-20)


def build_atlas(ctx: Context, out: DeviceOutput, colors, atlas_size=4096, stream=None):
    """The demo's atlas Builder on the device output of a batch (paths in
    index order).  colors: uint8 tensor [n_paths, 4] on the device.  Returns
    (atlas uint8 [S, S], vertices uint8 [n, 12], indices int32 [m]) tensors;
    indices hold u32 values."""
    torch = _torch()
    co = out.c()
    nv, ni = ctypes.c_int64(), ctypes.c_int64()
    cp = colors.data_ptr() if colors.numel() else None
    st = ctx.lib.ochre_build_atlas(ctx.h, ctypes.byref(co), out.n_paths, cp, None, atlas_size, None, 0, None, 0,
                                   ctypes.byref(nv), ctypes.byref(ni), _stream_ptr(stream))
    if st not in (0, -4):
        _lib.check(st, "ochre_build_atlas")
    atlas = torch.empty((atlas_size, atlas_size), dtype=torch.uint8, device="cuda")
    verts = torch.empty((max(nv.value, 1), ATLAS_VERTEX_BYTES), dtype=torch.uint8, device="cuda")
    idx = torch.empty((max(ni.value, 1),), dtype=torch.int32, device="cuda")
    _lib.check(ctx.lib.ochre_build_atlas(ctx.h, ctypes.byref(co), out.n_paths, cp, atlas.data_ptr(), atlas_size,
                                         verts.data_ptr(), nv.value, idx.data_ptr(), ni.value, ctypes.byref(nv),
                                         ctypes.byref(ni), _stream_ptr(stream)), "ochre_build_atlas")
    return atlas, verts[:nv.value], idx[:ni.value]


def run_to_host(ctx, batch, mode="plain", tile_capacity=None, span_capacity=None):
    """Convenience: upload, rasterise (growing the outputs once if needed)
    and return the per-path host arrays.  Modes: "plain", "batched" (the
    ordered batched kernel alone), "dense" (the dense kernels in front of it),
    and the same for strokes (batch.widths): "stroke_plain",
    "stroke_batched", "stroke_dense", and for draw lists (batch.strokes,
    batch.widths): "draw_plain", "draw_batched", "draw_dense".  The
    context's dense setting is restored afterwards."""
    d = dense_params(ctx)
    if mode in ("batched", "dense", "stroke_batched", "stroke_dense", "draw_batched", "draw_dense"):
        set_dense_params(ctx, mode in ("dense", "stroke_dense", "draw_dense"), d["dense_groups_per_cu"],
                         d["dense_waves_per_group"])
    try:
        return _run_to_host(ctx, batch, mode, tile_capacity, span_capacity)
    finally:
        set_dense_params(ctx, bool(d["dense"]), d["dense_groups_per_cu"], d["dense_waves_per_group"])


def _run_to_host(ctx, batch, mode, tile_capacity, span_capacity):
    torch = _torch()
    db = upload(batch)
    tc = tile_capacity or max(1024, batch.n_paths * 64)
    sc = span_capacity or max(1024, batch.n_paths * 8)
    if mode.startswith("stroke"):
        wd = torch.from_numpy(np.ascontiguousarray(batch.widths, np.float32)).to("cuda")
        fn = lambda c, d, o: stroke(c, d, wd, o, batched=mode != "stroke_plain")  # noqa: E731
    elif mode.startswith("draw"):
        wd = torch.from_numpy(np.ascontiguousarray(batch.widths, np.float32)).to("cuda")
        fl = torch.from_numpy(np.ascontiguousarray(batch.strokes, np.uint8)).to("cuda")
        fn = lambda c, d, o: draw(c, d, fl, wd, o, batched=mode != "draw_plain")  # noqa: E731
    else:
        fn = rasterize_plain if mode == "plain" else rasterize_batched
    for _ in range(2):
        out = DeviceOutput(batch.n_paths, tc, sc)
        try:
            fn(ctx, db, out)
            break
        except _lib.OchreError as e:
            if e.status != -4:
                raise
            tc, sc = out.n_tiles, out.n_spans
    torch.cuda.synchronize()
    return out.to_host()


def batched_params(ctx: Context):
    w, g, tc, gr = (ctypes.c_int() for _ in range(4))
    _lib.check(ctx.lib.ochre_context_get_batched_params(ctx.h, ctypes.byref(w), ctypes.byref(g), ctypes.byref(tc),
                                                        ctypes.byref(gr)), "get_batched_params")
    d = {"waves_per_cu": w.value, "paths_per_pop": g.value, "tiles_per_chunk": tc.value, "graph": gr.value}
    d.update(dense_params(ctx))
    return d


def dense_params(ctx: Context):
    en, gpc, wpg = (ctypes.c_int() for _ in range(3))
    _lib.check(ctx.lib.ochre_context_get_dense_params(ctx.h, ctypes.byref(en), ctypes.byref(gpc), ctypes.byref(wpg)),
               "get_dense_params")
    sg, lg = ctypes.c_int(), ctypes.c_int()
    _lib.check(ctx.lib.ochre_context_get_dense_small(ctx.h, ctypes.byref(sg)), "get_dense_small")
    _lib.check(ctx.lib.ochre_context_get_dense_large(ctx.h, ctypes.byref(lg)), "get_dense_large")
    return {"dense": en.value, "dense_groups_per_cu": gpc.value, "dense_waves_per_group": wpg.value,
            "dense_small_groups_per_cu": sg.value, "dense_large": lg.value}


def set_pass_limits(ctx: Context, max_lines=0, max_entries=0):
    """Per-pass limits of the plain pipeline (0: default); a batch over them
    is rasterised as consecutive path ranges."""
    _lib.check(ctx.lib.ochre_context_set_pass_limits(ctx.h, int(max_lines), int(max_entries)), "set_pass_limits")


def set_dense_large(ctx: Context, enabled=True):
    """The 256-tile dense instance behind the 64-tile one."""
    _lib.check(ctx.lib.ochre_context_set_dense_large(ctx.h, 1 if enabled else 0), "set_dense_large")


def set_dense_small(ctx: Context, groups_per_cu=0):
    """The small-grid dense instance in front of the 64-tile one (0: off)."""
    _lib.check(ctx.lib.ochre_context_set_dense_small(ctx.h, groups_per_cu), "set_dense_small")


def set_dense_params(ctx: Context, enabled=True, groups_per_cu=4, waves_per_group=4):
    """The dense single-walk kernel (order-independent f32 sums, alpha within
    1 LSB of the ordered result) in front of the ordered kernel."""
    _lib.check(ctx.lib.ochre_context_set_dense_params(ctx.h, 1 if enabled else 0, groups_per_cu, waves_per_group),
               "set_dense_params")


def set_batched_params(ctx: Context, waves_per_cu, paths_per_pop, tiles_per_chunk, graph=True):
    _lib.check(ctx.lib.ochre_context_set_batched_params(ctx.h, waves_per_cu, paths_per_pop, tiles_per_chunk,
                                                        1 if graph else 0), "set_batched_params")


def autotune(ctx: Context, db: DeviceBatch, out: DeviceOutput, reps=3, cache_key=None, cache_path=None, stream=None):
    """Runs the native autotuner on this batch (or reuses a cached choice for
    `cache_key`) and leaves the best parameters set in the context.  Returns
    (best, all_measurements)."""
    import json
    import os
    path = cache_path or os.environ.get("OCHRE_AUTOTUNE_CACHE") or os.path.expanduser("~/.cache/ochre_amd/autotune.json")
    if cache_key is not None and os.path.exists(path):
        try:
            cached = json.load(open(path)).get(cache_key)
        except Exception:
            cached = None
        if cached:
            set_batched_params(ctx, cached["waves_per_cu"], cached["paths_per_pop"], cached["tiles_per_chunk"])
            if "dense" in cached:
                set_dense_params(ctx, bool(cached["dense"]), cached["dense_groups_per_cu"],
                                 cached["dense_waves_per_group"])
                set_dense_small(ctx, cached.get("dense_small_groups_per_cu", 0))
                set_dense_large(ctx, bool(cached.get("dense_large", 1)))
            return cached, []
    res = (_lib.BatchedTuningC * 64)()
    n = ctypes.c_int()
    cb, co = db.c(), out.c()
    _lib.check(ctx.lib.ochre_autotune_batched(ctx.h, ctypes.byref(cb), ctypes.byref(co), _stream_ptr(stream), reps,
                                              res, 64, ctypes.byref(n)), "ochre_autotune_batched")
    allm = [{"waves_per_cu": r.waves_per_cu, "paths_per_pop": r.paths_per_pop, "tiles_per_chunk": r.tiles_per_chunk,
             "dense_groups_per_cu": r.dense_groups_per_cu, "dense_waves_per_group": r.dense_waves_per_group,
             "dense_small_groups_per_cu": r.dense_small_groups_per_cu, "dense_large": r.dense_large,
             "ms": r.ms} for r in res[:n.value]]
    best = dict(batched_params(ctx))
    best["ms"] = min(m["ms"] for m in allm) if allm else None
    if cache_key is not None:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = json.load(open(path)) if os.path.exists(path) else {}
            data[cache_key] = best
            json.dump(data, open(path, "w"), indent=1)
        except Exception:
            pass
    return best, allm


def autotune_cached(ctx: Context, db: DeviceBatch, out: DeviceOutput, reps=3, cache_path=None, stream=None):
    """The library's cached autotuner: applies the stored parameters for this
    device and batch signature, or tunes and stores them (appending to
    `cache_path` when given).  Returns True when it tuned."""
    tuned = ctypes.c_int()
    cb, co = db.c(), out.c()
    path = cache_path.encode() if cache_path else None
    _lib.check(ctx.lib.ochre_autotune_batched_cached(ctx.h, ctypes.byref(cb), ctypes.byref(co), _stream_ptr(stream),
                                                     reps, path, ctypes.byref(tuned)), "ochre_autotune_batched_cached")
    return bool(tuned.value)


def coverage_adjoint(ctx: Context, db: DeviceBatch, target: DeviceOutput, stream=None):
    """Loss per path and gradients w.r.t. the path-space points and the
    transforms (torch tensors on the device)."""
    torch = _torch()
    loss = torch.empty((max(db.n_paths, 1),), dtype=torch.float32, device="cuda")
    dp = torch.empty((max(db.n_points, 1), 2), dtype=torch.float32, device="cuda")
    dx = torch.empty((max(db.n_paths, 1), 6), dtype=torch.float32, device="cuda")
    cb = db.c()
    ct = target.target()
    _lib.check(ctx.lib.ochre_coverage_adjoint(ctx.h, ctypes.byref(cb), ctypes.byref(ct), loss.data_ptr(),
                                              dp.data_ptr(), dx.data_ptr(), _stream_ptr(stream)),
               "ochre_coverage_adjoint")
    return loss[:db.n_paths], dp[:db.n_points], dx[:db.n_paths]
