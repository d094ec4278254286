"""The drop-in Rasterizer's finish paths against the oracle Rasterizer, event
for event: a single fill or a single stroke (one upload, one launch), and
finish_many over a sequence mixing single fills, single strokes, empty
rasterizers and one with several fills in the middle (the demo's loop,
examples/svg.rs:140-156, finished into one builder)."""
import numpy as np
import pytest

import oracle_ref
from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def _same(events, ref, what):
    assert len(events) == len(ref), (what, len(events), len(ref))
    for a, b in zip(events, ref):
        assert a[:3] == b[:3], (what, a[:3], b[:3])
        if a[0] == "tile":
            assert np.array_equal(a[3], b[3]), what
        else:
            assert a[3] == b[3], what


def _oracle_events(kinds, batch):
    ref = []
    for k, (kind, i) in enumerate(kinds):
        o = oracle_ref.Rasterizer()
        if kind == "fill":
            v, p, w, xf = batch.path(i)
            o.fill(v, p, w, xf)
        elif kind == "stroke":
            v, p, w, xf = batch.path(i)
            o.stroke(v, p, float(batch.widths[i]), w, xf)
        elif kind == "two":
            for j in (i, i + 1):
                v, p, w, xf = batch.path(j)
                o.fill(v, p, w, xf)
        ref += o.finish()
    return ref


def _make(kinds, batch):
    from ochre_amd import Rasterizer
    rs = []
    for kind, i in kinds:
        r = Rasterizer()
        if kind == "fill":
            v, p, w, xf = batch.path(i)
            r.fill_arrays(v, p, w, xf)
        elif kind == "stroke":
            v, p, w, xf = batch.path(i)
            r.stroke_arrays(v, p, float(batch.widths[i]), w, xf)
        elif kind == "two":
            for j in (i, i + 1):
                v, p, w, xf = batch.path(j)
                r.fill_arrays(v, p, w, xf)
        rs.append(r)
    return rs


def test_single_fills_one_by_one(gpu_ctx):
    from ochre_amd import EventRecorder
    b = W.concat([W.readme_path(), W.glyph_run(40, seed=51), W.random_paths(40, 52)])
    kinds = [("fill", i) for i in range(b.n_paths)]
    events = []
    for r in _make(kinds, b):
        rec = EventRecorder()
        r.finish(rec)
        events += rec.events
    _same(events, _oracle_events(kinds, b), "single fills")


def test_single_strokes_one_by_one(gpu_ctx):
    """A Rasterizer holding one stroke: its outline, expanded on the device, is
    finished as a single fill is."""
    from ochre_amd import EventRecorder
    b = W.with_widths(W.concat([W.glyph_run(30, seed=56), W.random_paths(20, 57)]), 0.5, 3.0, 10)
    kinds = [("stroke", i) for i in range(b.n_paths)]
    events = []
    for r in _make(kinds, b):
        rec = EventRecorder()
        r.finish(rec)
        events += rec.events
    _same(events, _oracle_events(kinds, b), "single strokes")


def _limit_paths():
    """Single fills at the single-fill kernel's limits, on either side: more
    points than it stages in LDS; tile grids 64 and 66 tiles wide; more than
    128 occupied tiles in a grid that fits; more than 2,048 lines."""
    rng = np.random.default_rng(58)
    paths = []

    def polygon(n, cx, cy, r, jitter):
        t = 2 * np.pi * np.arange(n) / n
        rr = r * (1.0 + jitter * rng.standard_normal(n))
        pts = np.stack([cx + rr * np.cos(t), cy + rr * np.sin(t)], 1)
        return np.array([W.MOVE] + [W.LINE] * (n - 1) + [W.CLOSE], np.uint8), pts

    paths.append(polygon(1700, 200.3, 150.7, 60.0, 0.02))   # 1,700 points, about 50 occupied tiles
    for w in (503.0, 521.0):  # slivers in one tile row, grids 64 and 66 tiles wide
        paths.append((np.array([W.MOVE, W.LINE, W.LINE, W.LINE, W.CLOSE], np.uint8),
                      np.array([(3.5, 1.5), (3.5 + w, 2.0), (3.5 + w, 5.5), (3.5, 6.0)], np.float32)))
    zig = [(10.0 + 12.0 * k, 10.0 + (400.0 if k % 2 else 0.0)) for k in range(33)]  # >128 occupied tiles
    paths.append((np.array([W.MOVE] + [W.LINE] * 32 + [W.CLOSE], np.uint8), np.array(zig, np.float32)))
    paths.append(polygon(2300, 128.0, 128.0, 90.0, 0.01))  # over 2,048 lines
    return W.from_paths(paths, [1, 0, 0, 1, 0, 0])


def test_single_fills_at_the_kernel_limits(gpu_ctx):
    from ochre_amd import EventRecorder
    b = _limit_paths()
    kinds = [("fill", i) for i in range(b.n_paths)]
    events = []
    for r in _make(kinds, b):
        rec = EventRecorder()
        r.finish(rec)
        events += rec.events
    _same(events, _oracle_events(kinds, b), "limit fills")


def test_finish_many_matches_one_by_one_and_oracle(gpu_ctx):
    from ochre_amd import EventRecorder, finish_many
    b = W.with_widths(W.concat([W.glyph_run(60, seed=53), W.random_paths(40, 54)]), 0.5, 3.0, 9)
    rng = np.random.default_rng(55)
    kinds = []
    for i in range(b.n_paths - 1):
        r = rng.random()
        kinds.append(("stroke", i) if r < 0.3 else ("empty", i) if r < 0.35 else ("fill", i))
    kinds.insert(37, ("two", 10))  # several fills: finished on its own, in place
    rec = EventRecorder()
    finish_many(_make(kinds, b), rec)
    print(f"[finish_many] {len(kinds)} rasterizers, {len(rec.events)} TileBuilder calls", flush=True)
    _same(rec.events, _oracle_events(kinds, b), "finish_many")
