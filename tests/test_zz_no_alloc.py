"""The allocation-free batched call (ochre_rasterize_batched_noalloc) through
the C ABI, against ochre_rasterize_batched with the dense kernels on and
against the oracle: identical tile and span sets, alpha within 1 LSB (the
dense kernels' contract), on two calls in a row, the second with a smaller
input (the shared device array keeps the first call's contents); and its two
limits."""
import ctypes

import numpy as np
import pytest

import oracle_ref
from ochre_amd import _lib
from ochre_amd import workloads as W
from parity import compare

pytestmark = pytest.mark.gpu


def noalloc_to_host(ctx, batch, tiles, spans):
    """(status, DeviceOutput) of one noalloc call; the per-path ranges start
    zeroed, so a path the call does not write reads as empty."""
    import torch
    from ochre_amd.device import DeviceOutput, upload
    db = upload(batch)
    out = DeviceOutput(batch.n_paths, tiles, spans)
    for t in (out.path_tile_start, out.path_tile_count, out.path_span_start, out.path_span_count):
        t.zero_()
    cb, co = db.c(), out.c()
    st = ctx.lib.ochre_rasterize_batched_noalloc(ctx.h, ctypes.byref(cb), ctypes.byref(co), None)
    torch.cuda.synchronize()
    out.n_tiles, out.n_spans = co.n_tiles, co.n_spans
    return st, out


def test_noalloc_matches_batched_twice(gpu_ctx):
    from ochre_amd.device import run_to_host
    first = W.glyph_run(3000, seed=21)
    second = W.concat([W.glyph_run(400, seed=22), W.random_paths(60, 24, coord_scale=24.0)])
    for name, b in (("first", first), ("second (smaller)", second)):
        ref = oracle_ref.run_batch(b)
        print(f"[noalloc] {name}: {b.n_paths} paths, {ref.n_tiles} tiles", flush=True)
        st, out = noalloc_to_host(gpu_ctx, b, ref.n_tiles + 64, ref.n_spans + 64)
        assert st == 0, _lib.STATUS_NAMES.get(st, st)
        assert (out.n_tiles, out.n_spans) == (ref.n_tiles, ref.n_spans)
        got = out.to_host()
        batched = run_to_host(gpu_ctx, b, "dense", tile_capacity=ref.n_tiles + 64, span_capacity=ref.n_spans + 64)
        s1 = compare(ref, got, max_lsb=1)
        # against the existing entry point: the same sets, alpha within 1 LSB
        d = np.abs(got[1].astype(np.int16) - batched[1].astype(np.int16))
        assert all(np.array_equal(a, b2) for i, (a, b2) in enumerate(zip(got, batched)) if i != 1)
        assert int(d.max()) <= 1
        print(f"[noalloc] {name}: vs oracle {s1}; vs batched max {int(d.max())} LSB", flush=True)


def test_noalloc_limits(gpu_ctx):
    # a path over the 256-tile instance (a blob of radius 300-400 px): the
    # glyphs before it are written, its range is not, and the call fails
    glyphs = W.glyph_run(50, seed=23)
    blob = W.star_blobs(n=1, seed=3, nmin=1e3, nmax=2e3, rmin=300.0, rmax=400.0)
    b = W.concat([glyphs, blob])
    ref = oracle_ref.run_batch(glyphs)
    print(f"[noalloc] limits: 50 glyphs + one blob, {ref.n_tiles} glyph tiles", flush=True)
    st, out = noalloc_to_host(gpu_ctx, b, ref.n_tiles + 64, ref.n_spans + 64)
    assert st == -7, _lib.STATUS_NAMES.get(st, st)
    txy, ta, sxy, sw, pto, pso = out.to_host()
    assert pto[-1] == pto[-2] and pso[-1] == pso[-2]  # the blob's range is empty
    compare(ref, (txy, ta, sxy, sw, pto[:-1], pso[:-1]), max_lsb=1)
    # more paths than the fixed array holds: refused before anything runs
    n = (1 << 20) + 1
    big = W.PathBatch(verbs=np.zeros(n, np.uint8), verb_off=np.arange(n + 1, dtype=np.int64),
                      points=np.ones((n, 2), np.float32), point_off=np.arange(n + 1, dtype=np.int64),
                      weights=np.zeros(0, np.float32), weight_off=np.zeros(n + 1, np.int64),
                      xforms=np.tile(np.array([1, 0, 0, 1, 0, 0], np.float32), (n, 1)))
    st, out = noalloc_to_host(gpu_ctx, big, 16, 16)
    assert st == -7 and out.n_tiles == 0
