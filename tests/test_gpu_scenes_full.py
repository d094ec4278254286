"""The batched-scenes configuration at full size (256 scenes, 148,756 paths,
206 M tiles): too large for the oracle, so the two device pipelines check each
other through size-independent properties.  The plain pipeline is bit-exact
against the oracle on every smaller case (test_gpu_parity.py); here the
batched launch (dense instances, the ordered kernel, and the plain pipeline
split into path ranges for the blobs) must produce, path by path, the same
tile and span counts, the same tile positions in the same order, the same
spans, and alphas within 1 LSB of the plain result."""
import numpy as np
import pytest

from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def _owner(torch, start, count, total):
    """Path of every output slot: the ranges tile [0, total) without gaps."""
    order = torch.argsort(start)
    owner = torch.repeat_interleave(order, count[order].long())
    assert owner.numel() == total
    return owner


def test_scenes_full_plain_vs_batched(gpu_ctx):
    import torch
    from ochre_amd.device import DeviceOutput, rasterize_batched, rasterize_plain, upload
    b = W.scenes()
    print(f"[scenes-full] {b.name} verbs={b.verbs.size}", flush=True)
    db = upload(b)
    outs = {}
    for name, fn in (("plain", rasterize_plain), ("batched", rasterize_batched)):
        out = DeviceOutput(b.n_paths, 1 << 20, 1 << 16)
        try:
            fn(gpu_ctx, db, out)
        except Exception as e:
            assert getattr(e, "status", None) == -4, e
            out = DeviceOutput(b.n_paths, out.n_tiles, out.n_spans)
            fn(gpu_ctx, db, out)
        torch.cuda.synchronize()
        outs[name] = out
        print(f"[scenes-full] {name}: tiles={out.n_tiles} spans={out.n_spans}", flush=True)
    p, q = outs["plain"], outs["batched"]
    assert p.n_tiles == q.n_tiles and p.n_spans == q.n_spans
    assert torch.equal(p.path_tile_count, q.path_tile_count)
    assert torch.equal(p.path_span_count, q.path_span_count)
    T, S = p.n_tiles, p.n_spans
    # slot k of a path in the batched output <-> slot k of the same path in the plain one
    own = _owner(torch, q.path_tile_start, q.path_tile_count, T)
    exact = 0
    worst = 0
    chunk = 1 << 23
    for c0 in range(0, T, chunk):
        c1 = min(T, c0 + chunk)
        o = own[c0:c1]
        k = torch.arange(c0, c1, device="cuda") - q.path_tile_start[o]
        src = p.path_tile_start[o] + k
        assert torch.equal(q.tile_xy[c0:c1], p.tile_xy[src])
        d = (q.tile_alpha[c0:c1].int() - p.tile_alpha[src].int()).abs()
        worst = max(worst, int(d.max()))
        exact += int((d == 0).sum())
    frac = exact / (64 * T)
    print(f"[scenes-full] tiles identical in position; alpha max diff {worst} LSB, bit-exact fraction {frac:.7f}",
          flush=True)
    assert worst <= 1 and frac >= 0.999
    own_s = _owner(torch, q.path_span_start, q.path_span_count, S)
    k = torch.arange(S, device="cuda") - q.path_span_start[own_s]
    src = p.path_span_start[own_s] + k
    assert torch.equal(q.span_xy[:S], p.span_xy[src])
    assert torch.equal(q.span_w[:S], p.span_w[src])
