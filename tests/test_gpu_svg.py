"""The reference demo's SVG path (examples/svg.rs:103-164) end to end on the
device: the three documents it ships are parsed into their draw lists (a
fill and/or a stroke per path element, in document order), every draw is
rasterised as its own Rasterizer through the draw-list entry points, and the
atlas Builder is run over the result.  The oracle makes the same calls one
Rasterizer at a time (svg.rs:140-156).  Plain and the ordered batched kernel
are bit-exact; the dense kernels keep identical tile and span sets with alpha
within 1 LSB; the atlas, vertices and indices are byte-identical for plain."""
import numpy as np
import pytest

import oracle_ref
from ochre_amd import svg
from parity import compare
from test_svg import golden

pytestmark = pytest.mark.gpu

DOCS = ["Ghostscript_Tiger.svg.gz", "lorem-ipsum.svg.gz", "calabi-yau.svg.xz"]
_REF = {}


def _doc(name):
    if name not in _REF:
        b = svg.draw_batch(svg.parse(golden(name)), name=name)
        _REF[name] = (b, oracle_ref.run_batch(b))
    return _REF[name]


@pytest.mark.parametrize("name", DOCS)
@pytest.mark.parametrize("mode", ["draw_plain", "draw_batched", "draw_dense"])
def test_svg_draws_match_oracle(gpu_ctx, name, mode):
    from ochre_amd.device import run_to_host
    b, ref = _doc(name)
    assert ref.status == 0
    n_strokes = int(b.strokes.sum())
    print(f"[svg/{mode}] {name}: {b.n_paths} draws ({n_strokes} strokes), {ref.n_tiles} tiles", flush=True)
    got = run_to_host(gpu_ctx, b, mode, tile_capacity=ref.n_tiles + 64, span_capacity=ref.n_spans + 64)
    st = compare(ref, got, max_lsb=1 if mode == "draw_dense" else 0)
    print(f"[svg/{mode}] {name}: {st}", flush=True)
    if mode == "draw_dense":
        assert st["bit_exact_fraction"] >= 0.999


@pytest.mark.parametrize("name", DOCS[:2])
def test_svg_atlas_matches_oracle_builder(gpu_ctx, name):
    import torch
    from ochre_amd.device import DeviceOutput, build_atlas, draw, upload
    b, ref = _doc(name)
    print(f"[svg/atlas] {name}: {b.n_paths} draws", flush=True)
    st, a_ref, v_ref, i_ref = oracle_ref.atlas_batch(b, b.colors, 4096)
    assert st == 0
    out = DeviceOutput(b.n_paths, ref.n_tiles + 64, ref.n_spans + 64)
    fl = torch.from_numpy(b.strokes).to("cuda")
    wd = torch.from_numpy(b.widths).to("cuda")
    draw(gpu_ctx, upload(b), fl, wd, out, batched=False)
    atlas, verts, idx = build_atlas(gpu_ctx, out, torch.from_numpy(b.colors).to("cuda"), 4096)
    torch.cuda.synchronize()
    v = verts.cpu().numpy().reshape(-1).view(oracle_ref.ATLAS_VERTEX)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), i_ref)
    assert np.array_equal(v, v_ref)
    assert np.array_equal(atlas.cpu().numpy(), a_ref)
    print(f"[svg/atlas] {name}: {v.size // 4} quads, atlas identical", flush=True)


def test_draw_list_with_conic_fills(gpu_ctx):
    """A draw list whose fills carry conics (weights pass through to the fill)
    next to strokes, against the oracle."""
    from ochre_amd import workloads as W
    from ochre_amd.device import run_to_host
    b = W.random_paths(120, 13)
    assert b.weights.size > 0
    b = W.with_widths(b, 0.5, 3.0, 6)
    b.strokes = (np.arange(b.n_paths) % 3 == 1).astype(np.uint8)
    ref = oracle_ref.run_batch(b)
    for mode in ("draw_plain", "draw_batched"):
        st = compare(ref, run_to_host(gpu_ctx, b, mode), max_lsb=0)
        print(f"[draw/{mode}] conic fills + strokes: {st}", flush=True)
