"""The device atlas Builder (examples/svg.rs:15-88 on the device output)
against the oracle Builder fed by one oracle Rasterizer per path: atlas,
vertices and indices byte-identical, through the C ABI."""
import numpy as np
import pytest

import oracle_ref
from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def device_atlas(ctx, batch, colors, size, mode):
    import torch
    from ochre_amd.device import (DeviceOutput, build_atlas, dense_params, rasterize_batched, rasterize_plain,
                                  set_dense_params, stroke, upload)
    db = upload(batch)
    ref = oracle_ref.run_batch(batch, stroke=mode.startswith("stroke"))
    out = DeviceOutput(batch.n_paths, ref.n_tiles + 64, ref.n_spans + 64)
    if mode == "plain":
        rasterize_plain(ctx, db, out)
    elif mode == "batched":
        d = dense_params(ctx)
        set_dense_params(ctx, False, d["dense_groups_per_cu"], d["dense_waves_per_group"])
        try:
            rasterize_batched(ctx, db, out)
        finally:
            set_dense_params(ctx, True, d["dense_groups_per_cu"], d["dense_waves_per_group"])
    else:
        wd = torch.from_numpy(batch.widths).to("cuda")
        stroke(ctx, db, wd, out, batched=False)
    col = torch.from_numpy(colors).to("cuda")
    atlas, verts, idx = build_atlas(ctx, out, col, size)
    torch.cuda.synchronize()
    return atlas.cpu().numpy(), verts.cpu().numpy().reshape(-1).view(oracle_ref.ATLAS_VERTEX), \
        idx.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("mode", ["plain", "batched", "stroke_plain"])
def test_atlas_matches_oracle_builder(gpu_ctx, mode):
    b = W.concat([W.glyph_run(1500, seed=8), W.random_paths(100, 9)])
    b.name = "glyphs+random"
    if mode.startswith("stroke"):
        b = W.with_widths(b, 0.25, 3.0, 4)
    colors = np.random.default_rng(5).integers(0, 256, (b.n_paths, 4), dtype=np.uint8)
    print(f"[atlas/{mode}] {b.name} paths={b.n_paths}", flush=True)
    st, a_ref, v_ref, i_ref = oracle_ref.atlas_batch(b, colors, 4096, stroke=mode.startswith("stroke"))
    assert st == 0
    a, v, i = device_atlas(gpu_ctx, b, colors, 4096, mode)
    assert np.array_equal(i, i_ref)
    assert np.array_equal(v, v_ref)
    assert np.array_equal(a, a_ref)
    print(f"[atlas/{mode}] {v.size // 4} quads, atlas identical", flush=True)


def test_atlas_capacity(gpu_ctx):
    import torch
    from ochre_amd import _lib
    from ochre_amd.device import DeviceOutput, build_atlas, rasterize_plain, upload
    b = W.glyph_run(50, seed=2)
    out = DeviceOutput(b.n_paths, 4096, 1024)
    rasterize_plain(gpu_ctx, upload(b), out)
    col = torch.zeros((b.n_paths, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.OchreError) as e:
        build_atlas(gpu_ctx, out, col, 64)   # 64 slots, fewer than the tiles
    assert e.value.status == -4
