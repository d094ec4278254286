"""The optional pre-quantisation coverage output (ochre_batch_output's
tile_coverage) is written for every tile of a call, whichever kernel made the
tile: the dense kernels, the ordered one-wave kernel, and the plain
pipeline's tile stages (ordered, and unordered under the dense kernels) that
take the paths the batched kernels hand on.  The coverage is
min(|accum + area|, 1), the value rasterizer.rs:235 scales by 256, so its
quantisation must give the tile's alpha byte, and the batched call's
coverage must agree with the plain pipeline's to the dense kernels' 1 LSB."""
import numpy as np
import pytest

from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def _batch():
    # glyphs, and star blobs over every batched kernel's limits (handed on to
    # the plain pipeline), and the README path
    return W.concat([W.glyph_run(150, seed=61), W.star_blobs(n=2, seed=62, nmin=1e3, nmax=2e3, rmin=150, rmax=300),
                     W.readme_path(), W.random_paths(30, 63)])


def _render(ctx, batch, mode):
    import torch
    from ochre_amd.device import (DeviceOutput, dense_params, rasterize_batched, rasterize_plain,
                                  set_dense_params, upload)
    db = upload(batch)
    out = DeviceOutput(batch.n_paths, 200_000, 40_000, coverage=True)
    out.tile_cov.fill_(float("nan"))  # anything left unwritten shows
    if mode == "plain":
        rasterize_plain(ctx, db, out)
    else:
        d = dense_params(ctx)
        set_dense_params(ctx, mode == "batched_dense", d["dense_groups_per_cu"], d["dense_waves_per_group"])
        try:
            rasterize_batched(ctx, db, out)
        finally:
            set_dense_params(ctx, d["dense"], d["dense_groups_per_cu"], d["dense_waves_per_group"])
    torch.cuda.synchronize()
    return out


def _check_quantisation(out, what):
    n = out.n_tiles
    cov = out.tile_cov[:n].cpu().numpy()
    alpha = out.tile_alpha[:n].cpu().numpy()
    assert np.isfinite(cov).all(), f"{what}: coverage left unwritten in {int((~np.isfinite(cov)).any(1).sum())} tiles"
    assert (cov >= 0).all() and (cov <= 1).all(), what
    q = np.minimum(cov * np.float32(256), np.float32(255)).astype(np.uint8)
    assert np.array_equal(q, alpha), what


@pytest.mark.parametrize("mode", ["plain", "batched_ordered", "batched_dense"])
def test_coverage_for_every_tile(gpu_ctx, mode):
    out = _render(gpu_ctx, _batch(), mode)
    print(f"[coverage] {mode}: {out.n_tiles} tiles, {out.n_spans} spans", flush=True)
    _check_quantisation(out, mode)


def test_coverage_batched_matches_plain(gpu_ctx):
    b = _batch()
    ref = _render(gpu_ctx, b, "plain")
    for mode in ("batched_ordered", "batched_dense"):
        out = _render(gpu_ctx, b, mode)
        assert out.n_tiles == ref.n_tiles, mode
        for i in range(b.n_paths):
            ta, tc = int(ref.path_tile_start[i]), int(ref.path_tile_count[i])
            tb, tcb = int(out.path_tile_start[i]), int(out.path_tile_count[i])
            assert tc == tcb, (mode, i)
            assert np.array_equal(ref.tile_xy[ta:ta + tc].cpu().numpy(), out.tile_xy[tb:tb + tc].cpu().numpy())
            d = np.abs(ref.tile_cov[ta:ta + tc].cpu().numpy() - out.tile_cov[tb:tb + tc].cpu().numpy())
            if mode == "batched_ordered":
                assert d.max(initial=0.0) == 0.0, (mode, i)  # the ordered kernels sum as the plain pipeline does
            else:
                assert d.max(initial=0.0) <= 1.0 / 256.0, (mode, i)
