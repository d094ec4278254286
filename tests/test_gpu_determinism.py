"""Run-to-run determinism.  The plain pipeline and the ordered batched kernel
sum each pixel's f32 increments in emission order (the sort is stable, the
ordered commit follows emission indices), so two runs give byte-identical
output; the dense kernels add in arrival order with LDS atomics, so their
alpha may differ by 1 LSB between runs while tile and span sets stay equal
(the reference's own order is that of sort_unstable_by_key,
rasterizer.rs:209)."""
import numpy as np
import pytest

from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def _batch():
    b = W.concat([W.glyph_run(2000, seed=31), W.random_paths(200, 32), W.star_blobs(n=2, seed=33, nmax=3e3)])
    b.name = "glyphs+random+blobs"
    return b


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_ordered_modes_are_byte_identical_run_to_run(gpu_ctx, mode):
    from ochre_amd.device import run_to_host
    b = _batch()
    runs = [run_to_host(gpu_ctx, b, mode) for _ in range(3)]
    for k, (a, c) in enumerate(zip(runs[0], runs[1])):
        assert np.array_equal(a, c), (mode, k)
    for k, (a, c) in enumerate(zip(runs[0], runs[2])):
        assert np.array_equal(a, c), (mode, k)
    print(f"[determinism/{mode}] {b.n_paths} paths, {runs[0][0].shape[0]} tiles: 3 runs byte-identical", flush=True)


def test_dense_mode_variance_is_within_one_lsb(gpu_ctx):
    from ochre_amd.device import run_to_host
    b = _batch()
    runs = [run_to_host(gpu_ctx, b, "dense") for _ in range(3)]
    worst, differing = 0, 0
    for r in runs[1:]:
        for k in (0, 2, 3, 4, 5):  # tile/span coordinates, widths and ranges
            assert np.array_equal(runs[0][k], r[k]), k
        d = np.abs(runs[0][1].astype(np.int16) - r[1].astype(np.int16))
        worst = max(worst, int(d.max()) if d.size else 0)
        differing = max(differing, int((d != 0).sum()))
    print(f"[determinism/dense] alpha bytes differing between runs: at most {differing} of {runs[0][1].size}, "
          f"by at most {worst} LSB", flush=True)
    assert worst <= 1
