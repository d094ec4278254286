"""Parity of device stroking (flatten + stroke expansion on the GPU, then the
plain or batched fill) with the oracle's Rasterizer::stroke, through the C
ABI: identical tile and span sets; alpha bit-exact for the plain pipeline
and the ordered batched kernel, within 1 LSB for the dense kernels
(arrival-order f32 sums), whose bit-exact fraction must stay >= 0.999."""
import numpy as np
import pytest

import oracle_ref
from ochre_amd import workloads as W
from parity import compare

pytestmark = pytest.mark.gpu

MODES = ["stroke_plain", "stroke_batched", "stroke_dense"]


def check(ctx, batch, mode):
    from ochre_amd.device import run_to_host
    print(f"[{mode}] {batch.name} paths={batch.n_paths}", flush=True)
    ref = oracle_ref.run_batch(batch, stroke=True)
    assert ref.status == 0
    got = run_to_host(ctx, batch, mode)
    st = compare(ref, got, max_lsb=1 if mode == "stroke_dense" else 0)
    print(f"[{mode}] {batch.name}: {st}", flush=True)
    if mode == "stroke_dense":
        assert st["bit_exact_fraction"] >= 0.999
    return st


@pytest.mark.parametrize("mode", MODES)
def test_readme_path_stroked(gpu_ctx, mode):
    check(gpu_ctx, W.with_widths(W.readme_path(), 4.0, 4.0001, 0), mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_paths_stroked(gpu_ctx, seed, mode):
    # every verb kind, Close mid-contour (the sticky `closed`), contours
    # without a Move, repeated points, widths from 0 to 6 path units
    b = W.with_widths(W.random_paths(300, 40 + seed), 0.0, 6.0, seed)
    b.widths[::17] = 0.0
    check(gpu_ctx, b, mode)


@pytest.mark.parametrize("mode", MODES)
def test_stroked_glyphs(gpu_ctx, mode):
    check(gpu_ctx, W.stroked_glyphs(2000, seed=6), mode)


def test_stroke_edge_cases(gpu_ctx):
    M, L, Q, Z = W.MOVE, W.LINE, W.QUAD, W.CLOSE
    paths = [
        ([], np.zeros((0, 2), np.float32), []),                          # empty
        ([M], np.array([(5, 5)], np.float32), []),                         # single point
        ([M, L], np.array([(5, 5), (5, 5)], np.float32), []),              # zero length
        ([Z, Z], np.zeros((0, 2), np.float32), []),                        # only closes
        ([M, L, L, Z], np.array([(0, 0), (10, 0), (0, 0)], np.float32), []),   # folds back
        ([M, L, M, L], np.array([(0, 0), (10, 0), (20, 0), (30, 5)], np.float32), []),
        ([L, L, Q], np.array([(4, 4), (9, 1), (12, 12), (2, 9)], np.float32), []),  # no Move
        ([M, L, L], np.array([(-20, -20), (30, 0), (10, 40)], np.float32), []),   # negative coordinates
    ]
    xf = np.tile(np.array([1.5, 0.5, -0.25, 1.25, 20.0, 24.0], np.float32), (len(paths), 1))
    b = W.from_paths(paths, xf)
    b.widths = np.array([2.0, 3.0, 1.0, 2.0, 4.0, 1.5, 2.5, -3.0], np.float32)  # a negative width too
    b.name = "stroke_edge_cases"
    for mode in MODES:
        check(gpu_ctx, b, mode)


def test_stroke_rejects_inconsistent_input(gpu_ctx):
    import torch
    from ochre_amd import _lib
    from ochre_amd.device import DeviceOutput, stroke, upload
    b = W.random_paths(4, 3)
    db = upload(b)
    db.point_off = db.point_off.clone()
    db.point_off[1:] += 1          # path 0 claims one point more than its verbs use
    db.n_points += 1
    db.points = torch.cat([db.points, torch.zeros(2, device=db.points.device)])
    wd = torch.ones(b.n_paths, device="cuda")
    out = DeviceOutput(b.n_paths, 4096, 1024)
    with pytest.raises(_lib.OchreError) as e:
        stroke(gpu_ctx, db, wd, out, batched=True)
    assert e.value.status == -1


def test_rasterizer_api_stroke_mixed_with_fills(gpu_ctx):
    """stroke() and fill() calls accumulate into one path in call order."""
    from ochre_amd import EventRecorder, Rasterizer
    xf = np.array([1.25, 0.5, -0.5, 1.5, 40.0, 16.0], np.float32)
    v1 = np.array([W.MOVE, W.QUAD, W.CUBIC, W.CLOSE], np.uint8)
    p1 = np.array([(10, 10), (60, 0), (50, 50), (30, 70), (0, 60), (12, 30)], np.float32)
    v2 = np.array([W.MOVE, W.LINE, W.CONIC, W.LINE], np.uint8)
    p2 = np.array([(100, 20), (150, 40), (170, 90), (120, 100), (90, 60)], np.float32)
    r = Rasterizer()
    o = oracle_ref.Rasterizer()
    r.stroke_arrays(v1, p1, 3.0, [], xf); o.stroke(v1, p1, 3.0, (), xf)
    r.fill_arrays(v2, p2, [0.8], None); o.fill(v2, p2, [0.8], None)
    r.stroke_arrays(v2, p2, 1.5, [0.8], xf); o.stroke(v2, p2, 1.5, [0.8], xf)
    rec = EventRecorder()
    r.finish(rec)
    ref = o.finish()
    assert [e[:3] for e in rec.events] == [e[:3] for e in ref]
    for a, b in zip(rec.events, ref):
        if a[0] == "tile":
            assert np.array_equal(a[3], b[3])
        else:
            assert a[3] == b[3]


def test_stroke_flatten_limit_is_an_error(gpu_ctx):
    """A quadratic whose step underflows (the reference never finishes
    flattening it) returns OCHRE_ERR_FLATTEN_LIMIT, as the oracle does."""
    import torch
    from ochre_amd import _lib
    from ochre_amd.device import DeviceOutput, stroke, upload
    paths = [([W.MOVE, W.QUAD], np.array([(0, 0), (3e37, 1e37), (1, 1)], np.float32), [])]
    b = W.from_paths(paths, np.array([[1, 0, 0, 1, 0, 0]], np.float32))
    b.widths = np.ones(1, np.float32)
    assert oracle_ref.run_batch(b, stroke=True).status != 0
    out = DeviceOutput(1, 1024, 1024)
    with pytest.raises(_lib.OchreError) as e:
        stroke(gpu_ctx, upload(b), torch.ones(1, device="cuda"), out, batched=False)
    assert e.value.status == -5
