"""Parity of the device pipelines with the oracle, through the C ABI:
plain (five kernels), batched (ordered persistent kernel) and dense (the
single-walk kernel with order-independent f32 sums in front of it).  Tile
and span sets must be identical; alpha bit-exact for plain and batched and
within 1 LSB for dense, whose bit-exact fraction is printed."""
import numpy as np
import pytest

import oracle_ref
from ochre_amd import workloads as W
from parity import compare

pytestmark = pytest.mark.gpu


MODES = ["plain", "batched", "dense"]


def check(ctx, batch, mode, max_lsb=1, min_exact=None):
    from ochre_amd.device import run_to_host
    print(f"[{mode}] {batch.name} paths={batch.n_paths}", flush=True)
    ref = oracle_ref.run_batch(batch)
    got = run_to_host(ctx, batch, mode)
    if mode == "dense":
        # arrival-order f32 sums: 1 LSB where a sum sits on a quantisation
        # boundary, which is rare
        max_lsb = max(max_lsb, 1)
        min_exact = max(min_exact or 0.0, 0.999)
    st = compare(ref, got, max_lsb=max_lsb)
    print(f"[{mode}] {batch.name}: {st}", flush=True)
    if min_exact is not None:
        assert st["bit_exact_fraction"] >= min_exact
    return st


@pytest.mark.parametrize("mode", MODES)
def test_readme_path(gpu_ctx, mode):
    st = check(gpu_ctx, W.readme_path(), mode, max_lsb=0)
    assert st["tiles"] == 100 and st["spans"] == 23


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_paths(gpu_ctx, seed, mode):
    check(gpu_ctx, W.random_paths(300, seed), mode, max_lsb=0)


@pytest.mark.parametrize("mode", MODES)
def test_random_paths_identity_integer(gpu_ctx, mode):
    # integer and half-integer coordinates exercise DDA ties and zero-length steps
    check(gpu_ctx, W.random_paths(300, 7, integer_prob=1.0, xform="identity"), mode, max_lsb=0)


@pytest.mark.parametrize("mode", MODES)
def test_glyph_run_full(gpu_ctx, mode):
    check(gpu_ctx, W.glyph_run(), mode, max_lsb=0)


@pytest.mark.parametrize("mode", MODES)
def test_star_blobs_small(gpu_ctx, mode):
    # blobs exceed a wave's tile grid: the batched call hands them to the plain pipeline
    b = W.star_blobs()
    order = np.argsort(np.diff(b.verb_off))
    check(gpu_ctx, b.subset(order[:3]), mode, max_lsb=0)


_SCENES_REF = {}


def scenes_small():
    """The batched-scenes generator at 8 scenes (1,426 paths, 79 blobs of
    1k-5k cubics, 1.8 M tiles): every structural trait of the full config
    (log-uniform paths per scene, blobs handed on to the plain pipeline,
    rotated and sheared transforms) at a size the oracle renders in seconds."""
    b = W.scenes(n_scenes=8, seed=12, nmax=5e3)
    if "ref" not in _SCENES_REF:
        _SCENES_REF["ref"] = oracle_ref.run_batch(b)
    return b, _SCENES_REF["ref"]


@pytest.mark.parametrize("mode", MODES)
def test_scenes_small(gpu_ctx, mode):
    from ochre_amd.device import run_to_host
    b, ref = scenes_small()
    print(f"[{mode}] {b.name}", flush=True)
    got = run_to_host(gpu_ctx, b, mode, tile_capacity=ref.n_tiles + 64, span_capacity=ref.n_spans + 64)
    st = compare(ref, got, max_lsb=1 if mode == "dense" else 0)
    print(f"[{mode}] {b.name}: {st}", flush=True)
    if mode == "dense":
        assert st["bit_exact_fraction"] >= 0.999


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("limits", [(600000, 0), (0, 4000000), (300000, 2500000)])
def test_plain_split_into_path_ranges(gpu_ctx, mode, limits):
    """Per-pass limits lowered so that the plain pipeline (run directly, or
    on the paths the batched kernels hand on) cuts the batch into path
    ranges: the output must still be the oracle's, with per-path ranges
    pointing into one contiguous tile and span array (in "dense" mode the
    handed-on paths take the unordered tile stage: 1 LSB).  The batch has 9.6 M
    lines (at most 251 k in one path) and at most 68 M sort entries (1.9 M in
    one path), so every limit here cuts it into several ranges while each
    path still fits one pass.  The first attempt runs out of output room
    (capacity 1024), so the ranges after the overflow only count."""
    from ochre_amd.device import run_to_host, set_pass_limits
    b, ref = scenes_small()
    print(f"[{mode}] {b.name} limits={limits}", flush=True)
    set_pass_limits(gpu_ctx, *limits)
    try:
        got = run_to_host(gpu_ctx, b, mode, tile_capacity=1024, span_capacity=1024)  # grown once, after -4
    finally:
        set_pass_limits(gpu_ctx, 0, 0)
    st = compare(ref, got, max_lsb=1 if mode == "dense" else 0)
    print(f"[{mode}] split {limits}: {st}", flush=True)
    if mode == "dense":
        assert st["bit_exact_fraction"] >= 0.999


def test_plain_single_path_over_pass_limits(gpu_ctx):
    """A single path over the per-pass limits cannot be split: TOO_LARGE."""
    from ochre_amd import _lib
    from ochre_amd.device import run_to_host, set_pass_limits
    b = W.star_blobs()
    b = b.subset(np.argsort(np.diff(b.verb_off))[:1])
    set_pass_limits(gpu_ctx, 1000, 0)
    try:
        with pytest.raises(_lib.OchreError) as e:
            run_to_host(gpu_ctx, b, "plain")
    finally:
        set_pass_limits(gpu_ctx, 0, 0)
    assert e.value.status == -7


@pytest.mark.parametrize("mode", MODES)
def test_adjoint_glyphs(gpu_ctx, mode):
    check(gpu_ctx, W.adjoint_glyphs(), mode, max_lsb=0)


@pytest.mark.parametrize("mode", MODES)
def test_edge_cases(gpu_ctx, mode):
    M, L, Q, C, K, Z = W.MOVE, W.LINE, W.QUAD, W.CUBIC, W.CONIC, W.CLOSE
    paths = [
        ([], []),                                            # empty: one zero tile at (0, 0)
        ([M], [(5, 5)]),                                     # a lone Move
        ([M, M, M], [(1, 1), (9, 9), (20, 3)]),              # Moves only: closing lines of zero length
        ([L, L], [(30, 10), (10, 30)]),                      # no Move: starts at the origin
        ([M, L, Z, L, Z], [(3, 3), (40, 3), (3, 40)]),       # Close is a no-op
        ([M, L, L, L], [(8, 8), (72, 8), (72, 72), (8, 72)]),  # integer square
        ([M, L, L, L], [(8.5, 8.5), (40.5, 8.5), (40.5, 40.5), (8.5, 40.5)]),
        ([M, L, L, L, M, L, L, L], [(0, 0), (64, 0), (64, 64), (0, 64), (16, 16), (48, 16), (48, 48), (16, 48)]),
        ([M, L, L, L, M, L, L, L], [(0, 0), (64, 0), (64, 64), (0, 64), (16, 16), (16, 48), (48, 48), (48, 16)]),
        ([M, L, L], [(-40.25, -3.5), (-3.75, -60.5), (-20, 5)]),  # negative coordinates
        ([M, Q, Q], [(10, 10), (10, 10), (10, 10), (50, 10), (90, 10)]),  # degenerate quadratics
        ([M, C, Z], [(10, 10), (10, 10), (10, 10), (10, 10)]),  # a point-cubic
        ([M, K, K], [(0, 0), (50, 0), (50, 50), (0, 50), (0, 0)]),  # conics
        ([M, L, L, L], [(31990, 31990), (32100, 31995), (32200, 32300), (31000, 31000)]),  # past i16
        ([M, L, L], [(100, 100), (100.0000001, 300), (100, 500)]),  # nearly vertical
        ([M, L, L], [(0.5, 0.5), (4000.5, 1.5), (3.5, 3000.25)]),  # long lines
    ]
    weights = {12: [0.7, 2.5]}
    items = []
    for i, (v, p) in enumerate(paths):
        items.append((v, np.asarray(p, np.float32).reshape(-1, 2), weights.get(i, [])))
    b = W.from_paths(items, W.IDENTITY)
    b.name = "edge_cases"
    check(gpu_ctx, b, mode, max_lsb=0)


def test_rasterizer_api_event_order(gpu_ctx):
    """The host Rasterizer replays tiles and spans in the reference's order."""
    from ochre_amd import EventRecorder, Rasterizer
    for batch in (W.readme_path(), W.random_paths(20, 11)):
        for i in range(batch.n_paths):
            v, p, w, xf = batch.path(i)
            r = Rasterizer()
            r.fill_arrays(v, p, w, xf)
            rec = EventRecorder()
            r.finish(rec)
            o = oracle_ref.Rasterizer()
            o.fill(v, p, w, xf)
            ref = o.finish()
            assert len(rec.events) == len(ref), (batch.name, i)
            for a, b in zip(rec.events, ref):
                assert a[:3] == b[:3], (batch.name, i, a[:3], b[:3])
                if a[0] == "tile":
                    assert np.array_equal(a[3], b[3])
                else:
                    assert a[3] == b[3]


def test_rasterizer_api_single_large_fill(gpu_ctx):
    """A single fill goes through the ordered batched kernel; a path over its
    limits (here about 80 x 80 tiles) is handed to the plain pipeline inside the
    same call, and the events still match the reference's exactly."""
    from ochre_amd import EventRecorder, Rasterizer
    b = W.star_blobs(n=1, seed=5, nmin=1500, nmax=1500, rmin=300.0, rmax=300.0)
    v, p, w, xf = b.path(0)
    r = Rasterizer()
    r.fill_arrays(v, p, w, xf)
    rec = EventRecorder()
    r.finish(rec)
    o = oracle_ref.Rasterizer()
    o.fill(v, p, w, xf)
    ref = o.finish()
    tiles = sum(1 for e in ref if e[0] == "tile")
    print(f"[api-large-fill] events={len(ref)} tiles={tiles}", flush=True)
    assert tiles > 32 * 32 // 4
    assert len(rec.events) == len(ref)
    for a, c in zip(rec.events, ref):
        assert a[:3] == c[:3]
        if a[0] == "tile":
            assert np.array_equal(a[3], c[3])
        else:
            assert a[3] == c[3]


def test_rasterizer_api_raw_commands_and_multiple_fills(gpu_ctx):
    """move_to/line_to/command (untransformed) and fill (transformed)
    segments accumulate into one path, as in the reference."""
    from ochre_amd import EventRecorder, Rasterizer
    xf = np.array([1.5, 0.25, -0.5, 1.25, 30.0, 12.0], np.float32)
    r = Rasterizer()
    o = oracle_ref.Rasterizer()
    for rr in (r, o):
        rr.move_to(10.5, 10.25) if rr is o else rr.move_to(__import__("ochre_amd").Vec2(10.5, 10.25))
    r.line_to(__import__("ochre_amd").Vec2(60.0, 14.0)); o.line_to(60.0, 14.0)
    v = np.array([W.QUAD, W.CUBIC], np.uint8)
    p = np.array([(70, 40), (40, 70), (20, 60), (5, 50), (12, 30)], np.float32)
    r.fill_arrays(v, p, [], xf); o.fill(v, p, [], xf)
    from ochre_amd import PathCmd, Vec2
    r.command(PathCmd.Conic(Vec2(80, 80), Vec2(20, 90), 1.7)); o.command(W.CONIC, [(80, 80), (20, 90)], 1.7)
    v2 = np.array([W.MOVE, W.LINE, W.LINE], np.uint8)
    p2 = np.array([(100, 100), (140, 105), (120, 150)], np.float32)
    r.fill_arrays(v2, p2, [], None); o.fill(v2, p2, [], None)
    rec = EventRecorder()
    r.finish(rec)
    ref = o.finish()
    assert [e[:3] for e in rec.events] == [e[:3] for e in ref]
    for a, b in zip(rec.events, ref):
        if a[0] == "tile":
            assert np.array_equal(a[3], b[3])


def test_batched_params_and_graph_replays(gpu_ctx):
    """Every combination of persistent waves per CU, paths per queue pop and
    graph replay gives the oracle's result, on several consecutive replays
    with host work and synchronisation between them."""
    import ctypes

    import torch

    from ochre_amd.device import DeviceOutput, rasterize_batched, upload
    batch = W.concat([W.glyph_run(n=1500, seed=9), W.random_paths(200, 4)])
    batch.name = "glyphs+random"
    ref = oracle_ref.run_batch(batch)
    db = upload(batch)
    lib = gpu_ctx.lib
    # (dense, groups/CU, waves/group, small-grid groups/CU, 256-tile instance, waves/CU, paths/pop,
    #  tiles/chunk, graph)
    configs = [(0, 0, 0, 0, 0, 2, 1, 32, 1), (0, 0, 0, 0, 0, 6, 1, 16, 1), (0, 0, 0, 0, 0, 8, 4, 32, 1),
               (0, 0, 0, 0, 0, 4, 16, 16, 0), (0, 0, 0, 0, 0, 1, 2, 32, 0), (1, 4, 4, 8, 1, 6, 2, 32, 1),
               (1, 2, 2, 0, 0, 6, 2, 32, 1), (1, 1, 4, 4, 1, 2, 1, 16, 0), (1, 3, 4, 6, 0, 8, 1, 32, 1),
               (1, 4, 4, 0, 1, 1, 1, 32, 1), (1, 4, 4, 0, 0, 1, 2, 32, 1), (1, 2, 8, 0, 1, 2, 2, 32, 1),
               (1, 1, 8, 4, 0, 1, 1, 32, 0)]
    try:
        for dense, dg, dw, sg, lg, waves, gran, tc, graph in configs:
            assert lib.ochre_context_set_dense_params(gpu_ctx.h, dense, max(dg, 1), dw or 4) == 0
            assert lib.ochre_context_set_dense_small(gpu_ctx.h, sg) == 0
            assert lib.ochre_context_set_dense_large(gpu_ctx.h, lg) == 0
            assert lib.ochre_context_set_batched_params(gpu_ctx.h, waves, gran, tc, graph) == 0
            for rep in range(3):
                out = DeviceOutput(batch.n_paths, ref.n_tiles + 16, ref.n_spans + 16)
                rasterize_batched(gpu_ctx, db, out)
                torch.cuda.synchronize()
                st = compare(ref, out.to_host(), max_lsb=1 if dense else 0)
                print(f"[batched] dense={dense}/{dg}/{dw}/{sg}/{lg} waves={waves} gran={gran} tc={tc} graph={graph} "
                      f"rep={rep}: {st}", flush=True)
                if dense:
                    assert st["bit_exact_fraction"] >= 0.999
                _ = np.sum(np.arange(1000))  # host work between replays
    finally:
        lib.ochre_context_set_batched_params(gpu_ctx.h, 6, 2, 32, 1)
        lib.ochre_context_set_dense_params(gpu_ctx.h, 1, 4, 4)
        lib.ochre_context_set_dense_small(gpu_ctx.h, 0)
        lib.ochre_context_set_dense_large(gpu_ctx.h, 1)


@pytest.mark.parametrize("mode", MODES)
def test_non_terminating_line_returns_flatten_limit(gpu_ctx, mode):
    """The close of M(inf,3) L(40,5) L(20,40) never reaches t = 1 in the
    reference's DDA.  Every mode reaches the plain pipeline's count walk (the
    dense and ordered batched kernels hand a path with non-finite points on
    before walking), which stops after 2^20 steps: OCHRE_ERR_FLATTEN_LIMIT,
    the status the oracle's guard gives for the same path."""
    from ochre_amd import _lib
    from ochre_amd.device import run_to_host
    b = W.from_paths([([W.MOVE, W.LINE, W.LINE], np.array([(np.inf, 3), (40, 5), (20, 40)], np.float32))],
                     W.IDENTITY)
    assert oracle_ref.run_batch(b).status == 2
    print(f"[{mode}] non-terminating close", flush=True)
    with pytest.raises(_lib.OchreError) as e:
        run_to_host(gpu_ctx, b, mode)
    assert e.value.status == -5


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_run_to_run_determinism(gpu_ctx, mode):
    """The exact modes are reproducible: two runs of the same batch give
    byte-identical per-path output (the ordered modes sum each pixel in
    emission order, rasterizer.rs:209-228).  The batch mixes glyphs, fuzz paths
    and a blob the batched call hands to the plain pipeline.  (The dense mode
    sums in arrival order and is not reproducible bit for bit; it is checked
    against the oracle within 1 LSB instead.)"""
    from ochre_amd.device import run_to_host
    blobs = W.star_blobs()
    small = blobs.subset(np.argsort(np.diff(blobs.verb_off))[:1])
    b = W.concat([W.glyph_run(n=3000, seed=21), W.random_paths(300, 22), small])
    b.name = "determinism"
    print(f"[{mode}] determinism: {b.n_paths} paths", flush=True)
    first = run_to_host(gpu_ctx, b, mode)
    second = run_to_host(gpu_ctx, b, mode)
    for x, y in zip(first, second):
        assert x.dtype == y.dtype and x.shape == y.shape and x.tobytes() == y.tobytes()


def test_readme_fixture_through_the_rasterizer_api(gpu_ctx):
    """The drop-in Rasterizer on the README path prints, in the reference's
    examples/basic.rs format, exactly the committed fixture
    tests/golden/readme_path_basic.txt."""
    import os

    from ochre_amd import EventRecorder, Rasterizer
    from parity import basic_rs_text
    v, p, w, xf = W.readme_path().path(0)
    r = Rasterizer()
    r.fill_arrays(v, p, w, xf)
    rec = EventRecorder()
    r.finish(rec)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "readme_path_basic.txt")
    assert basic_rs_text(rec.events) == open(path).read()
