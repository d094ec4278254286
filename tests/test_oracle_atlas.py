"""Pins the oracle's atlas Builder (examples/svg.rs:15-88) against an
independent Python replay of the Builder over the oracle Rasterizer's own
TileBuilder call sequence, plus the Builder's invariants: slot 0 solid, tile
k in slot k + 1 (row-major, wrapping at atlas_size / 8), quads of 4 vertices
and 6 indices per call in call order, everything else of the atlas zero, and
the overflow report when the atlas is full."""
import numpy as np
import pytest

import oracle_ref as O
from ochre_amd import workloads as W


def replay(batch, colors, size, stroke=False):
    """Python Builder over the oracle's per-path call sequence."""
    atlas = np.zeros((size, size), np.uint8)
    atlas[:8, :8] = 255
    verts, idx = [], []
    row, col = 0, 1
    slots = size // 8
    for i in range(batch.n_paths):
        v, p, w, xf = batch.path(i)
        r = O.Rasterizer()
        if stroke:
            r.stroke(v, p, float(batch.widths[i]), w, xf)
        else:
            r.fill(v, p, w, xf)
        c = tuple(int(x) for x in colors[i])
        for e in r.finish():
            base = len(verts)
            if e[0] == "tile":
                _, x, y, a = e
                u1, u2, v1, v2 = col * 8, (col + 1) * 8, row * 8, (row + 1) * 8
                verts += [((x, y), (u1, v1), c), ((x + 8, y), (u2, v1), c), ((x + 8, y + 8), (u2, v2), c),
                          ((x, y + 8), (u1, v2), c)]
                atlas[row * 8:row * 8 + 8, col * 8:col * 8 + 8] = a.reshape(8, 8)
                col += 1
                if col == slots:
                    col, row = 0, row + 1
            else:
                _, x, y, wd = e
                verts += [((x, y), (0, 0), c), ((x + wd, y), (0, 0), c), ((x + wd, y + 8), (0, 0), c),
                          ((x, y + 8), (0, 0), c)]
            idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    va = np.zeros(len(verts), O.ATLAS_VERTEX)
    for k, (pos, uv, c) in enumerate(verts):
        va[k] = (pos, uv, c)
    return atlas, va, np.asarray(idx, np.uint32)


def colors_for(n, seed):
    return np.random.default_rng(seed).integers(0, 256, (n, 4), dtype=np.uint8)


@pytest.mark.parametrize("name", ["readme", "random", "glyphs"])
def test_builder_matches_replay(name):
    b = {"readme": W.readme_path, "random": lambda: W.random_paths(40, 3),
         "glyphs": lambda: W.glyph_run(60, seed=3)}[name]()
    col = colors_for(b.n_paths, 1)
    st, atlas, verts, idx = O.atlas_batch(b, col, 512)
    assert st == 0
    a2, v2, i2 = replay(b, col, 512)
    assert np.array_equal(atlas, a2)
    assert np.array_equal(verts, v2)
    assert np.array_equal(idx, i2)


def test_builder_layout_invariants():
    b = W.glyph_run(30, seed=4)
    col = colors_for(b.n_paths, 2)
    st, atlas, verts, idx = O.atlas_batch(b, col, 256)
    assert st == 0
    ref = O.run_batch(b)
    T, S = ref.n_tiles, ref.n_spans
    assert verts.size == 4 * (T + S) and idx.size == 6 * (T + S)
    assert (atlas[:8, :8] == 255).all()
    # tile k (in path order) sits in slot k + 1
    slots = 256 // 8
    for k in range(T):
        s = k + 1
        blk = atlas[(s // slots) * 8:(s // slots) * 8 + 8, (s % slots) * 8:(s % slots) * 8 + 8]
        assert np.array_equal(blk.reshape(-1), ref.tile_alpha[k])
    used = np.zeros_like(atlas, bool)
    for s in range(T + 1):
        used[(s // slots) * 8:(s // slots) * 8 + 8, (s % slots) * 8:(s % slots) * 8 + 8] = True
    assert not atlas[~used].any()
    # quads: indices 0,1,2,0,2,3 from each call's base; spans untextured
    q = idx.reshape(-1, 6).astype(np.int64) - 4 * np.arange(T + S)[:, None]
    assert (q == np.array([0, 1, 2, 0, 2, 3])).all()
    untextured = (verts["uv"] == 0).all(axis=1).reshape(-1, 4).all(axis=1)
    assert untextured.sum() == S


def test_stroked_batch_builder():
    b = W.with_widths(W.random_paths(20, 6), 0.5, 4.0, 1)
    col = colors_for(b.n_paths, 3)
    st, atlas, verts, idx = O.atlas_batch(b, col, 512, stroke=True)
    assert st == 0
    a2, v2, i2 = replay(b, col, 512, stroke=True)
    assert np.array_equal(atlas, a2) and np.array_equal(verts, v2) and np.array_equal(idx, i2)


def test_full_atlas_is_reported():
    b = W.glyph_run(30, seed=4)
    col = colors_for(b.n_paths, 2)
    st, *_ = O.atlas_batch(b, col, 32)   # 16 slots, one solid
    assert st & 0x100
