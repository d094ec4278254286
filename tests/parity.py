"""Comparison of device output with the oracle (test infrastructure)."""
from __future__ import annotations

import numpy as np


def compare(ref, got, max_lsb=1):
    """ref: oracle RasterOutput; got: (tile_xy, tile_alpha, span_xy, span_w,
    path_tile_off, path_span_off) in path order.  Asserts identical tile and
    span sets (per path, in order) and alpha within max_lsb; returns stats."""
    txy, ta, sxy, sw, pto, pso = got
    assert np.array_equal(pto, ref.path_tile_off), _first_diff("tiles per path", np.diff(pto), np.diff(ref.path_tile_off))
    assert np.array_equal(pso, ref.path_span_off), _first_diff("spans per path", np.diff(pso), np.diff(ref.path_span_off))
    assert np.array_equal(txy, ref.tile_xy), _first_row_diff("tile coordinates", txy, ref.tile_xy, pto)
    assert np.array_equal(sxy, ref.span_xy), _first_row_diff("span coordinates", sxy, ref.span_xy, pso)
    assert np.array_equal(sw, ref.span_w), _first_diff("span widths", sw, ref.span_w)
    d = np.abs(ta.astype(np.int16) - ref.tile_alpha.astype(np.int16))
    mx = int(d.max()) if d.size else 0
    if mx > max_lsb:
        t = int(np.argmax(d.max(axis=1)))
        p = int(np.searchsorted(pto, t, side="right") - 1)
        raise AssertionError(f"alpha differs by {mx} LSB at tile {t} (path {p}, xy {tuple(txy[t])}):\n"
                             f"got\n{ta[t].reshape(8, 8)}\nref\n{ref.tile_alpha[t].reshape(8, 8)}")
    exact = float((d == 0).mean()) if d.size else 1.0
    return {"tiles": int(txy.shape[0]), "spans": int(sxy.shape[0]), "max_lsb": mx, "bit_exact_fraction": exact}


def _first_diff(what, a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return f"{what}: shapes {a.shape} vs {b.shape}"
    idx = np.nonzero(a != b)[0]
    i = int(idx[0]) if idx.size else -1
    return f"{what}: first difference at {i}: got {a[i] if i >= 0 else None} ref {b[i] if i >= 0 else None}"


def _first_row_diff(what, a, b, off):
    if a.shape != b.shape:
        return f"{what}: shapes {a.shape} vs {b.shape}"
    rows = np.nonzero((a != b).any(axis=1))[0]
    i = int(rows[0])
    p = int(np.searchsorted(off, i, side="right") - 1)
    return f"{what}: first difference at {i} (path {p}): got {a[i]} ref {b[i]}"


def path_digests(txy, ta, sxy, sw, pto, pso):
    """Per-path SHA-256 digests of an output in TileBuilder order: `sets`
    covers the tile coordinates, span coordinates and span widths (the
    integer part of the contract), `alpha` the tiles' alpha bytes.  Returns a
    list of (n_tiles, n_spans, sets_hex, alpha_hex)."""
    import hashlib
    txy = np.ascontiguousarray(txy, np.int16)
    ta = np.ascontiguousarray(ta, np.uint8)
    sxy = np.ascontiguousarray(sxy, np.int16)
    sw = np.ascontiguousarray(sw, np.uint16)
    out = []
    for i in range(len(pto) - 1):
        t0, t1, s0, s1 = int(pto[i]), int(pto[i + 1]), int(pso[i]), int(pso[i + 1])
        h = hashlib.sha256()
        h.update(txy[t0:t1].tobytes())
        h.update(b"|")
        h.update(sxy[s0:s1].tobytes())
        h.update(b"|")
        h.update(sw[s0:s1].tobytes())
        a = hashlib.sha256(ta[t0:t1].tobytes())
        out.append((t1 - t0, s1 - s0, h.hexdigest(), a.hexdigest()))
    return out


def input_digests(batch):
    """Per-path SHA-256 of a batch's input (verbs, points, weights and
    transform bytes): shows that a generator reproduced the paths a fixture
    was made from."""
    import hashlib
    out = []
    for i in range(batch.n_paths):
        v, p, w, xf = batch.path(i)
        h = hashlib.sha256()
        for a in (v, p, w, xf):
            h.update(np.ascontiguousarray(a).tobytes())
            h.update(b"|")
        out.append(h.hexdigest())
    return out


def basic_rs_text(events):
    """TileBuilder events as the reference's examples/basic.rs:6-19 prints
    them: "tile at (x, y):" and eight rows of "{:3} " alphas indented by two
    spaces, or "span at (x, y), width w"."""
    lines = []
    for e in events:
        if e[0] == "tile":
            _, x, y, a = e
            a = np.asarray(a, np.uint8).reshape(8, 8)
            lines.append(f"tile at ({x}, {y}):")
            for row in a:
                lines.append("  " + "".join(f"{int(v):3d} " for v in row))
        else:
            _, x, y, w = e
            lines.append(f"span at ({x}, {y}), width {w}")
    return "\n".join(lines) + "\n"
