"""Pins the CPU oracle with checks that do not restate its formulas:

* f64 brute-force coverage: each pixel's exact winding integral computed by
  Sutherland-Hodgman clipping of every contour to the pixel square, against
  the tiles + spans the oracle emits (alpha = min(|cov|, 1) * 256, truncated,
  within 1 LSB for f32 rounding);
* flattening geometry: emitted points lie on the f64 curve, in order, the
  last equals the end point, chords stay within the tolerance, and segment
  counts match ceil(1/dt) of the f64 step size;
* closed forms: the README path flattens to 23 + 47 = 70 lines; integer
  squares give 255 inside and 0 outside; half-pixel edges give 128;
* quirks of the reference: an empty path emits one zero tile at (0, 0);
  Close is a no-op; repeated fill() calls accumulate into one path; a path
  without a leading Move starts at the origin.
"""
import math

import numpy as np
import pytest

import oracle_ref as O
from ochre_amd import workloads as W

M, L, Q, C, K, Z = W.MOVE, W.LINE, W.QUAD, W.CUBIC, W.CONIC, W.CLOSE


# ------------------------------------------------------------ brute force

def clip_area(poly, x0, y0):
    """Signed area of the closed polygon `poly` (f64 [n, 2]) clipped to the
    unit square at (x0, y0) (Sutherland-Hodgman; preserves the winding
    integral for self-intersecting input)."""
    pts = list(map(tuple, poly))
    for axis, bound, keep_ge in ((0, x0, True), (0, x0 + 1, False), (1, y0, True), (1, y0 + 1, False)):
        if not pts:
            return 0.0
        out = []
        n = len(pts)
        for i in range(n):
            a, b = pts[i - 1], pts[i]
            ina = (a[axis] >= bound) if keep_ge else (a[axis] <= bound)
            inb = (b[axis] >= bound) if keep_ge else (b[axis] <= bound)
            if inb:
                if not ina:
                    out.append(_intersect(a, b, axis, bound))
                out.append(b)
            elif ina:
                out.append(_intersect(a, b, axis, bound))
        pts = out
    if len(pts) < 3:
        return 0.0
    p = np.asarray(pts)
    return 0.5 * float(np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1]))


def _intersect(a, b, axis, bound):
    t = (bound - a[axis]) / (b[axis] - a[axis])
    return (a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1]))


def exact_alpha(contours, x0, y0, w, h):
    cov = np.zeros((h, w))
    for poly in contours:
        poly = np.asarray(poly, np.float64)
        lo = np.floor(poly.min(0)).astype(int)
        hi = np.floor(poly.max(0)).astype(int)
        for py in range(max(lo[1], y0), min(hi[1], y0 + h - 1) + 1):
            for px in range(max(lo[0], x0), min(hi[0], x0 + w - 1) + 1):
                cov[py - y0, px - x0] += clip_area(poly, px, py)
    # Rasterizer coverage is positive for clockwise-in-y-down winding; the
    # nonzero rule takes |.|, so the sign convention does not matter here.
    return cov


def render(events, x0, y0, w, h):
    img = np.zeros((h, w), np.int32)
    covered = np.zeros((h, w), bool)
    for e in events:
        if e[0] == "tile":
            _, x, y, a = e
            a = a.reshape(8, 8)
            for r in range(8):
                for c in range(8):
                    yy, xx = y + r - y0, x + c - x0
                    if 0 <= yy < h and 0 <= xx < w:
                        img[yy, xx] = a[r, c]
                        covered[yy, xx] = True
        else:
            _, x, y, wd = e
            for r in range(8):
                for c in range(wd):
                    yy, xx = y + r - y0, x + c - x0
                    if 0 <= yy < h and 0 <= xx < w:
                        img[yy, xx] = 255
                        covered[yy, xx] = True
    return img, covered


def rasterize_polygons(contours):
    """Rasterizes closed polygons (lines only) through the oracle."""
    r = O.Rasterizer()
    for poly in contours:
        verbs = [M] + [L] * (len(poly) - 1)
        r.fill(verbs, np.asarray(poly, np.float32))
    return r.finish()


def check_against_brute_force(contours, margin=9):
    f32 = [np.asarray(p, np.float32).astype(np.float64) for p in contours]
    allp = np.concatenate(f32)
    x0 = int(math.floor(allp[:, 0].min())) - margin
    y0 = int(math.floor(allp[:, 1].min())) - margin
    w = int(math.ceil(allp[:, 0].max())) + margin - x0
    h = int(math.ceil(allp[:, 1].max())) + margin - y0
    events = rasterize_polygons(contours)
    img, covered = render(events, x0, y0, w, h)
    cov = np.abs(exact_alpha(f32, x0, y0, w, h))
    want = np.minimum(np.floor(cov * 256.0), 255).astype(np.int32)
    diff = np.abs(img - want)
    assert diff.max() <= 1, (diff.max(), np.argwhere(diff > 1)[:5])
    # everything with coverage is inside a tile or a span
    assert not np.any((cov > 1e-6) & ~covered)
    # span pixels are fully covered
    for e in events:
        if e[0] == "span":
            _, x, y, wd = e
            assert np.all(cov[y - y0:y - y0 + 8, x - x0:x - x0 + wd] > 1.0 - 1e-5)
    return events, diff


def star_polygon(rng, n, cx, cy, r0, r1):
    th = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = rng.uniform(r0, r1, n)
    return np.stack([cx + r * np.cos(th), cy + r * np.sin(th)], 1)


@pytest.mark.parametrize("seed", range(6))
def test_simple_polygons_match_exact_coverage(seed):
    rng = np.random.default_rng(seed)
    poly = star_polygon(rng, int(rng.integers(3, 12)), rng.uniform(-20, 20), rng.uniform(-20, 20), 3, 18)
    check_against_brute_force([poly])


@pytest.mark.parametrize("seed", range(4))
def test_self_intersecting_polygons_match_winding_integral(seed):
    rng = np.random.default_rng(100 + seed)
    poly = rng.uniform(-15, 25, (int(rng.integers(4, 9)), 2))
    check_against_brute_force([poly])


def test_holes_and_double_winding():
    outer = [(0, 0), (40, 0), (40, 40), (0, 40)]
    hole = [(10.25, 10.5), (10.25, 29.75), (30.5, 29.75), (30.5, 10.5)]   # opposite orientation
    same = [(5.5, 5.5), (25.5, 5.5), (25.5, 25.5), (5.5, 25.5)]           # winding 2 where overlapping
    check_against_brute_force([outer, hole])
    events, _ = check_against_brute_force([outer, same])
    # winding 2 saturates to 255 (nonzero rule with abs)
    assert any(e[0] == "span" for e in events)


def test_integer_square_ring_and_spans():
    events, _ = check_against_brute_force([[(8, 8), (72, 8), (72, 72), (8, 72)]])
    tiles = {(e[1], e[2]): e[3] for e in events if e[0] == "tile"}
    spans = [tuple(e) for e in events if e[0] == "span"]
    # Edges on x = 8 / x = 72 and y = 8 / y = 72 touch tile columns and rows
    # 1 and 9 (x = 72 lands in pixel 72 with zero net coverage): a ring of 32
    # tiles; rows y = 16..64 have one span from x = 16 of width 56.
    ring = {(8 * i, 8 * j) for i in range(1, 10) for j in range(1, 10) if i in (1, 9) or j in (1, 9)}
    assert set(tiles) == ring
    assert spans == [("span", 16, y, 56) for y in range(16, 72, 8)]
    for (x, y), a in tiles.items():
        inside = x <= 64 and y <= 64   # the tile lies within [8, 72)^2
        assert (a == 255).all() if inside else not a.any()


def test_half_pixel_edges_are_128():
    events, _ = check_against_brute_force([[(8.5, 8.5), (40.5, 8.5), (40.5, 40.5), (8.5, 40.5)]])
    img, _ = render(events, 0, 0, 48, 48)
    assert img[20, 8] == 128 and img[20, 40] == 128 and img[8, 20] == 128 and img[40, 20] == 128
    assert img[8, 8] == 64  # quarter corner
    assert img[20, 20] == 255 and img[20, 7] == 0


# -------------------------------------------------------------- flattening

def bezier(ctrl, t):
    ctrl = np.asarray(ctrl, np.float64)
    n = len(ctrl) - 1
    pts = np.zeros((len(t), 2))
    for k in range(n + 1):
        b = math.comb(n, k) * (1 - t) ** (n - k) * t ** k
        pts += b[:, None] * ctrl[k]
    return pts


def conic(ctrl, w, t):
    p0, p1, p2 = (np.asarray(c, np.float64) for c in ctrl)
    a, b, c = (1 - t) ** 2, 2 * t * (1 - t) * w, t ** 2
    return (a[:, None] * p0 + b[:, None] * p1 + c[:, None] * p2) / (a + b + c)[:, None]


def project(curve_pts, ts, q):
    d = np.linalg.norm(curve_pts - q, axis=1)
    i = int(np.argmin(d))
    return ts[i], d[i]


@pytest.mark.parametrize("kind,seed", [(Q, s) for s in range(5)] + [(C, s) for s in range(5)] + [(K, s) for s in range(4)])
def test_flattened_points_lie_on_curve_within_tolerance(kind, seed):
    rng = np.random.default_rng(1000 + seed)
    npts = {Q: 2, C: 3, K: 2}[kind]
    ctrl = rng.uniform(0, 60, (npts + 1, 2)).astype(np.float32)
    w = float(rng.uniform(0.3, 3.0)) if kind == K else None
    verbs = [M, kind]
    st, (v, p, _) = O.flatten_path(verbs, ctrl, [w] if kind == K else [], 0.1)
    assert st == 0
    assert v[0] == M and np.all(v[1:] == L)
    pts = p[1:].astype(np.float64)
    c64 = ctrl.astype(np.float64)
    ts = np.linspace(0, 1, 200001)
    curve = conic(c64, w, ts) if kind == K else bezier(c64, ts)
    assert np.array_equal(p[-1], ctrl[-1])           # ends exactly at the end point
    prev_t = 0.0
    chain = [c64[0]] + list(pts)
    tvals = [0.0]
    for q in pts:
        t, d = project(curve, ts, q)
        assert d < 2e-3, d                           # on the curve
        assert t >= prev_t - 1e-4                    # in order
        prev_t = t
        tvals.append(t)
    # Chord deviation bound.  Quadratics: dt^2/8 * |B''| = tol exactly for the
    # reference's step size.  Cubics: the textbook bound dt^2/8 * max|B''|
    # (B'' is linear, so its maximum is at an end); the reference's step size
    # does not guarantee the tolerance itself.  Conics: the midpoint test only
    # bounds the error loosely.
    if kind == Q:
        bound = 0.1
    elif kind == C:
        d1 = np.linalg.norm(c64[0] - 2 * c64[1] + c64[2])
        d2 = np.linalg.norm(c64[1] - 2 * c64[2] + c64[3])
        a = -c64[0] + 3 * c64[1] - 3 * c64[2] + c64[3]
        b = 3 * (c64[0] - 2 * c64[1] + c64[2])
        dt = math.sqrt(math.sqrt(8) * 0.1 / max(np.linalg.norm(b), np.linalg.norm(a + b)))
        bound = dt * dt / 8 * 6 * max(d1, d2)
    else:
        bound = 0.25
    for k in range(1, len(chain)):
        a, b = chain[k - 1], chain[k]
        t0, t1 = tvals[k - 1], tvals[k]
        sub = np.linspace(t0, t1, 64)
        seg = conic(c64, w, sub) if kind == K else bezier(c64, sub)
        ab = b - a
        n = np.linalg.norm(ab)
        dd = seg - a
        dist = np.abs(ab[0] * dd[:, 1] - ab[1] * dd[:, 0]) / n if n > 0 else np.linalg.norm(dd, axis=1)
        assert dist.max() <= bound * 1.01 + 2e-3
    if kind != K:
        # segment count: the t loop runs ceil(1/dt) times (f64 dt, f32 rounding of t allows +-1)
        if kind == Q:
            dd = np.linalg.norm(c64[0] - 2 * c64[1] + c64[2])
            dt = math.sqrt(0.4 / dd)
        else:
            a = -c64[0] + 3 * c64[1] - 3 * c64[2] + c64[3]
            b = 3 * (c64[0] - 2 * c64[1] + c64[2])
            dt = math.sqrt(math.sqrt(8) * 0.1 / max(np.linalg.norm(b), np.linalg.norm(a + b)))
        assert abs(len(pts) - math.ceil(1 / dt)) <= 1


def test_readme_path_flattens_to_70_lines():
    r = O.Rasterizer()
    b = W.readme_path()
    v, p, w, xf = b.path(0)
    r.fill(v, p, w, xf)
    assert r.stats()["lines"] == 70  # 23 (quadratic) + 47 (cubic); the cubic ends at the start
    ev = r.finish()
    assert sum(e[0] == "tile" for e in ev) == 100 and sum(e[0] == "span" for e in ev) == 23


def test_degenerate_curves_emit_one_line():
    st, (v, p, _) = O.flatten_path([M, Q], [(0, 0), (5, 5), (10, 10)], [], 0.1)
    assert len(v) == 2 and np.array_equal(p[-1], [10, 10])
    st, (v, p, _) = O.flatten_path([M, C], [(3, 3)] * 4, [], 0.1)
    assert len(v) == 2


# ------------------------------------------------------------------ quirks

def test_empty_path_emits_one_zero_tile_at_origin():
    ev = O.Rasterizer().finish()
    assert len(ev) == 1 and ev[0][:3] == ("tile", 0, 0) and not ev[0][3].any()
    r = O.Rasterizer()
    r.fill([M, M], [(50, 50), (90, 90)])  # Moves only: no increments
    ev = r.finish()
    assert len(ev) == 1 and ev[0][:3] == ("tile", 0, 0)


def test_close_is_a_no_op():
    a = O.Rasterizer()
    a.fill([M, L, L, Z, L], [(3, 3), (40, 5), (20, 40), (50, 50)])
    b = O.Rasterizer()
    b.fill([M, L, L, L], [(3, 3), (40, 5), (20, 40), (50, 50)])
    ea, eb = a.finish(), b.finish()
    assert [e[:3] for e in ea] == [e[:3] for e in eb]
    assert all(np.array_equal(x[3], y[3]) for x, y in zip(ea, eb) if x[0] == "tile")


def test_fills_accumulate_into_one_path():
    pts = np.array([(3, 3), (40, 5), (20, 40), (60, 60), (90, 62), (70, 95)], np.float32)
    a = O.Rasterizer()
    a.fill([M, L, L], pts[:3])
    a.fill([M, L, L], pts[3:])
    b = O.Rasterizer()
    b.fill([M, L, L, M, L, L], pts)
    ea, eb = a.finish(), b.finish()
    assert [e[:3] for e in ea] == [e[:3] for e in eb]


def test_path_without_move_starts_at_origin():
    a = O.Rasterizer()
    a.fill([L, L], [(30, 10), (10, 30)])
    b = O.Rasterizer()
    b.fill([M, L, L], [(0, 0), (30, 10), (10, 30)])
    assert [e[:3] for e in a.finish()] == [e[:3] for e in b.finish()]


def test_transform_is_applied_before_flattening():
    # a quadratic scaled by 4 flattens with the device-pixel tolerance: more lines
    a = O.Rasterizer()
    a.fill([M, Q], [(0, 0), (10, 20), (20, 0)])
    b = O.Rasterizer()
    b.fill([M, Q], [(0, 0), (10, 20), (20, 0)], xform=[4, 0, 0, 4, 0, 0])
    assert b.stats()["lines"] > a.stats()["lines"]


def test_batch_runner_matches_single_rasterizers():
    b = W.random_paths(40, 3)
    res = O.run_batch(b)
    for i in range(b.n_paths):
        v, p, w, xf = b.path(i)
        r = O.Rasterizer()
        r.fill(v, p, w, xf)
        ev = [e for e in r.finish() if e[0] == "tile"]
        t0, t1 = res.path_tile_off[i], res.path_tile_off[i + 1]
        assert len(ev) == t1 - t0
        for k, e in enumerate(ev):
            assert tuple(res.tile_xy[t0 + k]) == (e[1], e[2])
            assert np.array_equal(res.tile_alpha[t0 + k], e[3])


# ------------------------------------------------------- guards and Q4 pins

def test_non_terminating_line_stops_at_walk_limit():
    """The close of M(inf,3) L(40,5) L(20,40) joins (20,40) to (inf,3): its
    crossing step |1/dx| is 0, so the reference's DDA (rasterizer.rs:97-136)
    never reaches t = 1.  The oracle stops after 2^20 increments with the
    flatten-limit status, as the device walks do, instead of growing its
    increment list until the allocation fails."""
    b = W.from_paths([([M, L, L], np.array([(np.inf, 3), (40, 5), (20, 40)], np.float32))], W.IDENTITY)
    res = O.run_batch(b)
    assert res.status == 2  # ORC_ERR_FLATTEN_LIMIT
    assert res.stats["increments"] <= (1 << 20) + 64


def _events(verbs, pts):
    r = O.Rasterizer()
    r.fill(verbs, pts)
    return r.finish()


def test_zero_length_step_at_a_tile_corner_adds_tiles():
    """Hand-derived: the segment (7.5,7.5)-(8.5,8.5) crosses the pixel corner
    (8,8), which is also a tile corner.  dtdx = dtdy = 1, so row_t1 = col_t1
    = 0.5: the tie steps x first (rasterizer.rs:108), and the next increment,
    at pixel (8,7) with t0 = t1 = 0.5, has zero length.  Its close back to
    (7.5,7.5) does the same at pixel (7,8).  Increments: (7,7) area 1/8 height
    1/2, (8,7) zero, (8,8) area 3/8 height 1/2; then (8,8) area -3/8 height
    -1/2, (7,8) zero, (7,7) area -1/8 height -1/2.  Every pixel sums to 0, so
    the output is four all-zero tiles in row-major order; the two off-diagonal
    ones exist only because of the zero-length increments.  Crossings:
    (tx 1, ty 0, +1) and (tx 0, ty 0, -1) sum to 0, and no row has a gap, so
    there are no spans."""
    ev = _events([M, L], [(7.5, 7.5), (8.5, 8.5)])
    assert [e[:3] for e in ev] == [("tile", 0, 0), ("tile", 8, 0), ("tile", 0, 8), ("tile", 8, 8)]
    assert all(not e[3].any() for e in ev)


def test_negative_start_on_an_integer_coordinate_adds_a_tile():
    """Hand-derived: the horizontal segment (8,4.5) -> (6.5,4.5) starts on
    x = 8 and moves left, so col_t1 = dtdx * (8 - 8) = -0 (rasterizer.rs:91-92)
    and the first increment, at pixel (8,4) with t0 = 0, t1 = -0, has zero
    length.  The rest (pixels (7,4), (6,4)) and the close back to (8,4.5)
    (pixels (6,4), (7,4); its last step consumes t = 1 and ends the walk
    before pixel (8,4)) have height 0.  So the output is two all-zero tiles,
    (0,0) and (8,0); tile (8,0) comes from the -0 increment alone."""
    ev = _events([M, L], [(8.0, 4.5), (6.5, 4.5)])
    assert [e[:3] for e in ev] == [("tile", 0, 0), ("tile", 8, 0)]
    assert all(not e[3].any() for e in ev)


def _seg_tiles(a, b):
    """Tiles whose open interior the f64 segment a-b passes through
    (Liang-Barsky clipping against each candidate tile square)."""
    out = set()
    lo = np.floor(np.minimum(a, b) / 8).astype(int)
    hi = np.floor(np.maximum(a, b) / 8).astype(int)
    d = b - a
    for ty in range(lo[1], hi[1] + 1):
        for tx in range(lo[0], hi[0] + 1):
            t0, t1 = 0.0, 1.0
            ok = True
            for k, (mn, mx) in enumerate(((8 * tx, 8 * tx + 8), (8 * ty, 8 * ty + 8))):
                if d[k] == 0.0:
                    if not (mn < a[k] < mx):
                        ok = False
                    continue
                u0, u1 = (mn - a[k]) / d[k], (mx - a[k]) / d[k]
                t0, t1 = max(t0, min(u0, u1)), min(t1, max(u0, u1))
            if ok and t1 - t0 > 1e-12:
                out.add((tx, ty))
    return out


def _winding(poly, x, y):
    """Nonzero winding number of the closed f64 polygon around (x, y)."""
    w = 0
    n = len(poly)
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
        if y0 <= y < y1 and (x1 - x0) * (y - y0) - (x - x0) * (y1 - y0) > 0:
            w += 1
        elif y1 <= y < y0 and (x1 - x0) * (y - y0) - (x - x0) * (y1 - y0) < 0:
            w -= 1
    return w


def test_readme_tiles_and_spans_from_geometry():
    """The README path's tile and span sets, derived without the DDA: the
    tiles are those whose interior the flattened outline passes through (f64
    clipping; the path's coordinates hit no tile corner), and a gap between
    two tiles of a row is a span exactly when the outline winds around the
    gap's pixels (f64 winding number at the first gap pixel's centre).  This
    pins the 100 tiles and 23 spans the oracle emits."""
    b = W.readme_path()
    v, p, w, xf = b.path(0)
    st, (fv, fp, _) = O.flatten_path(v, p, w, 0.1)
    assert st == 0
    assert fv[0] == M and (fv[1:-1] == L).all() and fv[-1] == Z  # one contour: Move, 70 Lines, Close
    poly = fp.astype(np.float64)
    assert np.array_equal(poly[0], poly[-1])  # the cubic ends exactly at the start
    poly = poly[:-1]
    tiles = set()
    for i in range(len(poly)):
        tiles |= _seg_tiles(poly[i], poly[(i + 1) % len(poly)])
    ev = _events(v, p)
    got_tiles = [(e[1] // 8, e[2] // 8) for e in ev if e[0] == "tile"]
    assert sorted(got_tiles, key=lambda t: (t[1], t[0])) == got_tiles
    assert set(got_tiles) == tiles and len(tiles) == 100
    spans = []
    for ty in sorted({t[1] for t in tiles}):
        row = sorted(t[0] for t in tiles if t[1] == ty)
        for a, c in zip(row, row[1:]):
            if c > a + 1 and _winding(poly, 8 * (a + 1) + 0.5, 8 * ty + 0.5) != 0:
                spans.append((8 * (a + 1), 8 * ty, 8 * (c - a - 1)))
    got_spans = [tuple(e[1:]) for e in ev if e[0] == "span"]
    assert got_spans == spans and len(spans) == 23
