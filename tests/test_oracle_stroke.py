"""Pins the oracle's stroking (Rasterizer::stroke, rasterizer.rs:169-171 =
fill(stroke(flatten(path)), transform), path.rs:114-274) without restating
its f32 formulas:

* the outline's exact f64 geometry, built from offset lines and their
  intersections, rasterised by the brute-force winding integral of
  test_oracle and compared with the oracle's tiles and spans (1 LSB):
    - an open segment is a butt-capped rectangle;
    - a closed polygon is a ring whose joins are miters (the intersection of
      the two offset edges) where the normals turn by at most 120 degrees and
      bevels beyond; the contour's last vertex is bevelled on the outer side
      and its first on the inner side (the offset walk starts there with no
      previous normal);
    - an open polyline is one contour: the forward offset, then the reverse
      offset, each ending with the join against the chord that closes the
      polyline (the reference's cap);
* invariants: repeated points do not change the outline; a Close makes every
  later contour of the path closed (`closed` is never reset); stroking equals
  filling the path-space flatten + stroke; the transform is applied after
  stroking (a scale maps a stroke of width w to a fill of the scaled outline).
"""
import math

import numpy as np
import pytest

import oracle_ref as O
from ochre_amd import workloads as W
from test_oracle import exact_alpha, render

M, L, Q, C, K, Z = W.MOVE, W.LINE, W.QUAD, W.CUBIC, W.CONIC, W.CLOSE


def normal(a, b):
    t = np.asarray(b, np.float64) - np.asarray(a, np.float64)
    n = np.array([-t[1], t[0]])
    ln = math.hypot(n[0], n[1])
    return n / ln if ln > 0 else np.zeros(2)


def join(v, na, nb, s):
    """Outline points at vertex v between edges with unit normals na -> nb,
    offset by s (signed half width): the miter (intersection of the two
    offset lines) when the normals turn by <= 120 degrees, else the bevel."""
    v = np.asarray(v, np.float64)
    if not na.any() or not nb.any():
        return [v + s * (na + nb)]
    c = float(na @ nb)
    if c < -0.5:
        return [v + s * na, v + s * nb]
    ta, tb = np.array([na[1], -na[0]]), np.array([nb[1], -nb[0]])
    pa, pb = v + s * na, v + s * nb
    A = np.stack([ta, -tb], 1)
    if abs(np.linalg.det(A)) < 1e-12:
        return [pa]
    t = np.linalg.solve(A, pb - pa)
    return [pa + t[0] * ta]


def closed_outline(c, w):
    c = [np.asarray(p, np.float64) for p in c]
    m, d = len(c), 0.5 * w
    n = [normal(c[k], c[(k + 1) % m]) for k in range(m)]
    outer = [c[m - 1] + d * n[m - 1]]
    for k in range(m - 1):
        outer += join(c[k], n[k - 1], n[k], d)
    outer.append(c[m - 1] + d * n[m - 2])
    inner = [c[0] - d * n[m - 1]]
    for k in range(m - 1, 0, -1):
        inner += join(c[k], -n[k], -n[k - 1], d)
    inner.append(c[0] - d * n[0])
    return [np.array(outer), np.array(inner)]


def open_outline(c, w):
    c = [np.asarray(p, np.float64) for p in c]
    m, d = len(c), 0.5 * w
    n = [normal(c[k], c[k + 1]) for k in range(m - 1)]
    ch = normal(c[m - 1], c[0])
    pts = [c[0] + d * n[0]]
    for k in range(1, m - 1):
        pts += join(c[k], n[k - 1], n[k], d)
    pts += join(c[m - 1], n[m - 2], ch, d)
    pts.append(c[m - 1] - d * n[m - 2])
    for k in range(m - 2, 0, -1):
        pts += join(c[k], -n[k], -n[k - 1], d)
    pts += join(c[0], -n[0], -ch, d)
    return [np.array(pts)]


def oracle_stroke_events(verbs, pts, width, xform=None, weights=()):
    r = O.Rasterizer()
    assert r.stroke(verbs, np.asarray(pts, np.float32), width, weights, xform) == 0
    return r.finish()


def compare(events, contours, margin=10):
    allp = np.concatenate(contours)
    x0 = int(math.floor(allp[:, 0].min())) - margin
    y0 = int(math.floor(allp[:, 1].min())) - margin
    w = int(math.ceil(allp[:, 0].max())) + margin - x0
    h = int(math.ceil(allp[:, 1].max())) + margin - y0
    img, covered = render(events, x0, y0, w, h)
    cov = np.abs(exact_alpha(contours, x0, y0, w, h))
    want = np.minimum(np.floor(cov * 256.0), 255).astype(np.int32)
    diff = np.abs(img - want)
    assert diff.max() <= 1, (diff.max(), np.argwhere(diff > 1)[:5])
    assert not np.any((cov > 1e-6) & ~covered)
    return img


def f32(pts):
    return np.asarray(pts, np.float32).astype(np.float64)


@pytest.mark.parametrize("seed", range(6))
def test_open_segment_is_a_butt_capped_rectangle(seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(10, 60, 2)
    th = rng.uniform(0, 2 * math.pi)
    b = a + rng.uniform(5, 40) * np.array([math.cos(th), math.sin(th)])
    w = float(np.float32(rng.uniform(0.5, 8.0)))
    p = f32([a, b])
    n = normal(p[0], p[1])
    rect = np.array([p[0] + 0.5 * w * n, p[1] + 0.5 * w * n, p[1] - 0.5 * w * n, p[0] - 0.5 * w * n])
    compare(oracle_stroke_events([M, L], p, w), [rect])


def test_axis_aligned_segment_fills_whole_pixels():
    img = compare(oracle_stroke_events([M, L], [(1, 4), (17, 4)], 2.0), [np.array([(1, 3), (17, 3), (17, 5), (1, 5)],
                                                                                  np.float64)])
    assert (img == 255).sum() == 16 * 2


def regular_polygon(rng, m, r):
    c = rng.uniform(30, 50, 2)
    ph = rng.uniform(0, 2 * math.pi)
    return f32([c + r * np.array([math.cos(ph + 2 * math.pi * k / m), math.sin(ph + 2 * math.pi * k / m)])
                for k in range(m)])


@pytest.mark.parametrize("m,seed", [(4, 0), (5, 1), (6, 2), (8, 3), (4, 4)])
def test_closed_regular_polygon_is_a_mitered_ring(m, seed):
    rng = np.random.default_rng(100 + seed)
    c = regular_polygon(rng, m, rng.uniform(8, 20))
    if seed == 4:
        c = c[::-1].copy()   # the other orientation
    w = float(np.float32(rng.uniform(1.0, 5.0)))
    compare(oracle_stroke_events([M] + [L] * (m - 1) + [Z], c, w), closed_outline(c, w))


def star(rng, m, r0, r1):
    c = rng.uniform(40, 60, 2)
    ph = rng.uniform(0, 2 * math.pi)
    pts = []
    for k in range(2 * m):
        r = r1 if k % 2 == 0 else r0
        a = ph + math.pi * k / m
        pts.append(c + r * np.array([math.cos(a), math.sin(a)]))
    return f32(pts)


def turn_cosines(c, closed):
    m = len(c)
    ks = range(m) if closed else range(1, m - 1)
    return [float(normal(c[k - 1], c[k]) @ normal(c[k], c[(k + 1) % m])) for k in ks]


@pytest.mark.parametrize("seed", range(4))
def test_sharp_vertices_are_bevelled(seed):
    rng = np.random.default_rng(200 + seed)
    while True:
        c = star(rng, int(rng.integers(3, 6)), rng.uniform(4, 8), rng.uniform(20, 30))
        cs = turn_cosines(c, True)
        if min(abs(x + 0.5) for x in cs) > 0.05 and min(cs) < -0.5:
            break
    w = float(np.float32(rng.uniform(1.0, 3.0)))
    compare(oracle_stroke_events([M] + [L] * (len(c) - 1) + [Z], c, w), closed_outline(c, w))


@pytest.mark.parametrize("apex_deg", [42.0, 48.0, 53.0, 57.0, 64.0, 75.0])
def test_bevel_threshold_is_a_120_degree_turn(apex_deg):
    # isosceles triangles: the apex turns by 180 - apex degrees, so apexes
    # below 60 degrees are bevelled and wider ones mitred
    rng = np.random.default_rng(int(apex_deg))
    h = rng.uniform(25, 35)
    half = h * math.tan(math.radians(apex_deg) / 2)
    ph = rng.uniform(0, 2 * math.pi)
    rot = np.array([[math.cos(ph), -math.sin(ph)], [math.sin(ph), math.cos(ph)]])
    c = f32([np.array([50.0, 50.0]) + rot @ p for p in ([0.0, -h / 2], [half, h / 2], [-half, h / 2])])
    cs = turn_cosines(c, True)
    assert min(abs(x + 0.5) for x in cs) > 0.02
    w = float(np.float32(rng.uniform(1.5, 3.0)))
    compare(oracle_stroke_events([M, L, L, Z], c, w), closed_outline(c, w))


@pytest.mark.parametrize("seed", range(5))
def test_open_polyline_caps_join_the_closing_chord(seed):
    rng = np.random.default_rng(300 + seed)
    while True:
        m = int(rng.integers(3, 7))
        c = f32(rng.uniform(20, 80, (m, 2)))
        chord_turns = [float(normal(c[-2], c[-1]) @ normal(c[-1], c[0])),
                       float(normal(c[1], c[0]) @ normal(c[0], c[-1]))]
        seg = [np.linalg.norm(c[k + 1] - c[k]) for k in range(m - 1)]
        if (min(abs(x + 0.5) for x in turn_cosines(c, False) + chord_turns) > 0.05 and min(seg) > 3
                and max(abs(x) for x in turn_cosines(c, False) + chord_turns) < 0.995):
            break
    w = float(np.float32(rng.uniform(1.0, 4.0)))
    compare(oracle_stroke_events([M] + [L] * (m - 1), c, w), open_outline(c, w))


def test_repeated_points_do_not_change_the_outline():
    c = f32([(10, 10), (40, 12), (35, 40), (12, 30)])
    dup = np.concatenate([c[:1], c[:2], c[1:2], c[2:], c[3:]])
    a = oracle_stroke_events([M] + [L] * 3 + [Z], c, 3.0)
    b = oracle_stroke_events([M] + [L] * (len(dup) - 1) + [Z], dup, 3.0)
    assert _same(a, b)


def _same(a, b):
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if x[:3] != y[:3]:
            return False
        if x[0] == "tile" and not np.array_equal(x[3], y[3]):
            return False
        if x[0] == "span" and x[3] != y[3]:
            return False
    return True


def test_close_makes_later_contours_closed():
    a = f32([(10, 10), (30, 10), (30, 30)])
    b = f32([(50, 10), (70, 15), (60, 35)])
    pts = np.concatenate([a, b])
    sticky = oracle_stroke_events([M, L, L, Z, M, L, L], pts, 2.0)
    explicit = oracle_stroke_events([M, L, L, Z, M, L, L, Z], pts, 2.0)
    assert _same(sticky, explicit)
    # without the first Close, the second contour is open and its outline
    # differs (the open cap)
    open_ = oracle_stroke_events([M, L, L, M, L, L], pts, 2.0)
    assert not _same(open_, explicit)
    compare(explicit, closed_outline(a, 2.0) + closed_outline(b, 2.0))


def test_stroke_is_fill_of_flatten_then_stroke_in_path_space():
    rng = np.random.default_rng(7)
    for _ in range(5):
        verbs = [M, Q, C, K, L, Z, M, C]
        pts = rng.uniform(0, 4, (1 + 2 + 3 + 2 + 1 + 1 + 3, 2)).astype(np.float32)
        ws = [float(rng.uniform(0.3, 2.0))]
        xf = np.array([9.0, 1.5, -2.0, 11.0, 20.0, 30.0], np.float32)
        st, (fv, fp, fw) = O.flatten_path(verbs, pts, ws, 0.1)
        assert st == 0 and set(fv.tolist()) <= {M, L, Z} and fw.size == 0
        st, (sv, sp, _) = O.stroke_path(fv, fp, 0.25)
        assert st == 0
        r = O.Rasterizer()
        assert r.fill(sv, sp, (), xf) == 0
        want = r.finish()
        got = oracle_stroke_events(verbs, pts, 0.25, xf, ws)
        assert _same(got, want)


def test_transform_applies_after_stroking():
    # a 4x scale of a width-1 stroke covers the 4x-scaled outline
    c = f32([(3, 3), (12, 4), (10, 11), (4, 9)])
    xf = np.array([4, 0, 0, 4, 5, 7], np.float32)
    ev = oracle_stroke_events([M, L, L, L, Z], c, 1.0, xf)
    outline = [o * 4.0 + np.array([5.0, 7.0]) for o in closed_outline(c, 1.0)]
    compare(ev, outline)


def test_curves_are_flattened_in_path_space():
    # flattening happens before the transform: the same quadratic stroked at
    # scale 1 and at scale 8 draws the same number of lines, and that number
    # follows from the path-space step dt = sqrt(4 * 0.1 / |p0 - 2 c + p2|)
    verbs, pts = [M, Q], np.array([(0, 0), (20, 30), (40, 0)], np.float32)
    st, (fv, _, _) = O.flatten_path(verbs, pts, (), 0.1)
    assert st == 0
    dd = np.linalg.norm(pts[0].astype(np.float64) - 2 * pts[1] + pts[2])
    assert int((fv == L).sum()) == math.ceil(1.0 / math.sqrt(0.4 / dd))
    r1, r8 = O.Rasterizer(), O.Rasterizer()
    assert r1.stroke(verbs, pts, 1.0) == 0
    assert r8.stroke(verbs, pts, 1.0, (), np.array([8, 0, 0, 8, 0, 0], np.float32)) == 0
    n1, n8 = r1.stats()["lines"], r8.stats()["lines"]
    r1.finish(), r8.finish()
    assert n1 == n8 > 2 * int((fv == L).sum())
