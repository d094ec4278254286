"""SVG ingestion (host code of libochre_hip.so; no GPU needed).

The fixtures tests/golden/*.svg.{gz,xz} are the three documents the
reference ships for its demo (examples/res/Ghostscript_Tiger.svg,
lorem-ipsum.svg, calabi-yau.svg), compressed.  The reference reads them with
usvg 0.13, which is not available here, so the parser is pinned by
hand-derived known answers, closed forms (degree elevation, points on the
ellipse, matrix products) and invariants of the real documents checked with
an independent reading of their XML (xml.etree), not by restating the
parser's rules."""
import gzip
import lzma
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from ochre_amd import svg
from ochre_amd._lib import OchreError

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M, L, C, Z = 0, 1, 3, 5


def golden(name):
    path = os.path.join(GOLDEN, name)
    opener = gzip.open if name.endswith(".gz") else lzma.open
    with opener(path, "rb") as f:
        return f.read()


def one_path(d, extra=""):
    doc = svg.parse(f'<svg xmlns="http://www.w3.org/2000/svg"><path {extra} d="{d}"/></svg>')
    assert doc.n_paths == 1
    return doc, doc.path(0)


def cmds(v, p):
    """[(verb, [points...]), ...] with f64 points."""
    out, k = [], 0
    for verb in v:
        n = {M: 1, L: 1, C: 3, Z: 0}[int(verb)]
        out.append((int(verb), [tuple(map(float, p[k + j])) for j in range(n)]))
        k += n
    return out


def f32(*xs):
    return [float(np.float32(x)) for x in xs]


def test_tiger_first_path_by_hand():
    """path8 of the tiger, decoded by hand: a relative moveto, a smooth cubic
    after a moveto (first control point = current point), a cubic and an
    implicit second cubic whose first control point is written as 0,0."""
    doc, (v, p) = one_path("m-122.3,84.285s0.1,1.894-0.73,1.875c-0.82-0.019-17.27-48.094-37.8-45.851,"
                           "0,0,17.78-7.353,38.53,43.976z")
    got = cmds(v, p)
    x0, y0 = -122.3, 84.285
    e1 = (x0 - 0.73, y0 + 1.875)
    e2 = (e1[0] - 37.8, e1[1] - 45.851)
    want = [
        (M, [(x0, y0)]),
        (C, [(x0, y0), (x0 + 0.1, y0 + 1.894), e1]),
        (C, [(e1[0] - 0.82, e1[1] - 0.019), (e1[0] - 17.27, e1[1] - 48.094), e2]),
        (C, [e2, (e2[0] + 17.78, e2[1] - 7.353), (e2[0] + 38.53, e2[1] + 43.976)]),
        (Z, []),
    ]
    assert [c[0] for c in got] == [c[0] for c in want]
    for (_, gp), (_, wp) in zip(got, want):
        for a, b in zip(gp, wp):
            assert a == tuple(f32(*b)), (a, b)
    # the draws: a default (black) fill, no stroke, identity transform
    assert doc.n_draws == 1 and doc.draw_stroke[0] == 0
    assert doc.draw_color[0].tolist() == [0, 0, 0, 255]
    assert doc.draw_xform[0].tolist() == [1, 0, 0, 1, 0, 0]


def test_command_forms_by_hand():
    # H/V relative and absolute, implicit lineto after moveto, exponents, numbers
    # run together ("1.5.5" is 1.5 and .5), comma/space separators
    _, (v, p) = one_path("M10 20 30,40 h5 v-5 H1e1 V.5 l1.5.5 L-2E+1-3e-1 z")
    assert cmds(v, p) == [
        (M, [(10, 20)]), (L, [(30, 40)]), (L, [(35, 40)]), (L, [(35, 35)]), (L, [(10, 35)]),
        (L, [(10, 0.5)]), (L, [(11.5, 1.0)]), (L, [(-20, float(np.float32(-0.3)))]), (Z, []),
    ]


def test_relative_moveto_after_close_is_from_the_subpath_start():
    _, (v, p) = one_path("M10 10 L20 10 L20 20 z m5 5 l1 0")
    assert cmds(v, p)[4:] == [(M, [(15, 15)]), (L, [(16, 15)])]
    # without a close it is from the current point
    _, (v, p) = one_path("M10 10 L20 10 L20 20 m5 5")
    assert cmds(v, p)[3] == (M, [(25, 25)])


def test_repeated_close_is_kept_once_and_errors_end_the_data():
    _, (v, p) = one_path("M0 0 L1 0 L1 1 z z Z L 5")
    assert [int(x) for x in v] == [M, L, L, Z]
    doc = svg.parse('<svg><path d="L1 1 L2 2"/></svg>')  # no leading moveto: no path
    assert doc.n_paths == 0
    doc = svg.parse('<svg><path d="M1 1"/></svg>')  # fewer than two segments: dropped
    assert doc.n_paths == 0


def test_smooth_cubic_reflection():
    _, (v, p) = one_path("M0 0 C1 2 3 4 5 5 S9 8 10 10 s1 1 2 0")
    c = cmds(v, p)
    assert c[2] == (C, [(7, 6), (9, 8), (10, 10)])  # reflection of (3, 4) about (5, 5)
    assert c[3] == (C, [(11, 12), (11, 11), (12, 10)])  # reflection of (9, 8) about (10, 10)


def _bez(ctrl, t):
    n = len(ctrl) - 1
    return sum(math.comb(n, k) * (1 - t) ** (n - k) * t ** k * np.array(ctrl[k], float) for k in range(n + 1))


def test_quadratics_are_exact_degree_elevations():
    _, (v, p) = one_path("M1 2 Q7 9 13 -4 T20 3 t5 5")
    c = cmds(v, p)
    assert [x[0] for x in c] == [M, C, C, C]
    quads = [((1, 2), (7, 9), (13, -4)), ((13, -4), (19, -17), (20, 3)), ((20, 3), (21, 23), (25, 8))]
    start = (1.0, 2.0)
    for (_, cp), q in zip(c[1:], quads):
        cub = [start] + cp
        for t in np.linspace(0, 1, 9):
            assert np.allclose(_bez(cub, t), _bez(q, t), atol=2e-5)
        start = cp[-1]


@pytest.mark.parametrize("arc,rx,ry,rot", [
    ("A30 20 0 0 1 60 10", 30, 20, 0), ("a25 25 0 1 0 40 0", 25, 25, 0), ("A40 15 30 1 1 55 35", 40, 15, 30)])
def test_arcs_end_exactly_and_stay_on_the_ellipse(arc, rx, ry, rot):
    doc, (v, p) = one_path("M10 10 " + arc)
    c = cmds(v, p)
    assert all(x[0] == C for x in c[1:]) and len(c) >= 3
    start = np.array(c[0][1][0])
    end = np.array(c[-1][1][2])
    m = re.findall(r"-?[\d.]+", arc)
    tgt = np.array([float(m[-2]), float(m[-1])]) + (start if arc[0] == "a" else 0)
    assert np.allclose(end, tgt, atol=1e-4)
    # the ellipse through the end points (SVG implementation notes F.6.5):
    # every segment end lies on it, and each cubic stays within 0.1 of it
    phi = math.radians(rot)
    R = np.array([[math.cos(phi), -math.sin(phi)], [math.sin(phi), math.cos(phi)]])
    ends = [np.array(x[1][2]) for x in c[1:]]

    def on_ellipse(q, ctr):
        u = R.T @ (q - ctr)
        return (u[0] / rx) ** 2 + (u[1] / ry) ** 2

    # centre: least squares over the segment ends (radii are not scaled up here)
    from scipy.optimize import least_squares
    ctr = least_squares(lambda z: [on_ellipse(q, z) - 1 for q in [start] + ends], start + 1).x
    for q in ends:
        assert abs(on_ellipse(q, ctr) - 1) < 1e-5
    prev = start
    for x in c[1:]:
        cub = [prev] + [np.array(q) for q in x[1]]
        for t in (0.25, 0.5, 0.75):
            u = R.T @ (_bez(cub, t) - ctr)
            r = math.hypot(u[0] / rx, u[1] / ry)
            assert abs(r - 1) * min(rx, ry) < 0.1
        prev = cub[-1]


def _affine(a, b, c, d, e, f):
    return np.array([[a, c, e], [b, d, f], [0, 0, 1]], float)


def test_transform_lists_compose_left_to_right_and_nest():
    t1 = "translate(10 20) rotate(30) scale(2, 3)"
    t2 = "matrix(1 0.5 -0.25 1 4 5) skewX(10)"
    doc = svg.parse(f'<svg><g transform="{t1}"><g transform="{t2}"><path transform="translate(-3)" '
                    f'd="M0 0 L1 1"/></g></g></svg>')
    r = math.radians(30)
    want = (_affine(1, 0, 0, 1, 10, 20) @ _affine(math.cos(r), math.sin(r), -math.sin(r), math.cos(r), 0, 0)
            @ _affine(2, 0, 0, 3, 0, 0) @ _affine(1, 0.5, -0.25, 1, 4, 5)
            @ _affine(1, 0, math.tan(math.radians(10)), 1, 0, 0) @ _affine(1, 0, 0, 1, -3, 0))
    xf = doc.draw_xform[0]
    # the C ABI transform: row-major Mat2x2 (m00 m01; m10 m11) + offset
    assert np.allclose(xf, [want[0, 0], want[0, 1], want[1, 0], want[1, 1], want[0, 2], want[1, 2]], rtol=1e-6)


def test_a_kept_group_restarts_the_transform():
    # a group with opacity is kept in usvg's tree; the demo applies only the
    # path's transform relative to it (node.transform())
    doc = svg.parse('<svg><g transform="scale(4)"><g opacity="0.5" transform="translate(7 7)">'
                    '<path transform="translate(1 2)" d="M0 0 L1 1"/></g></g></svg>')
    assert doc.draw_xform[0].tolist() == [1, 0, 0, 1, 1, 2]


def test_paint_inheritance_defaults_and_colours():
    doc = svg.parse(
        '<svg><g fill="#0a0" stroke="red" stroke-width="2">'
        '<path d="M0 0 L1 1"/>'                                   # fill #00aa00 + stroke red w 2
        '<path fill="none" d="M0 0 L1 1"/>'                       # stroke only
        '<path style="fill:rgb(10%,20,300);stroke:none" d="M0 0 L1 1"/>'  # style overrides
        '<path stroke-width="0" fill="url(#g)" d="M0 0 L1 1"/>'   # nothing drawn
        '<path fill="SteelBlue" stroke-opacity="0.5" d="M0 0 L1 1"/>'
        '</g><path d="M0 0 L1 1"/></svg>')                         # default: black fill, no stroke
    draws = list(zip(doc.draw_path.tolist(), doc.draw_stroke.tolist(), doc.draw_width.tolist(),
                     [tuple(c) for c in doc.draw_color.tolist()]))
    assert draws == [
        (0, 0, 0.0, (0, 0xaa, 0, 255)), (0, 1, 2.0, (255, 0, 0, 255)),
        (1, 1, 2.0, (255, 0, 0, 255)),
        (2, 0, 0.0, (26, 20, 255, 255)),
        (4, 0, 0.0, (0x46, 0x82, 0xb4, 255)), (4, 1, 2.0, (255, 0, 0, 128)),
        (5, 0, 0.0, (0, 0, 0, 255)),
    ]


def test_shapes():
    doc = svg.parse('<svg><rect x="1" y="2" width="10" height="5"/><line x1="0" y1="0" x2="3" y2="4"/>'
                    '<polygon points="0,0 4,0 4,3"/><polyline points="1 1 2 2 3 1"/>'
                    '<circle cx="10" cy="10" r="5"/><ellipse cx="0" cy="0" rx="8" ry="4"/></svg>')
    assert doc.n_paths == 6
    assert cmds(*doc.path(0)) == [(M, [(1, 2)]), (L, [(11, 2)]), (L, [(11, 7)]), (L, [(1, 7)]), (Z, [])]
    assert cmds(*doc.path(1)) == [(M, [(0, 0)]), (L, [(3, 4)])]
    assert cmds(*doc.path(2)) == [(M, [(0, 0)]), (L, [(4, 0)]), (L, [(4, 3)]), (Z, [])]
    assert cmds(*doc.path(3)) == [(M, [(1, 1)]), (L, [(2, 2)]), (L, [(3, 1)])]
    for i, (cx, cy, rx, ry) in ((4, (10, 10, 5, 5)), (5, (0, 0, 8, 4))):
        c = cmds(*doc.path(i))
        assert c[0] == (M, [(cx + rx, cy)]) and c[-1] == (Z, [])
        for _, pts in c[1:-1]:
            x, y = pts[-1]
            assert abs(((x - cx) / rx) ** 2 + ((y - cy) / ry) ** 2 - 1) < 1e-6


def test_not_an_svg_document():
    for bad in (b"", b"<html></html>", b"<svg><path d='M0 0'>", b"not xml"):
        with pytest.raises(OchreError):
            svg.parse(bad)


def _etree_paths(text):
    """Path elements in document order with their inherited fill / stroke /
    stroke-width and the nesting of transforms, read with xml.etree."""
    root = ET.fromstring(text)
    out = []

    def walk(el, fill, stroke, sw, depth_ts):
        tag = el.tag.split("}")[-1]
        fill = el.get("fill", fill)
        stroke = el.get("stroke", stroke)
        sw = el.get("stroke-width", sw)
        ts = depth_ts + ([el.get("transform")] if el.get("transform") else [])
        if tag == "path":
            out.append((el.get("d"), fill, stroke, float(sw), ts))
        elif tag in ("svg", "g"):
            for ch in el:
                walk(ch, fill, stroke, sw, ts)

    walk(root, "#000", "none", "1", [])
    return out


def _segments(d):
    """Segments the path data of one path makes: command letters with their
    implicit repetitions (arguments per command), repeated Z counted once."""
    n_args = dict(M=2, L=2, H=1, V=1, C=6, S=4, Q=4, T=2, A=7, Z=0)
    toks = re.findall(r"[A-Za-z]|[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", d)
    segs, i, last = [], 0, None
    while i < len(toks):
        cmd = toks[i].upper()
        i += 1
        j = i
        while j < len(toks) and not toks[j].isalpha():
            j += 1
        args = j - i
        i = j
        if cmd == "Z":
            if last != "Z":
                segs.append("Z")
            last = "Z"
            continue
        k = args // n_args[cmd]
        segs += [cmd] + ["L" if cmd == "M" else cmd] * (k - 1)
        last = cmd
    return segs


@pytest.mark.parametrize("name,n_paths", [("Ghostscript_Tiger.svg.gz", 240), ("lorem-ipsum.svg.gz", 1133),
                                          ("calabi-yau.svg.xz", 99)])
def test_reference_documents(name, n_paths):
    text = golden(name)
    doc = svg.parse(text)
    ref = _etree_paths(text)
    assert doc.n_paths == len(ref) == n_paths
    # segments per path, verb by verb
    for i, (d, fill, stroke, sw, ts) in enumerate(ref):
        v, p = doc.path(i)
        want = {"M": M, "L": L, "H": L, "V": L, "C": C, "S": C, "Q": C, "T": C, "Z": Z}
        assert [int(x) for x in v] == [want[s] for s in _segments(d)], i
        assert int(v[0]) == M
    # draws: a fill per path whose fill is not "none", a stroke per path with
    # a stroke colour and a positive width, fill first
    exp = []
    for i, (d, fill, stroke, sw, ts) in enumerate(ref):
        if fill != "none":
            exp.append((i, 0))
        if stroke != "none" and sw > 0:
            exp.append((i, 1))
    assert list(zip(doc.draw_path.tolist(), doc.draw_stroke.tolist())) == exp
    assert np.all(doc.draw_width[doc.draw_stroke == 1] > 0)


def test_tiger_transform_and_extent():
    """The tiger: every path under the one matrix(...) group (the inner groups
    carry paint only), every on-curve point inside the 900 x 900 viewBox once
    transformed (control points may lie outside)."""
    doc = svg.parse(golden("Ghostscript_Tiger.svg.gz"))
    assert doc.view_box == (0.0, 0.0, 900.0, 900.0)
    s, e, f = 1.7656463, 324.90716, 255.00942
    assert np.all(doc.draw_xform == np.array(f32(s, 0, 0, s, e, f), np.float32))
    ends = []
    for i in range(doc.n_paths):
        ends += [pts[-1] for verb, pts in cmds(*doc.path(i)) if pts]
    world = np.array(ends) * s + np.array([e, f])
    assert world.min() >= 0 and world.max() <= 900
    # 240 paths; fills and strokes as the groups give them
    assert doc.n_draws == int((doc.draw_stroke == 0).sum() + (doc.draw_stroke == 1).sum())
