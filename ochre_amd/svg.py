"""SVG ingestion (ctypes over ochre_svg_* in libochre_hip.so).

`parse(text)` returns the document's paths and the draw calls the
reference's demo makes from it (examples/svg.rs:103-164): one fill and/or
one stroke per path element, in document order.  `draw_batch(doc)` lays the
draws out as one draw-list batch (one path per draw, its own transform,
`strokes`, `widths` and `colors` set), ready for `device.draw` /
`run_to_host(..., "draw_*")` and `device.build_atlas`, which reproduce the
demo's sequence of Rasterizers finishing into one Builder.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .workloads import PathBatch


class SvgDrawC(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int64), ("stroke", ctypes.c_int32), ("width", ctypes.c_float),
                ("color", ctypes.c_uint8 * 4), ("transform", _lib.Transform)]


class SvgSummaryC(ctypes.Structure):
    _fields_ = [("n_paths", ctypes.c_int64), ("n_verbs", ctypes.c_int64), ("n_points", ctypes.c_int64),
                ("n_draws", ctypes.c_int64), ("n_skipped", ctypes.c_int64), ("view_box", ctypes.c_float * 4),
                ("width", ctypes.c_float), ("height", ctypes.c_float)]


@dataclass
class SvgDocument:
    verbs: np.ndarray       # u8 [V]
    verb_off: np.ndarray    # i64 [n_paths + 1]
    points: np.ndarray      # f32 [P, 2]
    point_off: np.ndarray   # i64 [n_paths + 1]
    draw_path: np.ndarray   # i64 [n_draws]
    draw_stroke: np.ndarray  # u8 [n_draws]
    draw_width: np.ndarray  # f32 [n_draws]
    draw_color: np.ndarray  # u8 [n_draws, 4]
    draw_xform: np.ndarray  # f32 [n_draws, 6]: m00 m01 m10 m11 ox oy
    view_box: tuple
    size: tuple
    n_skipped: int

    @property
    def n_paths(self) -> int:
        return int(self.verb_off.shape[0] - 1)

    @property
    def n_draws(self) -> int:
        return int(self.draw_path.shape[0])

    def path(self, i: int):
        """(verbs, points) of path i."""
        return (self.verbs[self.verb_off[i]:self.verb_off[i + 1]],
                self.points[self.point_off[i]:self.point_off[i + 1]])


def _copy(ptr, dtype, count):
    if count == 0 or not ptr:
        return np.zeros(0, dtype)
    n = count * np.dtype(dtype).itemsize
    return np.frombuffer(ctypes.string_at(ptr, n), dtype=dtype).copy()


def parse(text) -> SvgDocument:
    """Parses an SVG document (str or bytes).  Raises OchreError on a text
    that is not an <svg> document."""
    lib = _lib.load()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    _lib.check(lib.ochre_svg_parse(data, len(data), ctypes.byref(h)), "ochre_svg_parse")
    try:
        info = SvgSummaryC()
        _lib.check(lib.ochre_svg_info(h, ctypes.byref(info)), "ochre_svg_info")
        n, nd = info.n_paths, info.n_draws
        draws = (SvgDrawC * max(nd, 1)).from_buffer_copy(
            ctypes.string_at(lib.ochre_svg_draws(h), max(nd, 1) * ctypes.sizeof(SvgDrawC))) if nd else []
        return SvgDocument(
            verbs=_copy(lib.ochre_svg_verbs(h), np.uint8, info.n_verbs),
            verb_off=_copy(lib.ochre_svg_verb_offsets(h), np.int64, n + 1),
            points=_copy(lib.ochre_svg_points(h), np.float32, 2 * info.n_points).reshape(-1, 2),
            point_off=_copy(lib.ochre_svg_point_offsets(h), np.int64, n + 1),
            draw_path=np.array([d.path for d in draws[:nd]], np.int64),
            draw_stroke=np.array([d.stroke for d in draws[:nd]], np.uint8),
            draw_width=np.array([d.width for d in draws[:nd]], np.float32),
            draw_color=np.array([list(d.color) for d in draws[:nd]], np.uint8).reshape(-1, 4),
            draw_xform=np.array([list(d.transform.m) + list(d.transform.offset) for d in draws[:nd]],
                                np.float32).reshape(-1, 6),
            view_box=tuple(info.view_box), size=(info.width, info.height), n_skipped=info.n_skipped)
    finally:
        lib.ochre_svg_free(h)


def parse_file(path) -> SvgDocument:
    with open(path, "rb") as f:
        return parse(f.read())


def draw_batch(doc: SvgDocument, name="svg") -> PathBatch:
    """The document's draws as one draw-list batch, in draw order (a path
    element with both a fill and a stroke appears twice)."""
    from .workloads import from_paths
    parts = []
    for k in range(doc.n_draws):
        v, p = doc.path(int(doc.draw_path[k]))
        parts.append((v, p, np.zeros(0, np.float32)))
    b = from_paths(parts, doc.draw_xform)
    b.strokes = doc.draw_stroke.copy()
    b.widths = doc.draw_width.copy()
    b.colors = doc.draw_color.copy()
    b.name = name
    return b
