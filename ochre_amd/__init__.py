"""ochre_amd -- MI355X (gfx950) implementation of ochre's rasterisation path.

The Python surface mirrors the reference crate's public API
(glowcoil/ochre src/lib.rs:47-53): ``Vec2``, ``Mat2x2``, ``Transform``
(src/geom.rs), ``PathCmd`` (src/path.rs:5-12), ``TileBuilder``, ``Rasterizer``
and ``TILE_SIZE`` (src/rasterizer.rs).  ``Rasterizer.finish`` runs the HIP
pipeline through the C ABI (include/ochre.h); there is no CPU fallback.

The throughput interface (device-resident batches of paths) is in
``ochre_amd.device``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import OchreError

TILE_SIZE = 8  # rasterizer.rs:4

MOVE, LINE, QUADRATIC, CUBIC, CONIC, CLOSE = 0, 1, 2, 3, 4, 5


@dataclass(frozen=True)
class Vec2:
    """geom.rs:4-129 (host-side value type)."""
    x: float
    y: float


@dataclass(frozen=True)
class Mat2x2:
    """Row-major 2x2 matrix (geom.rs:131-200)."""
    m: tuple

    @staticmethod
    def new(a, b, c, d):
        return Mat2x2((float(a), float(b), float(c), float(d)))

    @staticmethod
    def id():
        return Mat2x2((1.0, 0.0, 0.0, 1.0))

    @staticmethod
    def scale(s):
        return Mat2x2((float(s), 0.0, 0.0, float(s)))

    @staticmethod
    def rotate(angle):
        a = np.float32(angle)
        c, s = float(np.cos(a)), float(np.sin(a))
        return Mat2x2((c, s, -s, c))

    def __mul__(self, o):
        a = np.asarray(self.m, np.float32)
        if isinstance(o, Mat2x2):
            b = np.asarray(o.m, np.float32)
            return Mat2x2(tuple(float(v) for v in (
                a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3],
                a[2] * b[0] + a[3] * b[2], a[2] * b[1] + a[3] * b[3])))
        if isinstance(o, Vec2):
            x, y = np.float32(o.x), np.float32(o.y)
            return Vec2(float(a[0] * x + a[1] * y), float(a[2] * x + a[3] * y))
        return NotImplemented


@dataclass(frozen=True)
class Transform:
    """Affine transform (geom.rs:202-263)."""
    matrix: Mat2x2
    offset: Vec2

    @staticmethod
    def id():
        return Transform(Mat2x2.id(), Vec2(0.0, 0.0))

    @staticmethod
    def translate(x, y):
        return Transform(Mat2x2.id(), Vec2(float(x), float(y)))

    @staticmethod
    def scale(s):
        return Transform(Mat2x2.scale(s), Vec2(0.0, 0.0))

    @staticmethod
    def rotate(angle):
        return Transform(Mat2x2.rotate(angle), Vec2(0.0, 0.0))

    def then(self, t: "Transform") -> "Transform":
        o = t.matrix * self.offset
        return Transform(t.matrix * self.matrix,
                         Vec2(float(np.float32(o.x) + np.float32(t.offset.x)),
                              float(np.float32(o.y) + np.float32(t.offset.y))))

    def as_array(self) -> np.ndarray:
        return np.array([*self.matrix.m, self.offset.x, self.offset.y], np.float32)

    def _c(self):
        t = _lib.Transform()
        a = self.as_array()
        for i in range(4):
            t.m[i] = float(a[i])
        t.offset[0], t.offset[1] = float(a[4]), float(a[5])
        return t


class PathCmd:
    """PathCmd (path.rs:5-12).  Construct with the named constructors."""
    __slots__ = ("verb", "points", "weight")

    def __init__(self, verb, points=(), weight=0.0):
        self.verb = verb
        self.points = tuple(points)
        self.weight = float(weight)

    @staticmethod
    def Move(p):
        return PathCmd(MOVE, (p,))

    @staticmethod
    def Line(p):
        return PathCmd(LINE, (p,))

    @staticmethod
    def Quadratic(c, p):
        return PathCmd(QUADRATIC, (c, p))

    @staticmethod
    def Cubic(c1, c2, p):
        return PathCmd(CUBIC, (c1, c2, p))

    @staticmethod
    def Conic(c, p, w):
        return PathCmd(CONIC, (c, p), w)

    Close = None  # replaced below


PathCmd.Close = PathCmd(CLOSE)


def encode_path(path):
    """Structure-of-arrays encoding of a list of PathCmd: (verbs u8,
    points f32 [n, 2], weights f32)."""
    verbs = np.fromiter((c.verb for c in path), np.uint8, len(path))
    pts = [(p.x, p.y) for c in path for p in c.points]
    points = np.asarray(pts, np.float32).reshape(-1, 2)
    weights = np.asarray([c.weight for c in path if c.verb == CONIC], np.float32)
    return verbs, points, weights


class TileBuilder:
    """TileBuilder (rasterizer.rs:12-22)."""

    def tile(self, x: int, y: int, data: np.ndarray) -> None:  # data: u8[64], row-major
        raise NotImplementedError

    def span(self, x: int, y: int, width: int) -> None:
        raise NotImplementedError


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


class Rasterizer:
    """Rasterizer (rasterizer.rs:39-268), backed by the gfx950 pipeline."""

    def __init__(self):
        self._lib = _lib.load()
        self._h = self._lib.ochre_rasterizer_new()
        if not self._h:
            raise OchreError(-6, "Rasterizer.new")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ochre_rasterizer_free(h)
            self._h = None

    def move_to(self, p: Vec2):
        _lib.check(self._lib.ochre_rasterizer_move_to(self._h, p.x, p.y), "move_to")

    def line_to(self, p: Vec2):
        _lib.check(self._lib.ochre_rasterizer_line_to(self._h, p.x, p.y), "line_to")

    def command(self, cmd: PathCmd):
        pts = np.asarray([(p.x, p.y) for p in cmd.points], np.float32).reshape(-1)
        _lib.check(self._lib.ochre_rasterizer_command(self._h, cmd.verb, _ptr(pts), cmd.weight), "command")

    @staticmethod
    def _arrays(verbs, points, weights, transform):
        v = np.ascontiguousarray(verbs, np.uint8)
        p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        if transform is None:
            tc = Transform.id()._c()
        elif isinstance(transform, Transform):
            tc = transform._c()
        else:
            a = np.asarray(transform, np.float32)
            tc = _lib.Transform()
            for i in range(4):
                tc.m[i] = float(a[i])
            tc.offset[0], tc.offset[1] = float(a[4]), float(a[5])
        return v, p, w, tc

    def fill_arrays(self, verbs, points, weights=(), transform=None):
        v, p, w, tc = self._arrays(verbs, points, weights, transform)
        _lib.check(self._lib.ochre_rasterizer_fill(self._h, _ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size,
                                                   ctypes.byref(tc)), "fill")

    def fill(self, path, transform: Transform):
        verbs, points, weights = encode_path(path)
        self.fill_arrays(verbs, points, weights, transform)

    def stroke_arrays(self, verbs, points, width, weights=(), transform=None):
        v, p, w, tc = self._arrays(verbs, points, weights, transform)
        _lib.check(self._lib.ochre_rasterizer_stroke(self._h, _ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size,
                                                     float(width), ctypes.byref(tc)), "stroke")

    def stroke(self, path, width: float, transform: Transform):
        """Rasterizer::stroke (rasterizer.rs:169-171): butt caps, miter joins."""
        verbs, points, weights = encode_path(path)
        self.stroke_arrays(verbs, points, width, weights, transform)

    def finish(self, builder: TileBuilder):
        """Rasterises on the GPU and replays the results into `builder`.
        Consumes the rasterizer."""
        def on_tile(_u, x, y, a):
            builder.tile(int(x), int(y), np.ctypeslib.as_array(a, shape=(64,)).copy())

        def on_span(_u, x, y, w):
            builder.span(int(x), int(y), int(w))

        tf, sf = _lib.TILE_FN(on_tile), _lib.SPAN_FN(on_span)
        tb = _lib.TileBuilderC(None, ctypes.cast(tf, ctypes.c_void_p), ctypes.cast(sf, ctypes.c_void_p))
        h, self._h = self._h, None
        _lib.check(self._lib.ochre_rasterizer_finish(h, ctypes.byref(tb)), "finish")


def finish_many(rasterizers, builder: TileBuilder):
    """Finishes the rasterizers into one builder, in order (the demo's loop
    of one Rasterizer per fill or stroke, examples/svg.rs:140-156); those that
    hold one fill or one stroke are rasterised together in one device call.
    Consumes every rasterizer."""
    lib = _lib.load()

    def on_tile(_u, x, y, a):
        builder.tile(int(x), int(y), np.ctypeslib.as_array(a, shape=(64,)).copy())

    def on_span(_u, x, y, w):
        builder.span(int(x), int(y), int(w))

    tf, sf = _lib.TILE_FN(on_tile), _lib.SPAN_FN(on_span)
    tb = _lib.TileBuilderC(None, ctypes.cast(tf, ctypes.c_void_p), ctypes.cast(sf, ctypes.c_void_p))
    hs = []
    for r in rasterizers:
        hs.append(r._h)
        r._h = None
    arr = (ctypes.c_void_p * max(len(hs), 1))(*hs)
    _lib.check(lib.ochre_rasterizer_finish_many(arr, len(hs), ctypes.byref(tb)), "finish_many")


class EventRecorder(TileBuilder):
    """A TileBuilder that records the call sequence."""

    def __init__(self):
        self.events = []

    def tile(self, x, y, data):
        self.events.append(("tile", x, y, data))

    def span(self, x, y, width):
        self.events.append(("span", x, y, width))


__all__ = ["TILE_SIZE", "Vec2", "Mat2x2", "Transform", "PathCmd", "TileBuilder", "Rasterizer",
           "EventRecorder", "OchreError", "encode_path", "finish_many"]
