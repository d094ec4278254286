"""ctypes binding of the CPU oracle (oracle/libochre_oracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# OCHRE_ORACLE_SO selects another build of the same sources (the sanitizer
# build, oracle/_san/libochre_oracle.so, in tests/test_oracle_sanitize.py).
ORACLE_SO = os.environ.get("OCHRE_ORACLE_SO") or os.path.join(ORACLE_DIR, "libochre_oracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ctypes.CDLL(ORACLE_SO)
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    lib.orc_run_batch.restype = P
    lib.orc_run_batch.argtypes = [i64, P, P, P, P, P, P, P, P, P]
    lib.orc_result_status.argtypes = [P]
    lib.orc_result_sizes.argtypes = [P, P, P]
    lib.orc_result_copy.argtypes = [P] * 7
    lib.orc_result_stats.argtypes = [P, P, P, P]
    lib.orc_result_free.argtypes = [P]
    lib.orc_new.restype = P
    lib.orc_free.argtypes = [P]
    lib.orc_move_to.argtypes = [P, ctypes.c_float, ctypes.c_float]
    lib.orc_line_to.argtypes = [P, ctypes.c_float, ctypes.c_float]
    lib.orc_command.argtypes = [P, ctypes.c_int, P, ctypes.c_float]
    lib.orc_fill.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t, P]
    lib.orc_stroke.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t,
                               ctypes.c_float, P]
    lib.orc_finish.argtypes = [P, P, P, P]
    lib.orc_stats.argtypes = [P, P, P, P]
    lib.orc_flatten_path.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t,
                                     ctypes.c_float, P]
    lib.orc_stroke_path.argtypes = [P, ctypes.c_size_t, P, ctypes.c_size_t, ctypes.c_float, P]
    lib.orc_path_free.argtypes = [P]
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


class RasterOutput:
    """Tiles and spans of a batch, in TileBuilder call order per path."""

    def __init__(self, tile_xy, tile_alpha, span_xy, span_w, path_tile_off, path_span_off,
                 status=0, stats=None):
        self.tile_xy = tile_xy            # i16 [T, 2]
        self.tile_alpha = tile_alpha      # u8 [T, 64] (row-major 8x8)
        self.span_xy = span_xy            # i16 [S, 2]
        self.span_w = span_w              # u16 [S]
        self.path_tile_off = path_tile_off
        self.path_span_off = path_span_off
        self.status = status
        self.stats = stats or {}

    @property
    def n_tiles(self):
        return int(self.tile_xy.shape[0])

    @property
    def n_spans(self):
        return int(self.span_xy.shape[0])


def _stroke_args(batch, stroke):
    """(widths, stroke_flags) for the oracle: stroke=True strokes every path
    with batch.widths; a batch carrying `strokes` (a draw list) strokes the
    paths flagged 1 and fills the others."""
    flags = getattr(batch, "strokes", None)
    if flags is not None:
        return (np.ascontiguousarray(batch.widths, np.float32), np.ascontiguousarray(flags, np.uint8))
    if stroke:
        return np.ascontiguousarray(batch.widths, np.float32), None
    return None, None


def run_batch(batch, stroke=False) -> RasterOutput:
    lib = load()
    v = np.ascontiguousarray(batch.verbs, np.uint8)
    vo = np.ascontiguousarray(batch.verb_off, np.int64)
    p = np.ascontiguousarray(batch.points, np.float32)
    po = np.ascontiguousarray(batch.point_off, np.int64)
    w = np.ascontiguousarray(batch.weights, np.float32)
    wo = np.ascontiguousarray(batch.weight_off, np.int64)
    xf = np.ascontiguousarray(batch.xforms, np.float32)
    widths, flags = _stroke_args(batch, stroke)
    res = lib.orc_run_batch(batch.n_paths, _ptr(v), _ptr(vo), _ptr(p), _ptr(po),
                            _ptr(w), _ptr(wo), _ptr(xf), _ptr(widths), _ptr(flags))
    try:
        nt = ctypes.c_int64()
        ns = ctypes.c_int64()
        lib.orc_result_sizes(res, ctypes.byref(nt), ctypes.byref(ns))
        T, S = nt.value, ns.value
        txy = np.zeros((T, 2), np.int16)
        ta = np.zeros((T, 64), np.uint8)
        sxy = np.zeros((S, 2), np.int16)
        sw = np.zeros(S, np.uint16)
        pto = np.zeros(batch.n_paths + 1, np.int64)
        pso = np.zeros(batch.n_paths + 1, np.int64)
        lib.orc_result_copy(res, _ptr(txy), _ptr(ta), _ptr(sxy), _ptr(sw), _ptr(pto), _ptr(pso))
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib.orc_result_stats(res, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        status = lib.orc_result_status(res)
        return RasterOutput(txy, ta, sxy, sw, pto, pso, status,
                            {"lines": a.value, "increments": b.value, "tile_increments": c.value})
    finally:
        lib.orc_result_free(res)


TILE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int16, ctypes.c_int16,
                           ctypes.POINTER(ctypes.c_uint8))
SPAN_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int16, ctypes.c_int16, ctypes.c_uint16)


ATLAS_VERTEX = np.dtype([("pos", "<i2", 2), ("uv", "<u2", 2), ("col", "u1", 4)])


def atlas_batch(batch, colors, atlas_size=4096, stroke=False):
    """The demo's atlas Builder (examples/svg.rs:15-88) fed by one oracle
    Rasterizer per path, in path order.  Returns (status, atlas [S, S] u8,
    vertices (ATLAS_VERTEX), indices u32)."""
    lib = load()
    fn = lib.orc_atlas_batch
    fn.restype = ctypes.c_int
    P = ctypes.c_void_p
    fn.argtypes = [ctypes.c_int64, P, P, P, P, P, P, P, P, P, P, ctypes.c_int, P, P, ctypes.c_size_t, P,
                   ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    v = np.ascontiguousarray(batch.verbs, np.uint8)
    vo = np.ascontiguousarray(batch.verb_off, np.int64)
    p = np.ascontiguousarray(batch.points, np.float32)
    po = np.ascontiguousarray(batch.point_off, np.int64)
    w = np.ascontiguousarray(batch.weights, np.float32)
    wo = np.ascontiguousarray(batch.weight_off, np.int64)
    xf = np.ascontiguousarray(batch.xforms, np.float32)
    widths, flags = _stroke_args(batch, stroke)
    col = np.ascontiguousarray(colors, np.uint8).reshape(-1, 4)
    atlas = np.empty((atlas_size, atlas_size), np.uint8)
    nv, ni = ctypes.c_size_t(), ctypes.c_size_t()
    args = [batch.n_paths, _ptr(v), _ptr(vo), _ptr(p), _ptr(po), _ptr(w), _ptr(wo), _ptr(xf),
            _ptr(widths) if widths is not None else None, _ptr(flags) if flags is not None else None, _ptr(col),
            atlas_size, _ptr(atlas)]
    fn(*args, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(ni))
    verts = np.zeros(nv.value, ATLAS_VERTEX)
    idx = np.zeros(ni.value, np.uint32)
    st = fn(*args, _ptr(verts) if verts.size else None, nv.value, _ptr(idx) if idx.size else None, ni.value,
            ctypes.byref(nv), ctypes.byref(ni))
    return st, atlas, verts, idx


class Rasterizer:
    """Oracle Rasterizer with the reference's method set (rasterizer.rs:48-268)."""

    def __init__(self):
        self._lib = load()
        self._h = self._lib.orc_new()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_free(self._h)
            self._h = None

    def move_to(self, x, y):
        self._lib.orc_move_to(self._h, x, y)

    def line_to(self, x, y):
        self._lib.orc_line_to(self._h, x, y)

    def command(self, verb, pts=(), weight=0.0):
        a = np.asarray(pts, np.float32).reshape(-1)
        return self._lib.orc_command(self._h, verb, _ptr(a) if a.size else None, weight)

    def fill(self, verbs, points, weights=(), xform=None):
        v = np.asarray(verbs, np.uint8)
        p = np.asarray(points, np.float32).reshape(-1, 2)
        w = np.asarray(weights, np.float32).reshape(-1)
        xf = np.asarray(xform if xform is not None else [1, 0, 0, 1, 0, 0], np.float32)
        return self._lib.orc_fill(self._h, _ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size, _ptr(xf))

    def stroke(self, verbs, points, width, weights=(), xform=None):
        v = np.asarray(verbs, np.uint8)
        p = np.asarray(points, np.float32).reshape(-1, 2)
        w = np.asarray(weights, np.float32).reshape(-1)
        xf = np.asarray(xform if xform is not None else [1, 0, 0, 1, 0, 0], np.float32)
        return self._lib.orc_stroke(self._h, _ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size,
                                    width, _ptr(xf))

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._lib.orc_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"lines": a.value, "increments": b.value, "tile_increments": c.value}

    def finish(self):
        """Returns the TileBuilder event list: ('tile', x, y, alpha[64]) and
        ('span', x, y, width), in call order."""
        events = []

        def on_tile(_u, x, y, a):
            events.append(("tile", x, y, np.ctypeslib.as_array(a, shape=(64,)).copy()))

        def on_span(_u, x, y, w):
            events.append(("span", x, y, w))

        tf, sf = TILE_FN(on_tile), SPAN_FN(on_span)
        self._lib.orc_finish(self._h, None, ctypes.cast(tf, ctypes.c_void_p), ctypes.cast(sf, ctypes.c_void_p))
        self._lib.orc_free(self._h)
        self._h = None
        return events


class _CPath(ctypes.Structure):
    _fields_ = [("verbs", ctypes.POINTER(ctypes.c_uint8)), ("points", ctypes.POINTER(ctypes.c_float)),
                ("weights", ctypes.POINTER(ctypes.c_float)),
                ("n_verbs", ctypes.c_size_t), ("n_points", ctypes.c_size_t), ("n_weights", ctypes.c_size_t),
                ("cap_verbs", ctypes.c_size_t), ("cap_points", ctypes.c_size_t),
                ("cap_weights", ctypes.c_size_t)]


def _take_path(cp):
    v = np.ctypeslib.as_array(cp.verbs, shape=(cp.n_verbs,)).copy() if cp.n_verbs else np.zeros(0, np.uint8)
    p = (np.ctypeslib.as_array(cp.points, shape=(cp.n_points * 2,)).copy().reshape(-1, 2)
         if cp.n_points else np.zeros((0, 2), np.float32))
    w = np.ctypeslib.as_array(cp.weights, shape=(cp.n_weights,)).copy() if cp.n_weights else np.zeros(0, np.float32)
    return v, p, w


def flatten_path(verbs, points, weights=(), tolerance=0.1):
    """Free `flatten` (path.rs:114-144)."""
    lib = load()
    v = np.asarray(verbs, np.uint8)
    p = np.asarray(points, np.float32).reshape(-1, 2)
    w = np.asarray(weights, np.float32).reshape(-1)
    cp = _CPath()
    st = lib.orc_flatten_path(_ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size, tolerance, ctypes.byref(cp))
    try:
        return st, _take_path(cp)
    finally:
        lib.orc_path_free(ctypes.byref(cp))


def stroke_path(verbs, points, width):
    """Free `stroke` (path.rs:152-274)."""
    lib = load()
    v = np.asarray(verbs, np.uint8)
    p = np.asarray(points, np.float32).reshape(-1, 2)
    cp = _CPath()
    st = lib.orc_stroke_path(_ptr(v), v.size, _ptr(p), p.shape[0], width, ctypes.byref(cp))
    try:
        return st, _take_path(cp)
    finally:
        lib.orc_path_free(ctypes.byref(cp))


# ------------------------------------------------------ f64 coverage adjoint

def _bind_adjoint(lib):
    if getattr(lib, "_adj_bound", False):
        return
    P = ctypes.c_void_p
    S = ctypes.c_size_t
    lib.orc64_path_loss.restype = ctypes.c_int
    lib.orc64_path_loss.argtypes = [P, S, P, S, P, S, P, P, P, S, P, P, S, P, P, P, P, P, S, P]
    lib._adj_bound = True


class Target:
    """A render used as the loss target: tiles (xy, f32 coverage) + spans."""

    def __init__(self, txy, cov, sxy, sw):
        self.txy = np.ascontiguousarray(txy, np.int16).reshape(-1, 2)
        self.cov = np.ascontiguousarray(cov, np.float32).reshape(-1, 64)
        self.sxy = np.ascontiguousarray(sxy, np.int16).reshape(-1, 2)
        self.sw = np.ascontiguousarray(sw, np.uint16).reshape(-1)


EMPTY_TARGET = None


def path_loss64(verbs, points, weights, xf, target=None, grad=False, want_cov=False, cov_cap=4096):
    """f64 forward coverage, loss against `target` and (grad=True) the
    gradient w.r.t. the path-space points and the transform.  Returns
    (status, loss, dpts, dxf, (tile_xy, cov))."""
    lib = load()
    _bind_adjoint(lib)
    v = np.ascontiguousarray(verbs, np.uint8)
    p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 2))
    w = np.ascontiguousarray(weights, np.float64).reshape(-1)
    x = np.ascontiguousarray(xf, np.float64).reshape(6)
    t = target or Target(np.zeros((0, 2)), np.zeros((0, 64)), np.zeros((0, 2)), np.zeros(0))
    loss = ctypes.c_double()
    dp = np.zeros_like(p) if grad else None
    dx = np.zeros(6) if grad else None
    oxy = np.zeros((cov_cap, 2), np.int16) if want_cov else None
    oc = np.zeros((cov_cap, 64)) if want_cov else None
    nt = ctypes.c_size_t()
    st = lib.orc64_path_loss(_ptr(v), v.size, _ptr(p), p.shape[0], _ptr(w), w.size, _ptr(x),
                             _ptr(t.txy), _ptr(t.cov), t.txy.shape[0], _ptr(t.sxy), _ptr(t.sw), t.sw.size,
                             ctypes.byref(loss), _ptr(dp) if grad else None, _ptr(dx) if grad else None,
                             _ptr(oxy) if want_cov else None, _ptr(oc) if want_cov else None, cov_cap,
                             ctypes.byref(nt))
    cov = (oxy[:nt.value], oc[:nt.value]) if want_cov else None
    return st, loss.value, dp, dx, cov
