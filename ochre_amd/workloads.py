"""Seeded synthetic path workloads.

Each generator returns a :class:`PathBatch`: the structure-of-arrays path
encoding the C ABI takes (one verb byte per ``PathCmd``, the commands' points
as f32 (x, y) pairs, conic weights, and one affine transform per path).  The
verb codes are the ``PathCmd`` discriminants of the reference
(src/path.rs:5-12): Move=0, Line=1, Quadratic=2, Cubic=3, Conic=4, Close=5.

Workloads:
  readme_path()   the doctest path of src/lib.rs:35-42 on an 800x600 canvas.
  glyph_run()     10,000 glyph-like paths on a 4096x4096 canvas: 1-4 closed
                  contours of 8-40 segments (40% line, 40% quadratic, 20% cubic),
                  em-box control points scaled to 12-48 px and translated.
  star_blobs()    200 self-intersecting blobs: one contour of log-uniform
                  1k-100k cubics on a radius-noised star (radius 200-2000 px,
                  sigma 15% of radius) centred on a 4096x4096 canvas.
  scenes()        256 independent 2048x2048 scenes of log-uniform 1-5000 paths
                  each (95% glyphs, 5% blobs of 1k-20k cubics) with a random
                  affine transform per path.
  adjoint_glyphs() 4,096 paths of one contour of 16 cubics (32-64 px) on a
                  1024x1024 canvas, for the coverage adjoint.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MOVE, LINE, QUAD, CUBIC, CONIC, CLOSE = 0, 1, 2, 3, 4, 5
VERB_POINTS = np.array([1, 1, 2, 3, 2, 0], dtype=np.int64)

SEED_BASE = 0x6F63687265000000  # "ochre"


@dataclasses.dataclass
class PathBatch:
    verbs: np.ndarray        # u8 [V]
    verb_off: np.ndarray     # i64 [n+1]
    points: np.ndarray       # f32 [P, 2]
    point_off: np.ndarray    # i64 [n+1]
    weights: np.ndarray      # f32 [W]
    weight_off: np.ndarray   # i64 [n+1]
    xforms: np.ndarray       # f32 [n, 6]: m00 m01 m10 m11 ox oy (row-major Mat2x2 + offset)
    widths: np.ndarray | None = None   # f32 [n] stroke widths (stroke mode)
    scene: np.ndarray | None = None    # i32 [n] scene id (batched scenes)
    strokes: np.ndarray | None = None  # u8 [n] draw list: 1 stroke (with widths), 0 fill
    colors: np.ndarray | None = None   # u8 [n, 4] draw colours (SVG draw lists)
    name: str = ""

    @property
    def n_paths(self) -> int:
        return int(self.verb_off.shape[0] - 1)

    def path(self, i: int):
        """(verbs, points, weights, xform) of path i."""
        v = self.verbs[self.verb_off[i]:self.verb_off[i + 1]]
        p = self.points[self.point_off[i]:self.point_off[i + 1]]
        w = self.weights[self.weight_off[i]:self.weight_off[i + 1]]
        return v, p, w, self.xforms[i]

    def subset(self, idx) -> "PathBatch":
        """A new batch holding the listed paths, in the listed order."""
        idx = np.asarray(idx, dtype=np.int64)
        parts = [self.path(int(i)) for i in idx]
        b = from_paths([(v, p, w) for v, p, w, _ in parts], self.xforms[idx])
        if self.widths is not None:
            b.widths = self.widths[idx].copy()
        if self.scene is not None:
            b.scene = self.scene[idx].copy()
        if self.strokes is not None:
            b.strokes = self.strokes[idx].copy()
        if self.colors is not None:
            b.colors = self.colors[idx].copy()
        b.name = self.name + "[subset]"
        return b

    def algorithmic_input_bytes(self) -> int:
        """Bytes a path batch occupies as the C ABI reads it: verbs (1 B),
        points (8 B), weights (4 B), transform (24 B) per path."""
        return int(self.verbs.size + 8 * self.points.shape[0] + 4 * self.weights.size
                   + 24 * self.n_paths)


def _offsets(counts) -> np.ndarray:
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(np.asarray(counts, dtype=np.int64), out=off[1:])
    return off


def from_paths(paths, xforms) -> PathBatch:
    """Build a batch from a list of (verbs, points[, weights]) tuples."""
    vs, ps, ws = [], [], []
    for item in paths:
        v = np.asarray(item[0], dtype=np.uint8)
        p = np.asarray(item[1], dtype=np.float32).reshape(-1, 2)
        w = np.asarray(item[2] if len(item) > 2 else [], dtype=np.float32).reshape(-1)
        vs.append(v)
        ps.append(p)
        ws.append(w)
    xf = np.asarray(xforms, dtype=np.float32).reshape(-1, 6)
    if xf.shape[0] == 1 and len(paths) > 1:
        xf = np.repeat(xf, len(paths), axis=0)
    return PathBatch(
        verbs=np.concatenate(vs) if vs else np.zeros(0, np.uint8),
        verb_off=_offsets([len(v) for v in vs]),
        points=np.concatenate(ps).astype(np.float32) if ps else np.zeros((0, 2), np.float32),
        point_off=_offsets([len(p) for p in ps]),
        weights=np.concatenate(ws).astype(np.float32) if ws else np.zeros(0, np.float32),
        weight_off=_offsets([len(w) for w in ws]),
        xforms=np.ascontiguousarray(xf),
    )


def concat(batches) -> PathBatch:
    """Concatenate batches (path order preserved)."""
    verbs = np.concatenate([b.verbs for b in batches])
    points = np.concatenate([b.points for b in batches])
    weights = np.concatenate([b.weights for b in batches])
    def cat_off(name):
        out = [np.zeros(1, np.int64)]
        base = 0
        for b in batches:
            o = getattr(b, name)
            out.append(o[1:] + base)
            base += int(o[-1])
        return np.concatenate(out)
    res = PathBatch(verbs=verbs, verb_off=cat_off("verb_off"), points=points,
                    point_off=cat_off("point_off"), weights=weights,
                    weight_off=cat_off("weight_off"),
                    xforms=np.concatenate([b.xforms for b in batches]))
    if all(b.scene is not None for b in batches):
        res.scene = np.concatenate([b.scene for b in batches])
    return res


IDENTITY = np.array([1, 0, 0, 1, 0, 0], dtype=np.float32)


def readme_path() -> PathBatch:
    """The doctest path (src/lib.rs:35-42, examples/basic.rs:26-31)."""
    verbs = [MOVE, QUAD, CUBIC, CLOSE]
    pts = [(400, 300), (500, 200), (400, 100), (350, 150), (100, 250), (400, 300)]
    b = from_paths([(verbs, pts)], IDENTITY)
    b.name = "readme_path"
    return b


def _glyph_geometry(rng, n, seg_min=8, seg_max=40, contours_max=4):
    """Verbs/points (em-box coordinates) of n glyph-like paths."""
    nc = rng.integers(1, contours_max + 1, n)
    C = int(nc.sum())
    nseg = rng.integers(seg_min, seg_max + 1, C)
    S = int(nseg.sum())
    u = rng.random(S)
    seg_type = np.where(u < 0.4, LINE, np.where(u < 0.8, QUAD, CUBIC)).astype(np.uint8)
    cid = np.repeat(np.arange(C), nseg)
    seg_start = _offsets(nseg)
    vstart = seg_start[:-1] + 2 * np.arange(C)          # Move position of each contour
    V = S + 2 * C
    verbs = np.empty(V, dtype=np.uint8)
    verbs[vstart] = MOVE
    verbs[np.arange(S) + 2 * cid + 1] = seg_type
    verbs[vstart + nseg + 1] = CLOSE
    npts = VERB_POINTS[verbs]
    P = int(npts.sum())
    points = rng.random((P, 2)).astype(np.float32)
    # per-path offsets
    path_of_contour = np.repeat(np.arange(n), nc)
    verbs_per_path = np.bincount(path_of_contour, weights=nseg + 2, minlength=n).astype(np.int64)
    verb_off = _offsets(verbs_per_path)
    pts_cum = _offsets(npts)
    point_off = pts_cum[verb_off]
    return verbs, verb_off, points, point_off


def _empty_weights(n):
    return np.zeros(0, np.float32), np.zeros(n + 1, np.int64)


def glyph_run(n=10_000, seed=1, canvas=4096.0, smin=12.0, smax=48.0) -> PathBatch:
    rng = np.random.default_rng(SEED_BASE + seed)
    verbs, verb_off, points, point_off = _glyph_geometry(rng, n)
    s = rng.uniform(smin, smax, n).astype(np.float32)
    off = (rng.random((n, 2)) * (canvas - s[:, None])).astype(np.float32)
    xf = np.zeros((n, 6), np.float32)
    xf[:, 0] = s
    xf[:, 3] = s
    xf[:, 4:] = off
    w, wo = _empty_weights(n)
    return PathBatch(verbs, verb_off, points, point_off, w, wo, xf, name=f"glyph_run(n={n})")


def stroked_glyphs(n=10_000, seed=5, canvas=4096.0, smin=12.0, smax=48.0, wmin=0.75, wmax=3.0) -> PathBatch:
    """The glyph geometry in pixel units (as SVG strokes are authored: the
    path-space flattening tolerance is 0.1 px), stroked with widths of
    0.75-3 px and placed by a translation."""
    rng = np.random.default_rng(SEED_BASE + seed)
    verbs, verb_off, points, point_off = _glyph_geometry(rng, n)
    s = rng.uniform(smin, smax, n).astype(np.float32)
    counts = np.diff(point_off)
    points = (points * np.repeat(s, counts)[:, None]).astype(np.float32)
    off = (rng.random((n, 2)) * (canvas - s[:, None])).astype(np.float32)
    xf = np.zeros((n, 6), np.float32)
    xf[:, 0] = 1.0
    xf[:, 3] = 1.0
    xf[:, 4:] = off
    w, wo = _empty_weights(n)
    b = PathBatch(verbs, verb_off, points, point_off, w, wo, xf, name=f"stroked_glyphs(n={n})")
    b.widths = rng.uniform(wmin, wmax, n).astype(np.float32)
    return b


def with_widths(batch: PathBatch, lo: float, hi: float, seed: int) -> PathBatch:
    """The batch with stroke widths drawn from [lo, hi) (path units)."""
    rng = np.random.default_rng(SEED_BASE + 2000 + seed)
    b = dataclasses.replace(batch)
    b.widths = rng.uniform(lo, hi, batch.n_paths).astype(np.float32)
    b.name = f"{batch.name}+widths[{lo},{hi})"
    return b


def _blob_points(rng, N, R, cx, cy):
    M = 3 * N
    theta = 2.0 * math.pi * np.arange(M) / M
    z = rng.standard_normal(M)
    r = R * (1.0 + 0.15 * z)
    pts = np.empty((M + 1, 2), np.float64)
    pts[:M, 0] = cx + r * np.cos(theta)
    pts[:M, 1] = cy + r * np.sin(theta)
    pts[M] = pts[0]
    return pts.astype(np.float32)


def _blob_verbs(N):
    v = np.full(N + 2, CUBIC, np.uint8)
    v[0] = MOVE
    v[-1] = CLOSE
    return v


def star_blobs(n=200, seed=2, canvas=4096.0, nmin=1e3, nmax=1e5, rmin=200.0, rmax=2000.0) -> PathBatch:
    rng = np.random.default_rng(SEED_BASE + seed)
    Ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), n))).astype(np.int64)
    Rs = rng.uniform(rmin, rmax, n)
    paths = []
    for i in range(n):
        paths.append((_blob_verbs(int(Ns[i])), _blob_points(rng, int(Ns[i]), Rs[i], canvas / 2, canvas / 2)))
    b = from_paths(paths, IDENTITY[None, :].repeat(n, 0))
    b.name = f"star_blobs(n={n})"
    return b


def _affine(rng, n, scale):
    th = rng.uniform(0.0, 2.0 * math.pi, n)
    sh = rng.uniform(-0.2, 0.2, n)
    c, s = np.cos(th), np.sin(th)
    # Rot(th) . Shear(sh) . Scale(scale)
    m00 = c * scale
    m01 = (c * sh - s) * scale
    m10 = s * scale
    m11 = (s * sh + c) * scale
    return m00, m01, m10, m11


def scenes(n_scenes=256, seed=3, canvas=2048.0, max_paths=5000, blob_frac=0.05,
           nmin=1e3, nmax=2e4, rmin=100.0, rmax=1000.0, radius_scale_cap=1024.0) -> PathBatch:
    """Batched uneven scenes.  Blob radius is scene-relative (100-1000 px) and
    the per-path scale of a blob is capped so that radius*scale <= 1024 px;
    with the literal 200-2000 px radius at scale up to 4 a single scene would
    emit on the order of 1e9 tiles."""
    rng = np.random.default_rng(SEED_BASE + seed)
    counts = np.rint(np.exp(rng.uniform(0.0, math.log(max_paths), n_scenes))).astype(np.int64)
    n = int(counts.sum())
    scene = np.repeat(np.arange(n_scenes, dtype=np.int32), counts)
    is_blob = rng.random(n) < blob_frac
    gidx = np.nonzero(~is_blob)[0]
    bidx = np.nonzero(is_blob)[0]
    ng, nb = len(gidx), len(bidx)
    gverbs, gvoff, gpts, gpoff = _glyph_geometry(rng, ng)
    # glyph transform: affine(scale 0.5-4) . (s * I), s in 12-48
    gs = rng.uniform(12.0, 48.0, ng)
    gsc = rng.uniform(0.5, 4.0, ng)
    m = _affine(rng, ng, gsc)
    gx = np.stack([m[0] * gs, m[1] * gs, m[2] * gs, m[3] * gs,
                   rng.uniform(0, canvas, ng), rng.uniform(0, canvas, ng)], 1).astype(np.float32)
    # blobs: centred at the origin in path space, then affine + offset
    Ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nb))).astype(np.int64)
    Rs = rng.uniform(rmin, rmax, nb)
    bsc = rng.uniform(0.5, 1.0, nb) * np.minimum(4.0, radius_scale_cap / Rs)
    mb = _affine(rng, nb, bsc)
    bx = np.stack([mb[0], mb[1], mb[2], mb[3], rng.uniform(0, canvas, nb), rng.uniform(0, canvas, nb)],
                  1).astype(np.float32)
    # assemble in path order
    per_path = [None] * n
    for k, i in enumerate(gidx):
        per_path[i] = (gverbs[gvoff[k]:gvoff[k + 1]], gpts[gpoff[k]:gpoff[k + 1]])
    for k, i in enumerate(bidx):
        per_path[i] = (_blob_verbs(int(Ns[k])), _blob_points(rng, int(Ns[k]), Rs[k], 0.0, 0.0))
    xf = np.empty((n, 6), np.float32)
    xf[gidx] = gx
    xf[bidx] = bx
    b = from_paths(per_path, xf)
    b.scene = scene
    b.name = f"scenes(n_scenes={n_scenes}, paths={n})"
    return b


def adjoint_glyphs(n=4096, seed=4, canvas=1024.0, smin=32.0, smax=64.0, n_cubics=16) -> PathBatch:
    rng = np.random.default_rng(SEED_BASE + seed)
    V = n_cubics + 2
    verbs = np.tile(_blob_verbs(n_cubics), n)
    pts = rng.random((n * (1 + 3 * n_cubics), 2)).astype(np.float32)
    s = rng.uniform(smin, smax, n).astype(np.float32)
    off = (rng.random((n, 2)) * (canvas - s[:, None])).astype(np.float32)
    xf = np.zeros((n, 6), np.float32)
    xf[:, 0] = s
    xf[:, 3] = s
    xf[:, 4:] = off
    w, wo = _empty_weights(n)
    b = PathBatch(verbs, np.arange(n + 1, dtype=np.int64) * V, pts,
                  np.arange(n + 1, dtype=np.int64) * (1 + 3 * n_cubics), w, wo, xf,
                  name=f"adjoint_glyphs(n={n})")
    return b


def random_paths(n, seed, max_cmds=24, coord_scale=64.0, conic=True, close_prob=0.15,
                 no_move_prob=0.1, integer_prob=0.2, xform="random") -> PathBatch:
    """Fuzz paths: every verb kind, Close in the middle of contours, paths that
    start without a Move, integer and half-integer coordinates (DDA ties),
    negative coordinates, empty paths and a random affine transform."""
    rng = np.random.default_rng(SEED_BASE + 1000 + seed)
    paths, xfs = [], []
    kinds = [LINE, QUAD, CUBIC, CONIC] if conic else [LINE, QUAD, CUBIC]
    for _ in range(n):
        k = int(rng.integers(0, max_cmds + 1))
        verbs, pts, ws = [], [], []
        snap = rng.random() < integer_prob
        def pt():
            p = (rng.random(2) - 0.3) * coord_scale
            if snap:
                p = np.round(p * 2.0) / 2.0
            return p
        for j in range(k):
            if j == 0 and rng.random() >= no_move_prob:
                v = MOVE
            else:
                r = rng.random()
                if r < close_prob:
                    v = CLOSE
                elif r < close_prob + 0.1:
                    v = MOVE
                else:
                    v = kinds[int(rng.integers(0, len(kinds)))]
            verbs.append(v)
            for _ in range(int(VERB_POINTS[v])):
                pts.append(pt())
            if v == CONIC:
                ws.append(float(rng.uniform(0.2, 3.0)))
        paths.append((verbs, np.asarray(pts, np.float32).reshape(-1, 2), ws))
        if xform == "identity":
            xfs.append(IDENTITY)
        else:
            th = rng.uniform(0, 2 * math.pi)
            sc = rng.uniform(0.5, 2.0)
            xfs.append(np.array([sc * math.cos(th), -sc * math.sin(th), sc * math.sin(th), sc * math.cos(th),
                                 rng.uniform(-16, 48), rng.uniform(-16, 48)], np.float32))
    b = from_paths(paths, np.stack(xfs))
    b.name = f"random_paths(n={n}, seed={seed})"
    return b


WORKLOADS = {
    "readme_path": readme_path,
    "glyph_run": glyph_run,
    "star_blobs": star_blobs,
    "scenes": scenes,
    "adjoint_glyphs": adjoint_glyphs,
    "stroked_glyphs": stroked_glyphs,
}
