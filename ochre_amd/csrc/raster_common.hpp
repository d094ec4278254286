// Device code shared by the plain and the batched pipelines: command ->
// line enumeration (path.rs:41-109, rasterizer.rs:61-69,145-165,181-183),
// the per-path tile grid, and the DDA walk (rasterizer.rs:72-140).
#pragma once

#include "device_common.hpp"

namespace ochre {

// Resolved command: its transformed points and the rasterizer state
// (`last`, `first`) it starts from (rasterizer.rs:145-165).
struct Cmd {
    int verb;
    int32_t path;
    bool last_in_path;
    bool bad;
    V2 pt[3];
    float w;
    V2 start, first;
};

// Enumerates the lines a command feeds to line_to (rasterizer.rs:61-69,
// 145-157 and the finish-time close, :181-183), in order.
template <class Emit>
__device__ inline bool command_lines(const Cmd& c, Emit&& emit) {
    bool ok = true;
    V2 last = c.start;
    V2 first = c.first;
    switch (c.verb) {
    case MOVE:
        if (!veq(last, first)) emit(last, first);
        last = c.pt[0];
        first = c.pt[0];
        break;
    case LINE:
        emit(last, c.pt[0]);
        last = c.pt[0];
        break;
    case QUAD: {
        float dt = quad_dt(last, c.pt[0], c.pt[1]);
        float t = 0.0f;
        uint32_t n = 0;
        V2 prev = last;
        while (t < 1.0f) {
            t = fminf(t + dt, 1.0f);
            V2 q = quad_at(t, last, c.pt[0], c.pt[1]);
            emit(prev, q);
            prev = q;
            if (++n >= FLATTEN_LIMIT) {
                ok = false;
                break;
            }
        }
        last = prev;
        break;
    }
    case CUBIC: {
        float dt = cubic_dt(last, c.pt[0], c.pt[1], c.pt[2]);
        float t = 0.0f;
        uint32_t n = 0;
        V2 prev = last;
        while (t < 1.0f) {
            t = fminf(t + dt, 1.0f);
            V2 q = cubic_at(t, last, c.pt[0], c.pt[1], c.pt[2]);
            emit(prev, q);
            prev = q;
            if (++n >= FLATTEN_LIMIT) {
                ok = false;
                break;
            }
        }
        last = prev;
        break;
    }
    case CONIC: {
        V2 prev = last;
        ok = conic_walk(last, c.pt[0], c.pt[1], c.w, [&](V2 q) {
            emit(prev, q);
            prev = q;
        });
        last = prev;
        break;
    }
    default:  // CLOSE: a no-op in the rasterizer (rasterizer.rs:154)
        break;
    }
    if (c.last_in_path && !veq(last, first)) emit(last, first);
    return ok;
}

// command_lines' count of lines, visiting (add) points whose box holds every
// line's end points: a quadratic's or cubic's control points instead of its
// evaluated points (their convex hull holds the curve; the grid's margin takes
// the evaluation's rounding), so a curve's lines are counted by the
// reference's t loop alone; a conic is flattened as command_lines does.
template <class Add>
__device__ inline bool command_count(const Cmd& c, uint32_t& n, Add&& add) {
    bool ok = true;
    V2 last = c.start;
    V2 first = c.first;
    switch (c.verb) {
    case MOVE:
        if (!veq(last, first)) {
            n++;
            add(last);
            add(first);
        }
        last = c.pt[0];
        first = c.pt[0];
        break;
    case LINE:
        n++;
        add(last);
        add(c.pt[0]);
        last = c.pt[0];
        break;
    case QUAD:
    case CUBIC: {
        const bool cubic = c.verb == CUBIC;
        const float dt = cubic ? cubic_dt(last, c.pt[0], c.pt[1], c.pt[2]) : quad_dt(last, c.pt[0], c.pt[1]);
        bool limited = false;
        n += t_steps(dt, &limited);
        if (limited) ok = false;
        add(last);
        add(c.pt[0]);
        add(c.pt[1]);
        if (cubic) add(c.pt[2]);
        last = cubic ? cubic_at(1.0f, last, c.pt[0], c.pt[1], c.pt[2]) : quad_at(1.0f, last, c.pt[0], c.pt[1]);
        break;
    }
    case CONIC: {
        V2 prev = last;
        ok = conic_walk(last, c.pt[0], c.pt[1], c.w, [&](V2 q) {
            n++;
            add(prev);
            add(q);
            prev = q;
        });
        last = prev;
        break;
    }
    default:
        break;
    }
    if (c.last_in_path && !veq(last, first)) {
        n++;
        add(last);
        add(first);
    }
    return ok;
}

constexpr float HUGE_COORD = 30000.0f;
enum BoxFlags : uint32_t { BOX_FULL = 1u, BOX_LINES = 2u };

struct BoxAcc {
    float minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
    bool full = false;
    __device__ void add(V2 p) {
        if (!(fabsf(p.x) <= HUGE_COORD) || !(fabsf(p.y) <= HUGE_COORD)) {
            full = true;
            return;
        }
        minx = fminf(minx, p.x);
        miny = fminf(miny, p.y);
        maxx = fmaxf(maxx, p.x);
        maxy = fmaxf(maxy, p.y);
    }
};

// Per-pixel sums of the unordered tile stage in fixed point, units of 2^-30,
// held as integer-valued f64: an increment's |area|, |height| <= 1 plus a few
// ulps for lines with finite coordinates within +-30000, so the sums of 2^22 of
// them stay below 2^53 units, where every f64 sum of integers is exact; the
// sum is then independent of arrival order, and the decoded f32 value is the
// correctly rounded sum of the increments (each truncated at 2^-30).  The
// conversion is three f64 operations (f32 to f64, scale, truncate), where a
// 64-bit integer one needs ten.
constexpr double FX_SCALE = 1073741824.0;       // 2^30
constexpr double FX_UNIT = 9.313225746154785e-10;  // 2^-30

__device__ inline double fx_of(float v) { return trunc((double)v * FX_SCALE); }
__device__ inline float fx_to_f(double q) { return (float)(q * FX_UNIT); }

// ------------------------------------------------------ (b) DDA binning
//
// Walks Rasterizer::line_to (rasterizer.rs:72-140) for one line and reports
// tile runs (maximal runs of consecutive increments in one tile) and
// tile-row crossings (TileIncrement, :123-131).  A run record holds the DDA
// state one step before the run's first increment (or "line start"), which
// is enough to replay the run bit-exactly later.
struct RunRec {
    uint32_t line;  // bit 31: run starts at the line's first increment
    uint32_t xy;    // x | y << 16 (i16) of the state before the entering step
    float row_t1, col_t1;
};

// Steps per line before the walk gives up: a line spanning the whole i16
// pixel range needs ~2^17; more only arise from coordinates so large that
// the reference would exhaust memory pushing increments.
constexpr uint32_t WALK_LIMIT = 1u << 20;

template <class Sink>
__device__ inline void walk_line(V2 last, V2 point, Sink& sink) {
    if (veq(point, last)) return;
    Dda d;
    dda_init(d, last, point);
    int16_t ty_prev = div8(d.y);
    int16_t cur_tx = div8(d.x), cur_ty = ty_prev;
    sink.run(cur_tx, cur_ty, true, d);
    for (uint32_t steps = 0;; steps++) {
        if (steps >= WALK_LIMIT) {
            sink.limit();
            break;
        }
        Dda before = d;
        bool term = dda_step(d);
        if (term) {
            d.x = f2i16(floorf(point.x));
            d.y = f2i16(floorf(point.y));
        }
        int16_t ty = div8(d.y);
        if (ty != ty_prev) {
            sink.crossing(div8(d.x), ty_prev < ty ? ty_prev : ty, (int8_t)(int16_t)(ty - ty_prev));
            ty_prev = ty;
        }
        if (term) break;
        int16_t tx = div8(d.x);
        if (tx != cur_tx || ty != cur_ty) {
            cur_tx = tx;
            cur_ty = ty;
            sink.run(tx, ty, false, before);
        }
    }
}


// Lines that line_counts and line_tiles cover: finite, within +-30000, at
// most 4096 px per axis (no i16 wrap, far below the walk's step limit).
__device__ inline bool chain_walk_ok(V2 last, V2 point) {
    return fabsf(last.x) <= HUGE_COORD && fabsf(last.y) <= HUGE_COORD && fabsf(point.x) <= HUGE_COORD &&
           fabsf(point.y) <= HUGE_COORD && fabsf(point.x - last.x) <= 4096.0f && fabsf(point.y - last.y) <= 4096.0f;
}

// walk_line at tile granularity (rasterizer.rs:72-140): the tiles the walk
// visits, in order (tile(tx, ty)), its tile-row crossings, the end-point
// snap's included (cross(tx, ty, sign)), and its increments (`incs`), without
// the per-pixel steps.  Each axis' crossing parameters form a chain of their
// own whatever the interleaving (c, min(c + step, 1), ...); the step that
// moves an axis into its next tile is k values along its chain (k from the
// pixel's place in its tile, then 8), and of two such steps the one with the
// smaller parameter comes first, the x step on a tie (dda_step takes y only
// when row_t1 < col_t1).  Every value below 1 is taken before the walk ends,
// so the increments are the chains' values below 1, plus the last one.  Only
// for the lines chain_walk_ok covers; returns false, visiting nothing, for
// the others.
template <class Tile, class Cross>
__device__ inline bool line_tiles(V2 last, V2 point, uint32_t& incs, Tile&& tile, Cross&& cross) {
    incs = 0;
    if (!chain_walk_ok(last, point)) return false;
    if (veq(point, last)) return true;  // rasterizer.rs:73
    Dda d;
    dda_init(d, last, point);
    int32_t tx = (int32_t)d.x >> 3, ty = (int32_t)d.y >> 3;
    tile(tx, ty);
    auto first_k = [](int32_t p, int dir) { return dir > 0 ? 8 - (p & 7) : (p & 7) + 1; };
    auto advance = [](float c, float step, int k) {
        for (int m = 1; m < k; m++) c = fminf(c + step, 1.0f);
        return c;
    };
    float ex = d.x_dir ? advance(d.col_t1, d.x_step, first_k(d.x, d.x_dir)) : 1.0f;
    float ey = d.y_dir ? advance(d.row_t1, d.y_step, first_k(d.y, d.y_dir)) : 1.0f;
    for (;;) {
        const bool xe = ex < 1.0f, ye = ey < 1.0f;
        if (xe && (!ye || ex <= ey)) {
            tx += d.x_dir;
            tile(tx, ty);
            ex = advance(fminf(ex + d.x_step, 1.0f), d.x_step, 8);
        } else if (ye) {
            const int32_t nty = ty + d.y_dir;
            cross(tx, nty < ty ? nty : ty, (int)d.y_dir);
            ty = nty;
            tile(tx, ty);
            ey = advance(fminf(ey + d.y_step, 1.0f), d.y_step, 8);
        } else {
            break;
        }
    }
    uint32_t nx = 0, ny = 0;
    for (float c = d.col_t1; c < 1.0f; c = fminf(c + d.x_step, 1.0f)) nx++;
    for (float c = d.row_t1; c < 1.0f; c = fminf(c + d.y_step, 1.0f)) ny++;
    incs = nx + ny + 1u;
    const int32_t ty_end = (int32_t)f2i16(floorf(point.y)) >> 3;  // the snap (rasterizer.rs:118-131)
    if (ty_end != ty) cross((int32_t)f2i16(floorf(point.x)) >> 3, ty_end < ty ? ty_end : ty, (int)(ty_end - ty));
    return true;
}

// walk_line's runs and tile-row crossings (for sinks that take no increments)
// at tile granularity, for the lines chain_walk_ok covers: line_tiles' chains,
// with the other axis' chain advanced to each tile event, which gives the
// state one step before the run the event starts (the x step is taken on a tie:
// the y steps before an x event are those with a smaller parameter, the x steps
// before a y event those with a smaller or equal one).  Emits the runs and
// crossings in walk_line's order.  Returns false, emitting nothing, for the
// lines it does not cover.
template <class Sink>
__device__ inline bool walk_line_tiles(V2 last, V2 point, Sink& sink) {
    if (!chain_walk_ok(last, point)) return false;
    if (veq(point, last)) return true;  // rasterizer.rs:73
    Dda d;
    dda_init(d, last, point);
    const int32_t x0 = d.x, y0 = d.y, xd = d.x_dir, yd = d.y_dir;
    int32_t tx = x0 >> 3, ty = y0 >> 3;
    sink.run((int16_t)tx, (int16_t)ty, true, d);
    auto first_k = [](int32_t p, int32_t dir) { return (uint32_t)(dir > 0 ? 8 - (p & 7) : (p & 7) + 1); };
    auto value_at = [](float c, uint32_t i, uint32_t target, float step) {
        for (; i < target; i++) c = fminf(c + step, 1.0f);
        return c;
    };
    uint32_t ix = 0, iy = 0;  // steps taken along each axis, and the next step's parameter
    float cx = d.col_t1, cy = d.row_t1;
    uint32_t exi = xd ? first_k(x0, xd) - 1u : 0u, eyi = yd ? first_k(y0, yd) - 1u : 0u;  // the events' step indices
    Dda b = d;
    for (;;) {
        const float exv = xd ? value_at(cx, ix, exi, d.x_step) : INFINITY;
        const float eyv = yd ? value_at(cy, iy, eyi, d.y_step) : INFINITY;
        const bool xe = exv < 1.0f, ye = eyv < 1.0f;
        if (xe && (!ye || exv <= eyv)) {
            while (cy < exv) {
                cy = fminf(cy + d.y_step, 1.0f);
                iy++;
            }
            b.x = (int16_t)(x0 + xd * (int32_t)exi);
            b.y = (int16_t)(y0 + yd * (int32_t)iy);
            b.row_t1 = cy;
            b.col_t1 = exv;
            tx += xd;
            sink.run((int16_t)tx, (int16_t)ty, false, b);
            ix = exi + 1u;
            cx = fminf(exv + d.x_step, 1.0f);
            exi += 8u;
        } else if (ye) {
            while (cx <= eyv) {
                cx = fminf(cx + d.x_step, 1.0f);
                ix++;
            }
            b.x = (int16_t)(x0 + xd * (int32_t)ix);
            b.y = (int16_t)(y0 + yd * (int32_t)eyi);
            b.row_t1 = eyv;
            b.col_t1 = cx;
            const int32_t nty = ty + yd;
            sink.crossing((int16_t)tx, (int16_t)(nty < ty ? nty : ty), (int8_t)yd);
            ty = nty;
            sink.run((int16_t)tx, (int16_t)ty, false, b);
            iy = eyi + 1u;
            cy = fminf(eyv + d.y_step, 1.0f);
            eyi += 8u;
        } else {
            break;
        }
    }
    const int32_t ty_end = (int32_t)f2i16(floorf(point.y)) >> 3;  // the snap (rasterizer.rs:118-131)
    if (ty_end != ty)
        sink.crossing((int16_t)((int32_t)f2i16(floorf(point.x)) >> 3), (int16_t)(ty_end < ty ? ty_end : ty),
                      (int8_t)(int16_t)(ty_end - ty));
    return true;
}

// Tile grid of a path from the bounding box (pixel floors) of all its line
// end points, widened by the DDA's worst-case overshoot: the crossing
// parameters are summed sequentially (rasterizer.rs:110,114), so after k
// steps they may be off by k * 2^-24, i.e. up to span^2 * 2^-24 pixels, plus
// one pixel of floor().  Paths with huge or non-finite coordinates get the
// whole i16 range (the DDA's coordinates wrap in i16).
__host__ __device__ inline void grid_from_box(bool has_lines, bool full, int32_t minx, int32_t miny, int32_t maxx,
                                     int32_t maxy, int32_t* tx0, int32_t* ty0, int32_t* w, int32_t* h) {
    *tx0 = *ty0 = *w = *h = 0;
    if (!has_lines) return;
    if (full) {
        *tx0 = *ty0 = -4096;
        *w = *h = 8192;
        return;
    }
    int64_t span = ((int64_t)maxx - minx > (int64_t)maxy - miny ? (int64_t)maxx - minx : (int64_t)maxy - miny) + 2;
    int64_t margin = 2 + (int64_t)ceil((double)span * (double)span * 1.2e-7);
    int64_t px0 = (int64_t)minx - margin, py0 = (int64_t)miny - margin;
    int64_t px1 = (int64_t)maxx + margin, py1 = (int64_t)maxy + margin;
    if (px0 < -32768) px0 = -32768;
    if (py0 < -32768) py0 = -32768;
    if (px1 > 32767) px1 = 32767;
    if (py1 > 32767) py1 = 32767;
    *tx0 = (int32_t)(px0 >> 3);
    *ty0 = (int32_t)(py0 >> 3);
    *w = (int32_t)((px1 >> 3) - *tx0 + 1);
    *h = (int32_t)((py1 >> 3) - *ty0 + 1);
}

}  // namespace ochre
