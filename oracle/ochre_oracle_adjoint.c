/*
 * ochre_oracle_adjoint.c -- f64 restatement of the reference's coverage and
 * its adjoint.  TEST INFRASTRUCTURE ONLY (the checker for the device
 * adjoint; never shipped or measured as the product).
 *
 * Forward: the reference rasteriser (src/path.rs:41-109 flattening,
 * src/rasterizer.rs:72-140 DDA, :221-265 accumulation, carry and spans)
 * carried out in double precision, keeping the pre-quantisation coverage
 * cov = min(|accum + area|, 1) of every tile pixel (rasterizer.rs:235).
 *
 * Loss of one path: sum over the pixels of the union of both renders' tiles
 * of (cov - target)^2, where a pixel outside a render's tiles has coverage 1
 * inside one of its spans and 0 elsewhere.
 *
 * Adjoint: reverse-mode derivative of that loss with respect to the path's
 * control points (path space) and its Transform (m00 m01 m10 m11 ox oy):
 *   dL/dv = 2 (cov - target) * sign(v) where |v| < 1, else 0;
 *   v(x) = carry + sum_{x' < x} h(x') + a(x) along each pixel row of a tile
 *   row, so dL/da(q) = g(q), dL/dh(q) = sum of g strictly right of q;
 *   each DDA increment's (area, height) is differentiated with respect to the
 *   line end points through its clip points (pixel-boundary crossings);
 *   line end points are curve evaluations at t_k = k * dt, differentiated
 *   through the Bernstein weights and through dt (path.rs:50,63);
 *   points are world = M * local + offset (geom.rs:260-262).
 * Conic flattening (dyadic t, no dt) is differentiated through its rational
 * point formula; weights themselves are not differentiated.
 * Discrete structure: how many segments a curve flattens to, and where a
 * conic subdivides, is decided exactly as the reference decides it, in f32
 * from the f32 world points (path.rs:50-57,63-73,86-103); the f64 forward
 * then evaluates those segments' end points at t_k = k * dt in double
 * precision.  The loss is therefore the f64 value of the reference's own
 * piecewise-smooth coverage function, and its kinks (a segment count
 * changing) sit where the reference's do.
 *
 * Parity pinning: the forward's u8 quantisation of cov matches the f32
 * oracle within 1 LSB (tests/test_oracle_adjoint.py); the adjoint is checked
 * against central finite differences of this f64 forward.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ochre_oracle.h"

#define TILE 8
#define TOL 0.1
#define FLAT_LIMIT (1 << 16)
#define CONIC_DEPTH 32

typedef struct {
    double x, y;
} d2;

static d2 mk(double x, double y) { d2 r = {x, y}; return r; }
static d2 add(d2 a, d2 b) { return mk(a.x + b.x, a.y + b.y); }
static d2 sub(d2 a, d2 b) { return mk(a.x - b.x, a.y - b.y); }
static d2 smul(double s, d2 a) { return mk(s * a.x, s * a.y); }
static double dot(d2 a, d2 b) { return a.x * b.x + a.y * b.y; }
static double len(d2 a) { return sqrt(dot(a, a)); }
static d2 lerp(double t, d2 a, d2 b) { return add(smul(1.0 - t, a), smul(t, b)); }
static int veq(d2 a, d2 b) { return a.x == b.x && a.y == b.y; }
static int16_t f2i16(double v) {
    if (v != v) return 0;
    if (v >= 32767.0) return 32767;
    if (v <= -32768.0) return -32768;
    return (int16_t)v;
}
static int16_t sgn16(double v) {
    if (v != v) return 0;
    return signbit(v) ? -1 : 1;
}
static int16_t div8(int16_t v) { return (int16_t)(v >> 3); }

static int verb_np(int v) {
    switch (v) {
    case ORC_MOVE: case ORC_LINE: return 1;
    case ORC_QUAD: case ORC_CONIC: return 2;
    case ORC_CUBIC: return 3;
    case ORC_CLOSE: return 0;
    default: return -1;
    }
}

/* ------------------------------------------------------- point sources
 * Every line end point is a function of the path's world points:
 *   SRC_ORIGIN  (0, 0)
 *   SRC_POINT   world point g
 *   SRC_QUAD    quadratic of command points (s, c, e) at t = k * dt
 *   SRC_CUBIC   cubic (s, c1, c2, e) at t = k * dt
 *   SRC_CONIC   conic (s, c, e, weight) at a fixed dyadic t
 * where s is the start point (a world point index or -1 for the origin). */
enum { SRC_ORIGIN = 0, SRC_POINT = 1, SRC_QUAD = 2, SRC_CUBIC = 3, SRC_CONIC = 4 };

typedef struct {
    int type;
    int64_t s, c1, c2, e;  /* world point indices (-1: origin) */
    int k;                 /* step index (t = k * dt) */
    double t;              /* parameter actually used */
    double w;              /* conic weight */
} source;

typedef struct {
    d2 p0, p1;
    source a, b;
} line;

typedef struct {
    line* v;
    size_t n, cap;
    int status;
} lines;

static void push_line(lines* L, d2 p0, d2 p1, source a, source b) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 256;
        L->v = (line*)realloc(L->v, L->cap * sizeof(line));
        if (!L->v) abort();
    }
    line* l = &L->v[L->n++];
    l->p0 = p0;
    l->p1 = p1;
    l->a = a;
    l->b = b;
}

static source src_point(int64_t g) {
    source s;
    memset(&s, 0, sizeof(s));
    s.type = g < 0 ? SRC_ORIGIN : SRC_POINT;
    s.s = g;
    return s;
}

static d2 wpt(const d2* W, int64_t g) { return g < 0 ? mk(0, 0) : W[g]; }

/* f32 mirror for the discrete decisions (path.rs:49-104 in f32) */
typedef struct {
    float x, y;
} f2;
static f2 fmk(float x, float y) { f2 r = {x, y}; return r; }
static f2 fadd(f2 a, f2 b) { return fmk(a.x + b.x, a.y + b.y); }
static f2 fsub(f2 a, f2 b) { return fmk(a.x - b.x, a.y - b.y); }
static f2 fsmul(float s, f2 a) { return fmk(s * a.x, s * a.y); }
static float flen(f2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static f2 flerp(float t, f2 a, f2 b) { return fadd(fsmul(1.0f - t, a), fsmul(t, b)); }
static f2 wpt32(const f2* W, int64_t g) { return g < 0 ? fmk(0.0f, 0.0f) : W[g]; }
static float quad_dt32(f2 s, f2 c, f2 e) { return sqrtf((4.0f * 0.1f) / flen(fadd(fsub(s, fsmul(2.0f, c)), e))); }
static float cubic_dt32(f2 s, f2 c1, f2 c2, f2 e) {
    f2 a = fadd(fsub(fadd(fsmul(-1.0f, s), fsmul(3.0f, c1)), fsmul(3.0f, c2)), e);
    f2 b = fsmul(3.0f, fadd(fsub(s, fsmul(2.0f, c1)), c2));
    float conc = fmaxf(flen(b), flen(fadd(a, b)));
    return sqrtf((sqrtf(8.0f) * 0.1f) / conc);
}
/* Number of segments of the reference's loop `while t < 1 { t = min(t + dt, 1) }`
 * in f32 (0 when it would exceed the flattening limit). */
static int steps32(float dt) {
    float t = 0.0f;
    int n = 0;
    while (t < 1.0f) {
        t = fminf(t + dt, 1.0f);
        if (++n >= FLAT_LIMIT) return 0;
    }
    return n;
}
static f2 conic_at32(float t, f2 s, f2 c, f2 e, float w) {
    f2 wc = fsmul(w, c);
    f2 p01 = flerp(t, s, wc), p12 = flerp(t, wc, e);
    float den = (1.0f - t) * (1.0f - t) + 2.0f * t * (1.0f - t) * w + t * t;
    return fsmul(1.0f / den, flerp(t, p01, p12));
}

static double quad_dt(d2 s, d2 c, d2 e) { return sqrt((4.0 * TOL) / len(add(sub(s, smul(2.0, c)), e))); }
static double cubic_dt(d2 s, d2 c1, d2 c2, d2 e) {
    d2 a = add(sub(add(smul(-1.0, s), smul(3.0, c1)), smul(3.0, c2)), e);
    d2 b = smul(3.0, add(sub(s, smul(2.0, c1)), c2));
    double lb = len(b), lab = len(add(a, b));
    return sqrt((sqrt(8.0) * TOL) / (lb > lab ? lb : lab));
}
static d2 quad_at(double t, d2 s, d2 c, d2 e) { return lerp(t, lerp(t, s, c), lerp(t, c, e)); }
static d2 cubic_at(double t, d2 s, d2 c1, d2 c2, d2 e) {
    d2 p01 = lerp(t, s, c1), p12 = lerp(t, c1, c2), p23 = lerp(t, c2, e);
    return lerp(t, lerp(t, p01, p12), lerp(t, p12, p23));
}
static d2 conic_at(double t, d2 s, d2 c, d2 e, double w) {
    d2 wc = smul(w, c);
    d2 p = lerp(t, lerp(t, s, wc), lerp(t, wc, e));
    double den = (1.0 - t) * (1.0 - t) + 2.0 * t * (1.0 - t) * w + t * t;
    return smul(1.0 / den, p);
}

typedef struct {
    lines* L;
    d2 prev;
    source prev_src;
    int64_t s, c, e;
    double w;
    const f2* W32;
    int depth_err;
} conic_state;

/* flatten_conic (path.rs:75-104): subdivision decided in f32 (p0f, p1f are
 * the f32 chord ends), points evaluated in f64 */
static void conic_rec(conic_state* st, const d2* W, double t0, double t1, d2 p0, d2 p1, f2 p0f, f2 p1f,
                      int depth) {
    double t = 0.5 * (t0 + t1);
    d2 S = wpt(W, st->s), C = W[st->c], E = W[st->e];
    d2 mid = conic_at(t, S, C, E, st->w);
    const f2 midf = conic_at32((float)t, wpt32(st->W32, st->s), st->W32[st->c], st->W32[st->e], (float)st->w);
    const float err = flen(fsub(midf, fsmul(0.5f, fadd(p0f, p1f))));
    if (err > 0.1f && depth < CONIC_DEPTH) {
        conic_rec(st, W, t0, t, p0, mid, p0f, midf, depth + 1);
        conic_rec(st, W, t, t1, mid, p1, midf, p1f, depth + 1);
    } else {
        if (err > 0.1f) st->depth_err = 1;
        source sm;
        memset(&sm, 0, sizeof(sm));
        sm.type = SRC_CONIC;
        sm.s = st->s;
        sm.c1 = st->c;
        sm.e = st->e;
        sm.t = t;
        sm.w = st->w;
        push_line(st->L, st->prev, mid, st->prev_src, sm);
        st->prev = mid;
        st->prev_src = sm;
        source se = src_point(st->e);
        push_line(st->L, st->prev, p1, st->prev_src, se);
        st->prev = p1;
        st->prev_src = se;
    }
}

/* Lines fed to line_to for one path (rasterizer.rs:61-69,145-165,181-183),
 * in order, with the sources of their end points.  World points W[g]. */
static int build_lines(const uint8_t* verbs, size_t nv, const d2* W, const f2* W32, size_t np,
                       const double* weights, size_t nw, lines* L) {
    int64_t g = 0, wi = 0;
    int64_t last = -1, first = -1; /* world point index of last / first (-1 origin) */
    d2 lastp = mk(0, 0), firstp = mk(0, 0);
    for (size_t i = 0; i < nv; i++) {
        int v = verbs[i];
        int k = verb_np(v);
        if (k < 0 || g + k > (int64_t)np) return ORC_ERR_ARG;
        switch (v) {
        case ORC_MOVE:
            if (!veq(lastp, firstp)) push_line(L, lastp, firstp, src_point(last), src_point(first));
            last = first = g;
            lastp = firstp = W[g];
            break;
        case ORC_LINE:
            if (!veq(W[g], lastp)) push_line(L, lastp, W[g], src_point(last), src_point(g));
            last = g;
            lastp = W[g];
            break;
        case ORC_QUAD: {
            d2 S = lastp, C = W[g], E = W[g + 1];
            const int n32 = steps32(quad_dt32(wpt32(W32, last), W32[g], W32[g + 1]));
            if (n32 == 0) { L->status |= ORC_ERR_FLATTEN_LIMIT; }
            double dt = quad_dt(S, C, E), t = 0.0;
            int kk = 0;
            d2 prev = S;
            source ps = src_point(last);
            while (kk < n32) {
                kk++;
                t = kk == n32 ? 1.0 : fmin((double)kk * dt, 1.0);
                d2 q = quad_at(t, S, C, E);
                source qs;
                memset(&qs, 0, sizeof(qs));
                if (t >= 1.0) {
                    qs = src_point(g + 1);
                } else {
                    qs.type = SRC_QUAD;
                    qs.s = last;
                    qs.c1 = g;
                    qs.e = g + 1;
                    qs.k = kk;
                    qs.t = t;
                }
                if (!veq(q, prev)) push_line(L, prev, q, ps, qs);
                prev = q;
                ps = qs;
                if (kk >= FLAT_LIMIT) { L->status |= ORC_ERR_FLATTEN_LIMIT; break; }
            }
            last = g + 1;
            lastp = W[g + 1];
            break;
        }
        case ORC_CUBIC: {
            d2 S = lastp, C1 = W[g], C2 = W[g + 1], E = W[g + 2];
            const int n32 = steps32(cubic_dt32(wpt32(W32, last), W32[g], W32[g + 1], W32[g + 2]));
            if (n32 == 0) { L->status |= ORC_ERR_FLATTEN_LIMIT; }
            double dt = cubic_dt(S, C1, C2, E), t = 0.0;
            int kk = 0;
            d2 prev = S;
            source ps = src_point(last);
            while (kk < n32) {
                kk++;
                t = kk == n32 ? 1.0 : fmin((double)kk * dt, 1.0);
                d2 q = cubic_at(t, S, C1, C2, E);
                source qs;
                memset(&qs, 0, sizeof(qs));
                if (t >= 1.0) {
                    qs = src_point(g + 2);
                } else {
                    qs.type = SRC_CUBIC;
                    qs.s = last;
                    qs.c1 = g;
                    qs.c2 = g + 1;
                    qs.e = g + 2;
                    qs.k = kk;
                    qs.t = t;
                }
                if (!veq(q, prev)) push_line(L, prev, q, ps, qs);
                prev = q;
                ps = qs;
                if (kk >= FLAT_LIMIT) { L->status |= ORC_ERR_FLATTEN_LIMIT; break; }
            }
            last = g + 2;
            lastp = W[g + 2];
            break;
        }
        case ORC_CONIC: {
            if (wi >= (int64_t)nw) return ORC_ERR_ARG;
            conic_state st;
            st.L = L;
            st.prev = lastp;
            st.prev_src = src_point(last);
            st.s = last;
            st.c = g;
            st.e = g + 1;
            st.w = weights[wi++];
            st.W32 = W32;
            st.depth_err = 0;
            conic_rec(&st, W, 0.0, 1.0, lastp, W[g + 1], wpt32(W32, last), W32[g + 1], 0);
            if (st.depth_err) L->status |= ORC_ERR_FLATTEN_LIMIT;
            last = g + 1;
            lastp = W[g + 1];
            break;
        }
        default:
            break; /* Close: a no-op */
        }
        g += k;
    }
    if (!veq(lastp, firstp)) push_line(L, lastp, firstp, src_point(last), src_point(first));
    return 0;
}

/* ------------------------------------------------------------------ DDA
 * One increment: pixel (x, y), the clip parameters t0 <= t1 and what bounds
 * them (0: line end, 1: a row boundary y = B, 2: a column boundary x = B). */
typedef struct {
    int16_t x, y;
    double t0, t1;
    int k0, k1;
    double b0, b1;
    double area, height;
    int virt;   /* a zero-length derivative carrier (see walk): no tile, no sums */
    int nograd; /* part of the forward, but no derivative (see walk) */
} incr;

typedef struct {
    incr* v;
    size_t n, cap;
} incrs;

typedef struct {
    int16_t tx, ty;
    int8_t sign;
} tinc;

typedef struct {
    tinc* v;
    size_t n, cap;
} tincs;

static void walk(d2 last, d2 point, incrs* I, tincs* T, int16_t* ty_prev) {
    if (veq(point, last)) return;
    int16_t xd = sgn16(point.x - last.x), yd = sgn16(point.y - last.y);
    double dtdx = 1.0 / (point.x - last.x), dtdy = 1.0 / (point.y - last.y);
    int16_t x = f2i16(floor(last.x)), y = f2i16(floor(last.y));
    double rt0 = 0, ct0 = 0, rt1, ct1;
    double rb1 = 0, cb1 = 0; /* boundary coordinates of the pending crossings */
    if (last.y == point.y) rt1 = INFINITY;
    else {
        double ny = point.y > last.y ? (double)(int16_t)(y + 1) : (double)y;
        rb1 = ny;
        rt1 = fmin(dtdy * (ny - last.y), 1.0);
    }
    if (last.x == point.x) ct1 = INFINITY;
    else {
        double nx = point.x > last.x ? (double)(int16_t)(x + 1) : (double)x;
        cb1 = nx;
        ct1 = fmin(dtdx * (nx - last.x), 1.0);
    }
    double xs = fabs(dtdx), ys = fabs(dtdy);
    int k_prev = 0;    /* kind of the bound of t0 */
    double b_prev = 0; /* its boundary coordinate */
    int nograd = 0;    /* set once the walk has passed the line's end (see below) */
    for (;;) {
        double t0 = fmax(rt0, ct0), t1 = fmin(rt1, ct1);
        int k1;
        double b1;
        if (rt1 < ct1) { k1 = rt1 >= 1.0 ? 0 : 1; b1 = rb1; }
        else { k1 = ct1 >= 1.0 ? 0 : 2; b1 = cb1; }
        d2 p0 = add(smul(1.0 - t0, last), smul(t0, point));
        d2 p1 = add(smul(1.0 - t1, last), smul(t1, point));
        double h = p1.y - p0.y;
        double right = (double)(int16_t)(x + 1);
        double area = 0.5 * h * ((right - p0.x) + (right - p1.x));
        if (I->n == I->cap) {
            I->cap = I->cap ? 2 * I->cap : 1024;
            I->v = (incr*)realloc(I->v, I->cap * sizeof(incr));
            if (!I->v) abort();
        }
        /* A line ending exactly on the next pixel boundary in its direction
         * of travel stops in the pixel before it, while the next line starts
         * beyond it (the floor of its start point).  Both lines take their
         * derivatives on the same side of that tie: this increment ends at
         * the boundary, and a zero-length increment in the pixel beyond
         * carries the end point's derivative.  (A corner tie keeps the end.) */
        int vx = 0, vy = 0, ends_here = 0;
        if (k1 == 0) {
            int on_y = yd > 0 && point.y == (double)(int16_t)(y + 1);
            int on_x = xd > 0 && point.x == (double)(int16_t)(x + 1);
            if (on_y != on_x) {
                if (on_y) { k1 = 1; b1 = (double)(int16_t)(y + 1); vy = 1; }
                else { k1 = 2; b1 = (double)(int16_t)(x + 1); vx = 1; }
            }
        } else if (k1 == 1 && yd < 0 && point.y == b1) {
            /* Going up onto a row boundary that holds the line's end point, the
             * crossing parameter can round below 1: the walk then steps into the
             * row above with a sliver, while the next line starts in this row.
             * On the same (floor) side, this increment ends at the end point and
             * the sliver carries no derivative. */
            k1 = 0;
            ends_here = 1;
        }
        incr* in = &I->v[I->n++];
        in->x = x;
        in->y = y;
        in->t0 = t0;
        in->t1 = t1;
        in->k0 = k_prev;
        in->b0 = b_prev;
        in->k1 = k1;
        in->b1 = b1;
        in->area = area;
        in->height = h;
        in->virt = 0;
        in->nograd = nograd;
        if (ends_here) nograd = 1;
        if ((vx || vy) && !nograd) {
            if (I->n == I->cap) {
                I->cap = 2 * I->cap;
                I->v = (incr*)realloc(I->v, I->cap * sizeof(incr));
                if (!I->v) abort();
            }
            incr* vi = &I->v[I->n++];
            vi->x = (int16_t)(x + vx);
            vi->y = (int16_t)(y + vy);
            vi->t0 = t1;
            vi->t1 = t1;
            vi->k0 = k1;
            vi->b0 = b1;
            vi->k1 = 0;
            vi->b1 = 0.0;
            vi->area = 0.0;
            vi->height = 0.0;
            vi->virt = 1;
            vi->nograd = 0;
        }
        if (rt1 < ct1) {
            rt0 = rt1;
            rt1 = fmin(rt1 + ys, 1.0);
            y = (int16_t)(y + yd);
            k_prev = k1;
            b_prev = rb1;
            rb1 += (double)yd;
        } else {
            ct0 = ct1;
            ct1 = fmin(ct1 + xs, 1.0);
            x = (int16_t)(x + xd);
            k_prev = k1;
            b_prev = cb1;
            cb1 += (double)xd;
        }
        if (rt0 == 1.0 || ct0 == 1.0) {
            x = f2i16(floor(point.x));
            y = f2i16(floor(point.y));
        }
        int16_t ty = div8(y);
        if (ty != *ty_prev) {
            if (T->n == T->cap) {
                T->cap = T->cap ? 2 * T->cap : 256;
                T->v = (tinc*)realloc(T->v, T->cap * sizeof(tinc));
                if (!T->v) abort();
            }
            T->v[T->n].tx = div8(x);
            T->v[T->n].ty = *ty_prev < ty ? *ty_prev : ty;
            T->v[T->n].sign = (int8_t)(int16_t)(ty - *ty_prev);
            T->n++;
            *ty_prev = ty;
        }
        if (rt0 == 1.0 || ct0 == 1.0) break;
    }
}

/* --------------------------------------------------------- tile model */

typedef struct {
    int16_t tx, ty;
    double a[64], h[64];  /* accumulated area / height */
    double v[64];         /* accum + area */
    double g[64], G[64];  /* dL/dv and dL/dheight */
} tile;

typedef struct {
    tile* t;
    size_t n;
    int16_t* span; /* (x, y, w) triples */
    size_t ns;
} render;

static int tile_cmp(const void* a, const void* b) {
    const tile* x = (const tile*)a;
    const tile* y = (const tile*)b;
    if (x->ty != y->ty) return x->ty < y->ty ? -1 : 1;
    if (x->tx != y->tx) return x->tx < y->tx ? -1 : 1;
    return 0;
}

static int tinc_cmp(const void* a, const void* b) {
    const tinc* x = (const tinc*)a;
    const tinc* y = (const tinc*)b;
    if (x->ty != y->ty) return x->ty < y->ty ? -1 : 1;
    if (x->tx != y->tx) return x->tx < y->tx ? -1 : 1;
    return 0;
}

static long find_tile(const render* R, int16_t tx, int16_t ty) {
    long lo = 0, hi = (long)R->n - 1;
    while (lo <= hi) {
        long m = (lo + hi) / 2;
        const tile* t = &R->t[m];
        if (t->ty == ty && t->tx == tx) return m;
        if (t->ty < ty || (t->ty == ty && t->tx < tx)) lo = m + 1;
        else hi = m - 1;
    }
    return -1;
}

/* Builds the tiles (in row-major order, increments summed in emission
 * order) and spans of one path from its increments. */
static void build_render(const incrs* I, tincs* T, render* R) {
    memset(R, 0, sizeof(*R));
    /* distinct tiles */
    size_t cap = 64;
    R->t = (tile*)malloc(cap * sizeof(tile));
    for (size_t i = 0; i < I->n; i++) {
        if (I->v[i].virt) continue;
        int16_t tx = div8(I->v[i].x), ty = div8(I->v[i].y);
        int seen = 0;
        for (size_t k = R->n; k-- > 0 && k + 64 >= R->n;)
            if (R->t[k].tx == tx && R->t[k].ty == ty) { seen = 1; break; }
        if (!seen) {
            /* full search (paths are small) */
            for (size_t k = 0; k < R->n; k++)
                if (R->t[k].tx == tx && R->t[k].ty == ty) { seen = 1; break; }
        }
        if (!seen) {
            if (R->n == cap) { cap *= 2; R->t = (tile*)realloc(R->t, cap * sizeof(tile)); }
            memset(&R->t[R->n], 0, sizeof(tile));
            R->t[R->n].tx = tx;
            R->t[R->n].ty = ty;
            R->n++;
        }
    }
    if (R->n == 0) { /* rasterizer.rs:193-198: one zero tile at (0, 0) */
        memset(&R->t[0], 0, sizeof(tile));
        R->n = 1;
        return;
    }
    qsort(R->t, R->n, sizeof(tile), tile_cmp);
    for (size_t i = 0; i < I->n; i++) {
        if (I->v[i].virt) continue;
        long k = find_tile(R, div8(I->v[i].x), div8(I->v[i].y));
        int p = (I->v[i].y & 7) * 8 + (I->v[i].x & 7);
        R->t[k].a[p] += I->v[i].area;
        R->t[k].h[p] += I->v[i].height;
    }
    /* carry along tile rows (rasterizer.rs:233-250) */
    double carry[8] = {0};
    for (size_t k = 0; k < R->n; k++) {
        if (k == 0 || R->t[k].ty != R->t[k - 1].ty) memset(carry, 0, sizeof(carry));
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) {
                R->t[k].v[y * 8 + x] = carry[y] + R->t[k].a[y * 8 + x];
                carry[y] += R->t[k].h[y * 8 + x];
            }
    }
    /* spans (rasterizer.rs:252-264) */
    if (T->n > 1) qsort(T->v, T->n, sizeof(tinc), tinc_cmp);
    size_t ti = 0, cs = 0;
    long wnd = 0;
    for (size_t k = 0; k + 1 < R->n; k++) {
        const tile* a = &R->t[k];
        const tile* b = &R->t[k + 1];
        if (b->ty == a->ty && b->tx > a->tx + 1) {
            while (ti < T->n && (T->v[ti].ty < a->ty || (T->v[ti].ty == a->ty && T->v[ti].tx <= a->tx)))
                wnd += T->v[ti++].sign;
            if (wnd != 0) {
                if (R->ns == cs) {
                    cs = cs ? 2 * cs : 16;
                    R->span = (int16_t*)realloc(R->span, cs * 3 * sizeof(int16_t));
                }
                R->span[3 * R->ns] = (int16_t)((a->tx + 1) * TILE);
                R->span[3 * R->ns + 1] = (int16_t)(a->ty * TILE);
                R->span[3 * R->ns + 2] = (int16_t)((b->tx - a->tx - 1) * TILE);
                R->ns++;
            }
        }
    }
}

static void free_render(render* R) {
    free(R->t);
    free(R->span);
    memset(R, 0, sizeof(*R));
}

/* Coverage of render R at pixel (px, py): a tile pixel, 1 inside a span,
 * else 0.  *in_tile reports whether the pixel is a tile pixel. */
static double coverage_at(const render* R, int px, int py, long* tile_idx) {
    long k = find_tile(R, (int16_t)(px >> 3), (int16_t)(py >> 3));
    if (tile_idx) *tile_idx = k;
    if (k >= 0) return fmin(fabs(R->t[k].v[(py & 7) * 8 + (px & 7)]), 1.0);
    for (size_t s = 0; s < R->ns; s++) {
        int x = R->span[3 * s], y = R->span[3 * s + 1], w = (uint16_t)R->span[3 * s + 2];
        if (py >= y && py < y + 8 && px >= x && px < x + w) return 1.0;
    }
    return 0.0;
}

/* Target render given as tiles (f32 coverage) + spans. */
typedef struct {
    const int16_t* txy;
    const float* cov;
    size_t nt;
    const int16_t* sxy;
    const uint16_t* sw;
    size_t ns;
} target;

static double target_at(const target* T, int px, int py, int* in_tile) {
    int16_t x = (int16_t)(px & ~7), y = (int16_t)(py & ~7);
    for (size_t k = 0; k < T->nt; k++)
        if (T->txy[2 * k] == x && T->txy[2 * k + 1] == y) {
            if (in_tile) *in_tile = 1;
            return T->cov[64 * k + (py & 7) * 8 + (px & 7)];
        }
    if (in_tile) *in_tile = 0;
    for (size_t s = 0; s < T->ns; s++) {
        int sx = T->sxy[2 * s], sy = T->sxy[2 * s + 1], w = T->sw[s];
        if (py >= sy && py < sy + 8 && px >= sx && px < sx + w) return 1.0;
    }
    return 0.0;
}

/* --------------------------------------------------- derivative helpers
 * d(clip point)/d(line end points): returns J[2][4] with columns
 * (x0, y0, x1, y1). */
static void clip_jacobian(int kind, double t, double b, d2 P0, d2 P1, double J[2][4]) {
    memset(J, 0, sizeof(double) * 8);
    if (kind == 0) {
        if (t <= 0.5) { J[0][0] = 1; J[1][1] = 1; }   /* t == 0: the start */
        else { J[0][2] = 1; J[1][3] = 1; }            /* t == 1: the end */
        return;
    }
    double dx = P1.x - P0.x, dy = P1.y - P0.y;
    if (kind == 1) { /* y = b: t = (b - y0)/dy, x = x0 + t dx */
        double tt = (b - P0.y) / dy;
        J[0][0] = 1.0 - tt;
        J[0][2] = tt;
        J[0][1] = dx * (tt - 1.0) / dy;
        J[0][3] = -dx * tt / dy;
    } else { /* x = b: t = (b - x0)/dx, y = y0 + t dy */
        double tt = (b - P0.x) / dx;
        J[1][1] = 1.0 - tt;
        J[1][3] = tt;
        J[1][0] = dy * (tt - 1.0) / dx;
        J[1][2] = -dy * tt / dx;
    }
}

/* Accumulates dL/d(world points) for a line end point with gradient gp. */
static void push_source(const source* s, d2 gp, const d2* W, d2* dW) {
    if (s->type == SRC_ORIGIN) return;
    if (s->type == SRC_POINT) { dW[s->s].x += gp.x; dW[s->s].y += gp.y; return; }
    d2 S = wpt(W, s->s);
    double t = s->t;
    if (s->type == SRC_QUAD) {
        d2 C = W[s->c1], E = W[s->e];
        double bs = (1 - t) * (1 - t), bc = 2 * t * (1 - t), be = t * t;
        if (s->s >= 0) { dW[s->s].x += bs * gp.x; dW[s->s].y += bs * gp.y; }
        dW[s->c1].x += bc * gp.x; dW[s->c1].y += bc * gp.y;
        dW[s->e].x += be * gp.x; dW[s->e].y += be * gp.y;
        /* through t = k * dt, dt = sqrt(0.4 / |D|), D = S - 2C + E */
        d2 dP = add(smul(2 * (1 - t), sub(C, S)), smul(2 * t, sub(E, C)));
        d2 D = add(sub(S, smul(2, C)), E);
        double dt = sqrt(0.4 / len(D));
        double coef = dot(gp, dP) * (double)s->k * (-0.5 * dt / dot(D, D));
        /* d dt = coef_dt * D . dD with dD = dS - 2 dC + dE */
        if (s->s >= 0) { dW[s->s].x += coef * D.x; dW[s->s].y += coef * D.y; }
        dW[s->c1].x += -2 * coef * D.x; dW[s->c1].y += -2 * coef * D.y;
        dW[s->e].x += coef * D.x; dW[s->e].y += coef * D.y;
        return;
    }
    if (s->type == SRC_CUBIC) {
        d2 C1 = W[s->c1], C2 = W[s->c2], E = W[s->e];
        double u = 1 - t;
        double bs = u * u * u, b1 = 3 * t * u * u, b2 = 3 * t * t * u, be = t * t * t;
        if (s->s >= 0) { dW[s->s].x += bs * gp.x; dW[s->s].y += bs * gp.y; }
        dW[s->c1].x += b1 * gp.x; dW[s->c1].y += b1 * gp.y;
        dW[s->c2].x += b2 * gp.x; dW[s->c2].y += b2 * gp.y;
        dW[s->e].x += be * gp.x; dW[s->e].y += be * gp.y;
        d2 dP = add(add(smul(3 * u * u, sub(C1, S)), smul(6 * t * u, sub(C2, C1))), smul(3 * t * t, sub(E, C2)));
        d2 a = add(sub(add(smul(-1, S), smul(3, C1)), smul(3, C2)), E);
        d2 b = smul(3, add(sub(S, smul(2, C1)), C2));
        d2 ab = add(a, b);
        double lb = len(b), lab = len(ab);
        double conc = lb > lab ? lb : lab;
        double dt = sqrt(sqrt(8.0) * TOL / conc);
        double coef = dot(gp, dP) * (double)s->k * (-0.5 * dt / conc);
        /* d conc = (v . dv)/|v| for the larger of b = 3(S - 2C1 + C2) and
         * a + b = 2S - 3C1 + E */
        if (lb > lab) {
            d2 n = smul(coef / lb, b);
            if (s->s >= 0) { dW[s->s].x += 3 * n.x; dW[s->s].y += 3 * n.y; }
            dW[s->c1].x += -6 * n.x; dW[s->c1].y += -6 * n.y;
            dW[s->c2].x += 3 * n.x; dW[s->c2].y += 3 * n.y;
        } else {
            d2 n = smul(coef / lab, ab);
            if (s->s >= 0) { dW[s->s].x += 2 * n.x; dW[s->s].y += 2 * n.y; }
            dW[s->c1].x += -3 * n.x; dW[s->c1].y += -3 * n.y;
            dW[s->e].x += n.x; dW[s->e].y += n.y;
        }
        return;
    }
    if (s->type == SRC_CONIC) {
        double u = 1 - t, w = s->w;
        double den = u * u + 2 * t * u * w + t * t;
        double bs = u * u / den, bc = 2 * t * u * w / den, be = t * t / den;
        if (s->s >= 0) { dW[s->s].x += bs * gp.x; dW[s->s].y += bs * gp.y; }
        dW[s->c1].x += bc * gp.x; dW[s->c1].y += bc * gp.y;
        dW[s->e].x += be * gp.x; dW[s->e].y += be * gp.y;
        (void)S;
    }
}

/* -------------------------------------------------------------- driver */

/* Loss (and, if dpts/dxf are non-null, its gradient) of one path.
 * points: path-space control points (pairs), xf: transform (6). */
int orc64_path_loss(const uint8_t* verbs, size_t nv, const double* points, size_t np, const double* weights,
                    size_t nw, const double* xf, const int16_t* t_xy, const float* t_cov, size_t t_nt,
                    const int16_t* t_sxy, const uint16_t* t_sw, size_t t_ns, double* loss, double* dpts,
                    double* dxf, int16_t* out_xy, double* out_cov, size_t out_cap, size_t* out_nt) {
    d2* W = (d2*)malloc((np + 1) * sizeof(d2));
    f2* W32 = (f2*)malloc((np + 1) * sizeof(f2));
    const float X32[6] = {(float)xf[0], (float)xf[1], (float)xf[2], (float)xf[3], (float)xf[4], (float)xf[5]};
    for (size_t i = 0; i < np; i++) {
        double x = points[2 * i], y = points[2 * i + 1];
        W[i] = add(mk(xf[0] * x + xf[1] * y, xf[2] * x + xf[3] * y), mk(xf[4], xf[5]));
        /* Transform::apply in f32 (geom.rs:170-179,260-262) */
        const float xs = (float)x, ys = (float)y;
        W32[i] = fadd(fmk(X32[0] * xs + X32[1] * ys, X32[2] * xs + X32[3] * ys), fmk(X32[4], X32[5]));
    }
    lines L;
    memset(&L, 0, sizeof(L));
    int st = build_lines(verbs, nv, W, W32, np, weights, nw, &L);
    free(W32);
    if (st) { free(W); free(L.v); return st; }
    incrs I;
    memset(&I, 0, sizeof(I));
    tincs T;
    memset(&T, 0, sizeof(T));
    size_t* line_of = NULL;
    size_t cap_lo = 0;
    int16_t ty_prev = 0;
    /* the chain of tile rows restarts at each Move (rasterizer.rs:68);
     * a line's start row is always its start pixel's row */
    for (size_t l = 0; l < L.n; l++) {
        size_t before = I.n;
        ty_prev = div8(f2i16(floor(L.v[l].p0.y)));
        walk(L.v[l].p0, L.v[l].p1, &I, &T, &ty_prev);
        if (I.n > cap_lo) {
            cap_lo = I.cap;
            line_of = (size_t*)realloc(line_of, cap_lo * sizeof(size_t));
        }
        for (size_t k = before; k < I.n; k++) line_of[k] = l;
    }
    render R;
    build_render(&I, &T, &R);
    target TG = {t_xy, t_cov, t_nt, t_sxy, t_sw, t_ns};
    /* loss over the union of tile pixels */
    double Ls = 0.0;
    for (size_t k = 0; k < R.n; k++)
        for (int p = 0; p < 64; p++) {
            int px = R.t[k].tx * 8 + (p & 7), py = R.t[k].ty * 8 + (p >> 3);
            double v = R.t[k].v[p];
            double cov = fmin(fabs(v), 1.0);
            double tv = target_at(&TG, px, py, NULL);
            Ls += (cov - tv) * (cov - tv);
            R.t[k].g[p] = fabs(v) < 1.0 ? 2.0 * (cov - tv) * (v < 0 ? -1.0 : 1.0) : 0.0;
        }
    for (size_t k = 0; k < t_nt; k++) {
        int16_t x = t_xy[2 * k], y = t_xy[2 * k + 1];
        if (find_tile(&R, (int16_t)(x >> 3), (int16_t)(y >> 3)) >= 0) continue;
        for (int p = 0; p < 64; p++) {
            double cf = coverage_at(&R, x + (p & 7), y + (p >> 3), NULL);
            double tv = t_cov[64 * k + p];
            Ls += (cf - tv) * (cf - tv);
        }
    }
    if (loss) *loss = Ls;
    if (out_nt) *out_nt = R.n;
    if (out_xy && out_cov)
        for (size_t k = 0; k < R.n && k < out_cap; k++) {
            out_xy[2 * k] = (int16_t)(R.t[k].tx * 8);
            out_xy[2 * k + 1] = (int16_t)(R.t[k].ty * 8);
            for (int p = 0; p < 64; p++) out_cov[64 * k + p] = fmin(fabs(R.t[k].v[p]), 1.0);
        }
    if (dpts || dxf) {
        /* G: sum of g strictly right of each pixel along its tile row */
        for (size_t k = R.n; k-- > 0;) {
            int row_end = (k + 1 == R.n) || R.t[k + 1].ty != R.t[k].ty;
            for (int y = 0; y < 8; y++) {
                double S = row_end ? 0.0 : R.t[k + 1].G[y * 8] + R.t[k + 1].g[y * 8];
                for (int x = 7; x >= 0; x--) {
                    R.t[k].G[y * 8 + x] = S;
                    S += R.t[k].g[y * 8 + x];
                }
            }
        }
        d2* dW = (d2*)calloc(np + 1, sizeof(d2));
        for (size_t i = 0; i < I.n; i++) {
            const incr* in = &I.v[i];
            if (in->nograd) continue;
            const line* ln = &L.v[line_of[i]];
            long k = find_tile(&R, div8(in->x), div8(in->y));
            if (k < 0) continue; /* a zero-length carrier outside the path's tiles */
            int p = (in->y & 7) * 8 + (in->x & 7);
            double gA = R.t[k].g[p], gH = R.t[k].G[p];
            if (gA == 0.0 && gH == 0.0) continue;
            double J0[2][4], J1[2][4];
            clip_jacobian(in->k0, in->t0, in->b0, ln->p0, ln->p1, J0);
            clip_jacobian(in->k1, in->t1, in->b1, ln->p0, ln->p1, J1);
            d2 P0 = add(smul(1.0 - in->t0, ln->p0), smul(in->t0, ln->p1));
            d2 P1 = add(smul(1.0 - in->t1, ln->p0), smul(in->t1, ln->p1));
            double h = P1.y - P0.y;
            double right = (double)(int16_t)(in->x + 1);
            double S = (right - P0.x) + (right - P1.x);
            double d[4];
            for (int j = 0; j < 4; j++) {
                double dh = J1[1][j] - J0[1][j];
                double da = 0.5 * dh * S + 0.5 * h * (-J0[0][j] - J1[0][j]);
                d[j] = gA * da + gH * dh;
            }
            push_source(&ln->a, mk(d[0], d[1]), W, dW);
            push_source(&ln->b, mk(d[2], d[3]), W, dW);
        }
        /* world = M * local + offset */
        if (dxf) memset(dxf, 0, 6 * sizeof(double));
        for (size_t i = 0; i < np; i++) {
            double lx = points[2 * i], ly = points[2 * i + 1];
            if (dpts) {
                dpts[2 * i] = xf[0] * dW[i].x + xf[2] * dW[i].y;
                dpts[2 * i + 1] = xf[1] * dW[i].x + xf[3] * dW[i].y;
            }
            if (dxf) {
                dxf[0] += dW[i].x * lx;
                dxf[1] += dW[i].x * ly;
                dxf[2] += dW[i].y * lx;
                dxf[3] += dW[i].y * ly;
                dxf[4] += dW[i].x;
                dxf[5] += dW[i].y;
            }
        }
        free(dW);
    }
    free_render(&R);
    free(I.v);
    free(T.v);
    free(line_of);
    free(L.v);
    free(W);
    return L.status;
}
