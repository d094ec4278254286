/*
 * ochre_oracle.c -- sequential CPU restatement of the reference rasteriser.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped or
 * measured as the product).  Citations are to the reference crate
 * glowcoil/ochre v0.2.0: src/geom.rs, src/path.rs, src/rasterizer.rs.
 *
 * Build with -ffp-contract=off and without -ffast-math: Rust never fuses
 * a*b+c, and its f32 '/' and sqrt are IEEE correctly rounded.
 */
#include "ochre_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TILE 8                 /* rasterizer.rs:4 */
#define TOLERANCE 0.1f         /* rasterizer.rs:6 */
#define FLATTEN_LIMIT (1 << 16) /* guard: the reference loops forever when dt == 0 */
#define CONIC_DEPTH_LIMIT 32    /* guard: the reference recurses without bound */
/* guard: increments per line_to.  A line whose crossing step is 0 (a finite
 * end point joined to an infinite one, e.g. the close of M(inf,3) L(40,5)
 * L(20,40)) never reaches t = 1 in the reference; the device walks give up
 * after the same number of steps with the same status. */
#define WALK_LIMIT (1u << 20)

/* ------------------------------------------------------------------ geom.rs */

typedef struct {
    float x, y;
} v2;

static inline v2 mk(float x, float y) { v2 r = {x, y}; return r; }
static inline v2 add(v2 a, v2 b) { return mk(a.x + b.x, a.y + b.y); }      /* geom.rs:73-82 */
static inline v2 sub(v2 a, v2 b) { return mk(a.x - b.x, a.y - b.y); }      /* geom.rs:91-100 */
static inline v2 smul(float s, v2 a) { return mk(s * a.x, s * a.y); }      /* geom.rs:120-129 */
static inline float dot(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }      /* geom.rs:19-21 */
static inline float length(v2 a) { return sqrtf(dot(a, a)); }              /* geom.rs:38-40 */
/* Vec2::lerp (geom.rs:50-52): (1 - t) * a + t * b */
static inline v2 lerp(float t, v2 a, v2 b) { return add(smul(1.0f - t, a), smul(t, b)); }
/* derived PartialEq (geom.rs:4): x == x && y == y */
static inline int veq(v2 a, v2 b) { return a.x == b.x && a.y == b.y; }
/* Transform::apply (geom.rs:260-262) with Mat2x2 * Vec2 (geom.rs:170-179) */
static inline v2 apply(const float* xf, v2 p) {
    v2 m = mk(xf[0] * p.x + xf[1] * p.y, xf[2] * p.x + xf[3] * p.y);
    return add(m, mk(xf[4], xf[5]));
}
/* Rust f32::min / f32::max ignore a NaN operand. */
static inline float rmin(float a, float b) { return fminf(a, b); }
static inline float rmax(float a, float b) { return fmaxf(a, b); }
/* `f as i16`: saturating, NaN -> 0 (Rust >= 1.45). */
static inline int16_t f2i16(float v) {
    if (v != v) return 0;
    if (v >= 32767.0f) return 32767;
    if (v <= -32768.0f) return -32768;
    return (int16_t)v;
}
/* `f.signum() as i16`: +1 for +0/+x, -1 for -0/-x, NaN -> 0. */
static inline int16_t signum_i16(float v) {
    if (v != v) return 0;
    return signbit(v) ? -1 : 1;
}
/* i16::wrapping_div_euclid(8): floor division. */
static inline int16_t div8(int16_t v) { return (int16_t)(v >> 3); }

/* -------------------------------------------------------- dynamic arrays */

typedef struct {
    int16_t x, y;
    float area, height;
} increment;  /* rasterizer.rs:25-30 */

typedef struct {
    int16_t tile_x, tile_y;
    int8_t sign;
} tile_increment;  /* rasterizer.rs:33-37 */

#define GROW(ptr, n, cap, T)                                              \
    do {                                                                  \
        if ((n) >= (cap)) {                                               \
            size_t nc = (cap) ? (cap) * 2 : 64;                           \
            while (nc <= (n)) nc *= 2;                                    \
            T* np = (T*)realloc((ptr), nc * sizeof(T));                   \
            if (!np) abort();                                             \
            (ptr) = np;                                                   \
            (cap) = nc;                                                   \
        }                                                                 \
    } while (0)

struct orc_rasterizer {
    increment* inc;
    size_t n_inc, cap_inc;
    tile_increment* tinc;
    size_t n_tinc, cap_tinc;
    v2 first, last;
    int16_t tile_y_prev;
    int64_t n_lines;
    int status;
};

/* Rasterizer::new (rasterizer.rs:50-58) */
orc_rasterizer* orc_new(void) {
    orc_rasterizer* r = (orc_rasterizer*)calloc(1, sizeof(orc_rasterizer));
    if (!r) abort();
    r->first = mk(0.0f, 0.0f);
    r->last = mk(0.0f, 0.0f);
    r->tile_y_prev = 0;
    return r;
}

void orc_free(orc_rasterizer* r) {
    if (!r) return;
    free(r->inc);
    free(r->tinc);
    free(r);
}

/* Rasterizer::line_to (rasterizer.rs:72-140) */
void orc_line_to(orc_rasterizer* r, float px, float py) {
    v2 point = mk(px, py);
    v2 last = r->last;
    if (!veq(point, last)) {
        r->n_lines++;
        int16_t x_dir = signum_i16(point.x - last.x);
        int16_t y_dir = signum_i16(point.y - last.y);
        float dtdx = 1.0f / (point.x - last.x);
        float dtdy = 1.0f / (point.y - last.y);
        int16_t x = f2i16(floorf(last.x));
        int16_t y = f2i16(floorf(last.y));
        float row_t0 = 0.0f;
        float col_t0 = 0.0f;
        float row_t1, col_t1;
        if (last.y == point.y) {
            row_t1 = INFINITY;
        } else {
            float next_y = point.y > last.y ? (float)(int16_t)(y + 1) : (float)y;
            row_t1 = rmin(dtdy * (next_y - last.y), 1.0f);
        }
        if (last.x == point.x) {
            col_t1 = INFINITY;
        } else {
            float next_x = point.x > last.x ? (float)(int16_t)(x + 1) : (float)x;
            col_t1 = rmin(dtdx * (next_x - last.x), 1.0f);
        }
        float x_step = fabsf(dtdx);
        float y_step = fabsf(dtdy);
        uint32_t steps = 0;

        for (;;) {
            float t0 = rmax(row_t0, col_t0);
            float t1 = rmin(row_t1, col_t1);
            v2 p0 = add(smul(1.0f - t0, last), smul(t0, point));
            v2 p1 = add(smul(1.0f - t1, last), smul(t1, point));
            float height = p1.y - p0.y;
            float right = (float)(int16_t)(x + 1);
            float area = 0.5f * height * ((right - p0.x) + (right - p1.x));

            GROW(r->inc, r->n_inc, r->cap_inc, increment);
            r->inc[r->n_inc].x = x;
            r->inc[r->n_inc].y = y;
            r->inc[r->n_inc].area = area;
            r->inc[r->n_inc].height = height;
            r->n_inc++;

            if (row_t1 < col_t1) {
                row_t0 = row_t1;
                row_t1 = rmin(row_t1 + y_step, 1.0f);
                y = (int16_t)(y + y_dir);
            } else {
                col_t0 = col_t1;
                col_t1 = rmin(col_t1 + x_step, 1.0f);
                x = (int16_t)(x + x_dir);
            }

            if (row_t0 == 1.0f || col_t0 == 1.0f) {
                x = f2i16(floorf(point.x));
                y = f2i16(floorf(point.y));
            }

            int16_t tile_y = div8(y);
            if (tile_y != r->tile_y_prev) {
                GROW(r->tinc, r->n_tinc, r->cap_tinc, tile_increment);
                r->tinc[r->n_tinc].tile_x = div8(x);
                r->tinc[r->n_tinc].tile_y = r->tile_y_prev < tile_y ? r->tile_y_prev : tile_y;
                r->tinc[r->n_tinc].sign = (int8_t)(int16_t)(tile_y - r->tile_y_prev);
                r->n_tinc++;
                r->tile_y_prev = tile_y;
            }

            if (row_t0 == 1.0f || col_t0 == 1.0f) break;
            if (++steps >= WALK_LIMIT) {
                r->status |= ORC_ERR_FLATTEN_LIMIT;
                break;
            }
        }
    }
    r->last = point;
}

/* Rasterizer::move_to (rasterizer.rs:61-69) */
void orc_move_to(orc_rasterizer* r, float px, float py) {
    if (!veq(r->last, r->first)) orc_line_to(r, r->first.x, r->first.y);
    r->first = mk(px, py);
    r->last = mk(px, py);
    r->tile_y_prev = div8(f2i16(floorf(py)));
}

/* ----------------------------------------------------------------- path.rs */

/* The flatten callback (path.rs:41): receives Move/Line/Close outputs. */
typedef void (*emit_fn)(void* ctx, int verb, v2 p);

typedef struct {
    v2 last, control, point;
    float weight, tolerance;
    emit_fn emit;
    void* ctx;
    int* status;
} conic_ctx;

/* flatten_conic (path.rs:76-101), recursive midpoint subdivision. */
static void flatten_conic(const conic_ctx* c, float t0, float t1, v2 p0, v2 p1, int depth) {
    float t = 0.5f * (t0 + t1);
    v2 wc = smul(c->weight, c->control);
    v2 p01 = lerp(t, c->last, wc);
    v2 p12 = lerp(t, wc, c->point);
    float denom = (1.0f - t) * (1.0f - t) + 2.0f * t * (1.0f - t) * c->weight + t * t;
    v2 midpoint = smul(1.0f / denom, lerp(t, p01, p12));
    float err = length(sub(midpoint, smul(0.5f, add(p0, p1))));
    if (err > c->tolerance && depth < CONIC_DEPTH_LIMIT) {
        flatten_conic(c, t0, t, p0, midpoint, depth + 1);
        flatten_conic(c, t, t1, midpoint, p1, depth + 1);
    } else {
        if (err > c->tolerance) *c->status |= ORC_ERR_FLATTEN_LIMIT;
        c->emit(c->ctx, ORC_LINE, midpoint);
        c->emit(c->ctx, ORC_LINE, p1);
    }
}

/* PathCmd::flatten (path.rs:41-109).  pts holds the command's points. */
static void flatten_cmd(int verb, const v2* pts, float weight, v2 last, float tolerance,
                        emit_fn emit, void* ctx, int* status) {
    switch (verb) {
    case ORC_MOVE:
        emit(ctx, ORC_MOVE, pts[0]);
        break;
    case ORC_LINE:
        emit(ctx, ORC_LINE, pts[0]);
        break;
    case ORC_QUAD: {
        v2 control = pts[0], point = pts[1];
        /* path.rs:50 */
        float dt = sqrtf((4.0f * tolerance) / length(add(sub(last, smul(2.0f, control)), point)));
        float t = 0.0f;
        int64_t n = 0;
        while (t < 1.0f) {
            t = rmin(t + dt, 1.0f);
            v2 p01 = lerp(t, last, control);
            v2 p12 = lerp(t, control, point);
            emit(ctx, ORC_LINE, lerp(t, p01, p12));
            if (++n >= FLATTEN_LIMIT) { *status |= ORC_ERR_FLATTEN_LIMIT; break; }
        }
        break;
    }
    case ORC_CUBIC: {
        v2 c1 = pts[0], c2 = pts[1], point = pts[2];
        /* path.rs:60-63 */
        v2 a = add(sub(add(smul(-1.0f, last), smul(3.0f, c1)), smul(3.0f, c2)), point);
        v2 b = smul(3.0f, add(sub(last, smul(2.0f, c1)), c2));
        float conc = rmax(length(b), length(add(a, b)));
        float dt = sqrtf((sqrtf(8.0f) * tolerance) / conc);
        float t = 0.0f;
        int64_t n = 0;
        while (t < 1.0f) {
            t = rmin(t + dt, 1.0f);
            v2 p01 = lerp(t, last, c1);
            v2 p12 = lerp(t, c1, c2);
            v2 p23 = lerp(t, c2, point);
            v2 p012 = lerp(t, p01, p12);
            v2 p123 = lerp(t, p12, p23);
            emit(ctx, ORC_LINE, lerp(t, p012, p123));
            if (++n >= FLATTEN_LIMIT) { *status |= ORC_ERR_FLATTEN_LIMIT; break; }
        }
        break;
    }
    case ORC_CONIC: {
        conic_ctx c;
        c.last = last;
        c.control = pts[0];
        c.point = pts[1];
        c.weight = weight;
        c.tolerance = tolerance;
        c.emit = emit;
        c.ctx = ctx;
        c.status = status;
        flatten_conic(&c, 0.0f, 1.0f, last, pts[1], 0);
        break;
    }
    case ORC_CLOSE:
        emit(ctx, ORC_CLOSE, mk(0.0f, 0.0f));
        break;
    default:
        *status |= ORC_ERR_ARG;
        break;
    }
}

static int verb_points(int verb) {
    switch (verb) {
    case ORC_MOVE: return 1;
    case ORC_LINE: return 1;
    case ORC_QUAD: return 2;
    case ORC_CUBIC: return 3;
    case ORC_CONIC: return 2;
    case ORC_CLOSE: return 0;
    default: return -1;
    }
}

/* The rasterizer's flatten callback (rasterizer.rs:146-156): Move -> move_to,
 * Line -> line_to, anything else (Close) ignored. */
static void raster_emit(void* ctx, int verb, v2 p) {
    orc_rasterizer* r = (orc_rasterizer*)ctx;
    if (verb == ORC_MOVE) orc_move_to(r, p.x, p.y);
    else if (verb == ORC_LINE) orc_line_to(r, p.x, p.y);
}

/* Rasterizer::command (rasterizer.rs:145-157) */
int orc_command(orc_rasterizer* r, int verb, const float* pts, float weight) {
    int np = verb_points(verb);
    if (np < 0) return ORC_ERR_ARG;
    v2 p[3];
    for (int i = 0; i < np; i++) p[i] = mk(pts[2 * i], pts[2 * i + 1]);
    int st = 0;
    flatten_cmd(verb, p, weight, r->last, TOLERANCE, raster_emit, r, &st);
    r->status |= st;
    return st;
}

/* Validate a verb/point/weight triple and count consumption. */
static int check_path(const uint8_t* verbs, size_t n_verbs, size_t n_points, size_t n_weights) {
    size_t np = 0, nw = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        int k = verb_points(verbs[i]);
        if (k < 0) return ORC_ERR_ARG;
        np += (size_t)k;
        if (verbs[i] == ORC_CONIC) nw++;
    }
    if (np != n_points || nw != n_weights) return ORC_ERR_ARG;
    return ORC_OK;
}

/* Rasterizer::fill (rasterizer.rs:161-165) with PathCmd::transform
 * (path.rs:16-37): every point is transformed, conic weights are not. */
int orc_fill(orc_rasterizer* r, const uint8_t* verbs, size_t n_verbs,
             const float* points, size_t n_points,
             const float* weights, size_t n_weights, const float* xf) {
    if (check_path(verbs, n_verbs, n_points, n_weights)) return ORC_ERR_ARG;
    size_t pi = 0, wi = 0;
    int st = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        int verb = verbs[i];
        int np = verb_points(verb);
        v2 p[3];
        for (int k = 0; k < np; k++) p[k] = apply(xf, mk(points[2 * (pi + k)], points[2 * (pi + k) + 1]));
        float w = 0.0f;
        if (verb == ORC_CONIC) w = weights[wi++];
        pi += (size_t)np;
        flatten_cmd(verb, p, w, r->last, TOLERANCE, raster_emit, r, &st);
    }
    r->status |= st;
    return st;
}

/* ---------------------------------------------------- path.rs:114-274 */

static void path_push(orc_path* o, int verb, const v2* pts, int np, float w) {
    GROW(o->verbs, o->n_verbs, o->cap_verbs, uint8_t);
    o->verbs[o->n_verbs++] = (uint8_t)verb;
    for (int k = 0; k < np; k++) {
        GROW(o->points, o->n_points * 2 + 1, o->cap_points, float);
        o->points[o->n_points * 2] = pts[k].x;
        o->points[o->n_points * 2 + 1] = pts[k].y;
        o->n_points++;
    }
    if (verb == ORC_CONIC) {
        GROW(o->weights, o->n_weights, o->cap_weights, float);
        o->weights[o->n_weights++] = w;
    }
}

void orc_path_free(orc_path* p) {
    free(p->verbs);
    free(p->points);
    free(p->weights);
    memset(p, 0, sizeof(*p));
}

static void path_emit(void* ctx, int verb, v2 p) {
    orc_path* o = (orc_path*)ctx;
    path_push(o, verb, &p, verb == ORC_CLOSE ? 0 : 1, 0.0f);
}

/* flatten (path.rs:114-144): flattens with `last` tracked from each
 * command's end point (Close leaves it unchanged). */
int orc_flatten_path(const uint8_t* verbs, size_t n_verbs, const float* points, size_t n_points,
                     const float* weights, size_t n_weights, float tolerance, orc_path* out) {
    if (check_path(verbs, n_verbs, n_points, n_weights)) return ORC_ERR_ARG;
    v2 last = mk(0.0f, 0.0f);
    size_t pi = 0, wi = 0;
    int st = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        int verb = verbs[i];
        int np = verb_points(verb);
        v2 p[3];
        for (int k = 0; k < np; k++) p[k] = mk(points[2 * (pi + k)], points[2 * (pi + k) + 1]);
        float w = verb == ORC_CONIC ? weights[wi++] : 0.0f;
        pi += (size_t)np;
        flatten_cmd(verb, p, w, last, tolerance, path_emit, out, &st);
        if (np > 0) last = p[np - 1];
    }
    return st;
}

/* join (path.rs:163-171) */
static void stroke_join(orc_path* o, float width, v2 prev_normal, v2 next_normal, v2 point) {
    float offset = 1.0f / (1.0f + dot(prev_normal, next_normal));
    if (fabsf(offset) > 2.0f) {
        v2 a = add(point, smul(0.5f * width, prev_normal));
        v2 b = add(point, smul(0.5f * width, next_normal));
        path_push(o, ORC_LINE, &a, 1, 0.0f);
        path_push(o, ORC_LINE, &b, 1, 0.0f);
    } else {
        v2 a = add(point, smul(0.5f * width * offset, add(prev_normal, next_normal)));
        path_push(o, ORC_LINE, &a, 1, 0.0f);
    }
}

/* offset (path.rs:174-215); contour points are pts[0..n). */
static void stroke_offset(orc_path* o, float width, const v2* pts, size_t n, int closed, int reverse) {
    v2 first_point = (closed == reverse) ? pts[0] : pts[n - 1];
    v2 prev_point = first_point;
    v2 prev_normal = mk(0.0f, 0.0f);
    size_t i = 0;
    for (;;) {
        v2 next_point;
        if (i < n) next_point = reverse ? pts[n - i - 1] : pts[i];
        else next_point = first_point;
        if (!veq(next_point, prev_point) || i == n) {
            v2 tangent = sub(next_point, prev_point);
            v2 normal = mk(-tangent.y, tangent.x);
            float len = length(normal);
            normal = (len == 0.0f) ? mk(0.0f, 0.0f) : smul(1.0f / len, normal);
            stroke_join(o, width, prev_normal, normal, prev_point);
            prev_point = next_point;
            prev_normal = normal;
        }
        i++;
        if (i > n) break;
    }
}

/* stroke (path.rs:152-274).  The input must be piecewise linear
 * (Move/Line/Close); the reference panics otherwise (path.rs:158,264-266),
 * here ORC_ERR_ARG is returned. */
int orc_stroke_path(const uint8_t* verbs, size_t n_verbs, const float* points, size_t n_points,
                    float width, orc_path* out) {
    size_t np_total = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        int v = verbs[i];
        if (v != ORC_MOVE && v != ORC_LINE && v != ORC_CLOSE) return ORC_ERR_ARG;
        np_total += (size_t)verb_points(v);
    }
    if (np_total != n_points) return ORC_ERR_ARG;
    /* Point of each command (Close has none; it is never read as a point
     * because contour ranges exclude it). */
    v2* cp = (v2*)malloc((n_verbs + 1) * sizeof(v2));
    size_t pi = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        if (verbs[i] == ORC_CLOSE) cp[i] = mk(0.0f, 0.0f);
        else { cp[i] = mk(points[2 * pi], points[2 * pi + 1]); pi++; }
    }
    size_t contour_start = 0, contour_end = 0;
    int closed = 0;
    size_t i = 0;
    for (;;) {
        int has = i < n_verbs;
        int verb = has ? verbs[i] : -1;
        if (verb == ORC_CLOSE) closed = 1;
        if (!has || verb == ORC_MOVE || verb == ORC_CLOSE) {
            if (contour_start != contour_end) {
                const v2* contour = cp + contour_start;
                size_t n = contour_end - contour_start;
                size_t base = out->n_verbs;
                stroke_offset(out, width, contour, n, closed, 0);
                out->verbs[base] = ORC_MOVE;  /* output[base] = Move(get_point(output[base])) */
                if (closed) path_push(out, ORC_CLOSE, NULL, 0, 0.0f);
                base = out->n_verbs;
                stroke_offset(out, width, contour, n, closed, 1);
                if (closed) out->verbs[base] = ORC_MOVE;
                path_push(out, ORC_CLOSE, NULL, 0, 0.0f);
            }
        }
        if (!has) break;
        if (verb == ORC_MOVE) {
            contour_start = contour_end;
            contour_end = contour_start + 1;
        } else if (verb == ORC_LINE) {
            contour_end += 1;
        } else { /* Close */
            contour_start = contour_end + 1;
            contour_end = contour_start;
            closed = 1;
        }
        i++;
    }
    free(cp);
    return ORC_OK;
}

/* Rasterizer::stroke (rasterizer.rs:169-171): flatten in path space, stroke,
 * then fill with the transform. */
int orc_stroke(orc_rasterizer* r, const uint8_t* verbs, size_t n_verbs,
               const float* points, size_t n_points,
               const float* weights, size_t n_weights, float width, const float* xf) {
    orc_path flat, stroked;
    memset(&flat, 0, sizeof(flat));
    memset(&stroked, 0, sizeof(stroked));
    int st = orc_flatten_path(verbs, n_verbs, points, n_points, weights, n_weights, TOLERANCE, &flat);
    if (st & ORC_ERR_ARG) { orc_path_free(&flat); return st; }
    st |= orc_stroke_path(flat.verbs, flat.n_verbs, flat.points, flat.n_points, width, &stroked);
    if (!(st & ORC_ERR_ARG))
        st |= orc_fill(r, stroked.verbs, stroked.n_verbs, stroked.points, stroked.n_points, NULL, 0, xf);
    orc_path_free(&flat);
    orc_path_free(&stroked);
    r->status |= st;
    return st;
}

/* --------------------------------------------------------------- finish */

typedef struct {
    int16_t tile_x, tile_y;
    size_t start, end;
} bin;

static int bin_cmp(const void* pa, const void* pb) {
    const bin* a = (const bin*)pa;
    const bin* b = (const bin*)pb;
    if (a->tile_y != b->tile_y) return a->tile_y < b->tile_y ? -1 : 1;
    if (a->tile_x != b->tile_x) return a->tile_x < b->tile_x ? -1 : 1;
    /* Stable: equal keys keep emission order (one of the orders the
     * reference's sort_unstable_by_key may produce). */
    if (a->start != b->start) return a->start < b->start ? -1 : 1;
    return 0;
}

static int tinc_cmp(const void* pa, const void* pb) {
    const tile_increment* a = (const tile_increment*)pa;
    const tile_increment* b = (const tile_increment*)pb;
    if (a->tile_y != b->tile_y) return a->tile_y < b->tile_y ? -1 : 1;
    if (a->tile_x != b->tile_x) return a->tile_x < b->tile_x ? -1 : 1;
    return 0;
}

/* Rasterizer::finish (rasterizer.rs:180-268) */
int orc_finish(orc_rasterizer* r, void* user, orc_tile_fn tile_cb, orc_span_fn span_cb) {
    if (!veq(r->last, r->first)) orc_line_to(r, r->first.x, r->first.y);

    /* Bin consecutive same-tile increments (rasterizer.rs:193-208). */
    size_t cap_bins = 64, n_bins = 0;
    bin* bins = (bin*)malloc(cap_bins * sizeof(bin));
    bin cur;
    cur.tile_x = 0;
    cur.tile_y = 0;
    cur.start = 0;
    cur.end = 0;
    if (r->n_inc > 0) {
        cur.tile_x = div8(r->inc[0].x);
        cur.tile_y = div8(r->inc[0].y);
    }
    for (size_t i = 0; i < r->n_inc; i++) {
        int16_t tx = div8(r->inc[i].x);
        int16_t ty = div8(r->inc[i].y);
        if (tx != cur.tile_x || ty != cur.tile_y) {
            GROW(bins, n_bins, cap_bins, bin);
            bins[n_bins++] = cur;
            cur.tile_x = tx;
            cur.tile_y = ty;
            cur.start = i;
            cur.end = i;
        }
        cur.end += 1;
    }
    GROW(bins, n_bins, cap_bins, bin);
    bins[n_bins++] = cur;
    qsort(bins, n_bins, sizeof(bin), bin_cmp);                       /* :209 */
    if (r->n_tinc > 1) qsort(r->tinc, r->n_tinc, sizeof(tile_increment), tinc_cmp); /* :211 */

    float areas[TILE * TILE], heights[TILE * TILE], prev[TILE], next[TILE];
    memset(areas, 0, sizeof(areas));
    memset(heights, 0, sizeof(heights));
    memset(prev, 0, sizeof(prev));
    memset(next, 0, sizeof(next));
    size_t ti = 0;
    long winding = 0;

    for (size_t i = 0; i < n_bins; i++) {
        bin b = bins[i];
        for (size_t k = b.start; k < b.end; k++) {                    /* :223-228 */
            const increment* inc = &r->inc[k];
            int x = inc->x & (TILE - 1);
            int y = inc->y & (TILE - 1);
            areas[y * TILE + x] += inc->area;
            heights[y * TILE + x] += inc->height;
        }
        int last_of_tile = (i + 1 == n_bins) || bins[i + 1].tile_x != b.tile_x || bins[i + 1].tile_y != b.tile_y;
        if (!last_of_tile) continue;

        uint8_t tile[TILE * TILE];                                    /* :231-239 */
        for (int y = 0; y < TILE; y++) {
            float accum = prev[y];
            for (int x = 0; x < TILE; x++) {
                float v = rmin(fabsf(accum + areas[y * TILE + x]) * 256.0f, 255.0f);
                tile[y * TILE + x] = (uint8_t)v;
                accum += heights[y * TILE + x];
            }
            next[y] = accum;
        }
        if (tile_cb) tile_cb(user, (int16_t)(b.tile_x * TILE), (int16_t)(b.tile_y * TILE), tile);

        memset(areas, 0, sizeof(areas));                               /* :243-250 */
        memset(heights, 0, sizeof(heights));
        if (i + 1 < n_bins && bins[i + 1].tile_y == b.tile_y) memcpy(prev, next, sizeof(prev));
        else memset(prev, 0, sizeof(prev));
        memset(next, 0, sizeof(next));

        if (i + 1 < n_bins && bins[i + 1].tile_y == b.tile_y && bins[i + 1].tile_x > b.tile_x + 1) { /* :252 */
            while (ti < r->n_tinc) {
                tile_increment t = r->tinc[ti];
                if (t.tile_y > b.tile_y || (t.tile_y == b.tile_y && t.tile_x > b.tile_x)) break;
                winding += t.sign;
                ti++;
            }
            if (winding != 0) {
                int16_t width = (int16_t)(bins[i + 1].tile_x - b.tile_x - 1);
                if (span_cb)
                    span_cb(user, (int16_t)((b.tile_x + 1) * TILE), (int16_t)(b.tile_y * TILE),
                            (uint16_t)((uint16_t)width * (uint16_t)TILE));
            }
        }
    }
    free(bins);
    return r->status;
}

void orc_stats(const orc_rasterizer* r, int64_t* lines, int64_t* increments, int64_t* tile_increments) {
    if (lines) *lines = r->n_lines;
    if (increments) *increments = (int64_t)r->n_inc;
    if (tile_increments) *tile_increments = (int64_t)r->n_tinc;
}

/* ---------------------------------------------------------- batch runner */

struct orc_result {
    int16_t* tile_xy;
    uint8_t* tile_alpha;
    size_t n_tiles, cap_tiles;
    int16_t* span_xy;
    uint16_t* span_w;
    size_t n_spans, cap_spans;
    int64_t* path_tile_off;
    int64_t* path_span_off;
    int64_t n_paths;
    int64_t lines, increments, tile_increments;
    int status;
};

static void res_tile(void* user, int16_t x, int16_t y, const uint8_t* a) {
    orc_result* r = (orc_result*)user;
    if (r->n_tiles >= r->cap_tiles) {
        size_t nc = r->cap_tiles ? r->cap_tiles * 2 : 1024;
        r->tile_xy = (int16_t*)realloc(r->tile_xy, nc * 2 * sizeof(int16_t));
        r->tile_alpha = (uint8_t*)realloc(r->tile_alpha, nc * 64);
        if (!r->tile_xy || !r->tile_alpha) abort();
        r->cap_tiles = nc;
    }
    r->tile_xy[2 * r->n_tiles] = x;
    r->tile_xy[2 * r->n_tiles + 1] = y;
    memcpy(r->tile_alpha + 64 * r->n_tiles, a, 64);
    r->n_tiles++;
}

static void res_span(void* user, int16_t x, int16_t y, uint16_t w) {
    orc_result* r = (orc_result*)user;
    if (r->n_spans >= r->cap_spans) {
        size_t nc = r->cap_spans ? r->cap_spans * 2 : 1024;
        r->span_xy = (int16_t*)realloc(r->span_xy, nc * 2 * sizeof(int16_t));
        r->span_w = (uint16_t*)realloc(r->span_w, nc * sizeof(uint16_t));
        if (!r->span_xy || !r->span_w) abort();
        r->cap_spans = nc;
    }
    r->span_xy[2 * r->n_spans] = x;
    r->span_xy[2 * r->n_spans + 1] = y;
    r->span_w[r->n_spans] = w;
    r->n_spans++;
}

orc_result* orc_run_batch(int64_t n_paths,
                          const uint8_t* verbs, const int64_t* verb_off,
                          const float* points, const int64_t* point_off,
                          const float* weights, const int64_t* weight_off,
                          const float* xforms, const float* widths, const uint8_t* stroke_flags) {
    orc_result* res = (orc_result*)calloc(1, sizeof(orc_result));
    if (!res) abort();
    res->n_paths = n_paths;
    res->path_tile_off = (int64_t*)calloc((size_t)n_paths + 1, sizeof(int64_t));
    res->path_span_off = (int64_t*)calloc((size_t)n_paths + 1, sizeof(int64_t));
    for (int64_t p = 0; p < n_paths; p++) {
        orc_rasterizer* r = orc_new();
        const uint8_t* v = verbs + verb_off[p];
        size_t nv = (size_t)(verb_off[p + 1] - verb_off[p]);
        const float* pts = points + 2 * point_off[p];
        size_t np = (size_t)(point_off[p + 1] - point_off[p]);
        const float* w = weights ? weights + (weight_off ? weight_off[p] : 0) : NULL;
        size_t nw = weight_off ? (size_t)(weight_off[p + 1] - weight_off[p]) : 0;
        int st;
        if (widths && (!stroke_flags || stroke_flags[p]))
            st = orc_stroke(r, v, nv, pts, np, w, nw, widths[p], xforms + 6 * p);
        else st = orc_fill(r, v, nv, pts, np, w, nw, xforms + 6 * p);
        res->status |= st;
        st = orc_finish(r, res, res_tile, res_span);
        res->status |= st;
        int64_t a, b, c;
        orc_stats(r, &a, &b, &c);
        res->lines += a;
        res->increments += b;
        res->tile_increments += c;
        orc_free(r);
        res->path_tile_off[p + 1] = (int64_t)res->n_tiles;
        res->path_span_off[p + 1] = (int64_t)res->n_spans;
    }
    return res;
}

int orc_result_status(const orc_result* res) { return res->status; }

void orc_result_sizes(const orc_result* res, int64_t* n_tiles, int64_t* n_spans) {
    *n_tiles = (int64_t)res->n_tiles;
    *n_spans = (int64_t)res->n_spans;
}

void orc_result_copy(const orc_result* res, int16_t* tile_xy, uint8_t* tile_alpha,
                     int16_t* span_xy, uint16_t* span_w,
                     int64_t* path_tile_off, int64_t* path_span_off) {
    if (tile_xy && res->n_tiles) memcpy(tile_xy, res->tile_xy, res->n_tiles * 2 * sizeof(int16_t));
    if (tile_alpha && res->n_tiles) memcpy(tile_alpha, res->tile_alpha, res->n_tiles * 64);
    if (span_xy && res->n_spans) memcpy(span_xy, res->span_xy, res->n_spans * 2 * sizeof(int16_t));
    if (span_w && res->n_spans) memcpy(span_w, res->span_w, res->n_spans * sizeof(uint16_t));
    if (path_tile_off) memcpy(path_tile_off, res->path_tile_off, ((size_t)res->n_paths + 1) * sizeof(int64_t));
    if (path_span_off) memcpy(path_span_off, res->path_span_off, ((size_t)res->n_paths + 1) * sizeof(int64_t));
}

void orc_result_stats(const orc_result* res, int64_t* lines, int64_t* increments, int64_t* tile_increments) {
    *lines = res->lines;
    *increments = res->increments;
    *tile_increments = res->tile_increments;
}

void orc_result_free(orc_result* res) {
    if (!res) return;
    free(res->tile_xy);
    free(res->tile_alpha);
    free(res->span_xy);
    free(res->span_w);
    free(res->path_tile_off);
    free(res->path_span_off);
    free(res);
}
