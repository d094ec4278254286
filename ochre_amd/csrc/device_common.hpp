// Device-side arithmetic of the ochre rasterisation path (gfx950).
//
// Every routine reproduces the reference's f32 arithmetic operation by
// operation (glowcoil/ochre src/geom.rs, src/path.rs, src/rasterizer.rs):
//   - no a*b+c contraction (the file is compiled with -ffp-contract=off and
//     the pragma below repeats it);
//   - IEEE '/' and sqrt (hipcc's default correctly-rounded f32 div/sqrt);
//   - NaN-ignoring min/max (fminf/fmaxf lower to v_min_f32/v_max_f32);
//   - saturating f32->i16 casts with NaN -> 0, wrapping i16 arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace ochre {

constexpr int TILE = 8;                  // rasterizer.rs:4
constexpr float TOLERANCE = 0.1f;        // rasterizer.rs:6
constexpr uint32_t FLATTEN_LIMIT = 1u << 16;  // lines per curve; the reference loops forever on dt == 0
constexpr int CONIC_DEPTH_LIMIT = 32;    // the reference recurses without bound

enum Verb : uint8_t { MOVE = 0, LINE = 1, QUAD = 2, CUBIC = 3, CONIC = 4, CLOSE = 5 };

// Points consumed per verb (path.rs:5-12).
__host__ __device__ inline int verb_points(int v) {
    return v == MOVE ? 1 : v == LINE ? 1 : v == QUAD ? 2 : v == CUBIC ? 3 : v == CONIC ? 2 : v == CLOSE ? 0 : -1;
}

struct V2 {
    float x, y;
};

__host__ __device__ inline V2 mk(float x, float y) { return V2{x, y}; }
__host__ __device__ inline V2 add(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }   // geom.rs:73-82
__host__ __device__ inline V2 sub(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }   // geom.rs:91-100
__host__ __device__ inline V2 smul(float s, V2 a) { return mk(s * a.x, s * a.y); }   // geom.rs:120-129
__host__ __device__ inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }   // geom.rs:19-21
__host__ __device__ inline float vlen(V2 a) { return sqrtf(dot(a, a)); }             // geom.rs:38-40
__host__ __device__ inline V2 lerp(float t, V2 a, V2 b) {                            // geom.rs:50-52
    return add(smul(1.0f - t, a), smul(t, b));
}
__host__ __device__ inline bool veq(V2 a, V2 b) { return a.x == b.x && a.y == b.y; } // derived PartialEq

// Transform::apply (geom.rs:260-262): Mat2x2 (row-major) * v + offset.
struct Xform {
    float m0, m1, m2, m3, ox, oy;
};
__host__ __device__ inline V2 apply(const Xform& t, V2 p) {
    V2 m = mk(t.m0 * p.x + t.m1 * p.y, t.m2 * p.x + t.m3 * p.y);
    return add(m, mk(t.ox, t.oy));
}

// `f as i16` (saturating, NaN -> 0).
__host__ __device__ inline int16_t f2i16(float v) {
    if (v != v) return 0;
    if (v >= 32767.0f) return 32767;
    if (v <= -32768.0f) return -32768;
    return (int16_t)(int)v;
}
// `f.signum() as i16`.
__host__ __device__ inline int16_t signum_i16(float v) {
    if (v != v) return 0;
    return signbit(v) ? (int16_t)-1 : (int16_t)1;
}
__host__ __device__ inline int16_t div8(int16_t v) { return (int16_t)(v >> 3); }
__host__ __device__ inline int16_t wrap16(int v) { return (int16_t)v; }

// ------------------------------------------------------------- flattening

// Quadratic dt (path.rs:50).
__device__ inline float quad_dt(V2 last, V2 c, V2 p) {
    return sqrtf((4.0f * TOLERANCE) / vlen(add(sub(last, smul(2.0f, c)), p)));
}
// Cubic dt (path.rs:60-63).
__device__ inline float cubic_dt(V2 last, V2 c1, V2 c2, V2 p) {
    V2 a = add(sub(add(smul(-1.0f, last), smul(3.0f, c1)), smul(3.0f, c2)), p);
    V2 b = smul(3.0f, add(sub(last, smul(2.0f, c1)), c2));
    float conc = fmaxf(vlen(b), vlen(add(a, b)));
    return sqrtf((sqrtf(8.0f) * TOLERANCE) / conc);
}
// Number of `t = min(t + dt, 1)` iterations of the flatten loops
// (path.rs:51-57, 64-73); *limited is set when the guard trips.
__device__ inline uint32_t t_steps(float dt, bool* limited) {
    float t = 0.0f;
    uint32_t n = 0;
    while (t < 1.0f) {
        t = fminf(t + dt, 1.0f);
        if (++n >= FLATTEN_LIMIT) {
            *limited = true;
            break;
        }
    }
    return n;
}
__device__ inline V2 quad_at(float t, V2 last, V2 c, V2 p) {
    V2 p01 = lerp(t, last, c);
    V2 p12 = lerp(t, c, p);
    return lerp(t, p01, p12);
}
__device__ inline V2 cubic_at(float t, V2 last, V2 c1, V2 c2, V2 p) {
    V2 p01 = lerp(t, last, c1);
    V2 p12 = lerp(t, c1, c2);
    V2 p23 = lerp(t, c2, p);
    V2 p012 = lerp(t, p01, p12);
    V2 p123 = lerp(t, p12, p23);
    return lerp(t, p012, p123);
}

// Conic midpoint subdivision (path.rs:76-101) as an explicit-stack DFS in
// the recursion's visiting order.  emit(point) is called for each output
// line end point (two per leaf).  Returns false if the depth guard tripped.
template <class Emit>
__device__ inline bool conic_walk(V2 last, V2 c, V2 p, float w, Emit&& emit) {
    struct Ent {
        float t;
        V2 pt;
        int depth;
    };
    Ent stack[CONIC_DEPTH_LIMIT + 2];
    int sp = 0;
    stack[sp++] = Ent{1.0f, p, 0};
    float t0 = 0.0f;
    V2 p0 = last;
    bool ok = true;
    V2 wc = smul(w, c);
    while (sp > 0) {
        Ent& top = stack[sp - 1];
        float t1 = top.t;
        V2 p1 = top.pt;
        int depth = top.depth;
        float t = 0.5f * (t0 + t1);
        V2 p01 = lerp(t, last, wc);
        V2 p12 = lerp(t, wc, p);
        float denom = (1.0f - t) * (1.0f - t) + 2.0f * t * (1.0f - t) * w + t * t;
        V2 mid = smul(1.0f / denom, lerp(t, p01, p12));
        float err = vlen(sub(mid, smul(0.5f, add(p0, p1))));
        if (err > TOLERANCE && depth < CONIC_DEPTH_LIMIT) {
            top.depth = depth + 1;
            stack[sp++] = Ent{t, mid, depth + 1};
        } else {
            if (err > TOLERANCE) ok = false;
            emit(mid);
            emit(p1);
            t0 = t1;
            p0 = p1;
            sp--;
        }
    }
    return ok;
}

// ------------------------------------------------------------- the DDA
//
// Rasterizer::line_to (rasterizer.rs:72-140) as a resumable state machine.
// The reference keeps row_t0/col_t0; t0 = max(row_t0, col_t0) is always the
// value consumed by the previous step (the merge of the two crossing
// sequences is monotone), and the loop ends exactly when a step consumes 1.0,
// so the state here is (x, y, row_t1, col_t1, t0).
struct Dda {
    V2 last, point;
    float x_step, y_step;
    float row_t1, col_t1, t0;
    int16_t x_dir, y_dir, x, y;
};

__device__ inline void dda_consts(Dda& d, V2 last, V2 point) {
    d.last = last;
    d.point = point;
    d.x_dir = signum_i16(point.x - last.x);
    d.y_dir = signum_i16(point.y - last.y);
    float dtdx = 1.0f / (point.x - last.x);
    float dtdy = 1.0f / (point.y - last.y);
    d.x_step = fabsf(dtdx);
    d.y_step = fabsf(dtdy);
}

// Full initial state (rasterizer.rs:73-95).
__device__ inline void dda_init(Dda& d, V2 last, V2 point) {
    dda_consts(d, last, point);
    float dtdx = 1.0f / (point.x - last.x);
    float dtdy = 1.0f / (point.y - last.y);
    d.x = f2i16(floorf(last.x));
    d.y = f2i16(floorf(last.y));
    d.t0 = 0.0f;
    if (last.y == point.y) {
        d.row_t1 = INFINITY;
    } else {
        float next_y = point.y > last.y ? (float)wrap16(d.y + 1) : (float)d.y;
        d.row_t1 = fminf(dtdy * (next_y - last.y), 1.0f);
    }
    if (last.x == point.x) {
        d.col_t1 = INFINITY;
    } else {
        float next_x = point.x > last.x ? (float)wrap16(d.x + 1) : (float)d.x;
        d.col_t1 = fminf(dtdx * (next_x - last.x), 1.0f);
    }
}

// The line's point at parameter t: (1 - t) * last + t * point (rasterizer.rs:100-101).
__device__ inline V2 dda_point(const Dda& d, float t) { return add(smul(1.0f - t, d.last), smul(t, d.point)); }

// The increment at the current pixel with its start point p0 carried from
// the previous step (rasterizer.rs:98-106), p0 left at its end point for the
// next.  With finite end points row_t1 and col_t1 are never NaN (each is
// clamped by fminf), so the step's t0 is exactly the previous increment's
// t1 — up to the sign of a zero t, which no sum or alpha can see; for other
// lines the start point is recomputed from t0.
__device__ inline void dda_increment_carried(const Dda& d, V2& p0, bool finite, float& area, float& height) {
    if (!finite) p0 = dda_point(d, d.t0);
    const V2 p1 = dda_point(d, fminf(d.row_t1, d.col_t1));
    height = p1.y - p0.y;
    const float right = (float)wrap16(d.x + 1);
    area = 0.5f * height * ((right - p0.x) + (right - p1.x));
    p0 = p1;
}

__device__ inline bool line_finite(V2 a, V2 b) { return isfinite(a.x) && isfinite(a.y) && isfinite(b.x) && isfinite(b.y); }

// The increment pushed at the current pixel (rasterizer.rs:98-106).
__device__ inline void dda_increment(const Dda& d, float& area, float& height) {
    float t1 = fminf(d.row_t1, d.col_t1);
    V2 p0 = add(smul(1.0f - d.t0, d.last), smul(d.t0, d.point));
    V2 p1 = add(smul(1.0f - t1, d.last), smul(t1, d.point));
    height = p1.y - p0.y;
    float right = (float)wrap16(d.x + 1);
    area = 0.5f * height * ((right - p0.x) + (right - p1.x));
}

// One step (rasterizer.rs:108-116); returns true when the loop terminates
// (rasterizer.rs:118,133).  The end-point snap (:118-121) is left to the
// caller because it only affects the tile-row crossing chain.
__device__ inline bool dda_step(Dda& d) {
    float c;
    if (d.row_t1 < d.col_t1) {
        c = d.row_t1;
        d.row_t1 = fminf(d.row_t1 + d.y_step, 1.0f);
        d.y = wrap16(d.y + d.y_dir);
    } else {
        c = d.col_t1;
        d.col_t1 = fminf(d.col_t1 + d.x_step, 1.0f);
        d.x = wrap16(d.x + d.x_dir);
    }
    d.t0 = c;
    return c == 1.0f;
}

// k steps of dda_step in select form (both candidate updates computed, one
// kept: the same f32 operations on the same values, so the same state), for
// replaying a walk up to one of its increments: the caller keeps k below the
// line's increment count, so no replayed step is the terminating one.
__device__ inline void dda_skip(Dda& d, uint32_t k) {
    for (uint32_t i = 0; i < k; i++) {
        const bool take_y = d.row_t1 < d.col_t1;
        const float c = take_y ? d.row_t1 : d.col_t1;
        const float r = fminf(d.row_t1 + d.y_step, 1.0f), q = fminf(d.col_t1 + d.x_step, 1.0f);
        d.row_t1 = take_y ? r : d.row_t1;
        d.col_t1 = take_y ? d.col_t1 : q;
        d.y = take_y ? wrap16(d.y + d.y_dir) : d.y;
        d.x = take_y ? d.x : wrap16(d.x + d.x_dir);
        d.t0 = c;
    }
}

}  // namespace ochre
