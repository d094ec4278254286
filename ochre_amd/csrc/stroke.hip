// Stroking on the device.  Rasterizer::stroke (rasterizer.rs:169-171) is
// fill(stroke(flatten(path, TOLERANCE), width), transform): the outline is
// built in path space and then filled with the path's transform.  Here two
// expansion passes build every path's outline in device memory:
//   flatten  path.rs:114-144 (PathCmd::flatten, path.rs:41-109): Move/Line/
//            Close out, `last` taken from each command's end point;
//   stroke   path.rs:152-274: per contour, a forward and a reverse offset
//            polyline with miter joins (a bevel where the miter would reach
//            beyond twice the half width);
// and the outline batch then goes through the fill pipelines unchanged.
// Each pass runs one wave per path, once to count and once to write, with a
// device scan between; the wave works through a path 64 commands (or
// indices) at a time and places each lane's output at positions from a wave
// scan, so the path's outline command sequence is the reference's exactly.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "pipeline.hpp"
#include "raster_common.hpp"
#include "scan.hpp"

namespace ochre {

namespace {

struct PathSrc {
    const uint8_t* verbs;
    const int64_t* verb_off;
    const float* points;
    const int64_t* point_off;
    const float* weights;
    const int64_t* weight_off;
};

// ------------------------------------------------------------ flatten
// flatten (path.rs:114-144), one wave per path, one lane per command (64 at
// a time).  A lane's outputs: Move -> M, Line -> L, Close -> Z, a curve ->
// its segments' end points as L (path.rs:49-104); `last` is the previous
// command's end point, which is the point before the command's own.
// Returns the error bits of the lane.
template <bool EMIT>
__global__ __launch_bounds__(WAVE) void k_sflatten(PathSrc in, const uint8_t* stroke_flags, int64_t n, int64_t* cnt,
                                                  const int64_t* off_v, const int64_t* off_p, uint8_t* verbs,
                                                  float* points, uint32_t* error) {
    const int64_t p = blockIdx.x;
    if (p >= n) return;
    const int lane = threadIdx.x;
    if (stroke_flags && !stroke_flags[p]) {  // a fill of a draw list: not flattened here
        if (!EMIT && lane == 0) cnt[p] = cnt[n + p] = 0;
        return;
    }
    const int64_t vb = in.verb_off[p], ve = in.verb_off[p + 1];
    const int64_t pb = in.point_off[p], pe = in.point_off[p + 1];
    const int64_t wb = in.weight_off[p], we = in.weight_off[p + 1];
    uint32_t cpts = 0, ccon = 0, err = 0;
    uint64_t ov = EMIT ? (uint64_t)off_v[p] : 0u, op = EMIT ? (uint64_t)off_p[p] : 0u;
    for (int64_t cb = vb; cb < ve; cb += WAVE) {
        const int64_t i = cb + lane;
        const bool valid = i < ve;
        const int verb = valid ? in.verbs[i] : (int)CLOSE;
        int k = verb_points(verb);
        bool bad = valid && k < 0;
        if (k < 0) k = 0;
        const uint32_t k_incl = wave_incl_scan((uint32_t)k, OpAdd());
        const uint32_t cc = (valid && verb == CONIC) ? 1u : 0u;
        const uint32_t cc_incl = wave_incl_scan(cc, OpAdd());
        const int64_t g = pb + cpts + (k_incl - (uint32_t)k);
        const int64_t wi = wb + ccon + (cc_incl - cc);
        if (valid && (g + k > pe || (cc && wi >= we))) bad = true;
        V2 P0 = mk(0.0f, 0.0f), P1 = P0, P2 = P0, S = P0;
        float w = 0.0f;
        if (valid && !bad) {
            if (k > 0) P0 = mk(in.points[2 * g], in.points[2 * g + 1]);
            if (k > 1) P1 = mk(in.points[2 * g + 2], in.points[2 * g + 3]);
            if (k > 2) P2 = mk(in.points[2 * g + 4], in.points[2 * g + 5]);
            if (g > pb) S = mk(in.points[2 * g - 2], in.points[2 * g - 1]);
            if (cc) w = in.weights[wi];
        }
        // outputs of this command: verbs (M/L/Z) and points
        uint32_t nv = 0;
        float dt = 0.0f;
        if (valid && !bad) {
            if (verb == MOVE || verb == LINE || verb == CLOSE) {
                nv = 1;
            } else if (verb == QUAD || verb == CUBIC) {
                dt = verb == QUAD ? quad_dt(S, P0, P1) : cubic_dt(S, P0, P1, P2);
                float t = 0.0f;
                while (t < 1.0f) {
                    t = fminf(t + dt, 1.0f);
                    if (++nv >= FLATTEN_LIMIT) {
                        err |= ERR_FLATTEN;
                        break;
                    }
                }
            } else {
                if (!conic_walk(S, P0, P1, w, [&](V2) { nv++; })) err |= ERR_FLATTEN;
            }
        }
        if (bad) err |= ERR_INPUT;
        const uint32_t npt = (valid && verb != CLOSE) ? nv : 0u;
        const uint32_t v_incl = wave_incl_scan(nv, OpAdd());
        const uint32_t p_incl = wave_incl_scan(npt, OpAdd());
        if (EMIT && valid && !bad && !err) {
            uint64_t vo = ov + (v_incl - nv), po = op + (p_incl - npt);
            auto put = [&](int vv, V2 q) {
                verbs[vo++] = (uint8_t)vv;
                points[2 * po] = q.x;
                points[2 * po + 1] = q.y;
                po++;
            };
            if (verb == MOVE || verb == LINE) {
                put(verb, P0);
            } else if (verb == CLOSE) {
                verbs[vo] = (uint8_t)CLOSE;
            } else if (verb == QUAD || verb == CUBIC) {
                float t = 0.0f;
                while (t < 1.0f) {
                    t = fminf(t + dt, 1.0f);
                    put(LINE, verb == QUAD ? quad_at(t, S, P0, P1) : cubic_at(t, S, P0, P1, P2));
                }
            } else {
                conic_walk(S, P0, P1, w, [&](V2 q) { put(LINE, q); });
            }
        }
        ov += __shfl(v_incl, WAVE - 1, WAVE);
        op += __shfl(p_incl, WAVE - 1, WAVE);
        cpts += __shfl(k_incl, WAVE - 1, WAVE);
        ccon += __shfl(cc_incl, WAVE - 1, WAVE);
    }
    if (cpts != (uint32_t)(pe - pb)) err |= ERR_INPUT;
    if (!EMIT && lane == 0) {
        cnt[p] = (int64_t)ov;
        cnt[n + p] = (int64_t)op;
    }
    if (err) atomicOr(error, err);
}

// ------------------------------------------------------------- stroke
// stroke (path.rs:152-274), one wave per path.  The flattened commands are
// cut into contours exactly where the reference processes them (a Move, a
// Close, the end); a contour holds Move/Line commands only, so its points
// are contiguous.  Each side of a contour (offset, path.rs:174-215) is a pass
// over i = 0..n, 64 indices at a time: index i is a join event when its point
// differs from the previous one (or i = n), the join uses the normal of the
// previous event (a ballot finds it), and each join writes one point (miter)
// or two (bevel) at positions from a wave scan.

__device__ inline V2 unit_normal(V2 a, V2 b) {
    const V2 tangent = sub(b, a);
    const V2 normal = mk(-tangent.y, tangent.x);
    const float len = vlen(normal);
    return len == 0.0f ? mk(0.0f, 0.0f) : smul(1.0f / len, normal);
}

struct StrokeOut {
    uint8_t* verbs;  // null: count only
    float* points;
    uint64_t v, q;   // cursors (uniform across the wave)
};

// One side of one contour of n >= 1 points at pts[0 .. n) (path.rs:174-215).
// `first_move`: the side's first output is a Move (else a Line).
template <bool EMIT>
__device__ void stroke_side(const float* pts, uint32_t n, bool closed, bool reverse, float width, bool first_move,
                            StrokeOut& o) {
    const int lane = threadIdx.x;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    auto at = [&](uint32_t k) { return mk(pts[2 * k], pts[2 * k + 1]); };
    const V2 first_point = closed == reverse ? at(0) : at(n - 1);
    V2 carry_pt = first_point;            // point of index i0 - 1 (for i0 = 0: first_point)
    V2 carry_n = mk(0.0f, 0.0f);          // normal of the last event so far
    bool first_out = true;
    for (uint32_t i0 = 0; i0 <= n; i0 += WAVE) {
        const uint32_t i = i0 + (uint32_t)lane;
        const bool in_range = i <= n;
        V2 pt = first_point;
        if (i < n) pt = at(reverse ? n - 1 - i : i);
        // previous point: lane - 1, or the carry for lane 0
        V2 prv = mk(__shfl_up(pt.x, 1, WAVE), __shfl_up(pt.y, 1, WAVE));
        if (lane == 0) prv = carry_pt;
        const bool event = in_range && (!veq(pt, prv) || i == n);
        const V2 nrm = event ? unit_normal(prv, pt) : mk(0.0f, 0.0f);
        // the previous event's normal
        const uint64_t ev = __ballot(event);
        const uint64_t before = ev & lt;
        const int src = before ? 63 - __clzll((long long)before) : 0;
        V2 pn = mk(__shfl(nrm.x, src, WAVE), __shfl(nrm.y, src, WAVE));
        if (!before) pn = carry_n;
        // join (path.rs:163-171) at the previous point
        V2 j0 = mk(0.0f, 0.0f), j1 = j0;
        uint32_t cnt = 0;
        if (event) {
            const float offset = 1.0f / (1.0f + dot(pn, nrm));
            if (fabsf(offset) > 2.0f) {
                j0 = add(prv, smul(0.5f * width, pn));
                j1 = add(prv, smul(0.5f * width, nrm));
                cnt = 2;
            } else {
                j0 = add(prv, smul(0.5f * width * offset, add(pn, nrm)));
                cnt = 1;
            }
        }
        const uint32_t c_incl = wave_incl_scan(cnt, OpAdd());
        const uint32_t total = __shfl(c_incl, WAVE - 1, WAVE);
        if (EMIT && cnt) {
            const uint64_t pos = o.q + (c_incl - cnt);
            const bool is_first = first_out && pos == o.q;
            o.verbs[o.v + (c_incl - cnt)] = (uint8_t)(is_first && first_move ? MOVE : LINE);
            o.points[2 * pos] = j0.x;
            o.points[2 * pos + 1] = j0.y;
            if (cnt == 2) {
                o.verbs[o.v + (c_incl - cnt) + 1] = (uint8_t)LINE;
                o.points[2 * pos + 2] = j1.x;
                o.points[2 * pos + 3] = j1.y;
            }
        }
        o.v += total;
        o.q += total;
        if (total) first_out = false;
        // carries for the next 64 indices
        const int lastl = WAVE - 1;
        carry_pt = mk(__shfl(pt.x, lastl, WAVE), __shfl(pt.y, lastl, WAVE));
        if (ev) {
            const int le = 63 - __clzll((long long)ev);
            carry_n = mk(__shfl(nrm.x, le, WAVE), __shfl(nrm.y, le, WAVE));
        }
    }
}

template <bool EMIT>
__device__ inline void put_close(StrokeOut& o) {
    if (EMIT && threadIdx.x == 0) o.verbs[o.v] = (uint8_t)CLOSE;
    o.v += 1;
}

template <bool EMIT>
__global__ __launch_bounds__(WAVE) void k_sstroke(const uint8_t* fverbs, const float* fpoints, const int64_t* foff_v,
                                                 const int64_t* foff_p, const float* widths, int64_t n, int64_t* cnt,
                                                 const int64_t* off_v, const int64_t* off_p, uint8_t* verbs,
                                                 float* points, const uint8_t* stroke_flags, PathSrc src,
                                                 const int64_t* off_w, float* weights) {
    const int64_t p = blockIdx.x;
    if (p >= n) return;
    const int lane = threadIdx.x;
    if (stroke_flags && !stroke_flags[p]) {
        // a fill of a draw list: its commands pass through to the fill unchanged
        const int64_t vb = src.verb_off[p], nv = src.verb_off[p + 1] - vb;
        const int64_t pb = src.point_off[p], np = src.point_off[p + 1] - pb;
        const int64_t wb = src.weight_off[p], nw = src.weight_off[p + 1] - wb;
        if (!EMIT) {
            if (lane == 0) {
                cnt[p] = nv;
                cnt[n + p] = np;
                cnt[2 * n + p] = nw;
            }
            return;
        }
        for (int64_t k = lane; k < nv; k += WAVE) verbs[off_v[p] + k] = src.verbs[vb + k];
        for (int64_t k = lane; k < 2 * np; k += WAVE) points[2 * off_p[p] + k] = src.points[2 * pb + k];
        for (int64_t k = lane; k < nw; k += WAVE) weights[off_w[p] + k] = src.weights[wb + k];
        return;
    }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint8_t* v = fverbs + foff_v[p];
    const float* pts = fpoints + 2 * foff_p[p];
    const uint32_t nv = (uint32_t)(foff_v[p + 1] - foff_v[p]);
    const float width = widths[p];
    StrokeOut o{EMIT ? verbs + off_v[p] : nullptr, EMIT ? points + 2 * off_p[p] : nullptr, 0, 0};
    // contour [cs, b) of point-bearing commands starting at point qs
    uint32_t cs = 0, qs = 0, closes = 0;
    bool closed = false;
    auto contour = [&](uint32_t b) {
        if (b <= cs) return;
        const uint32_t m = b - cs;
        const float* c = pts + 2 * qs;
        stroke_side<EMIT>(c, m, closed, false, width, true, o);
        if (closed) put_close<EMIT>(o);
        stroke_side<EMIT>(c, m, closed, true, width, closed, o);
        put_close<EMIT>(o);
    };
    for (uint32_t i0 = 0; i0 < nv; i0 += WAVE) {
        const uint32_t i = i0 + (uint32_t)lane;
        const int verb = i < nv ? v[i] : -1;
        const uint64_t bmask = __ballot(verb == MOVE || verb == CLOSE);
        const uint64_t cmask = __ballot(verb == CLOSE);
        uint64_t m = bmask;
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t b = i0 + (uint32_t)l;
            const bool is_close = (cmask >> l) & 1ull;
            const uint32_t closes_before = closes + (uint32_t)__popcll(cmask & ((1ull << l) - 1ull));
            if (is_close) closed = true;
            contour(b);
            // the next contour starts at the Move, or after the Close
            cs = is_close ? b + 1 : b;
            qs = cs - (closes_before + (is_close ? 1u : 0u));
        }
        closes += (uint32_t)__popcll(cmask);
    }
    contour(nv);
    (void)lt;
    if (!EMIT && lane == 0) {
        cnt[p] = (int64_t)o.v;
        cnt[n + p] = (int64_t)o.q;
        if (stroke_flags) cnt[2 * n + p] = 0;  // outlines hold no conics
    }
}

__global__ void k_zero_i64(int64_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

struct LoadI64 {
    const int64_t* p;
    __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct StoreI64 {
    int64_t* p;
    __device__ void operator()(int64_t i, int64_t excl, int64_t) const { p[i] = excl; }
};

// Exclusive offsets off[0 .. n] of cnt[0 .. n), the total landing in off[n].
hipError_t offsets_i64(Context* ctx, const int64_t* cnt, int64_t* off, int64_t n, hipStream_t s) {
    if (n <= SCAN_TILE) {
        hipLaunchKernelGGL((k_scan_single<int64_t, OpAdd, LoadI64, StoreI64>), dim3(1), dim3(SCAN_NT), 0, s,
                           LoadI64{cnt}, StoreI64{off}, n, OpAdd(), (int64_t)0, off + n);
        return hipGetLastError();
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipError_t e = ctx->scan_partials.ensure((size_t)nb * 16);
    if (e != hipSuccess) return e;
    int64_t* partials = ctx->scan_partials.as<int64_t>();
    hipLaunchKernelGGL((k_scan_reduce<int64_t, OpAdd, LoadI64>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s,
                       LoadI64{cnt}, n, partials, OpAdd(), (int64_t)0);
    hipLaunchKernelGGL((k_scan_partials<int64_t, OpAdd>), dim3(1), dim3(SCAN_NT), 0, s, partials, nb, OpAdd(),
                       (int64_t)0, off + n);
    hipLaunchKernelGGL((k_scan_down<int64_t, OpAdd, LoadI64, StoreI64>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s,
                       LoadI64{cnt}, StoreI64{off}, n, (const int64_t*)partials, OpAdd(), (int64_t)0);
    return hipGetLastError();
}

}  // namespace

int stroke_expand(Context* ctx, const SegInput& in, const float* widths, const uint8_t* stroke_flags,
                  StrokeOutline& o, hipStream_t s) {
    const int64_t n = in.n_paths;
    o = StrokeOutline{};
    if (n <= 0) return 0;
    if (in.n_segs != n || in.seg_raw || in.seg_path || in.path_seg_off || !widths) return -1;
    StrokeBuffers& B = ctx->stroke;
    HIPCHK_RET(B.cnt.ensure((size_t)n * 24));
    HIPCHK_RET(B.fv.ensure((size_t)(n + 1) * 8));
    HIPCHK_RET(B.fp.ensure((size_t)(n + 1) * 8));
    HIPCHK_RET(B.ov.ensure((size_t)(n + 1) * 8));
    HIPCHK_RET(B.op.ensure((size_t)(n + 1) * 8));
    HIPCHK_RET(B.ow.ensure((size_t)(n + 1) * 8));
    HIPCHK_RET(B.err.ensure(16));
    if (!B.host) HIPCHK_RET(hipHostMalloc(&B.host, 64));
    int64_t* h = reinterpret_cast<int64_t*>(B.host);
    PathSrc src{in.verbs, in.seg_verb_off, in.points, in.seg_point_off, in.weights, in.seg_weight_off};
    int64_t* cnt = B.cnt.as<int64_t>();
    uint32_t* err = B.err.as<uint32_t>();
    const uint32_t blocks = (uint32_t)n;  // one wave per path
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, B.err.as<int64_t>(), (int64_t)2);
    hipLaunchKernelGGL(k_sflatten<false>, dim3(blocks), dim3(WAVE), 0, s, src, stroke_flags, n, cnt, nullptr,
                       nullptr, nullptr, nullptr, err);
    HIPCHK_RET(offsets_i64(ctx, cnt, B.fv.as<int64_t>(), n, s));
    HIPCHK_RET(offsets_i64(ctx, cnt + n, B.fp.as<int64_t>(), n, s));
    HIPCHK_RET(hipMemcpyAsync(h, B.fv.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(h + 1, B.fp.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(h + 2, err, 4, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    if ((uint32_t)h[2] & ERR_INPUT) return -1;
    if ((uint32_t)h[2] & ERR_FLATTEN) return -5;
    const int64_t FV = h[0], FP = h[1];
    HIPCHK_RET(B.fverbs.ensure((size_t)std::max<int64_t>(FV, 1)));
    HIPCHK_RET(B.fpoints.ensure((size_t)std::max<int64_t>(FP, 1) * 8));
    hipLaunchKernelGGL(k_sflatten<true>, dim3(blocks), dim3(WAVE), 0, s, src, stroke_flags, n, cnt,
                       B.fv.as<int64_t>(), B.fp.as<int64_t>(), B.fverbs.as<uint8_t>(), B.fpoints.as<float>(), err);
    hipLaunchKernelGGL(k_sstroke<false>, dim3(blocks), dim3(WAVE), 0, s, B.fverbs.as<uint8_t>(),
                       B.fpoints.as<float>(), B.fv.as<int64_t>(), B.fp.as<int64_t>(), widths, n, cnt, nullptr,
                       nullptr, nullptr, nullptr, stroke_flags, src, nullptr, nullptr);
    HIPCHK_RET(offsets_i64(ctx, cnt, B.ov.as<int64_t>(), n, s));
    HIPCHK_RET(offsets_i64(ctx, cnt + n, B.op.as<int64_t>(), n, s));
    if (stroke_flags) HIPCHK_RET(offsets_i64(ctx, cnt + 2 * n, B.ow.as<int64_t>(), n, s));
    HIPCHK_RET(hipMemcpyAsync(h, B.ov.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(h + 1, B.op.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    if (stroke_flags) HIPCHK_RET(hipMemcpyAsync(h + 2, B.ow.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    const int64_t OV = h[0], OP = h[1], OW = stroke_flags ? h[2] : 0;
    HIPCHK_RET(B.overbs.ensure((size_t)std::max<int64_t>(OV, 1)));
    HIPCHK_RET(B.opoints.ensure((size_t)std::max<int64_t>(OP, 1) * 8));
    HIPCHK_RET(B.oweights.ensure((size_t)std::max<int64_t>(OW, 1) * 4));
    hipLaunchKernelGGL(k_sstroke<true>, dim3(blocks), dim3(WAVE), 0, s, B.fverbs.as<uint8_t>(),
                       B.fpoints.as<float>(), B.fv.as<int64_t>(), B.fp.as<int64_t>(), widths, n, cnt,
                       B.ov.as<int64_t>(), B.op.as<int64_t>(), B.overbs.as<uint8_t>(), B.opoints.as<float>(),
                       stroke_flags, src, B.ow.as<int64_t>(), B.oweights.as<float>());
    // stroke outlines hold no conics: with no fills every weight offset is zero
    if (!stroke_flags)
        hipLaunchKernelGGL(k_zero_i64, dim3((uint32_t)std::min<int64_t>((n + 256) / 256, 1024)), dim3(256), 0, s,
                           B.ow.as<int64_t>(), n + 1);
    HIPCHK_RET(hipGetLastError());
    o.in.n_segs = n;
    o.in.n_paths = n;
    o.in.verbs = B.overbs.as<uint8_t>();
    o.in.points = B.opoints.as<float>();
    o.in.weights = OW ? B.oweights.as<float>() : nullptr;
    o.in.seg_verb_off = B.ov.as<int64_t>();
    o.in.seg_point_off = B.op.as<int64_t>();
    o.in.seg_weight_off = B.ow.as<int64_t>();
    o.in.seg_xf = in.seg_xf;
    o.in.n_verbs = OV;
    o.in.n_points = OP;
    o.in.n_weights = OW;
    return 0;
}

int run_stroke(Context* ctx, const SegInput& in, const float* widths, const uint8_t* stroke_flags, OutputDesc& out,
               hipStream_t s, bool batched) {
    out.n_tiles = out.n_spans = 0;
    StrokeOutline o;
    int st = stroke_expand(ctx, in, widths, stroke_flags, o, s);
    if (st || in.n_paths <= 0) return st;
    if (batched)
        return run_batched(ctx, o.in, out, s, ctx->batched_waves_per_cu, ctx->batched_gran, ctx->batched_tc,
                           ctx->batched_graph);
    return run_plain(ctx, o.in, out, s);
}

}  // namespace ochre
