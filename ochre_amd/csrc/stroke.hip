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
// Each pass runs one lane per path, once to count and once to write, with a
// device scan between; a path's outline is produced by one lane in order, so
// its command sequence is the reference's exactly.
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

// Output cursor: counts when verbs == nullptr.
struct Emitter {
    uint8_t* verbs;
    float* points;
    int64_t v = 0, q = 0;
    __device__ void push(int verb, V2 p) {
        if (verbs) {
            verbs[v] = (uint8_t)verb;
            if (verb != CLOSE) {
                points[2 * q] = p.x;
                points[2 * q + 1] = p.y;
            }
        }
        v++;
        if (verb != CLOSE) q++;
    }
};

// flatten (path.rs:114-144).  Returns error bits: ERR_INPUT for an unknown
// verb or a path whose points/weights do not match its verbs, ERR_FLATTEN
// when a curve needs more than FLATTEN_LIMIT segments (the reference would
// not terminate).
__device__ uint32_t flatten_one(const PathSrc& in, int64_t p, Emitter& out) {
    const int64_t vb = in.verb_off[p], ve = in.verb_off[p + 1];
    int64_t g = in.point_off[p], wi = in.weight_off[p];
    const int64_t ge = in.point_off[p + 1], we = in.weight_off[p + 1];
    V2 last = mk(0.0f, 0.0f);
    for (int64_t i = vb; i < ve; i++) {
        const int verb = in.verbs[i];
        const int k = verb_points(verb);
        if (k < 0 || g + k > ge || (verb == CONIC && wi >= we)) return ERR_INPUT;
        V2 pt[3] = {mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};  // static indices: registers
        if (k > 0) pt[0] = mk(in.points[2 * g], in.points[2 * g + 1]);
        if (k > 1) pt[1] = mk(in.points[2 * (g + 1)], in.points[2 * (g + 1) + 1]);
        if (k > 2) pt[2] = mk(in.points[2 * (g + 2)], in.points[2 * (g + 2) + 1]);
        if (verb == MOVE || verb == LINE) {
            out.push(verb, pt[0]);
        } else if (verb == QUAD || verb == CUBIC) {
            const float dt = verb == QUAD ? quad_dt(last, pt[0], pt[1]) : cubic_dt(last, pt[0], pt[1], pt[2]);
            float t = 0.0f;
            uint32_t n = 0;
            while (t < 1.0f) {
                t = fminf(t + dt, 1.0f);
                out.push(LINE, verb == QUAD ? quad_at(t, last, pt[0], pt[1]) : cubic_at(t, last, pt[0], pt[1], pt[2]));
                if (++n >= FLATTEN_LIMIT) return ERR_FLATTEN;
            }
        } else if (verb == CONIC) {
            if (!conic_walk(last, pt[0], pt[1], in.weights[wi], [&](V2 r) { out.push(LINE, r); })) return ERR_FLATTEN;
            wi++;
        } else {
            out.push(CLOSE, mk(0.0f, 0.0f));
        }
        if (k == 1) last = pt[0];
        else if (k == 2) last = pt[1];
        else if (k == 3) last = pt[2];
        g += k;
    }
    return g == ge && wi == we ? 0u : (uint32_t)ERR_INPUT;
}

// join (path.rs:163-171)
template <class Sink>
__device__ inline void stroke_join(float width, V2 prev_n, V2 next_n, V2 point, Sink&& sink) {
    const float offset = 1.0f / (1.0f + dot(prev_n, next_n));
    if (fabsf(offset) > 2.0f) {
        sink(add(point, smul(0.5f * width, prev_n)));
        sink(add(point, smul(0.5f * width, next_n)));
    } else {
        sink(add(point, smul(0.5f * width * offset, add(prev_n, next_n))));
    }
}

// offset (path.rs:174-215) over a contour of n >= 1 points c[0..n).
template <class Sink>
__device__ inline void stroke_offset(const float* c, int64_t n, bool closed, bool reverse, float width, Sink&& sink) {
    auto at = [&](int64_t k) { return mk(c[2 * k], c[2 * k + 1]); };
    const V2 first_point = closed == reverse ? at(0) : at(n - 1);
    V2 prev_point = first_point;
    V2 prev_normal = mk(0.0f, 0.0f);
    for (int64_t i = 0; i <= n; i++) {
        const V2 next_point = i < n ? at(reverse ? n - i - 1 : i) : first_point;
        if (!veq(next_point, prev_point) || i == n) {
            const V2 tangent = sub(next_point, prev_point);
            const V2 normal = mk(-tangent.y, tangent.x);
            const float len = vlen(normal);
            const V2 unit = len == 0.0f ? mk(0.0f, 0.0f) : smul(1.0f / len, normal);
            stroke_join(width, prev_normal, unit, prev_point, sink);
            prev_point = next_point;
            prev_normal = unit;
        }
    }
}

// stroke (path.rs:152-274) of one flattened path.  A contour holds only Move
// and Line commands (a Close ends it and is skipped), so its points are
// contiguous: command cs of the polyline is point qs, and the contour
// [cs, ce) is points [qs, qs + ce - cs).  `closed` is never cleared once a
// Close has been seen, as in the reference.
__device__ void stroke_one(const uint8_t* verbs, const float* pts, int64_t nv, float width, Emitter& out) {
    int64_t cs = 0, ce = 0, qs = 0, q = 0;
    bool closed = false;
    for (int64_t i = 0;; i++) {
        const int verb = i < nv ? verbs[i] : -1;
        if (verb == CLOSE) closed = true;
        if ((verb < 0 || verb == MOVE || verb == CLOSE) && cs != ce) {
            const float* c = pts + 2 * qs;
            const int64_t n = ce - cs;
            bool first = true;
            stroke_offset(c, n, closed, false, width, [&](V2 r) {
                out.push(first ? MOVE : LINE, r);
                first = false;
            });
            if (closed) out.push(CLOSE, mk(0.0f, 0.0f));
            first = true;
            stroke_offset(c, n, closed, true, width, [&](V2 r) {
                out.push(first && closed ? MOVE : LINE, r);
                first = false;
            });
            out.push(CLOSE, mk(0.0f, 0.0f));
        }
        if (verb < 0) break;
        if (verb == MOVE) {
            cs = ce;
            ce = cs + 1;
            qs = q;
        } else if (verb == LINE) {
            ce += 1;
        } else {
            cs = ce + 1;
            ce = cs;
            closed = true;
            qs = q;
        }
        if (verb != CLOSE) q++;
    }
}

// Per-path sizes: cnt[0 .. n) verbs and cnt[n .. 2n) points.
__global__ void k_flatten_paths(PathSrc in, int64_t n, int64_t* cnt, const int64_t* off_v, const int64_t* off_p,
                                uint8_t* verbs, float* points, uint32_t* error) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    Emitter e{verbs ? verbs + off_v[p] : nullptr, verbs ? points + 2 * off_p[p] : nullptr};
    const uint32_t err = flatten_one(in, p, e);
    if (err) atomicOr(error, err);
    if (!verbs) {
        cnt[p] = e.v;
        cnt[n + p] = e.q;
    }
}

__global__ void k_stroke_paths(const uint8_t* fverbs, const float* fpoints, const int64_t* foff_v,
                               const int64_t* foff_p, const float* widths, int64_t n, int64_t* cnt,
                               const int64_t* off_v, const int64_t* off_p, uint8_t* verbs, float* points) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    Emitter e{verbs ? verbs + off_v[p] : nullptr, verbs ? points + 2 * off_p[p] : nullptr};
    stroke_one(fverbs + foff_v[p], fpoints + 2 * foff_p[p], foff_v[p + 1] - foff_v[p], widths[p], e);
    if (!verbs) {
        cnt[p] = e.v;
        cnt[n + p] = e.q;
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

int stroke_expand(Context* ctx, const SegInput& in, const float* widths, StrokeOutline& o, hipStream_t s) {
    const int64_t n = in.n_paths;
    o = StrokeOutline{};
    if (n <= 0) return 0;
    if (in.n_segs != n || in.seg_raw || in.seg_path || in.path_seg_off || !widths) return -1;
    StrokeBuffers& B = ctx->stroke;
    HIPCHK_RET(B.cnt.ensure((size_t)n * 16));
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
    const uint32_t blocks = (uint32_t)((n + 63) / 64);
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, s, B.err.as<int64_t>(), (int64_t)2);
    hipLaunchKernelGGL(k_flatten_paths, dim3(blocks), dim3(64), 0, s, src, n, cnt, nullptr, nullptr, nullptr,
                       nullptr, err);
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
    hipLaunchKernelGGL(k_flatten_paths, dim3(blocks), dim3(64), 0, s, src, n, cnt, B.fv.as<int64_t>(),
                       B.fp.as<int64_t>(), B.fverbs.as<uint8_t>(), B.fpoints.as<float>(), err);
    hipLaunchKernelGGL(k_stroke_paths, dim3(blocks), dim3(64), 0, s, B.fverbs.as<uint8_t>(), B.fpoints.as<float>(),
                       B.fv.as<int64_t>(), B.fp.as<int64_t>(), widths, n, cnt, nullptr, nullptr, nullptr, nullptr);
    HIPCHK_RET(offsets_i64(ctx, cnt, B.ov.as<int64_t>(), n, s));
    HIPCHK_RET(offsets_i64(ctx, cnt + n, B.op.as<int64_t>(), n, s));
    HIPCHK_RET(hipMemcpyAsync(h, B.ov.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(h + 1, B.op.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    const int64_t OV = h[0], OP = h[1];
    HIPCHK_RET(B.overbs.ensure((size_t)std::max<int64_t>(OV, 1)));
    HIPCHK_RET(B.opoints.ensure((size_t)std::max<int64_t>(OP, 1) * 8));
    hipLaunchKernelGGL(k_stroke_paths, dim3(blocks), dim3(64), 0, s, B.fverbs.as<uint8_t>(), B.fpoints.as<float>(),
                       B.fv.as<int64_t>(), B.fp.as<int64_t>(), widths, n, cnt, B.ov.as<int64_t>(),
                       B.op.as<int64_t>(), B.overbs.as<uint8_t>(), B.opoints.as<float>());
    // outlines hold no conics: every weight offset is zero
    hipLaunchKernelGGL(k_zero_i64, dim3((uint32_t)std::min<int64_t>((n + 256) / 256, 1024)), dim3(256), 0, s,
                       B.ow.as<int64_t>(), n + 1);
    HIPCHK_RET(hipGetLastError());
    o.in.n_segs = n;
    o.in.n_paths = n;
    o.in.verbs = B.overbs.as<uint8_t>();
    o.in.points = B.opoints.as<float>();
    o.in.weights = nullptr;
    o.in.seg_verb_off = B.ov.as<int64_t>();
    o.in.seg_point_off = B.op.as<int64_t>();
    o.in.seg_weight_off = B.ow.as<int64_t>();
    o.in.seg_xf = in.seg_xf;
    o.in.n_verbs = OV;
    o.in.n_points = OP;
    o.in.n_weights = 0;
    return 0;
}

int run_stroke(Context* ctx, const SegInput& in, const float* widths, OutputDesc& out, hipStream_t s, bool batched) {
    out.n_tiles = out.n_spans = 0;
    StrokeOutline o;
    int st = stroke_expand(ctx, in, widths, o, s);
    if (st || in.n_paths <= 0) return st;
    if (batched)
        return run_batched(ctx, o.in, out, s, ctx->batched_waves_per_cu, ctx->batched_gran, ctx->batched_tc,
                           ctx->batched_graph);
    return run_plain(ctx, o.in, out, s);
}

}  // namespace ochre
