// Batched pipeline: one persistent launch works through a queue of paths,
// one wave64 per path, everything for a path on one CU:
//   flatten   lane per command (64 at a time, wave scans carry the point /
//             conic / last-Move state), lines to per-wave scratch
//   count     lane per line, lanes refilled from the line list as they finish
//             (no lane idles while one long line is walked); marks occupied
//             tiles and sums tile-row crossings in LDS
//   allocate  tile and span counts -> one atomic per path for its output
//             range; spans written directly (integer winding)
//   emit      lane per line again, each increment (pixel, area, height)
//             written to scratch at its emission index
//   commit    64 increments at a time in emission order into LDS tile
//             accumulators; lanes hitting the same pixel are found by ballot
//             matching and commit in lane order, so every pixel's f32 sum is
//             taken in the reference's emission order (rasterizer.rs:221-228)
//   finalize  row carry (rasterizer.rs:233-250) and quantisation, 64 B stores
// Paths too large for a wave (grid > 32x32 tiles or > 8192 lines) are
// deferred to the plain pipeline by the host driver.
//
// With the dense mode on (the default), k_dense (dense.inc) runs first and
// takes every path whose tile grid fits its LDS grid in a single walk with
// order-independent accumulation; k_fused then works through the paths it
// handed on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "pipeline.hpp"
#include "raster_common.hpp"
#include "scan.hpp"

namespace ochre {

constexpr int F_LCAP = 8192;   // lines per path
constexpr int F_ICAP = 8192;   // increments per commit batch
constexpr int F_GRID = 32;     // tile grid side (256 px)
constexpr int F_TSTRIDE = 65;  // accumulator stride per tile (bank spread)
constexpr size_t F_SCRATCH = (size_t)F_LCAP * 16 + (size_t)(F_LCAP + 1) * 4 + 12 + (size_t)F_ICAP * 16;
constexpr uint32_t F_INVALID = 0xffffffffu;

struct FusedIn {
    int64_t n_paths;
    const uint8_t* verbs;
    const int64_t* verb_off;
    const float* points;
    const int64_t* point_off;
    const float* weights;
    const int64_t* weight_off;
    const Xform* xf;
    int64_t n_points, n_weights;
    // when set, the queue runs over ids[0 .. *n_ids) (paths handed on by
    // k_dense) instead of 0 .. n_paths
    const uint32_t* ids = nullptr;
    const uint32_t* n_ids = nullptr;
    // more handed-on paths than this go to the plain pipeline whole: one wave
    // per path is slow on the large paths the dense kernels hand on, and the
    // plain pipeline's fixed cost is repaid from a few hundred of them
    uint32_t defer_above = 0xffffffffu;
};

constexpr uint32_t F_DEFER_ABOVE = 512;  // handed-on paths beyond which k_fused defers them all

struct FusedOut {
    int16_t* tile_xy;
    uint8_t* tile_alpha;
    float* tile_cov;  // optional: pre-quantisation coverage min(|accum + area|, 1), 64 per tile
    int64_t tile_cap;
    int16_t* span_xy;
    uint16_t* span_w;
    int64_t span_cap;
    int64_t* pts;
    int32_t* ptc;
    int64_t* pss;
    int32_t* psc;
};

struct FusedCtl {
    uint32_t queue;     // path queue of the first kernel of the launch
    uint32_t deferred;  // paths handed to the plain pipeline
    uint32_t error;
    uint32_t dq;        // paths handed on by k_dense
    unsigned long long tiles;
    unsigned long long spans;
    uint32_t queue2;    // queue of k_fused over the handed-on paths
    uint32_t queue3;    // queue of the large-grid dense kernel
    uint32_t dq2;       // paths handed on by the second dense instance
    uint32_t queue4;    // queue of the third dense instance
    uint32_t dq3;       // paths handed on by the third dense instance
    uint32_t pad[51];
};

__global__ void k_fused_reset(FusedCtl* c) {
    if (threadIdx.x == 0) {
        c->queue = 0;
        c->deferred = 0;
        c->error = 0;
        c->dq = 0;
        c->tiles = 0;
        c->spans = 0;
        c->queue2 = 0;
        c->queue3 = 0;
        c->dq2 = 0;
        c->queue4 = 0;
        c->dq3 = 0;
    }
}

struct OpMax64 {
    __device__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};

__device__ inline V2 fload(const FusedIn& in, int64_t g, const Xform& X) {
    return apply(X, mk(in.points[2 * g], in.points[2 * g + 1]));
}

template <int TC>
struct LdsState {
    uint32_t occ[F_GRID];
    uint32_t row_rank[F_GRID + 1];
    uint32_t chunk_tile[TC];
    // tsign is dead once spans are written; the accumulators live after that
    union {
        int32_t tsign[F_GRID * F_GRID];
        struct {
            float acc_a[TC * F_TSTRIDE];
            float acc_h[TC * F_TSTRIDE];
            float cin[TC * 8];
        } a;
        struct {  // fixed-point area / height sums of half a chunk (the coverage adjoint)
            unsigned long long fa[TC / 2 * 64];
            unsigned long long fh[TC / 2 * 64];
        } x;
    } u;
};

// Rank of tile (tx, ty) (grid-local) among the path's occupied tiles in
// row-major order.
template <class S>
__device__ inline uint32_t tile_rank(const S& s, int tx, int ty) {
    return s.row_rank[ty] + (uint32_t)__popc(s.occ[ty] & ((1u << tx) - 1u));
}

// Two 64-line windows of the line list held in registers (lane i holds lines
// base + i and base + 64 + i), so refilling lanes read a line with __shfl
// instead of a dependent global load.
struct LineWindow {
    float4 w0, w1;
    uint32_t o0, o1;  // per-line increment offsets (emit walk)
    uint32_t base;
    __device__ void load(const float4* lines, const uint32_t* off, uint32_t L, uint32_t b, int lane) {
        base = b;
        const uint32_t i0 = b + lane, i1 = b + 64 + lane;
        w0 = i0 < L ? lines[i0] : make_float4(0, 0, 0, 0);
        w1 = i1 < L ? lines[i1] : make_float4(0, 0, 0, 0);
        if (off) {
            o0 = i0 < L ? off[i0] : 0u;
            o1 = i1 < L ? off[i1] : 0u;
        }
    }
    // Fetches line `my` (must be in [base, base + 128)) for every lane.
    __device__ float4 get(uint32_t my, uint32_t* o, bool with_off) const {
        const uint32_t src = my - base;
        const int sl = (int)(src & 63u);
        float4 a, b;
        a.x = __shfl(w0.x, sl, WAVE);
        a.y = __shfl(w0.y, sl, WAVE);
        a.z = __shfl(w0.z, sl, WAVE);
        a.w = __shfl(w0.w, sl, WAVE);
        b.x = __shfl(w1.x, sl, WAVE);
        b.y = __shfl(w1.y, sl, WAVE);
        b.z = __shfl(w1.z, sl, WAVE);
        b.w = __shfl(w1.w, sl, WAVE);
        if (with_off) {
            uint32_t x0 = __shfl(o0, sl, WAVE), x1 = __shfl(o1, sl, WAVE);
            *o = src < 64 ? x0 : x1;
        }
        return src < 64 ? a : b;
    }
    // Slides by 64 once `next` has passed the first window.
    __device__ void advance(const float4* lines, const uint32_t* off, uint32_t L, uint32_t next, int lane) {
        if (next >= base + 64) {
            w0 = w1;
            o0 = o1;
            base += 64;
            const uint32_t i1 = base + 64 + lane;
            w1 = i1 < L ? lines[i1] : make_float4(0, 0, 0, 0);
            if (off) o1 = i1 < L ? off[i1] : 0u;
        }
    }
};

// Adjoint inputs/outputs (see adjoint.hip).  The target of path p is its
// tiles [tt_start[p], +tt_count[p]) (row-major, 64 f32 coverages each) and
// spans [ts_start[p], +ts_count[p]).
struct AdjArgs {
    const int64_t* tt_start;
    const int32_t* tt_count;
    const int16_t* t_xy;
    const float* t_cov;
    const int64_t* ts_start;
    const int32_t* ts_count;
    const int16_t* ts_xy;
    const uint16_t* ts_w;
    float* loss;      // [n_paths]
    float* d_points;  // [n_points * 2]
    float* d_xf;      // [n_paths * 6]
};

constexpr int A_TCAP = 256;   // tiles per path
constexpr int A_VCAP = 1024;  // commands per path
constexpr int A_PCAP = 4096;  // points per path
constexpr int A_SCAP = 1024;  // spans per path
constexpr size_t A_OFF_TILES = 0;                                            // float[A_TCAP][128]: g, G
constexpr size_t A_OFF_CMD = A_OFF_TILES + (size_t)A_TCAP * 128 * 4;         // float[A_VCAP][12]
constexpr size_t A_OFF_CMDOFF = A_OFF_CMD + (size_t)A_VCAP * 12 * 4;         // uint32[A_VCAP + 1]
constexpr size_t A_OFF_DW = A_OFF_CMDOFF + (size_t)(A_VCAP + 4) * 4;         // float2[A_PCAP]
constexpr size_t A_OFF_PCMD = A_OFF_DW + (size_t)A_PCAP * 8;                 // uint32[A_PCAP]
constexpr size_t A_OFF_SPANS = A_OFF_PCMD + (size_t)A_PCAP * 4;              // uint32[A_SCAP][2]
constexpr size_t A_OFF_L64 = (A_OFF_SPANS + (size_t)A_SCAP * 8 + 255) & ~(size_t)255;  // double2[F_LCAP][2]
constexpr size_t A_SCRATCH = A_OFF_L64 + (size_t)F_LCAP * 32;

template <int TC>
__device__ void adjoint_backward(const FusedIn& in, const AdjArgs& adj, FusedCtl* ctl, uint8_t* sc, uint8_t* asc,
                                 LdsState<TC>& s, int64_t p, uint32_t L, uint32_t NT, uint32_t NS, int32_t tx0,
                                 int32_t ty0, int32_t GW, int32_t GH);

// The adjoint evaluates its forward in path-local device coordinates: the
// transform's offset less a tile-aligned origin (the first point's tile), so
// that f32 coordinates stay small.  At 1000 px an f32 coordinate carries
// ~6e-5 px of rounding, enough to move a pixel's |v| across the clamp at 1
// relative to the exact coverage; local coordinates cut that tenfold.  A
// whole-tile shift leaves the pixel and tile grid unchanged, and the
// derivatives with respect to the offset are those of the original one.
// (ox, oy) is the origin in tiles.
__device__ inline Xform adjoint_local_xform(const FusedIn& in, int64_t p, int32_t* ox, int32_t* oy) {
    Xform X = in.xf[p];
    *ox = *oy = 0;
    const int64_t pb = in.point_off[p];
    if (in.point_off[p + 1] > pb) {
        const V2 w = apply(X, mk(in.points[2 * pb], in.points[2 * pb + 1]));
        if (fabsf(w.x) < 1.0e6f && fabsf(w.y) < 1.0e6f) {
            *ox = (int32_t)floorf(w.x * 0.125f);
            *oy = (int32_t)floorf(w.y * 0.125f);
            X.ox -= (float)(*ox * TILE);
            X.oy -= (float)(*oy * TILE);
        }
    }
    return X;
}

template <int TC, bool ADJ>
__device__ void fused_path(const FusedIn& in, const FusedOut& out, FusedCtl* ctl, uint32_t* deferred,
                           uint8_t* sc, LdsState<TC>& s, int64_t p, const AdjArgs& adj, uint8_t* asc) {
    const int lane = threadIdx.x;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    float4* lines = reinterpret_cast<float4*>(sc);
    uint32_t* incoff = reinterpret_cast<uint32_t*>(sc + (size_t)F_LCAP * 16);
    float4* incs = reinterpret_cast<float4*>(sc + (size_t)F_LCAP * 16 + (size_t)(F_LCAP + 1) * 4 + 12);

    const int64_t vb = in.verb_off[p], ve = in.verb_off[p + 1];
    const int64_t pb = in.point_off[p], pe = in.point_off[p + 1];
    const int64_t wb = in.weight_off[p], we = in.weight_off[p + 1];
    int32_t aox = 0, aoy = 0;
    const Xform X = ADJ ? adjoint_local_xform(in, p, &aox, &aoy) : in.xf[p];

    // ---------------------------------------------------------- flatten
    uint32_t L = 0;
    uint32_t cpts = 0, ccon = 0;
    int64_t cmove = -1;
    bool overflow = false;
    bool bad = false, flat_bad = false;
    BoxAcc box;
    bool any_line = false;
    for (int64_t cb = vb; cb < ve; cb += WAVE) {
        const int64_t i = cb + lane;
        const bool valid = i < ve;
        int verb = valid ? in.verbs[i] : (int)CLOSE;
        int np = verb_points(verb);
        bool lane_bad = valid && np < 0;
        if (np < 0) np = 0;
        const uint32_t np_incl = wave_incl_scan((uint32_t)np, OpAdd());
        const uint32_t cc = (valid && verb == CONIC) ? 1u : 0u;
        const uint32_t cc_incl = wave_incl_scan(cc, OpAdd());
        const int64_t g = pb + cpts + (np_incl - (uint32_t)np);
        const int64_t mv = (valid && verb == MOVE) ? g : (int64_t)-1;
        const int64_t mv_incl = wave_incl_scan(mv, OpMax64());
        int64_t mv_excl = shfl_up_any(mv_incl, 1);
        if (lane == 0) mv_excl = -1;
        mv_excl = mv_excl > cmove ? mv_excl : cmove;
        Cmd c;
        c.verb = verb;
        c.path = (int32_t)p;
        c.last_in_path = (i == ve - 1);
        c.bad = false;
        if (valid && g + np > pe) lane_bad = true;
        if (valid && !lane_bad) {
            // static indices keep the command in registers
            if (np > 0) c.pt[0] = fload(in, g, X);
            if (np > 1) c.pt[1] = fload(in, g + 1, X);
            if (np > 2) c.pt[2] = fload(in, g + 2, X);
            c.w = 0.0f;
            if (verb == CONIC) {
                int64_t wi = wb + ccon + (cc_incl - 1);
                if (wi < we) c.w = in.weights[wi];
                else lane_bad = true;
            }
            c.start = g > pb ? fload(in, g - 1, X) : mk(0.0f, 0.0f);
            c.first = mv_excl >= 0 ? fload(in, mv_excl, X) : mk(0.0f, 0.0f);
        }
        // A certain overflow of the line list, from the curves' steps alone
        // (t advances by at most dt + 1e-7 per step of the flatten loop): the
        // path is handed on before any of its curves is evaluated here.
        uint32_t lower = 0;
        if (valid && !lane_bad && (verb == QUAD || verb == CUBIC)) {
            const float dt = verb == QUAD ? quad_dt(c.start, c.pt[0], c.pt[1])
                                          : cubic_dt(c.start, c.pt[0], c.pt[1], c.pt[2]);
            const float est = 0.99f / (dt + 1e-7f);
            if (est > 0.0f) lower = est >= (float)F_LCAP ? (uint32_t)F_LCAP + 1u : (uint32_t)est;
        }
        const uint32_t lower_tot = __shfl(wave_incl_scan(lower, OpAdd()), WAVE - 1, WAVE);
        if (L + lower_tot > (uint32_t)F_LCAP) {
            if (lane == 0) deferred[atomicAdd(&ctl->deferred, 1u)] = (uint32_t)p;  // (the adjoint's large path)
            return;
        }
        uint32_t n = 0;
        if (valid && !lane_bad) {
            bool ok = command_lines(c, [&](V2 a, V2 b) {
                n++;
                box.add(a);
                box.add(b);
            });
            if (!ok) flat_bad = true;
        }
        if (n) any_line = true;
        const uint32_t n_incl = wave_incl_scan(n, OpAdd());
        const uint32_t chunk_total = __shfl(n_incl, WAVE - 1, WAVE);
        if (__ballot(lane_bad)) bad = true;
        if (!overflow && L + chunk_total > (uint32_t)F_LCAP) overflow = true;
        if (!overflow && valid && !lane_bad && n) {
            uint32_t o = L + n_incl - n;
            command_lines(c, [&](V2 a, V2 b) { lines[o++] = make_float4(a.x, a.y, b.x, b.y); });
        }
        if (ADJ && valid && i - vb < A_VCAP) {
            reinterpret_cast<uint32_t*>(asc + A_OFF_CMDOFF)[i - vb] = L + n_incl - n;
            if (i == ve - 1) reinterpret_cast<uint32_t*>(asc + A_OFF_CMDOFF)[i - vb + 1] = L + n_incl;
            for (int k = 0; k < np; k++)
                if (g - pb + k < A_PCAP) reinterpret_cast<uint32_t*>(asc + A_OFF_PCMD)[g - pb + k] = (uint32_t)(i - vb);
        }
        L += chunk_total;
        cpts += __shfl(np_incl, WAVE - 1, WAVE);
        ccon += __shfl(cc_incl, WAVE - 1, WAVE);
        const int64_t last_mv = shfl_any(mv_incl, WAVE - 1);
        cmove = last_mv > cmove ? last_mv : cmove;
    }
    if (bad || (cpts != (uint32_t)(pe - pb))) {
        if (lane == 0) atomicOr(&ctl->error, (uint32_t)ERR_INPUT);
        return;
    }
    if (__ballot(flat_bad)) {
        if (lane == 0) atomicOr(&ctl->error, (uint32_t)ERR_FLATTEN);
        return;
    }
    // path bounding box of all line end points
    const bool has_lines = __ballot(any_line) != 0;
    const bool full = __ballot(box.full) != 0;
    float bminx = box.minx, bminy = box.miny, bmaxx = box.maxx, bmaxy = box.maxy;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        bminx = fminf(bminx, __shfl_xor(bminx, d, WAVE));
        bminy = fminf(bminy, __shfl_xor(bminy, d, WAVE));
        bmaxx = fmaxf(bmaxx, __shfl_xor(bmaxx, d, WAVE));
        bmaxy = fmaxf(bmaxy, __shfl_xor(bmaxy, d, WAVE));
    }
    int32_t tx0 = 0, ty0 = 0, GW = 0, GH = 0;
    if (has_lines && !full && bminx <= bmaxx)
        grid_from_box(true, false, (int32_t)floorf(bminx), (int32_t)floorf(bminy), (int32_t)floorf(bmaxx),
                      (int32_t)floorf(bmaxy), &tx0, &ty0, &GW, &GH);
    if (overflow || full || GW > F_GRID || GH > F_GRID ||
        (ADJ && (ve - vb > A_VCAP || pe - pb > A_PCAP))) {
        if (lane == 0) deferred[atomicAdd(&ctl->deferred, 1u)] = (uint32_t)p;  // (the adjoint's large path)
        return;
    }

    // ------------------------------------------------------- count walk
    for (int k = lane; k < F_GRID * F_GRID; k += WAVE) s.u.tsign[k] = 0;
    if (lane < F_GRID) s.occ[lane] = 0;
    __syncthreads();
    bool bbox_err = false, walk_err = false;
    {
        // a lane per line at tile granularity (line_tiles: the tiles, tile-row
        // crossings and increments from the DDA's two crossing-parameter
        // chains); a line it does not cover is walked pixel by pixel
        auto mark = [&](int32_t tx, int32_t ty) {
            const int lx = tx - tx0, ly = ty - ty0;
            if ((unsigned)lx < (unsigned)GW && (unsigned)ly < (unsigned)GH) atomicOr(&s.occ[ly], 1u << lx);
            else bbox_err = true;
        };
        auto cross = [&](int32_t tx, int32_t ty, int sign) {
            const int lx = tx - tx0, ly = ty - ty0;
            if ((unsigned)lx < (unsigned)GW && (unsigned)ly < (unsigned)GH) atomicAdd(&s.u.tsign[ly * F_GRID + lx], sign);
            else bbox_err = true;
        };
        for (uint32_t line = (uint32_t)lane; line < L; line += WAVE) {
            const float4 ln = lines[line];
            const V2 a = mk(ln.x, ln.y), point = mk(ln.z, ln.w);
            uint32_t cnt = 0;
            if (line_tiles(a, point, cnt, mark, cross)) {
                incoff[line] = cnt;
                continue;
            }
            if (veq(point, a)) {  // rasterizer.rs:73
                incoff[line] = 0;
                continue;
            }
            Dda d;
            dda_init(d, a, point);
            int16_t ty_prev = div8(d.y), cur_tx = div8(d.x), cur_ty = ty_prev;
            mark(cur_tx, cur_ty);
            for (uint32_t steps = 0;;) {
                cnt++;
                bool term = dda_step(d);
                if (term) {  // rasterizer.rs:118-121
                    d.x = f2i16(floorf(point.x));
                    d.y = f2i16(floorf(point.y));
                }
                const int16_t ty = div8(d.y);
                if (ty != ty_prev) {  // rasterizer.rs:123-131
                    cross(div8(d.x), ty_prev < ty ? ty_prev : ty, (int)(int8_t)(int16_t)(ty - ty_prev));
                    ty_prev = ty;
                }
                if (!term && ++steps >= WALK_LIMIT) {
                    walk_err = true;
                    term = true;
                }
                if (term) break;
                const int16_t tx = div8(d.x);
                if (tx != cur_tx || ty != cur_ty) {
                    cur_tx = tx;
                    cur_ty = ty;
                    mark(tx, ty);
                }
            }
            incoff[line] = cnt;
        }
    }
    __syncthreads();
    if (__ballot(bbox_err) || __ballot(walk_err)) {
        if (lane == 0) atomicOr(&ctl->error, __ballot(bbox_err) ? (uint32_t)ERR_BBOX : (uint32_t)ERR_FLATTEN);
        return;
    }

    // --------------------------------------------- tiles, spans, output
    // lane = tile row
    const bool row_lane = lane < GH;
    const uint32_t row_tiles = row_lane ? (uint32_t)__popc(s.occ[lane]) : 0u;
    const uint32_t rt_incl = wave_incl_scan(row_tiles, OpAdd());
    if (row_lane) s.row_rank[lane] = rt_incl - row_tiles;
    const uint32_t NT = __shfl(rt_incl, WAVE - 1, WAVE);
    int32_t row_sum = 0;
    if (row_lane)
        for (int x = 0; x < GW; x++) row_sum += s.u.tsign[lane * F_GRID + x];
    const int32_t rs_incl = wave_incl_scan(row_sum, OpAdd());
    const int32_t w_before = rs_incl - row_sum;  // winding of all crossings in earlier rows
    // spans (rasterizer.rs:252-264): between consecutive tiles of a row
    uint32_t row_spans = 0;
    if (row_lane) {
        int32_t wnd = w_before, w_prev = 0, prev = -1;
        const uint32_t o = s.occ[lane];
        for (int x = 0; x < GW; x++) {
            wnd += s.u.tsign[lane * F_GRID + x];
            if ((o >> x) & 1u) {
                if (prev >= 0 && x > prev + 1 && w_prev != 0) row_spans++;
                prev = x;
                w_prev = wnd;
            }
        }
    }
    const uint32_t rsp_incl = wave_incl_scan(row_spans, OpAdd());
    const uint32_t NS = __shfl(rsp_incl, WAVE - 1, WAVE);
    if (ADJ) {
        if (NT > (uint32_t)A_TCAP || NS > (uint32_t)A_SCAP) {
            if (lane == 0) deferred[atomicAdd(&ctl->deferred, 1u)] = (uint32_t)p;  // the large path
            return;
        }
        if (row_lane && row_spans) {
            uint32_t si = rsp_incl - row_spans;
            int32_t wnd = w_before, w_prev = 0, prev = -1;
            const uint32_t o = s.occ[lane];
            uint32_t* sp = reinterpret_cast<uint32_t*>(asc + A_OFF_SPANS);
            for (int x = 0; x < GW; x++) {
                wnd += s.u.tsign[lane * F_GRID + x];
                if ((o >> x) & 1u) {
                    if (prev >= 0 && x > prev + 1 && w_prev != 0) {
                        sp[2 * si] = (uint32_t)lane;                                  // grid row
                        sp[2 * si + 1] = (uint32_t)(prev + 1) | ((uint32_t)x << 16);  // tile columns [a, b)
                        si++;
                    }
                    prev = x;
                    w_prev = wnd;
                }
            }
        }
    }
    const uint32_t NT_out = NT ? NT : 1u;
    unsigned long long tb = 0, sb = 0;
    if (!ADJ && lane == 0) {
        tb = atomicAdd(&ctl->tiles, (unsigned long long)NT_out);
        sb = atomicAdd(&ctl->spans, (unsigned long long)NS);
        out.pts[p] = (int64_t)tb;
        out.ptc[p] = (int32_t)NT_out;
        out.pss[p] = (int64_t)sb;
        out.psc[p] = (int32_t)NS;
    }
    tb = __shfl(tb, 0, WAVE);
    sb = __shfl(sb, 0, WAVE);
    if (!ADJ && (tb + NT_out > (unsigned long long)out.tile_cap || sb + NS > (unsigned long long)out.span_cap)) {
        if (lane == 0) atomicOr(&ctl->error, (uint32_t)ERR_CAPACITY);
        return;
    }
    if (!ADJ && row_lane && row_spans) {
        uint64_t si = sb + (rsp_incl - row_spans);
        int32_t wnd = w_before, w_prev = 0, prev = -1;
        const uint32_t o = s.occ[lane];
        const int32_t ty = ty0 + lane;
        for (int x = 0; x < GW; x++) {
            wnd += s.u.tsign[lane * F_GRID + x];
            if ((o >> x) & 1u) {
                if (prev >= 0 && x > prev + 1 && w_prev != 0) {
                    out.span_xy[2 * si] = (int16_t)((tx0 + prev + 1) * TILE);
                    out.span_xy[2 * si + 1] = (int16_t)(ty * TILE);
                    out.span_w[si] = (uint16_t)((uint16_t)(x - prev - 1) * (uint16_t)TILE);
                    si++;
                }
                prev = x;
                w_prev = wnd;
            }
        }
    }
    if (NT == 0) {  // no increments: one all-zero tile at (0, 0) (rasterizer.rs:193-198)
        if (ADJ) {
            adjoint_backward<TC>(in, adj, ctl, sc, asc, s, p, L, 0, NS, tx0, ty0, GW, GH);
            return;
        }
        if (lane < 4) reinterpret_cast<uint4*>(out.tile_alpha + tb * 64)[lane] = make_uint4(0, 0, 0, 0);
        if (out.tile_cov)
            for (int k = lane; k < 64; k += WAVE) out.tile_cov[tb * 64 + k] = 0.0f;
        if (lane == 0) {
            out.tile_xy[2 * tb] = 0;
            out.tile_xy[2 * tb + 1] = 0;
        }
        return;
    }

    // ------------------------------------------- increment offsets
    {
        uint32_t run = 0;
        for (uint32_t b = 0; b < L; b += WAVE) {
            uint32_t j = b + lane;
            uint32_t c = j < L ? incoff[j] : 0u;
            uint32_t ci = wave_incl_scan(c, OpAdd());
            if (j < L) incoff[j] = run + ci - c;
            run += __shfl(ci, WAVE - 1, WAVE);
        }
        if (lane == 0) incoff[L] = run;
    }
    __syncthreads();

    // ------------------------------------- chunks of tiles: emit, commit
    float carry = 0.0f;
    int carry_row = -1;
    for (uint32_t r0 = 0; r0 < NT; r0 += TC) {
        const uint32_t nt = NT - r0 < (uint32_t)TC ? NT - r0 : (uint32_t)TC;
        if ((uint32_t)lane < nt) {
            uint32_t rank = r0 + lane;
            int ty = 0;
            while (ty + 1 < GH && s.row_rank[ty + 1] <= rank) ty++;
            uint32_t k = rank - s.row_rank[ty];
            uint32_t o = s.occ[ty];
            for (uint32_t q = 0; q < k; q++) o &= o - 1;
            int tx = __ffs(o) - 1;
            s.chunk_tile[lane] = ((uint32_t)ty << 8) | (uint32_t)tx;
        }
        for (int k = lane; k < TC * F_TSTRIDE; k += WAVE) {
            s.u.a.acc_a[k] = 0.0f;
            s.u.a.acc_h[k] = 0.0f;
        }
        __syncthreads();
        uint32_t l0 = 0;
        while (l0 < L) {
            // lines [l0, l1) whose increments fit the batch buffer
            const uint32_t base = incoff[l0];
            uint32_t lo = l0 + 1, hi = L;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (incoff[mid] - base <= (uint32_t)F_ICAP) lo = mid;
                else hi = mid - 1;
            }
            const uint32_t l1 = lo;
            const uint32_t n_inc = incoff[l1] - base;
            if (n_inc > (uint32_t)F_ICAP) {  // one line longer than the buffer: the grid cap rules it out
                if (lane == 0) atomicOr(&ctl->error, (uint32_t)ERR_BBOX);
                return;
            }
            // emit: lane per line, refilled as lines finish
            {
                uint32_t next = l0;
                bool active = false;
                Dda d;
                uint32_t pos = 0;
                int16_t cur_tx = 0, cur_ty = 0;
                uint32_t slot_base = F_INVALID;
                LineWindow win;
                win.load(lines, incoff, l1, l0, lane);
                while (next < l1 || __ballot(active)) {
                    const uint64_t need = __ballot(!active);
                    const uint32_t my = next + (uint32_t)__popcll(need & lt);
                    float4 ln = make_float4(0, 0, 0, 0);
                    uint32_t lo_off = 0;
                    if (need && next < l1) ln = win.get(my, &lo_off, true);
                    if (!active && my < l1) {
                        V2 a = mk(ln.x, ln.y), b = mk(ln.z, ln.w);
                        if (!veq(a, b)) {
                            dda_init(d, a, b);
                            pos = lo_off - base;
                            cur_tx = (int16_t)0x7fff;
                            cur_ty = (int16_t)0x7fff;
                            active = true;
                        }
                    }
                    next += (uint32_t)__popcll(need);
                    win.advance(lines, incoff, l1, next, lane);
                    if (active) {
                        int16_t tx = div8(d.x), ty = div8(d.y);
                        if (tx != cur_tx || ty != cur_ty) {
                            cur_tx = tx;
                            cur_ty = ty;
                            uint32_t rank = tile_rank(s, tx - tx0, ty - ty0);
                            slot_base = (rank - r0 < nt) ? (rank - r0) * F_TSTRIDE : F_INVALID;
                        }
                        float area, height;
                        dda_increment(d, area, height);
                        uint32_t key = slot_base == F_INVALID ? F_INVALID
                                                              : slot_base + (uint32_t)(((d.y & 7) << 3) | (d.x & 7));
                        incs[pos++] = make_float4(__uint_as_float(key), area, height, 0.0f);
                        if (dda_step(d)) active = false;
                    }
                }
            }
            __syncthreads();
            // commit in emission order
            const float4 inv = make_float4(__uint_as_float(F_INVALID), 0.0f, 0.0f, 0.0f);
            float4 pf0 = (uint32_t)lane < n_inc ? incs[lane] : inv;
            float4 pf1 = (uint32_t)(lane + WAVE) < n_inc ? incs[lane + WAVE] : inv;
            for (uint32_t cb = 0; cb < n_inc; cb += WAVE) {
                const float4 v = pf0;
                pf0 = pf1;
                const uint32_t jn = cb + 2 * WAVE + lane;
                pf1 = jn < n_inc ? incs[jn] : inv;
                const uint32_t key = __float_as_uint(v.x);
                const bool valid = key != F_INVALID;
                uint64_t match = __ballot(valid);
#pragma unroll
                for (int b = 0; b < 12; b++) {
                    const bool bit = (key >> b) & 1u;
                    const uint64_t m = __ballot(bit);
                    match &= bit ? m : ~m;
                }
                const int rank = __popcll(match & lt);
                for (int r = 0;; r++) {
                    if (valid && rank == r) {
                        s.u.a.acc_a[key] += v.y;
                        s.u.a.acc_h[key] += v.z;
                    }
                    if (!__ballot(valid && rank > r)) break;
                }
            }
            __syncthreads();
            l0 = l1;
        }
        // row carry, one lane per pixel row, sequential over the chunk's tiles
        if (lane < 8) {
            for (uint32_t t = 0; t < nt; t++) {
                int ty = (int)(s.chunk_tile[t] >> 8);
                if (ty != carry_row) {
                    carry = 0.0f;
                    carry_row = ty;
                }
                s.u.a.cin[t * 8 + lane] = carry;
#pragma unroll
                for (int x = 0; x < 8; x++) carry += s.u.a.acc_h[t * F_TSTRIDE + lane * 8 + x];
            }
        }
        __syncthreads();
        if (ADJ) {
            // keep v = accum + area of every pixel (rank order) for the loss
            if ((uint32_t)lane < nt) {
                const int t = lane;
                float* tv = reinterpret_cast<float*>(asc + A_OFF_TILES) + (size_t)(r0 + t) * 128;
#pragma unroll
                for (int y = 0; y < 8; y++) {
                    float accum = s.u.a.cin[t * 8 + y];
#pragma unroll
                    for (int x = 0; x < 8; x++) {
                        const int k = t * F_TSTRIDE + y * 8 + x;
                        tv[y * 8 + x] = accum + s.u.a.acc_a[k];
                        accum += s.u.a.acc_h[k];
                    }
                }
            }
        }
        // quantise and store (rasterizer.rs:231-241)
        if (!ADJ && (uint32_t)lane < nt) {
            const int t = lane;
            const int ty = (int)(s.chunk_tile[t] >> 8), tx = (int)(s.chunk_tile[t] & 0xffu);
            uint32_t words[16];
#pragma unroll
            for (int y = 0; y < 8; y++) {
                float accum = s.u.a.cin[t * 8 + y];
                uint32_t w0 = 0, w1 = 0;
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    const int k = t * F_TSTRIDE + y * 8 + x;
                    if (out.tile_cov)
                        out.tile_cov[(tb + r0 + (uint32_t)t) * 64 + y * 8 + x] = fminf(fabsf(accum + s.u.a.acc_a[k]), 1.0f);
                    float v = fminf(fabsf(accum + s.u.a.acc_a[k]) * 256.0f, 255.0f);
                    uint32_t b = (uint32_t)v;
                    if (x < 4) w0 |= b << (8 * x);
                    else w1 |= b << (8 * (x - 4));
                    accum += s.u.a.acc_h[k];
                }
                words[2 * y] = w0;
                words[2 * y + 1] = w1;
            }
            const uint64_t ti = tb + r0 + (uint32_t)t;
            uint4* dst = reinterpret_cast<uint4*>(out.tile_alpha + ti * 64);
            dst[0] = make_uint4(words[0], words[1], words[2], words[3]);
            dst[1] = make_uint4(words[4], words[5], words[6], words[7]);
            dst[2] = make_uint4(words[8], words[9], words[10], words[11]);
            dst[3] = make_uint4(words[12], words[13], words[14], words[15]);
            out.tile_xy[2 * ti] = (int16_t)((tx0 + tx) * TILE);
            out.tile_xy[2 * ti + 1] = (int16_t)((ty0 + ty) * TILE);
        }
        __syncthreads();
    }
    if (ADJ) adjoint_backward<TC>(in, adj, ctl, sc, asc, s, p, L, NT, NS, tx0, ty0, GW, GH);
}

template <int TC>
__global__ __launch_bounds__(WAVE) void k_fused(FusedIn in, FusedOut out, FusedCtl* ctl, uint32_t* deferred,
                                                uint8_t* scratch, uint32_t gran) {
    __shared__ LdsState<TC> s;
    uint8_t* sc = scratch + (size_t)blockIdx.x * F_SCRATCH;
    const int64_t n = in.ids ? (int64_t)*in.n_ids : in.n_paths;
    if (n == 0) return;  // nothing handed on (the common case behind the dense kernels)
    if (in.ids && n > (int64_t)in.defer_above) {
        for (int64_t q = (int64_t)blockIdx.x * WAVE + threadIdx.x; q < n; q += (int64_t)gridDim.x * WAVE)
            deferred[atomicAdd(&ctl->deferred, 1u)] = in.ids[q];
        return;
    }
    for (;;) {
        uint32_t base = 0;
        if (threadIdx.x == 0) base = atomicAdd(in.ids ? &ctl->queue2 : &ctl->queue, gran);
        base = __shfl(base, 0, WAVE);
        if ((int64_t)base >= n) break;
        const int64_t end = std::min<int64_t>((int64_t)base + gran, n);
        for (int64_t q = base; q < end; q++)
            fused_path<TC, false>(in, out, ctl, deferred, sc, s, in.ids ? (int64_t)in.ids[q] : q, AdjArgs{},
                                  nullptr);
    }
}

#include "adjoint.inc"
#include "adjoint_large.inc"
#include "dense.inc"
#include "single.inc"

// ---------------------------------------------- deferred paths -> plain

// Gathers the deferred paths into a compact segment input.
__global__ void k_gather_paths(FusedIn in, const uint32_t* ids, uint32_t n, const int64_t* dst_voff,
                               const int64_t* dst_poff, const int64_t* dst_woff, uint8_t* verbs, float* points,
                               float* weights, Xform* xf) {
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const int64_t p = ids[k];
        const int64_t vb = in.verb_off[p], nv = in.verb_off[p + 1] - vb;
        const int64_t pb = in.point_off[p], np = in.point_off[p + 1] - pb;
        const int64_t wb = in.weight_off[p], nw = in.weight_off[p + 1] - wb;
        for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) verbs[dst_voff[k] + i] = in.verbs[vb + i];
        for (int64_t i = threadIdx.x; i < 2 * np; i += blockDim.x) points[2 * dst_poff[k] + i] = in.points[2 * pb + i];
        for (int64_t i = threadIdx.x; i < nw; i += blockDim.x) weights[dst_woff[k] + i] = in.weights[wb + i];
        if (threadIdx.x == 0) xf[k] = in.xf[p];
    }
}

__global__ void k_scatter_ranges(const uint32_t* ids, uint32_t n, const int64_t* ts, const int32_t* tc,
                                 const int64_t* ss, const int32_t* scn, int64_t tbase, int64_t sbase, int64_t* pts,
                                 int32_t* ptc, int64_t* pss, int32_t* psc) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = ids[k];
    pts[p] = ts[k] + tbase;
    ptc[p] = tc[k];
    pss[p] = ss[k] + sbase;
    psc[p] = scn[k];
}

struct BatchedState {
    DBuf scratch, ascratch, ctl, deferred, dense_ids, dense_ids2, dense_ids3;
    DBuf large_cmd, large_grid;  // the large-path adjoint's scratch
    DBuf g_verbs, g_points, g_weights, g_voff, g_poff, g_woff, g_xf, g_ts, g_tc, g_ss, g_sc;
    int n_waves = 0;
    uint32_t gran = 1;
    // graph of (reset, fused) for the last argument set
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint8_t> graph_key;
    ~BatchedState() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int run_batched(Context* ctx, const SegInput& in, OutputDesc& out, hipStream_t s, int waves_per_cu, int gran,
                int tiles_per_chunk, bool use_graph) {
    if (!ctx->batched) ctx->batched = new BatchedState();
    BatchedState* B = reinterpret_cast<BatchedState*>(ctx->batched);
    const int64_t NP = in.n_paths;
    out.n_tiles = out.n_spans = 0;
    if (NP <= 0) return 0;
    int dev = 0;
    HIPCHK_RET(hipGetDevice(&dev));
    int cus = 0;
    HIPCHK_RET(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int n_waves = std::max(1, cus * std::max(1, waves_per_cu));
    HIPCHK_RET(B->scratch.ensure((size_t)n_waves * F_SCRATCH));
    HIPCHK_RET(B->ctl.ensure(sizeof(FusedCtl)));
    HIPCHK_RET(B->deferred.ensure((size_t)NP * 4));
    if (!ctx->host_pinned) HIPCHK_RET(hipHostMalloc(&ctx->host_pinned, 4096));
    FusedIn fi;
    fi.n_paths = NP;
    fi.verbs = in.verbs;
    fi.verb_off = in.seg_verb_off;
    fi.points = in.points;
    fi.point_off = in.seg_point_off;
    fi.weights = in.weights;
    fi.weight_off = in.seg_weight_off;
    fi.xf = in.seg_xf;
    fi.n_points = in.n_points;
    fi.n_weights = in.n_weights;
    FusedOut fo;
    fo.tile_xy = out.tile_xy;
    fo.tile_alpha = out.tile_alpha;
    fo.tile_cov = out.tile_cov;
    fo.tile_cap = out.tile_capacity;
    fo.span_xy = out.span_xy;
    fo.span_w = out.span_w;
    fo.span_cap = out.span_capacity;
    fo.pts = out.path_tile_start;
    fo.ptc = out.path_tile_count;
    fo.pss = out.path_span_start;
    fo.psc = out.path_span_count;
    FusedCtl* ctl = B->ctl.as<FusedCtl>();
    uint32_t* deferred = B->deferred.as<uint32_t>();
    uint8_t* scratch = B->scratch.as<uint8_t>();
    const uint32_t g = (uint32_t)std::max(1, gran);
    const bool tc16 = tiles_per_chunk <= 16;
    const bool dense = ctx->dense;
    const int dnw = ctx->dense_waves == 2 ? 2 : (ctx->dense_waves == 8 ? 8 : 4);
    const int n_groups = std::max(1, cus * std::max(1, ctx->dense_groups_per_cu));
    const int n_small = dense ? cus * std::max(0, ctx->dense_small_groups) : 0;
    const int n_large = (dense && ctx->dense_large) ? cus : 0;
    // The dense instances run one after another, each over the paths the one
    // before handed on: [small grids, 32 tiles] -> 64 tiles -> [256 tiles], then
    // the ordered kernel over what is left.
    enum { D_SMALL, D_MAIN, D_LARGE };
    int stages[3], n_stages = 0;
    if (dense) {
        if (n_small) stages[n_stages++] = D_SMALL;
        stages[n_stages++] = D_MAIN;
        if (n_large) stages[n_stages++] = D_LARGE;
    }
    uint32_t* lists[3] = {nullptr, nullptr, nullptr};
    uint32_t* counts[3] = {&ctl->dq, &ctl->dq2, &ctl->dq3};
    uint32_t* queues[3] = {&ctl->queue, &ctl->queue3, &ctl->queue4};
    FusedIn fin[4] = {fi, fi, fi, fi};  // stage i reads fin[i]; k_fused reads fin[n_stages]
    if (dense) {
        HIPCHK_RET(B->dense_ids.ensure((size_t)NP * 4));
        HIPCHK_RET(B->dense_ids2.ensure((size_t)NP * 4));
        if (n_stages > 2) HIPCHK_RET(B->dense_ids3.ensure((size_t)NP * 4));
        lists[0] = B->dense_ids.as<uint32_t>();
        lists[1] = B->dense_ids2.as<uint32_t>();
        lists[2] = B->dense_ids3.as<uint32_t>();
        for (int i = 1; i <= n_stages; i++) {
            fin[i].ids = lists[i - 1];
            fin[i].n_ids = counts[i - 1];
        }
        fin[n_stages].defer_above = F_DEFER_ABOVE;
    }
    const bool timing = ctx->timing;
    if (timing) {
        for (auto& e : ctx->ev)
            if (!e) HIPCHK_RET(hipEventCreate(&e));
    }
    // timing without a graph brackets the dense kernels and k_fused with events
    auto enqueue = [&](hipStream_t st, bool ev) {
        hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(64), 0, st, ctl);
        if (ev) (void)hipEventRecord(ctx->ev[0], st);
        for (int i = 0; i < n_stages; i++) {
            const FusedIn& f = fin[i];
            if (stages[i] == D_SMALL)
                hipLaunchKernelGGL((k_dense<2, 32>), dim3(n_small), dim3(2 * WAVE), 0, st, f, fo, ctl, queues[i],
                                   lists[i], counts[i]);
            else if (stages[i] == D_LARGE)
                hipLaunchKernelGGL((k_dense<8, 256>), dim3(n_large), dim3(8 * WAVE), 0, st, f, fo, ctl, queues[i],
                                   lists[i], counts[i]);
            else if (dnw == 2)
                hipLaunchKernelGGL((k_dense<2, 64>), dim3(n_groups), dim3(2 * WAVE), 0, st, f, fo, ctl, queues[i],
                                   lists[i], counts[i]);
            else if (dnw == 8)
                hipLaunchKernelGGL((k_dense<8, 64>), dim3(n_groups), dim3(8 * WAVE), 0, st, f, fo, ctl, queues[i],
                                   lists[i], counts[i]);
            else
                hipLaunchKernelGGL((k_dense<4, 64>), dim3(n_groups), dim3(4 * WAVE), 0, st, f, fo, ctl, queues[i],
                                   lists[i], counts[i]);
        }
        if (dense && ev) (void)hipEventRecord(ctx->ev[2], st);
        const FusedIn& f = fin[n_stages];
        if (tc16) hipLaunchKernelGGL(k_fused<16>, dim3(n_waves), dim3(WAVE), 0, st, f, fo, ctl, deferred, scratch, g);
        else hipLaunchKernelGGL(k_fused<32>, dim3(n_waves), dim3(WAVE), 0, st, f, fo, ctl, deferred, scratch, g);
        if (ev) (void)hipEventRecord(ctx->ev[1], st);
    };
    if (use_graph) {
        std::vector<uint8_t> key(sizeof(fi) + sizeof(fo) + 41);
        key.back() = tc16 ? 1 : 0;
        memcpy(key.data(), &fi, sizeof(fi));
        memcpy(key.data() + sizeof(fi), &fo, sizeof(fo));
        memcpy(key.data() + sizeof(fi) + sizeof(fo), &n_waves, 4);
        memcpy(key.data() + sizeof(fi) + sizeof(fo) + 4, &g, 4);
        memcpy(key.data() + sizeof(fi) + sizeof(fo) + 8, &scratch, 8);
        const int dkey = dense ? (dnw << 16) | n_groups : 0;
        memcpy(key.data() + sizeof(fi) + sizeof(fo) + 24, &dkey, 4);
        memcpy(key.data() + sizeof(fi) + sizeof(fo) + 28, &n_small, 4);
        memcpy(key.data() + sizeof(fi) + sizeof(fo) + 32, &n_large, 4);
        if (!B->exec || key != B->graph_key) {
            if (B->exec) (void)hipGraphExecDestroy(B->exec);
            if (B->graph) (void)hipGraphDestroy(B->graph);
            B->exec = nullptr;
            B->graph = nullptr;
            hipStream_t cs;
            HIPCHK_RET(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            HIPCHK_RET(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            enqueue(cs, false);
            hipError_t e = hipStreamEndCapture(cs, &B->graph);
            (void)hipStreamDestroy(cs);
            HIPCHK_RET(e);
            HIPCHK_RET(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
            B->graph_key = key;
        }
        if (timing) HIPCHK_RET(hipEventRecord(ctx->ev[0], s));
        HIPCHK_RET(hipGraphLaunch(B->exec, s));
        if (timing) HIPCHK_RET(hipEventRecord(ctx->ev[1], s));
    } else {
        enqueue(s, timing);
    }
    HIPCHK_RET(hipGetLastError());
    FusedCtl* hc = reinterpret_cast<FusedCtl*>(reinterpret_cast<uint8_t*>(ctx->host_pinned) + 1024);
    HIPCHK_RET(hipMemcpyAsync(hc, ctl, sizeof(FusedCtl), hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    if (timing) {
        // graph: the whole launch; direct: k_dense and k_fused separately
        float ms = 0;
        if (use_graph) {
            HIPCHK_RET(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            auto& t = ctx->timings["batched_graph"];
            t.first += ms;
            t.second += 1;
        } else if (dense) {
            HIPCHK_RET(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]));
            auto& t = ctx->timings["k_dense"];
            t.first += ms;
            t.second += 1;
            HIPCHK_RET(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[1]));
            auto& t2 = ctx->timings["k_fused"];
            t2.first += ms;
            t2.second += 1;
        } else {
            HIPCHK_RET(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            auto& t = ctx->timings["k_fused"];
            t.first += ms;
            t.second += 1;
        }
    }
    if (hc->error & ERR_INPUT) return -1;
    if (hc->error & ERR_FLATTEN) return -5;
    if (hc->error & ERR_BBOX) return -3;
    const int64_t Tf = (int64_t)hc->tiles, Sf = (int64_t)hc->spans;
    const uint32_t nd = hc->deferred;
    int64_t Tp = 0, Sp = 0;
    int st = 0;
    if (nd > 0) {
        // compact the deferred paths and run the plain pipeline on them
        std::vector<uint32_t> ids(nd);
        HIPCHK_RET(hipMemcpy(ids.data(), deferred, nd * 4, hipMemcpyDeviceToHost));
        std::sort(ids.begin(), ids.end());
        std::vector<int64_t> voff(NP + 1), poff(NP + 1), woff(NP + 1);
        HIPCHK_RET(hipMemcpy(voff.data(), in.seg_verb_off, (NP + 1) * 8, hipMemcpyDeviceToHost));
        HIPCHK_RET(hipMemcpy(poff.data(), in.seg_point_off, (NP + 1) * 8, hipMemcpyDeviceToHost));
        HIPCHK_RET(hipMemcpy(woff.data(), in.seg_weight_off, (NP + 1) * 8, hipMemcpyDeviceToHost));
        std::vector<int64_t> dv(nd + 1, 0), dp(nd + 1, 0), dw(nd + 1, 0);
        for (uint32_t k = 0; k < nd; k++) {
            uint32_t p = ids[k];
            dv[k + 1] = dv[k] + voff[p + 1] - voff[p];
            dp[k + 1] = dp[k] + poff[p + 1] - poff[p];
            dw[k + 1] = dw[k] + woff[p + 1] - woff[p];
        }
        HIPCHK_RET(B->g_verbs.ensure(std::max<int64_t>(dv[nd], 1)));
        HIPCHK_RET(B->g_points.ensure(std::max<int64_t>(dp[nd], 1) * 8));
        HIPCHK_RET(B->g_weights.ensure(std::max<int64_t>(dw[nd], 1) * 4));
        HIPCHK_RET(B->g_voff.ensure((nd + 1) * 8));
        HIPCHK_RET(B->g_poff.ensure((nd + 1) * 8));
        HIPCHK_RET(B->g_woff.ensure((nd + 1) * 8));
        HIPCHK_RET(B->g_xf.ensure(nd * sizeof(Xform)));
        HIPCHK_RET(B->g_ts.ensure(nd * 8));
        HIPCHK_RET(B->g_tc.ensure(nd * 4));
        HIPCHK_RET(B->g_ss.ensure(nd * 8));
        HIPCHK_RET(B->g_sc.ensure(nd * 4));
        HIPCHK_RET(hipMemcpyAsync(deferred, ids.data(), nd * 4, hipMemcpyHostToDevice, s));
        HIPCHK_RET(hipMemcpyAsync(B->g_voff.p, dv.data(), (nd + 1) * 8, hipMemcpyHostToDevice, s));
        HIPCHK_RET(hipMemcpyAsync(B->g_poff.p, dp.data(), (nd + 1) * 8, hipMemcpyHostToDevice, s));
        HIPCHK_RET(hipMemcpyAsync(B->g_woff.p, dw.data(), (nd + 1) * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_paths, dim3(std::min<uint32_t>(nd, 4096)), dim3(256), 0, s, fi, deferred, nd,
                           B->g_voff.as<int64_t>(), B->g_poff.as<int64_t>(), B->g_woff.as<int64_t>(),
                           B->g_verbs.as<uint8_t>(), B->g_points.as<float>(), B->g_weights.as<float>(),
                           B->g_xf.as<Xform>());
        HIPCHK_RET(hipGetLastError());
        SegInput si;
        si.n_segs = nd;
        si.n_paths = nd;
        si.verbs = B->g_verbs.as<uint8_t>();
        si.points = B->g_points.as<float>();
        si.weights = B->g_weights.as<float>();
        si.seg_verb_off = B->g_voff.as<int64_t>();
        si.seg_point_off = B->g_poff.as<int64_t>();
        si.seg_weight_off = B->g_woff.as<int64_t>();
        si.seg_xf = B->g_xf.as<Xform>();
        si.n_verbs = dv[nd];
        si.n_points = dp[nd];
        si.n_weights = dw[nd];
        OutputDesc od;
        const bool room = Tf <= out.tile_capacity && Sf <= out.span_capacity;
        od.tile_xy = out.tile_xy + (room ? 2 * Tf : 0);
        od.tile_alpha = out.tile_alpha + (room ? 64 * Tf : 0);
        od.tile_cov = out.tile_cov ? out.tile_cov + (room ? 64 * Tf : 0) : nullptr;
        od.tile_capacity = room ? out.tile_capacity - Tf : 0;
        od.span_xy = out.span_xy + (room ? 2 * Sf : 0);
        od.span_w = out.span_w + (room ? Sf : 0);
        od.span_capacity = room ? out.span_capacity - Sf : 0;
        od.path_tile_start = B->g_ts.as<int64_t>();
        od.path_tile_count = B->g_tc.as<int32_t>();
        od.path_span_start = B->g_ss.as<int64_t>();
        od.path_span_count = B->g_sc.as<int32_t>();
        // under the dense kernels' 1-LSB contract the handed-on paths take the
        // unordered tile stage; with them off the plain pipeline stays exact
        st = run_plain(ctx, si, od, s, dense);
        Tp = od.n_tiles;
        Sp = od.n_spans;
        if (st == 0) {
            hipLaunchKernelGGL(k_scatter_ranges, dim3((nd + 255) / 256), dim3(256), 0, s, deferred, nd,
                               od.path_tile_start, od.path_tile_count, od.path_span_start, od.path_span_count, Tf,
                               Sf, out.path_tile_start, out.path_tile_count, out.path_span_start,
                               out.path_span_count);
            HIPCHK_RET(hipGetLastError());
        } else if (st != -4) {
            return st;
        }
    }
    out.n_tiles = Tf + Tp;
    out.n_spans = Sf + Sp;
    if ((hc->error & ERR_CAPACITY) || out.n_tiles > out.tile_capacity || out.n_spans > out.span_capacity) return -4;
    return st;
}

// The large-path adjoint of path p (one path over the whole device; see
// adjoint_large.inc).  `hp`: 2 KiB of pinned host memory for the read-backs.
static int run_adjoint_large(const FusedIn& fi, const AdjArgs& a, int64_t p, hipStream_t s, BatchedState* B,
                             uint8_t* hp) {
    int64_t* off = reinterpret_cast<int64_t*>(hp);
    HIPCHK_RET(hipMemcpyAsync(off, fi.verb_off + p, 16, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(off + 2, fi.point_off + p, 16, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    const int64_t nv = off[1] - off[0], np = off[3] - off[2];
    if (nv < 0 || np < 0) return -1;
    // command scratch: cg, cmv, cw | cl | pcmd | slot | box, err, nt | acc
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o = 0;
    const size_t o_cg = o;
    o = al(o + (size_t)nv * 8);
    const size_t o_cmv = o;
    o = al(o + (size_t)nv * 8);
    const size_t o_cw = o;
    o = al(o + (size_t)nv * 8);
    const size_t o_cl = o;
    o = al(o + (size_t)(nv + 1) * 4);
    const size_t o_pcmd = o;
    o = al(o + (size_t)np * 4);
    const size_t o_slot = o;
    o = al(o + (size_t)nv * 48);
    const size_t o_ctl = o;  // box[4], err, nt
    o = al(o + 32);
    const size_t o_acc = o;
    o = al(o + 8 * 8);
    HIPCHK_RET(B->large_cmd.ensure(o));
    uint8_t* c = B->large_cmd.as<uint8_t>();
    LargeAdj A;
    memset(&A, 0, sizeof(A));
    A.cg = reinterpret_cast<int64_t*>(c + o_cg);
    A.cmv = reinterpret_cast<int64_t*>(c + o_cmv);
    A.cw = reinterpret_cast<int64_t*>(c + o_cw);
    A.cl = reinterpret_cast<uint32_t*>(c + o_cl);
    A.pcmd = reinterpret_cast<uint32_t*>(c + o_pcmd);
    A.slot = reinterpret_cast<float*>(c + o_slot);
    A.box = reinterpret_cast<unsigned*>(c + o_ctl);
    A.err = reinterpret_cast<uint32_t*>(c + o_ctl + 16);
    A.nt = reinterpret_cast<uint32_t*>(c + o_ctl + 20);
    A.acc = reinterpret_cast<double*>(c + o_acc);
    HIPCHK_RET(hipMemsetAsync(A.box, 0xff, 8, s));        // min x, min y
    HIPCHK_RET(hipMemsetAsync(A.box + 2, 0, 16, s));      // max x, max y, err, nt
    HIPCHK_RET(hipMemsetAsync(A.acc, 0, 8 * 8, s));
    const int T = 256;
    const unsigned gv = (unsigned)std::max<int64_t>(1, (nv + T - 1) / T);
    hipLaunchKernelGGL(k_al_cmds, dim3(1), dim3(1024), 0, s, fi, p, A);
    hipLaunchKernelGGL(k_al_count, dim3(gv), dim3(T), 0, s, fi, p, A);
    hipLaunchKernelGGL(k_al_scan, dim3(1), dim3(1024), 0, s, fi, p, A);
    HIPCHK_RET(hipGetLastError());
    uint32_t* hc = reinterpret_cast<uint32_t*>(hp + 64);
    HIPCHK_RET(hipMemcpyAsync(hc, A.cl + nv, 4, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipMemcpyAsync(hc + 1, A.box, 24, hipMemcpyDeviceToHost, s));  // box[4], err, nt
    HIPCHK_RET(hipStreamSynchronize(s));
    const uint32_t e1 = hc[5];
    if (e1 & ERR_INPUT) return -1;
    if (e1 & ERR_FLATTEN) return -5;
    if (e1 & ERR_CAPACITY) return -7;  // coordinates beyond the 16-bit pixel range
    const int64_t L = hc[0];
    if (L >= AL_MAX_LINES) return -7;
    int32_t tx0 = 0, ty0 = 0, GW = 0, GH = 0;
    if (L > 0)
        grid_from_box(true, false, (int32_t)floorf(unord_f(hc[1])), (int32_t)floorf(unord_f(hc[2])),
                      (int32_t)floorf(unord_f(hc[3])), (int32_t)floorf(unord_f(hc[4])), &tx0, &ty0, &GW, &GH);
    const int64_t tiles = (int64_t)GW * GH;
    if (tiles > AL_MAX_TILES) return -7;
    const int32_t ow = (GW + 31) / 32;
    // line and grid scratch
    o = 0;
    const size_t o_l32 = o;
    o = al(o + (size_t)L * 16);
    const size_t o_l64 = o;
    o = al(o + (size_t)L * 32);
    const size_t o_dl = o;
    o = al(o + (size_t)L * 16);
    const size_t o_occ = o;
    o = al(o + (size_t)ow * GH * 4);
    const size_t o_fa = o;
    o = al(o + (size_t)tiles * 64 * 8);
    const size_t o_fh = o;
    o = al(o + (size_t)tiles * 64 * 8);
    const size_t o_ts = o;
    o = al(o + (size_t)tiles * 4);
    const size_t o_span = o;
    o = al(o + (size_t)tiles);
    const size_t o_g = o;
    o = al(o + (size_t)tiles * 64 * 4);
    const size_t o_G = o;
    o = al(o + (size_t)tiles * 64 * 4);
    HIPCHK_RET(B->large_grid.ensure(std::max<size_t>(o, 256)));
    uint8_t* gbuf = B->large_grid.as<uint8_t>();
    A.l32 = reinterpret_cast<float4*>(gbuf + o_l32);
    A.l64 = reinterpret_cast<double2*>(gbuf + o_l64);
    A.dl = reinterpret_cast<float4*>(gbuf + o_dl);
    A.occ = reinterpret_cast<uint32_t*>(gbuf + o_occ);
    A.fa = reinterpret_cast<unsigned long long*>(gbuf + o_fa);
    A.fh = reinterpret_cast<unsigned long long*>(gbuf + o_fh);
    A.tsign = reinterpret_cast<int32_t*>(gbuf + o_ts);
    A.span = gbuf + o_span;
    A.g = reinterpret_cast<float*>(gbuf + o_g);
    A.G = reinterpret_cast<float*>(gbuf + o_G);
    A.tx0 = tx0;
    A.ty0 = ty0;
    A.GW = GW;
    A.GH = GH;
    A.ow = ow;
    if (o_g > o_occ) HIPCHK_RET(hipMemsetAsync(gbuf + o_occ, 0, o_g - o_occ, s));  // occ, sums, crossings, spans
    const unsigned gl = (unsigned)std::max<int64_t>(1, (L + T - 1) / T);
    const unsigned gp = (unsigned)std::max<int64_t>(1, (np + T - 1) / T);
    const unsigned gr = (unsigned)std::max<int64_t>(1, ((int64_t)GH * TILE + T - 1) / T);
    hipLaunchKernelGGL(k_al_lines, dim3(gv), dim3(T), 0, s, fi, p, A);
    if (L > 0) hipLaunchKernelGGL(k_al_fwd, dim3(gl), dim3(T), 0, s, L, A);
    hipLaunchKernelGGL(k_al_spans, dim3(1), dim3(1024), 0, s, A);
    if (GH > 0) hipLaunchKernelGGL(k_al_rows, dim3(gr), dim3(T), 0, s, fi, A, a, p);
    // target tiles: one thread each (their count is the caller's, bounded by its capacity)
    hipLaunchKernelGGL(k_al_target, dim3(4096), dim3(T), 0, s, fi, A, a, p);
    if (L > 0) hipLaunchKernelGGL(k_al_bwd, dim3(gl), dim3(T), 0, s, L, A);
    hipLaunchKernelGGL(k_al_cmd_bwd, dim3(gv), dim3(T), 0, s, fi, p, A);
    if (np > 0) hipLaunchKernelGGL(k_al_gather, dim3(gp), dim3(T), 0, s, fi, p, A, a);
    hipLaunchKernelGGL(k_al_final, dim3(1), dim3(1), 0, s, A, a, p);
    HIPCHK_RET(hipGetLastError());
    HIPCHK_RET(hipMemcpyAsync(hc, A.err, 4, hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    if (hc[0] & ERR_INPUT) return -1;
    if (hc[0] & ERR_FLATTEN) return -5;
    if (hc[0] & ERR_BBOX) return -3;
    return 0;
}

int run_adjoint(Context* ctx, const SegInput& in, const AdjTarget& tgt, float* loss, float* d_points, float* d_xf,
                hipStream_t s, int waves_per_cu, int gran, int tiles_per_chunk) {
    if (!ctx->batched) ctx->batched = new BatchedState();
    BatchedState* B = reinterpret_cast<BatchedState*>(ctx->batched);
    const int64_t NP = in.n_paths;
    if (NP <= 0) return 0;
    int dev = 0, cus = 0;
    HIPCHK_RET(hipGetDevice(&dev));
    HIPCHK_RET(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int n_waves = std::max(1, cus * std::max(1, waves_per_cu));
    HIPCHK_RET(B->scratch.ensure((size_t)n_waves * F_SCRATCH));
    HIPCHK_RET(B->ascratch.ensure((size_t)n_waves * A_SCRATCH));
    HIPCHK_RET(B->ctl.ensure(sizeof(FusedCtl)));
    if (!ctx->host_pinned) HIPCHK_RET(hipHostMalloc(&ctx->host_pinned, 4096));
    FusedIn fi;
    fi.n_paths = NP;
    fi.verbs = in.verbs;
    fi.verb_off = in.seg_verb_off;
    fi.points = in.points;
    fi.point_off = in.seg_point_off;
    fi.weights = in.weights;
    fi.weight_off = in.seg_weight_off;
    fi.xf = in.seg_xf;
    fi.n_points = in.n_points;
    fi.n_weights = in.n_weights;
    AdjArgs a;
    a.tt_start = tgt.tile_start;
    a.tt_count = tgt.tile_count;
    a.t_xy = tgt.tile_xy;
    a.t_cov = tgt.tile_cov;
    a.ts_start = tgt.span_start;
    a.ts_count = tgt.span_count;
    a.ts_xy = tgt.span_xy;
    a.ts_w = tgt.span_w;
    a.loss = loss;
    a.d_points = d_points;
    a.d_xf = d_xf;
    FusedCtl* ctl = B->ctl.as<FusedCtl>();
    const uint32_t g = (uint32_t)std::max(1, gran);
    hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(64), 0, s, ctl);
    if (ctx->timing) {
        if (!ctx->ev[0]) HIPCHK_RET(hipEventCreate(&ctx->ev[0]));
        if (!ctx->ev[1]) HIPCHK_RET(hipEventCreate(&ctx->ev[1]));
        HIPCHK_RET(hipEventRecord(ctx->ev[0], s));
    }
    HIPCHK_RET(B->deferred.ensure((size_t)NP * 4));
    uint32_t* large = B->deferred.as<uint32_t>();
    if (tiles_per_chunk <= 16)
        hipLaunchKernelGGL(k_adjoint<16>, dim3(n_waves), dim3(WAVE), 0, s, fi, a, ctl, B->scratch.as<uint8_t>(),
                           B->ascratch.as<uint8_t>(), g, large);
    else
        hipLaunchKernelGGL(k_adjoint<32>, dim3(n_waves), dim3(WAVE), 0, s, fi, a, ctl, B->scratch.as<uint8_t>(),
                           B->ascratch.as<uint8_t>(), g, large);
    HIPCHK_RET(hipGetLastError());
    if (ctx->timing) HIPCHK_RET(hipEventRecord(ctx->ev[1], s));
    FusedCtl* hc = reinterpret_cast<FusedCtl*>(reinterpret_cast<uint8_t*>(ctx->host_pinned) + 1024);
    HIPCHK_RET(hipMemcpyAsync(hc, ctl, sizeof(FusedCtl), hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    if (ctx->timing) {
        float ms = 0;
        HIPCHK_RET(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        auto& t = ctx->timings["k_adjoint"];
        t.first += ms;
        t.second += 1;
    }
    if (hc->error & ERR_INPUT) return -1;
    if (hc->error & ERR_FLATTEN) return -5;
    if (hc->error & ERR_BBOX) return -3;
    if (hc->error & ERR_CAPACITY) return -7;
    // the paths over one wave's limits, one at a time over the whole device
    const uint32_t n_large = hc->deferred;
    if (n_large) {
        std::vector<uint32_t> ids(n_large);
        HIPCHK_RET(hipMemcpy(ids.data(), large, (size_t)n_large * 4, hipMemcpyDeviceToHost));
        std::sort(ids.begin(), ids.end());
        uint8_t* hp = reinterpret_cast<uint8_t*>(ctx->host_pinned) + 2048;
        for (uint32_t id : ids) {
            const int st = run_adjoint_large(fi, a, (int64_t)id, s, B, hp);
            if (st) return st;
        }
    }
    return 0;
}

// ------------------------------------------- one path, no host round trip
//
// A Rasterizer's single fill: k_single (one workgroup of sixteen waves) resets
// the control block itself, writes the output and the counts wherever `out`
// and `ctl` (SMALL_CTL_BYTES) point, and nothing synchronises.  The drop-in
// finish points both at pinned host memory mapped into the device's address
// space, so that no copy follows the launch; after the caller's
// synchronisation small_batch_result reads the counts.  A path over the
// kernel's limits is only counted as deferred (the caller reruns it through
// run_batched, which handles it).
static_assert(sizeof(FusedCtl) == SMALL_CTL_BYTES, "control block size");

int launch_single_fill(const SegInput& in, const OutputDesc& out, void* ctl, hipStream_t s) {
    if (in.n_paths != 1) return -1;
    FusedIn fi;
    fi.n_paths = 1;
    fi.verbs = in.verbs;
    fi.verb_off = in.seg_verb_off;
    fi.points = in.points;
    fi.point_off = in.seg_point_off;
    fi.weights = in.weights;
    fi.weight_off = in.seg_weight_off;
    fi.xf = in.seg_xf;
    fi.n_points = in.n_points;
    fi.n_weights = in.n_weights;
    FusedOut fo;
    fo.tile_xy = out.tile_xy;
    fo.tile_alpha = out.tile_alpha;
    fo.tile_cov = out.tile_cov;
    fo.tile_cap = out.tile_capacity;
    fo.span_xy = out.span_xy;
    fo.span_w = out.span_w;
    fo.span_cap = out.span_capacity;
    fo.pts = out.path_tile_start;
    fo.ptc = out.path_tile_count;
    fo.pss = out.path_span_start;
    fo.psc = out.path_span_count;
    hipLaunchKernelGGL(k_single, dim3(1), dim3(S_NT), 0, s, fi, fo, reinterpret_cast<FusedCtl*>(ctl));
    HIPCHK_RET(hipGetLastError());
    return 0;
}

SmallBatchResult small_batch_result(const void* host_ctl) {
    const FusedCtl* hc = reinterpret_cast<const FusedCtl*>(host_ctl);
    SmallBatchResult r;
    r.tiles = (int64_t)hc->tiles;
    r.spans = (int64_t)hc->spans;
    r.deferred = hc->deferred;
    r.status = (hc->error & ERR_INPUT) ? -1 : (hc->error & ERR_FLATTEN) ? -5 : (hc->error & ERR_BBOX) ? -3
               : (hc->error & ERR_CAPACITY) ? -4 : 0;
    return r;
}

// ------------------------------------------------ the dense chain alone
//
// The dense 64-tile instance over every path, then the 256-tile instance over
// the paths it hands on, with control memory (DENSE_CTL_BYTES: queues,
// counters, error bits) and the handed-on list (n_paths entries) supplied by
// the caller: nothing is allocated.  Paths the 256-tile instance hands on are
// counted, not listed.
static_assert(sizeof(FusedCtl) == DENSE_CTL_BYTES, "control block size");

int launch_dense_chain(const SegInput& in, const OutputDesc& out, uint8_t* ctl_mem, uint32_t* ids, hipStream_t s) {
    const int64_t NP = in.n_paths;
    if (NP <= 0) return 0;
    int dev = 0, cus = 0;
    HIPCHK_RET(hipGetDevice(&dev));
    HIPCHK_RET(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    FusedCtl* ctl = reinterpret_cast<FusedCtl*>(ctl_mem);
    FusedIn fi;
    fi.n_paths = NP;
    fi.verbs = in.verbs;
    fi.verb_off = in.seg_verb_off;
    fi.points = in.points;
    fi.point_off = in.seg_point_off;
    fi.weights = in.weights;
    fi.weight_off = in.seg_weight_off;
    fi.xf = in.seg_xf;
    fi.n_points = in.n_points;
    fi.n_weights = in.n_weights;
    FusedOut fo;
    fo.tile_xy = out.tile_xy;
    fo.tile_alpha = out.tile_alpha;
    fo.tile_cov = out.tile_cov;
    fo.tile_cap = out.tile_capacity;
    fo.span_xy = out.span_xy;
    fo.span_w = out.span_w;
    fo.span_cap = out.span_capacity;
    fo.pts = out.path_tile_start;
    fo.ptc = out.path_tile_count;
    fo.pss = out.path_span_start;
    fo.psc = out.path_span_count;
    FusedIn f2 = fi;  // the 256-tile instance runs over the 64-tile one's hand-ons
    f2.ids = ids;
    f2.n_ids = &ctl->dq;
    hipLaunchKernelGGL(k_fused_reset, dim3(1), dim3(64), 0, s, ctl);
    hipLaunchKernelGGL((k_dense<4, 64>), dim3(cus * 4), dim3(4 * WAVE), 0, s, fi, fo, ctl, &ctl->queue, ids,
                       &ctl->dq);
    hipLaunchKernelGGL((k_dense<16, 256>), dim3(cus), dim3(16 * WAVE), 0, s, f2, fo, ctl, &ctl->queue3,
                       (uint32_t*)nullptr, &ctl->dq2);
    HIPCHK_RET(hipGetLastError());
    return 0;
}

int finish_dense_chain(const uint8_t* ctl_mem, OutputDesc& out, hipStream_t s) {
    FusedCtl hc;
    HIPCHK_RET(hipMemcpyAsync(&hc, ctl_mem, sizeof(FusedCtl), hipMemcpyDeviceToHost, s));
    HIPCHK_RET(hipStreamSynchronize(s));
    if (hc.error & ERR_INPUT) return -1;
    if (hc.error & ERR_FLATTEN) return -5;
    if (hc.error & ERR_BBOX) return -3;
    out.n_tiles = (int64_t)hc.tiles;
    out.n_spans = (int64_t)hc.spans;
    if ((hc.error & ERR_CAPACITY) || out.n_tiles > out.tile_capacity || out.n_spans > out.span_capacity) return -4;
    if (hc.dq2) return -7;
    return 0;
}

}  // namespace ochre
