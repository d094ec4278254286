// The plain five-stage pipeline (gfx950):
//   (a) flatten      k_flatten<false> counts lines per command, a device scan
//                    allocates slots, k_flatten<true> writes them
//   (b) binning      k_dda<false>/<true> walk each line's DDA and emit one
//                    sort entry per tile run and per tile-row crossing,
//                    keyed (path, tile_y, tile_x); k_rs_* radix-sort them
//   (c) accumulation k_tile_rows: one wave per tile row; a lane per run
//                    replays the row's runs, then the runs are added in
//                    sorted order, a lane per increment, into LDS sums
//   (d) row scan     the carry (rasterizer.rs:233-250) folded across the
//                    lanes of each pixel row, and the winding of the tile-row
//                    crossings (rasterizer.rs:252-261)
//   (e) compaction   k_tile_counts counts tiles and spans per tile row; they
//                    are written at the scanned offsets
// References: glowcoil/ochre src/path.rs:41-109, src/rasterizer.rs:61-268.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "pipeline.hpp"
#include "radix_sort.hpp"
#include "raster_common.hpp"
#include "scan.hpp"

namespace ochre {

// ----------------------------------------------------------------- utils

hipError_t DBuf::ensure(size_t need) {
    if (need <= bytes && p) return hipSuccess;
    size_t nb = std::max<size_t>(need + need / 4, 4096);
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        bytes = 0;
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&p, nb);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    bytes = nb;
    return hipSuccess;
}

DBuf::~DBuf() {
    if (p) (void)hipFree(p);
}

AtlasState::~AtlasState() {
    if (host) (void)hipHostFree(host);
}

StrokeBuffers::~StrokeBuffers() {
    if (host) (void)hipHostFree(host);
}

Context::~Context() {
    if (host_pinned) (void)hipHostFree(host_pinned);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

int map_hip_error(hipError_t e) {
    if (e == hipSuccess) return 0;
    if (e == hipErrorOutOfMemory) return -2;
    return -3;
}

hipError_t ensure_scan_capacity(Context* ctx, int64_t n) {
    int64_t blocks = std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
    return ctx->scan_partials.ensure((size_t)blocks * 16);
}

template <typename T, typename Op, typename Load, typename Store>
static hipError_t launch_scan(Context* ctx, int64_t n, Load load, Store store, Op op, T identity, T* total,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n <= SCAN_TILE) {
        hipLaunchKernelGGL((k_scan_single<T, Op, Load, Store>), dim3(1), dim3(SCAN_NT), 0, s, load, store, n, op,
                           identity, total);
        return hipGetLastError();
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if ((size_t)nb * sizeof(T) > ctx->scan_partials.bytes) return hipErrorInvalidValue;
    T* partials = ctx->scan_partials.as<T>();
    hipLaunchKernelGGL((k_scan_reduce<T, Op, Load>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s, load, n, partials, op,
                       identity);
    hipLaunchKernelGGL((k_scan_partials<T, Op>), dim3(1), dim3(SCAN_NT), 0, s, partials, nb, op, identity, total);
    hipLaunchKernelGGL((k_scan_down<T, Op, Load, Store>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s, load, store, n,
                       (const T*)partials, op, identity);
    return hipGetLastError();
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
static void fill(T* p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
}

static inline uint32_t bits_for(uint64_t n) {  // smallest b with (1 << b) >= n
    uint32_t b = 0;
    while (b < 63 && (1ull << b) < n) b++;
    return b;
}

// --------------------------------------------------- (a) command prologue

struct CmdAcc {
    uint32_t pts;
    uint32_t conics;
    int32_t move;
    uint32_t pad;
};
struct OpCmd {
    __device__ CmdAcc operator()(CmdAcc a, CmdAcc b) const {
        return CmdAcc{a.pts + b.pts, a.conics + b.conics, a.move > b.move ? a.move : b.move, 0u};
    }
};
struct LoadCmd {
    const uint8_t* verbs;
    __device__ CmdAcc operator()(int64_t i) const {
        int v = verbs[i];
        int np = verb_points(v);
        return CmdAcc{(uint32_t)(np < 0 ? 0 : np), v == CONIC ? 1u : 0u, v == MOVE ? (int32_t)i : -1, 0u};
    }
};
struct StoreCmd {
    CmdAcc* out;
    __device__ void operator()(int64_t i, CmdAcc excl, CmdAcc) const { out[i] = excl; }
};

// Largest index s in [0, n) with arr[s] <= v (arr non-decreasing, arr[0] <= v).
__device__ inline int64_t seg_search(const int64_t* arr, int64_t n, int64_t v) {
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (arr[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline V2 load_point(const SegInput& in, int64_t g, int64_t s) {
    V2 p = mk(in.points[2 * g], in.points[2 * g + 1]);
    if (in.seg_raw && in.seg_raw[s]) return p;
    return apply(in.seg_xf[s], p);
}

__device__ inline Cmd resolve_cmd(const SegInput& in, const CmdAcc* scan, int64_t i) {
    Cmd c;
    int64_t s = seg_search(in.seg_verb_off, in.n_segs, i);
    int64_t p = in.seg_path ? in.seg_path[s] : s;
    int64_t fs = in.path_seg_off ? in.path_seg_off[p] : p;
    int64_t ls = in.path_seg_off ? in.path_seg_off[p + 1] : p + 1;
    int64_t vb = in.seg_verb_off[fs], ve = in.seg_verb_off[ls];
    int64_t pb = in.seg_point_off[fs];
    CmdAcc a = scan[i];
    int64_t g = a.pts;
    c.path = (int32_t)p;
    c.verb = in.verbs[i];
    c.last_in_path = (i == ve - 1);
    int np = verb_points(c.verb);
    c.bad = np < 0;
    if (i == in.seg_verb_off[s] && (g != in.seg_point_off[s] || (int64_t)a.conics != in.seg_weight_off[s]))
        c.bad = true;
    if (np < 0) np = 0;
    if (g + np > in.n_points) {
        c.bad = true;
        np = 0;
    }
    // static indices keep the command in registers
    if (np > 0) c.pt[0] = load_point(in, g, s);
    if (np > 1) c.pt[1] = load_point(in, g + 1, s);
    if (np > 2) c.pt[2] = load_point(in, g + 2, s);
    c.w = 0.0f;
    if (c.verb == CONIC) {
        if ((int64_t)a.conics < in.n_weights) c.w = in.weights[a.conics];
        else c.bad = true;
    }
    if (g > pb) {
        int64_t sp = (g - 1 >= in.seg_point_off[s]) ? s : seg_search(in.seg_point_off, in.n_segs, g - 1);
        c.start = load_point(in, g - 1, sp);
    } else {
        c.start = mk(0.0f, 0.0f);
    }
    if (a.move >= vb) {
        int64_t ms = seg_search(in.seg_verb_off, in.n_segs, a.move);
        int64_t gm = scan[a.move].pts;
        if (gm < in.n_points) c.first = load_point(in, gm, ms);
        else c.bad = true;
    } else {
        c.first = mk(0.0f, 0.0f);
    }
    return c;
}

struct FlattenArgs {
    SegInput in;
    const CmdAcc* scan;
    uint32_t* cmd_lines;
    const uint32_t* cmd_line_off;
    float4* lines;
    uint32_t* line_path;
    PathBox* box;
    Counters* cnt;
};

// A command of the emit pass's block, for the threads that make its lines:
// its points, a curve's step, and command_lines' decomposition of its lines
// (a Move's implicit close or a Line's own line, a curve's segments, the
// path's final close).
struct FlatCmd {
    float2 start, p0, p1, p2, first;
    float dt;
    uint32_t info;  // verb | segments << 4 (up to FLATTEN_LIMIT) | own << 30 | close << 31
    uint32_t path;
};

template <bool EMIT>
__global__ __launch_bounds__(256) void k_flatten(FlattenArgs a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in_range = i < a.in.n_verbs;
    if (EMIT) {
        // A thread per command resolves it (a conic's lines it writes itself);
        // then a thread per line of the block's commands makes its line, so
        // that consecutive threads write consecutive lines.
        __shared__ FlatCmd fc[256];
        __shared__ uint32_t foff[257];
        const int t = threadIdx.x;
        const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
        const int nb = (int)min<int64_t>((int64_t)blockDim.x, a.in.n_verbs - i0);
        if (in_range) {
            const Cmd c = resolve_cmd(a.in, a.scan, i);
            const uint32_t o = a.cmd_line_off[i];
            foff[t] = o;
            if (t == nb - 1) foff[nb] = o + a.cmd_lines[i];
            FlatCmd f;
            f.start = make_float2(c.start.x, c.start.y);
            f.p0 = make_float2(c.pt[0].x, c.pt[0].y);
            f.p1 = make_float2(c.pt[1].x, c.pt[1].y);
            f.p2 = make_float2(c.pt[2].x, c.pt[2].y);
            f.first = make_float2(c.first.x, c.first.y);
            f.dt = 0.0f;
            f.path = (uint32_t)c.path;
            uint32_t own = 0, nseg = 0;
            V2 last = c.start, first = c.first;
            if (c.bad) {
                f.info = 0xfu;  // no lines (the count pass reported it)
            } else if (c.verb == CONIC) {
                uint32_t q = o;
                command_lines(c, [&](V2 p0, V2 p1) {
                    a.lines[q] = make_float4(p0.x, p0.y, p1.x, p1.y);
                    a.line_path[q] = (uint32_t)c.path;
                    q++;
                });
                f.info = (uint32_t)CONIC;
            } else {
                if (c.verb == MOVE) {
                    own = veq(last, first) ? 0u : 1u;
                    last = first = c.pt[0];
                } else if (c.verb == LINE) {
                    own = 1u;
                    last = c.pt[0];
                } else if (c.verb == QUAD || c.verb == CUBIC) {
                    const bool cubic = c.verb == CUBIC;
                    f.dt = cubic ? cubic_dt(last, c.pt[0], c.pt[1], c.pt[2]) : quad_dt(last, c.pt[0], c.pt[1]);
                    bool limited = false;
                    nseg = t_steps(f.dt, &limited);
                    last = cubic ? cubic_at(1.0f, last, c.pt[0], c.pt[1], c.pt[2]) : quad_at(1.0f, last, c.pt[0], c.pt[1]);
                }
                const uint32_t close = (c.last_in_path && !veq(last, first)) ? 1u : 0u;
                f.info = (uint32_t)c.verb | (nseg << 4) | (own << 30) | (close << 31);
            }
            fc[t] = f;
        }
        __syncthreads();
        if (nb <= 0) return;
        const uint32_t l0 = foff[0], l1 = foff[nb];
        for (uint32_t q = l0 + (uint32_t)t; q < l1; q += blockDim.x) {
            int lo = 0, hi = nb - 1;  // the last command with foff <= q (it has lines)
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (foff[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            const FlatCmd f = fc[lo];
            const int verb = (int)(f.info & 0xfu);
            if (verb == CONIC || verb == 0xf) continue;  // written above / none
            const uint32_t j = q - foff[lo], nseg = (f.info >> 4) & 0x3ffffffu, own = (f.info >> 30) & 1u;
            const V2 c0 = mk(f.start.x, f.start.y), fp = mk(f.first.x, f.first.y);
            const V2 c1 = mk(f.p0.x, f.p0.y), c2 = mk(f.p1.x, f.p1.y), c3 = mk(f.p2.x, f.p2.y);
            V2 pa, pb;
            if (verb == QUAD || verb == CUBIC) {
                const bool cubic = verb == CUBIC;
                if (j < nseg) {  // segment j: t accumulated the reference's way (path.rs:49-74)
                    float tt = 0.0f, tp = 0.0f;
                    for (uint32_t k = 0; k <= j; k++) {
                        tp = tt;
                        tt = fminf(tt + f.dt, 1.0f);
                    }
                    pa = j == 0 ? c0 : (cubic ? cubic_at(tp, c0, c1, c2, c3) : quad_at(tp, c0, c1, c2));
                    pb = cubic ? cubic_at(tt, c0, c1, c2, c3) : quad_at(tt, c0, c1, c2);
                } else {  // the path's final close from the curve's end
                    pa = cubic ? cubic_at(1.0f, c0, c1, c2, c3) : quad_at(1.0f, c0, c1, c2);
                    pb = fp;
                }
            } else if (own && j == 0) {  // a Move's implicit close, or a Line
                pa = c0;
                pb = verb == MOVE ? fp : c1;
            } else {  // the path's final close, from a Line's point or the command's start (a Close)
                pa = verb == LINE ? c1 : c0;
                pb = fp;
            }
            a.lines[q] = make_float4(pa.x, pa.y, pb.x, pb.y);
            a.line_path[q] = f.path;
        }
        return;
    }
    // count pass: lines per command, and a box holding the path's lines (from
    // the commands' control points: command_count).  A path's commands are
    // contiguous, so the wave reduces the box per path (segmented by path) and
    // one lane per path segment does the atomics.
    int32_t key = -1;
    uint32_t n = 0, flags = 0;
    BoxAcc box;
    if (in_range) {
        Cmd c = resolve_cmd(a.in, a.scan, i);
        key = c.path;
        if (c.bad) {
            atomicOr(&a.cnt->error, (uint32_t)ERR_INPUT);
        } else {
            const bool ok = command_count(c, n, [&](V2 q) { box.add(q); });
            if (!ok) atomicOr(&a.cnt->error, (uint32_t)ERR_FLATTEN);
        }
        a.cmd_lines[i] = n;
        if (n > 0) flags = box.full ? (uint32_t)(BOX_FULL | BOX_LINES) : (uint32_t)BOX_LINES;
    }
    int32_t mnx = INT32_MAX, mny = INT32_MAX, mxx = INT32_MIN, mxy = INT32_MIN;
    if (n > 0 && box.minx <= box.maxx) {
        mnx = (int32_t)floorf(box.minx);
        mny = (int32_t)floorf(box.miny);
        mxx = (int32_t)floorf(box.maxx);
        mxy = (int32_t)floorf(box.maxy);
    }
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int32_t k2 = __shfl_up(key, d, WAVE);
        const uint32_t f2 = __shfl_up(flags, d, WAVE);
        const int32_t a0 = __shfl_up(mnx, d, WAVE), a1 = __shfl_up(mny, d, WAVE);
        const int32_t a2 = __shfl_up(mxx, d, WAVE), a3 = __shfl_up(mxy, d, WAVE);
        if (lane >= d && k2 == key) {
            flags |= f2;
            mnx = min(mnx, a0);
            mny = min(mny, a1);
            mxx = max(mxx, a2);
            mxy = max(mxy, a3);
        }
    }
    const int32_t knext = __shfl_down(key, 1, WAVE);
    const bool seg_end = lane == WAVE - 1 || knext != key;
    if (key >= 0 && seg_end && flags) {
        PathBox* b = &a.box[key];
        atomicOr(&b->flags, flags);
        if (mnx <= mxx) {
            atomicMin(&b->minx, mnx);
            atomicMin(&b->miny, mny);
            atomicMax(&b->maxx, mxx);
            atomicMax(&b->maxy, mxy);
        }
    }
}

__global__ void k_path_grid(const PathBox* box, PathGrid* grid, uint32_t* rows, int64_t n_paths, Counters* cnt) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    PathBox b = box[p];
    PathGrid g;
    grid_from_box((b.flags & BOX_LINES) != 0, (b.flags & BOX_FULL) != 0, b.minx, b.miny, b.maxx, b.maxy, &g.tx0,
                  &g.ty0, &g.w, &g.h);
    if (b.flags & BOX_LINES) atomicMax(&cnt->max_w, (uint32_t)g.w);
    g.row_base = 0;
    grid[p] = g;
    rows[p] = (uint32_t)g.h + 1;  // + one row that carries the empty-path tile
}

struct LoadU32 {
    const uint32_t* p;
    __device__ uint32_t operator()(int64_t i) const { return p[i]; }
};
struct StoreU32 {
    uint32_t* p;
    __device__ void operator()(int64_t i, uint32_t excl, uint32_t) const { p[i] = excl; }
};
struct StoreRowBase {
    PathGrid* g;
    __device__ void operator()(int64_t i, uint32_t excl, uint32_t) const { g[i].row_base = excl; }
};

// In the unordered tile stage the run that starts a line has no record (its
// replay starts where the line does): its entry's value is the line; the
// value of a later run is its record's index with RT_REC set.  (Replaying
// every run from its line's start instead of keeping records measured slower:
// the extra steps cost more than the records' writes and gathers.)
constexpr uint32_t RT_REC = 0x40000000u;

// walk_line's run and entry counts (CountSink) without the walk, for the lines
// chain_walk_ok covers (rasterizer.rs:72-140).
// Each step moves one pixel along x or y, taking the smaller of the two
// crossing parameters, and the walk ends at the first parameter it takes
// that equals 1.  Each axis' parameters form a chain of their own (c, min(c +
// step, 1), ...) whatever the interleaving, so every value below 1 of both
// chains is taken before the end: the walk's last pixel before the end is its
// start moved by nx along x and ny along y (the chains' values below 1).  A
// step changes one tile coordinate by at most one, so the runs are 1 + the
// tile-x and tile-y changes, and the tile-row crossings are the tile-y changes
// plus one when the end-point snap (:118-121) lands in another tile row.
// Returns false for the lines it does not cover.
__device__ inline bool line_counts(V2 last, V2 point, uint32_t& runs, uint32_t& entries) {
    if (!chain_walk_ok(last, point)) return false;
    runs = entries = 0;
    if (veq(point, last)) return true;  // rasterizer.rs:73
    Dda d;
    dda_init(d, last, point);
    int32_t nx = 0, ny = 0;
    for (float c = d.col_t1; c < 1.0f; c = fminf(c + d.x_step, 1.0f)) nx++;
    for (float c = d.row_t1; c < 1.0f; c = fminf(c + d.y_step, 1.0f)) ny++;
    const int32_t x0 = d.x, y0 = d.y;
    const int32_t xl = x0 + (int32_t)d.x_dir * nx, yl = y0 + (int32_t)d.y_dir * ny;
    const uint32_t ty_changes = (uint32_t)abs((yl >> 3) - (y0 >> 3));
    runs = 1u + (uint32_t)abs((xl >> 3) - (x0 >> 3)) + ty_changes;
    const int32_t ty_end = (int32_t)f2i16(floorf(point.y)) >> 3;
    entries = runs + ty_changes + (ty_end != (yl >> 3) ? 1u : 0u);
    return true;
}

struct CountSink {
    uint32_t runs = 0, entries = 0;
    __device__ void run(int16_t, int16_t, bool, const Dda&) {
        runs++;
        entries++;
    }
    __device__ void crossing(int16_t, int16_t, int8_t) { entries++; }
    Counters* cnt;
    __device__ void limit() { atomicOr(&cnt->error, (uint32_t)ERR_FLATTEN); }
};

struct EmitSink {
    uint32_t line;
    uint32_t entry, run_idx;
    bool unordered;     // the unordered tile stage's entries (see RT_FIRST_RUN)
    PathGrid g;
    uint32_t xbits;
    uint32_t* keys;
    uint32_t* vals;
    RunRec* recs;
    Counters* cnt;
    __device__ bool key_of(int16_t tx, int16_t ty, uint32_t* k) {
        uint32_t lx = (uint32_t)((int32_t)tx - g.tx0);
        uint32_t ly = (uint32_t)((int32_t)ty - g.ty0);
        if (lx >= (uint32_t)g.w || ly >= (uint32_t)g.h) {
            atomicOr(&cnt->error, (uint32_t)ERR_BBOX);
            return false;
        }
        *k = ((g.row_base + ly) << xbits) | lx;
        return true;
    }
    __device__ void run(int16_t tx, int16_t ty, bool start, const Dda& d) {
        uint32_t k = 0;
        key_of(tx, ty, &k);
        keys[entry] = k;
        if (unordered && start) {
            vals[entry] = line;
            entry++;
            return;
        }
        vals[entry] = run_idx | (unordered ? RT_REC : 0u);
        RunRec r;
        r.line = line | (start ? 0x80000000u : 0u);
        r.xy = (uint32_t)(uint16_t)d.x | ((uint32_t)(uint16_t)d.y << 16);
        r.row_t1 = d.row_t1;
        r.col_t1 = d.col_t1;
        recs[run_idx] = r;
        entry++;
        run_idx++;
    }
    __device__ void limit() {}
    __device__ void crossing(int16_t tx, int16_t ty, int8_t sign) {
        uint32_t k = 0;
        key_of(tx, ty, &k);
        keys[entry] = k;
        vals[entry] = 0x80000000u | (uint32_t)(uint8_t)sign;
        entry++;
    }
};

struct DdaArgs {
    const float4* lines;
    const uint32_t* line_path;
    int64_t n_lines;
    uint64_t* line_cnt;        // count pass: (runs << 32) | entries per line, read by the emit pass
    const PathGrid* grid;
    uint32_t xbits;
    uint32_t* keys;
    uint32_t* vals;
    RunRec* recs;
    Counters* cnt;
    uint64_t* blk_entries;     // count pass: the entries of each block's lines
    uint64_t* blk_cnt;         // count pass: each block's (runs << 32) | entries; emit pass: exclusive offsets
    bool unordered;            // the unordered tile stage's entries (RT_REC)
};


// Line offsets: the count pass leaves each line's packed (runs, entries)
// count and each block's sum; the blocks' sums are scanned, and the emit pass
// adds a block-local scan of its lines' counts (no device-wide scan over the
// lines).  A block's packed sum cannot carry between the halves: 256 lines of
// at most WALK_LIMIT entries each.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_dda(DdaArgs a) {
    int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    __shared__ unsigned long long part[4];
    if (!EMIT) {
        unsigned long long e = 0, c = 0;
        if (l < a.n_lines) {
            float4 L = a.lines[l];
            CountSink s;
            s.cnt = a.cnt;
            if (!line_counts(mk(L.x, L.y), mk(L.z, L.w), s.runs, s.entries))
                walk_line(mk(L.x, L.y), mk(L.z, L.w), s);
            const uint32_t runs = (a.unordered && s.runs) ? s.runs - 1u : s.runs;  // no record for the first run
            c = ((uint64_t)runs << 32) | s.entries;
            a.line_cnt[l] = c;
            e = s.entries;
        }
        // the 64-bit total of the block's entries (k_sum_u64 adds the blocks):
        // the packed block offsets are scanned in 32-bit halves, which wrap
        // silently past 2^32
        for (int o = 32; o > 0; o >>= 1) {
            e += __shfl_xor(e, o);
            c += __shfl_xor(c, o);
        }
        __shared__ unsigned long long cpart[4];
        if ((threadIdx.x & 63) == 0) {
            part[threadIdx.x >> 6] = e;
            cpart[threadIdx.x >> 6] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            a.blk_entries[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
            a.blk_cnt[blockIdx.x] = cpart[0] + cpart[1] + cpart[2] + cpart[3];
        }
        return;
    }
    const uint64_t c = l < a.n_lines ? a.line_cnt[l] : 0ull;
    unsigned long long btotal;
    const uint64_t off = a.blk_cnt[blockIdx.x] +
                         block_excl_scan<256, unsigned long long>(c, OpAdd(), 0ull, part, &btotal);
    if (l >= a.n_lines) return;
    float4 L = a.lines[l];
    {
        EmitSink s;
        s.line = (uint32_t)l;
        s.entry = (uint32_t)off;
        s.run_idx = (uint32_t)(off >> 32);
        s.g = a.grid[a.line_path[l]];
        s.xbits = a.xbits;
        s.keys = a.keys;
        s.vals = a.vals;
        s.recs = a.recs;
        s.cnt = a.cnt;
        s.unordered = a.unordered;
        if (!walk_line_tiles(mk(L.x, L.y), mk(L.z, L.w), s)) walk_line(mk(L.x, L.y), mk(L.z, L.w), s);
    }
}

// *out += sum of a[0, n): a grid of at most 256 blocks, one atomic each.
__global__ __launch_bounds__(256) void k_sum_u64(const uint64_t* a, int64_t n, unsigned long long* out) {
    unsigned long long v = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += a[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __shared__ unsigned long long part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

struct LoadU64 {
    const uint64_t* p;
    __device__ uint64_t operator()(int64_t i) const { return p[i]; }
};
struct StoreU64 {
    uint64_t* p;
    __device__ void operator()(int64_t i, uint64_t excl, uint64_t) const { p[i] = excl; }
};

// ------------------------------------------------ groups and tile rows

struct LoadGroupStart {
    const uint32_t* keys;
    __device__ uint32_t operator()(int64_t i) const { return (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u; }
};
// Group tables from the sorted entries: each group's first entry and key,
// whether it holds a run, and its crossing sum.  A thread's chunk of
// consecutive entries is taken at once: one sign atomic and one run flag per
// group the chunk touches, instead of one per entry.
struct StoreGroup {
    using chunk_tag = void;
    const uint32_t* keys;
    const uint32_t* vals;
    uint32_t* gbegin;
    uint32_t* gkey;
    uint32_t* ghasrun;
    int32_t* gsign;
    int64_t n;
    __device__ void flush(uint32_t gid, int32_t sg, bool run) const {
        if (sg) atomicAdd(&gsign[gid], sg);
        if (run) ghasrun[gid] = 1u;
    }
    // items [base, base + SCAN_ITEMS) of flags v (1: a group starts), r = the
    // groups started before base
    __device__ void chunk(int64_t base, int64_t n_, uint32_t r, const uint32_t* v) const {
        uint32_t gid = 0;
        int32_t sg = 0;
        bool run = false, open = false;
        for (int j = 0; j < SCAN_ITEMS; j++) {
            const int64_t i = base + j;
            if (i >= n_) break;
            if (v[j]) {
                if (open) flush(gid, sg, run);
                gid = r;
                gbegin[gid] = (uint32_t)i;
                gkey[gid] = keys[i];
                sg = 0;
                run = false;
            } else if (!open) {
                gid = r - 1;  // the group continues from the previous chunk
            }
            open = true;
            r += v[j];
            const uint32_t val = vals[i];
            if (val >> 31)
                sg += (int8_t)(val & 0xffu);
            else
                run = true;
            if (i == n_ - 1) gbegin[gid + 1] = (uint32_t)n_;
        }
        if (open) flush(gid, sg, run);
    }
};

struct RowArgs {
    const uint32_t* gkey;
    const int32_t* gsign;
    const uint32_t* n_groups;
    uint32_t xbits;
    uint32_t* row_gbegin;
    uint32_t* row_gend;
    int32_t* row_sign;
};

__global__ void k_row_table(RowArgs a, int64_t cap) {
    const uint32_t G = *a.n_groups;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < G && g < cap;
         g += (int64_t)gridDim.x * blockDim.x) {
        uint32_t row = a.gkey[g] >> a.xbits;
        if (g == 0 || (a.gkey[g - 1] >> a.xbits) != row) a.row_gbegin[row] = (uint32_t)g;
        if (g == G - 1 || (a.gkey[g + 1] >> a.xbits) != row) a.row_gend[row] = (uint32_t)g + 1;
        int s = a.gsign[g];
        if (s) atomicAdd(&a.row_sign[row], s);
    }
}

// Per path: winding entering each tile row (the running sum of all earlier
// crossings, rasterizer.rs:253-260), the row -> path map, and the
// empty-path tile (rasterizer.rs:193-198: no increments still emits one
// all-zero tile at (0, 0)).
__global__ void k_path_rows(const PathGrid* grid, int64_t n_paths, const uint32_t* row_gbegin,
                            const uint32_t* row_gend, const int32_t* row_sign, int32_t* row_wb,
                            uint32_t* row_path, uint64_t* row_cnt) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    PathGrid g = grid[p];
    int32_t running = 0;
    bool nonempty = false;
    for (int r = 0; r < g.h; r++) {
        uint32_t row = g.row_base + r;
        row_wb[row] = running;
        running += row_sign[row];
        nonempty |= row_gend[row] > row_gbegin[row];
        row_path[row] = (uint32_t)p;
    }
    uint32_t filler = g.row_base + g.h;
    row_path[filler] = (uint32_t)p;
    row_cnt[filler] = nonempty ? 0ull : (1ull << 32);
}

// ----------------------------------------- (c)(d)(e) tile rows, one wave each

struct TileArgs {
    const PathGrid* grid;
    const uint32_t* row_path;
    const uint32_t* row_gbegin;
    const uint32_t* row_gend;
    const int32_t* row_wb;
    int64_t n_rows;
    const uint32_t* gbegin;
    const uint32_t* gkey;
    const uint32_t* ghasrun;
    const int32_t* gsign;
    uint32_t xbits;
    const uint32_t* vals;     // sorted entry values
    const RunRec* recs;
    const float4* lines;
    uint64_t* row_cnt;        // count pass output
    const uint64_t* row_off;  // emit pass input: (tiles << 32) | spans exclusive offsets
    int16_t* tile_xy;
    uint8_t* tile_alpha;
    float* tile_cov;          // optional pre-quantisation coverage, 64 per tile
    int16_t* span_xy;
    uint16_t* span_w;
    Counters* cnt;
};

constexpr int TG = 32;             // groups per chunk (one lane each)
constexpr int ROWS_PER_WAVE = 8;   // tile rows per wave

// Inclusive scan restarting at every lane with head set.
__device__ inline int32_t seg_incl_scan(int32_t v, bool head) {
    const int lane = lane_id();
    int f = head ? 1 : 0;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int32_t ov = __shfl_up(v, d, WAVE);
        int of = __shfl_up(f, d, WAVE);
        if (lane >= d && !f) v = ov + v;
        if (lane >= d) f |= of;
    }
    return v;
}

// Count pass of the tile stage: tiles and spans per tile row.  One wave per
// block of 8 consecutive tile rows; each chunk of up to 32 groups is handled
// one lane per group, the rows' running winding carried between chunks.
__global__ __launch_bounds__(WAVE) void k_tile_counts(TileArgs a) {
    // per row slot: running winding, tiles, spans, pending tile (for spans)
    __shared__ int32_t r_w[ROWS_PER_WAVE], r_pend_tx[ROWS_PER_WAVE], r_pend_w[ROWS_PER_WAVE];
    __shared__ uint32_t r_nt[ROWS_PER_WAVE], r_ns[ROWS_PER_WAVE], r_pend[ROWS_PER_WAVE];
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const uint32_t xmask = (1u << a.xbits) - 1;
    const int64_t n_blocks = (a.n_rows + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE;
    for (int64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
        const int64_t R0 = blk * ROWS_PER_WAVE;
        // lanes 0..7 describe the block's rows
        uint32_t gb = 0xffffffffu, ge = 0;
        bool real_row = false;
        if (lane < ROWS_PER_WAVE && R0 + lane < a.n_rows) {
            const int64_t row = R0 + lane;
            const uint32_t p = a.row_path[row];
            const PathGrid G = a.grid[p];
            real_row = row - (int64_t)G.row_base < G.h;  // not the empty-path row
            if (real_row) {
                uint32_t b = a.row_gbegin[row], e = a.row_gend[row];
                if (e > b) {
                    gb = b;
                    ge = e;
                }
                r_w[lane] = a.row_wb[row];
            }
            r_nt[lane] = 0;
            r_ns[lane] = 0;
            r_pend[lane] = 0;
        }
        uint32_t G0 = gb, G1 = ge;
#pragma unroll
        for (int d = 1; d < ROWS_PER_WAVE; d <<= 1) {
            G0 = min(G0, (uint32_t)__shfl_xor(G0, d, WAVE));
            G1 = max(G1, (uint32_t)__shfl_xor(G1, d, WAVE));
        }
        G0 = __shfl(G0, 0, WAVE);
        G1 = __shfl(G1, 0, WAVE);
        __syncthreads();
        for (uint32_t c = G0; c < G1 && G0 < G1; c += TG) {
            const uint32_t g = c + lane;
            const bool valid = lane < TG && g < G1;
            const uint32_t key = valid ? a.gkey[g] : 0u;
            const int64_t row = (int64_t)(key >> a.xbits);
            const int slot = valid ? (int)(row - R0) : ROWS_PER_WAVE;
            int32_t tx = 0;
            if (valid) {
                const PathGrid G = a.grid[a.row_path[row]];
                tx = G.tx0 + (int32_t)(key & xmask);
            }
            const bool hr = valid && a.ghasrun[g];
            const int32_t sg = valid ? a.gsign[g] : 0;
            const int prev_slot = __shfl_up(slot, 1, WAVE);
            const bool head = lane == 0 || prev_slot != slot;
            // lanes of my row slot
            uint64_t same = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const bool bit = (slot >> b) & 1;
                const uint64_t m = __ballot(bit);
                same &= bit ? m : ~m;
            }
            const uint64_t tiles = __ballot(hr);
            const int32_t wend = (valid ? r_w[slot] : 0) + seg_incl_scan(sg, head);
            const uint64_t my_tiles = tiles & same;
            // spans to the next tile of the same row (rasterizer.rs:252-264)
            const uint64_t above = my_tiles & ~le;
            const int nb = above ? (__ffsll((long long)above) - 1) : lane;
            const int32_t btx = __shfl(tx, nb, WAVE);
            const bool span_here = hr && above && btx > tx + 1 && wend != 0;
            // the first tile of a row in this chunk closes that row's pending tile
            const bool first_tile = hr && (my_tiles & lt) == 0;
            bool pend_span = false;
            if (first_tile && r_pend[slot]) pend_span = tx > r_pend_tx[slot] + 1 && r_pend_w[slot] != 0;
            const uint64_t spans = __ballot(span_here);
            const uint64_t pends = __ballot(pend_span);
            __syncthreads();
            // per-row state after this chunk: the last lane of each row segment
            const int next_slot = __shfl_down(slot, 1, WAVE);
            const bool tail = valid && (lane == WAVE - 1 || next_slot != slot || !(lane + 1 < TG && c + lane + 1 < G1));
            const int last = my_tiles ? 63 - __clzll((long long)my_tiles) : lane;
            const int32_t ltx = __shfl(tx, last, WAVE);
            const int32_t lw = __shfl(wend, last, WAVE);
            if (tail) {
                r_w[slot] = wend;
                r_nt[slot] += (uint32_t)__popcll(my_tiles);
                r_ns[slot] += (uint32_t)__popcll(spans & same) + (uint32_t)__popcll(pends & same);
                if (my_tiles) {
                    r_pend[slot] = 1;
                    r_pend_tx[slot] = ltx;
                    r_pend_w[slot] = lw;
                }
            }
            __syncthreads();
        }
        if (lane < ROWS_PER_WAVE && R0 + lane < a.n_rows && real_row)
            a.row_cnt[R0 + lane] = ((uint64_t)r_nt[lane] << 32) | r_ns[lane];
        __syncthreads();
    }
}

// Emit pass of the tile stage: one wave per tile row, walking the row's
// groups (bins, rasterizer.rs:213-265) in key order.  The row's sort entries
// are taken TR_BLOCK at a time and each of the first TR_BLOCK lanes replays
// one run into LDS: its increments in step order, and a 64-byte map from
// pixel to step (a run stays inside one 8x8 tile and the DDA never revisits a
// pixel, so a run has at most 15 increments, at distinct pixels).  Lane p then
// owns pixel p of the tile: for each run in sorted order it looks its pixel up
// in the run's map and adds the increment found there to its register sums,
// so every pixel's f32 sums follow the sorted order exactly
// (rasterizer.rs:221-228), with no read-modify-write through LDS.  At the end
// of a tile the row carry is folded across the 8 lanes of each pixel row in x
// order, the alphas quantised and stored (rasterizer.rs:230-250), and a span
// is written where the previous tile of the row is more than one tile to the
// left and the winding after it is non-zero (rasterizer.rs:252-264).
constexpr int TR_SLOTS = 16;  // increments per run slot
constexpr int TR_BLOCK = 32;  // runs replayed per block (6 KB of LDS per wave; 16 and 64 measured slower)
// (Padding these rows against the lane-strided bank conflicts of the replay
// was measured slower: the replay is a small part of the kernel.)
constexpr int TR_SSTR = TR_SLOTS;  // float2 per run slot row
constexpr int TR_MSTR = 64;        // bytes per run map row

__device__ inline uint32_t rl(uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); }

// The first tile row of a workgroup of a row kernel (rows first_row +
// k * gridDim.x).  Neighbouring tile rows gather the same lines and run
// records: runs of ROW_CHUNK consecutive rows go to the workgroups of one XCD
// (workgroups are dealt round-robin over the eight XCDs by id), so that the
// second gather hits that XCD's L2; the runs themselves go round-robin over
// the XCDs, which keeps their loads even.  Grids that are a multiple of
// 8 * ROW_CHUNK only (row_grid); others take rows in id order.
constexpr int ROW_CHUNK = 16;
__device__ inline int64_t xcd_first_row() {
    if (gridDim.x % (8u * ROW_CHUNK) != 0) return blockIdx.x;
    const int64_t j = blockIdx.x / 8u;  // the workgroup's place among its XCD's
    return ((j / ROW_CHUNK) * 8 + (int64_t)(blockIdx.x % 8u)) * ROW_CHUNK + j % ROW_CHUNK;
}
static inline uint32_t row_grid(int64_t rows, int64_t cap) {
    int64_t g = std::min<int64_t>(rows, cap);
    if (g >= 8 * ROW_CHUNK) g -= g % (8 * ROW_CHUNK);
    return (uint32_t)g;
}

__global__ __launch_bounds__(WAVE) void k_tile_rows(TileArgs a) {
    __shared__ float2 inc[TR_BLOCK * TR_SSTR];
    __shared__ __align__(16) uint8_t smap[TR_BLOCK * TR_MSTR];  // run j, pixel p -> step + 1 (0: not visited)
    const int lane = threadIdx.x;
    const uint32_t xmask = (1u << a.xbits) - 1;
    for (int64_t row = xcd_first_row(); row < a.n_rows; row += gridDim.x) {
        const PathGrid G = a.grid[a.row_path[row]];
        if (row - (int64_t)G.row_base >= G.h) continue;  // the empty-path row (k_path_out)
        const uint32_t gb = a.row_gbegin[row], ge = a.row_gend[row];
        if (ge <= gb) continue;
        const int32_t ty = G.ty0 + (int32_t)(row - (int64_t)G.row_base);
        const uint64_t off = a.row_off[row];
        const uint32_t row_tiles = (uint32_t)(off >> 32), row_spans = (uint32_t)off;
        const uint32_t E_hi = a.gbegin[ge];
        int32_t winding = a.row_wb[row];
        float carry = 0.0f;  // of the lane's pixel row
        uint32_t nt = 0, ns = 0;
        bool have_prev = false;
        int32_t prev_tx = 0, prev_w = 0;
        // groups gbase + lane: key, has runs, crossing sum, end of its entries
        uint32_t gbase = gb, gk = 0, ghr = 0, gend = 0;
        int32_t gsg = 0;
        auto load_groups = [&]() {
            const uint32_t g = gbase + (uint32_t)lane;
            if (g < ge) {
                gk = a.gkey[g];
                ghr = a.ghasrun[g];
                gsg = a.gsign[g];
                gend = a.gbegin[g + 1];
            }
        };
        load_groups();
        uint32_t e = a.gbegin[gb];
        uint32_t ebase = e + TR_BLOCK;  // no entry block loaded yet
        float pa = 0.0f, ph = 0.0f;     // this lane's pixel sums for the current tile
        for (uint32_t g = gb; g < ge; g++) {
            if (g - gbase >= (uint32_t)WAVE) {
                gbase = g;
                load_groups();
            }
            const int j = (int)(g - gbase);
            const uint32_t key = rl(gk, j), hr = rl(ghr, j), g_end = rl(gend, j);
            winding += (int32_t)rl((uint32_t)gsg, j);
            if (!hr) {  // tile-row crossings only
                e = g_end;
                continue;
            }
            const int32_t tx = G.tx0 + (int32_t)(key & xmask);
            if (have_prev && tx > prev_tx + 1 && prev_w != 0) {
                if (lane == 0) {
                    const uint32_t si = row_spans + ns;
                    a.span_xy[2 * si] = (int16_t)((prev_tx + 1) * TILE);
                    a.span_xy[2 * si + 1] = (int16_t)(ty * TILE);
                    a.span_w[si] = (uint16_t)((uint16_t)(tx - prev_tx - 1) * (uint16_t)TILE);
                }
                ns++;
            }
            while (e < g_end) {
                if (e - ebase >= (uint32_t)TR_BLOCK) {
                    // (c) replay the next TR_BLOCK entries, one run per lane
                    ebase = e;
                    __syncthreads();  // the previous block's maps are read
                    if (lane < TR_BLOCK) {
                        uint4* row_map = reinterpret_cast<uint4*>(smap + lane * TR_MSTR);
#pragma unroll
                        for (int k = 0; k < TR_MSTR / 16; k++) row_map[k] = make_uint4(0, 0, 0, 0);
                        const uint32_t ei = ebase + (uint32_t)lane;
                        const uint32_t v = ei < E_hi ? a.vals[ei] : 0x80000000u;
                        if (!(v >> 31)) {
                            const RunRec r = a.recs[v];
                            const float4 L = a.lines[r.line & 0x7fffffffu];
                            Dda d;
                            const V2 last = mk(L.x, L.y), point = mk(L.z, L.w);
                            if (r.line >> 31) {
                                dda_init(d, last, point);
                            } else {
                                dda_consts(d, last, point);
                                d.x = (int16_t)(r.xy & 0xffffu);
                                d.y = (int16_t)(r.xy >> 16);
                                d.row_t1 = r.row_t1;
                                d.col_t1 = r.col_t1;
                                d.t0 = 0.0f;
                                dda_step(d);
                            }
                            const int16_t rtx = div8(d.x), rty = div8(d.y);
                            const bool finite = line_finite(last, point);
                            V2 p0 = dda_point(d, d.t0);
                            bool ended = false;
                            for (int k = 0; k < TR_SLOTS; k++) {
                                float area, height;
                                dda_increment_carried(d, p0, finite, area, height);
                                inc[lane * TR_SSTR + k] = make_float2(area, height);
                                smap[lane * TR_MSTR + (((d.y & 7) << 3) | (d.x & 7))] = (uint8_t)(k + 1);
                                if (dda_step(d) || div8(d.x) != rtx || div8(d.y) != rty) {
                                    ended = true;
                                    break;
                                }
                            }
                            if (!ended) atomicOr(&a.cnt->error, (uint32_t)ERR_REPLAY);
                        }
                    }
                    __syncthreads();
                }
                // up to 8 runs of this block at a time, added in sorted order
                const uint32_t jr = e - ebase;
                const uint32_t m = min(min(8u, g_end - e), (uint32_t)TR_BLOCK - jr);
                uint32_t k8[8];
#pragma unroll
                for (int i = 0; i < 8; i++) k8[i] = (uint32_t)i < m ? smap[(jr + i) * TR_MSTR + lane] : 0u;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    if (k8[i]) {
                        const float2 v = inc[(jr + i) * TR_SSTR + k8[i] - 1];
                        pa = pa + v.x;
                        ph = ph + v.y;
                    }
                }
                e += m;
            }
            // (d)(e) lane = pixel (y = lane >> 3, x = lane & 7)
            const int x = lane & 7;
            float fold = carry, mine = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float hk = __shfl(ph, (lane & ~7) | k, WAVE);
                if (x == k) mine = fold;
                fold = fold + hk;
            }
            carry = fold;
            const float q = fminf(fabsf(mine + pa) * 256.0f, 255.0f);
            if (a.tile_cov) a.tile_cov[(size_t)(row_tiles + nt) * 64 + lane] = fminf(fabsf(mine + pa), 1.0f);
            pa = 0.0f;
            ph = 0.0f;
            const uint32_t ti = row_tiles + nt;
            a.tile_alpha[(size_t)ti * 64 + lane] = (uint8_t)(uint32_t)q;
            if (lane == 0) {
                a.tile_xy[2 * ti] = (int16_t)(tx * TILE);
                a.tile_xy[2 * ti + 1] = (int16_t)(ty * TILE);
            }
            nt++;
            have_prev = true;
            prev_tx = tx;
            prev_w = winding;
        }
    }
}

// ------------------------------------------- unordered tile stage (batched)
// For the large paths the batched pipeline hands on: instead of the radix
// sort, the sort entries are bucketed by (tile row, window of UW tiles) with
// no order inside a bucket, and each window's runs are replayed into LDS sums
// with float atomics, in arrival order — alpha within 1 LSB of the ordered
// result where a sum sits on a quantisation boundary, as in the dense kernels;
// the tile and span sets, which are integer, are exact.  One workgroup per tile
// row walks the row's windows left to right, carrying the winding, the last
// tile and the per-pixel-row height carry across them.
constexpr int UW = 32;   // tiles per window
constexpr int UWS = 65;  // accumulator slots per tile's sums (one of padding)
static_assert(UW <= WAVE && UW % 32 == 0, "a window is at most one wave of tiles, in 32-bit occupancy words");

struct LoadNWin {
    const PathGrid* g;
    __device__ uint32_t operator()(int64_t p) const { return (uint32_t)g[p].h * (((uint32_t)g[p].w + UW - 1) / UW); }
};

// row -> path and the row's first window bucket
__global__ void k_rows_setup(const PathGrid* grid, int64_t n_paths, const uint32_t* pwbase, uint32_t* row_path,
                             uint32_t* row_wbase) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    const PathGrid g = grid[p];
    const uint32_t nwin = ((uint32_t)g.w + UW - 1) / UW;
    for (int r = 0; r <= g.h; r++) {  // the last one is the empty-path row
        row_path[g.row_base + r] = (uint32_t)p;
        row_wbase[g.row_base + r] = pwbase[p] + (uint32_t)r * nwin;
    }
}

__device__ inline uint32_t entry_bucket(uint32_t key, uint32_t xbits, const uint32_t* row_wbase) {
    return row_wbase[key >> xbits] + ((key & ((1u << xbits) - 1u)) / UW);
}

// Bucketing: each workgroup takes BK_ITEMS entries per thread, consecutive
// entries on consecutive lanes.  The entries are in emission order (the runs
// and crossings of neighbouring lines along a contour), so consecutive ones
// mostly share a bucket: in each wave only the first lane of a run of equal
// buckets inserts the run's length into an LDS hash table (open addressing,
// room for every entry), and a bucket gets one global atomic per workgroup.
constexpr int BK_NT = 256;
constexpr int BK_ITEMS = 8;
constexpr int BK_HBITS = 11;
constexpr int BK_HS = 1 << BK_HBITS;  // == BK_NT * BK_ITEMS
constexpr uint32_t BK_EMPTY = 0xffffffffu;
static_assert(BK_HS >= BK_NT * BK_ITEMS, "the bucket hash must hold every entry of a workgroup");

struct BucketLds {
    uint32_t key[BK_HS];
    uint32_t cnt[BK_HS];
    uint32_t aux[BK_HS];  // histogram: the bucket's tile row; scatter: the workgroup's first slot in the bucket
    int32_t sgn[BK_HS];   // histogram: the bucket's crossing sum
};

__device__ inline uint64_t lanemask_le() { return lanemask_lt() | (1ull << (threadIdx.x & (WAVE - 1))); }

// Inserts n entries of bucket b; returns the bucket's slot and the first
// entry's rank within the workgroup.  Terminates: the table has room for
// every entry of the workgroup, hence for every distinct bucket.
__device__ inline uint32_t bucket_insert(BucketLds& h, uint32_t b, uint32_t n, uint32_t* rank) {
    uint32_t slot = (b * 2654435761u) >> (32 - BK_HBITS);
    for (;;) {
        const uint32_t old = atomicCAS(&h.key[slot], BK_EMPTY, b);
        if (old == BK_EMPTY || old == b) {
            *rank = atomicAdd(&h.cnt[slot], n);
            return slot;
        }
        slot = (slot + 1) & (BK_HS - 1);
    }
}

// One wave's 64 consecutive entries (the valid ones a prefix of the lanes):
// the slot of each entry's bucket and its rank within the workgroup.
__device__ inline void bucket_wave(BucketLds& h, bool valid, uint32_t b, uint32_t* slot, uint32_t* rank,
                                   uint32_t* head_row, uint32_t row) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t prev = __shfl_up(b, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != b);
    const uint64_t hm = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    uint32_t sl = 0, rk = 0;
    if (head) {
        const uint64_t above = hm & ~lanemask_le();
        const int next = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
        sl = bucket_insert(h, b, (uint32_t)(next - lane), &rk);
        if (head_row) head_row[sl] = row;
    }
    const int hl = valid ? 63 - __clzll((long long)(hm & lanemask_le())) : lane;
    *slot = __shfl(sl, hl, WAVE);
    *rank = __shfl(rk, hl, WAVE) + (uint32_t)(lane - hl);
}

template <bool HIST>
__device__ inline void bucket_init(BucketLds& h) {
    for (int k = threadIdx.x; k < BK_HS; k += BK_NT) {
        h.key[k] = BK_EMPTY;
        h.cnt[k] = 0;
        if (HIST) h.sgn[k] = 0;
    }
}

// Bucket sizes, and the crossing sum of every tile row.
__global__ __launch_bounds__(BK_NT) void k_bucket_hist(const uint32_t* keys, const uint32_t* vals, int64_t R,
                                                       uint32_t xbits, const uint32_t* row_wbase, uint32_t* bcnt,
                                                       int32_t* row_sign) {
    __shared__ BucketLds h;
    bucket_init<true>(h);
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * BK_NT * BK_ITEMS + threadIdx.x;
    uint32_t kk[BK_ITEMS], vv[BK_ITEMS];
#pragma unroll
    for (int u = 0; u < BK_ITEMS; u++) {
        const int64_t e = e0 + (int64_t)u * BK_NT;
        kk[u] = e < R ? keys[e] : 0u;
        vv[u] = e < R ? vals[e] : 0u;
    }
#pragma unroll
    for (int u = 0; u < BK_ITEMS; u++) {
        const bool valid = e0 + (int64_t)u * BK_NT < R;
        const uint32_t row = kk[u] >> xbits;
        const uint32_t b = valid ? entry_bucket(kk[u], xbits, row_wbase) : BK_EMPTY;
        uint32_t slot, rank;
        bucket_wave(h, valid, b, &slot, &rank, h.aux, row);
        if (valid && (vv[u] >> 31)) {
            const int sg = (int)(int8_t)(vv[u] & 0xffu);
            if (sg) atomicAdd(&h.sgn[slot], sg);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < BK_HS; k += BK_NT) {
        if (h.key[k] == BK_EMPTY) continue;
        atomicAdd(&bcnt[h.key[k]], h.cnt[k]);
        if (h.sgn[k]) atomicAdd(&row_sign[h.aux[k]], h.sgn[k]);
    }
}

// Scatter into the buckets: one cursor atomic per bucket per workgroup.
__global__ __launch_bounds__(BK_NT) void k_bucket_scatter(const uint32_t* keys, const uint32_t* vals, int64_t R,
                                                          uint32_t xbits, const uint32_t* row_wbase, uint32_t* bcur,
                                                          uint32_t* okeys, uint32_t* ovals) {
    __shared__ BucketLds h;
    bucket_init<false>(h);
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * BK_NT * BK_ITEMS + threadIdx.x;
    uint32_t kk[BK_ITEMS], vv[BK_ITEMS], slot[BK_ITEMS], rank[BK_ITEMS];
#pragma unroll
    for (int u = 0; u < BK_ITEMS; u++) {
        const int64_t e = e0 + (int64_t)u * BK_NT;
        kk[u] = e < R ? keys[e] : 0u;
        vv[u] = e < R ? vals[e] : 0u;
    }
#pragma unroll
    for (int u = 0; u < BK_ITEMS; u++) {
        const bool valid = e0 + (int64_t)u * BK_NT < R;
        const uint32_t b = valid ? entry_bucket(kk[u], xbits, row_wbase) : BK_EMPTY;
        bucket_wave(h, valid, b, &slot[u], &rank[u], nullptr, 0u);
        if (!valid) slot[u] = BK_EMPTY;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < BK_HS; k += BK_NT)
        if (h.key[k] != BK_EMPTY) h.aux[k] = atomicAdd(&bcur[h.key[k]], h.cnt[k]);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BK_ITEMS; u++) {
        if (slot[u] != BK_EMPTY) {
            const uint32_t o = h.aux[slot[u]] + rank[u];
            okeys[o] = kk[u];
            ovals[o] = vv[u];
        }
    }
}

// Per path: the winding entering each row, and the empty-path tile.
__global__ void k_path_rows_u(const PathGrid* grid, int64_t n_paths, const uint32_t* pwbase, const uint32_t* bbeg,
                              const int32_t* row_sign, int32_t* row_wb, uint64_t* row_cnt) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    const PathGrid g = grid[p];
    int32_t running = 0;
    for (int r = 0; r < g.h; r++) {
        row_wb[g.row_base + r] = running;
        running += row_sign[g.row_base + r];
    }
    const uint32_t nwin = ((uint32_t)g.w + UW - 1) / UW;
    const bool nonempty = bbeg[pwbase[p] + (uint32_t)g.h * nwin] > bbeg[pwbase[p]];
    row_cnt[g.row_base + g.h] = nonempty ? 0ull : (1ull << 32);
}

struct RowTileArgs {
    const PathGrid* grid;
    const uint32_t* row_path;
    const uint32_t* row_wbase;
    const int32_t* row_wb;
    int64_t n_rows;
    const uint32_t* bbeg;  // bucket starts (n_buckets + 1)
    const uint32_t* keys;  // bucketed entries
    const uint32_t* vals;
    uint32_t xbits;
    const RunRec* recs;
    const float4* lines;
    uint64_t* row_cnt;        // count pass output
    const uint64_t* row_off;  // emit pass input
    int16_t* tile_xy;
    uint8_t* tile_alpha;
    float* tile_cov;  // optional pre-quantisation coverage, 64 per tile
    int16_t* span_xy;
    uint16_t* span_w;
    Counters* cnt;
};

constexpr int RT_NW = 4;
constexpr int RT_NT = RT_NW * WAVE;
constexpr int RT_ROWS = TILE / RT_NW;  // pixel rows whose carry each wave takes
static_assert(RT_ROWS * RT_NW == TILE && RT_ROWS <= 2, "the carry takes one or two pixel rows per wave");
constexpr int RT_U = 2;  // entries per thread whose loads are issued together

// The window's per-pixel sums are fixed point in f64 slots (fx_of: units of
// 2^-30, integer-valued): the sums of the increments (|area|, |height| <= 1 +
// a few ulps, each truncated at 2^-30) are then exact, so that the decoded f32
// sum is the correctly rounded sum of the increments, independent of arrival
// order.  A window of a full-range path (coordinates that may be huge or
// non-finite) or with 2^22 entries or more keeps f32 atomics in the slots' low
// words instead.
constexpr uint32_t RT_FX_MAX_ENTRIES = 1u << 22;  // a pixel takes at most one increment per run: < 2^22 x (1 + ulps) < 2^23

template <bool EMIT>
struct RowLds {
    long long acc_a[EMIT ? UW * UWS : 2];  // tile t, pixel p at t * UWS + p
    long long acc_h[EMIT ? UW * UWS : 2];
    int32_t tsign[UW];
    uint32_t occ[UW / 32];
    uint8_t tlist[UW];  // the window's occupied tiles in x order
    float rtot[EMIT ? UW * 8 : 1];  // per (occupied tile, pixel row): height total, then incoming carry
    uint32_t nt, ns, nocc;
};

// The increments of a run from the DDA state at its first one (p0: that
// increment's start point), into the window's LDS sums; CARRY: every active
// lane's line is finite, so each increment's start point is the previous one's
// end point (see dda_increment_carried).
template <bool EMIT, bool CARRY>
__device__ inline void rt_replay_steps(RowLds<EMIT>& s, Dda& d, V2 p0, bool finite, uint32_t x, bool fx,
                                       Counters* cnt) {
    // the pixel within the run's tile: the run ends when it leaves 0..7 on
    // either axis (or the line ends)
    uint32_t lx = (uint32_t)(d.x & 7), ly = (uint32_t)(d.y & 7);
    const uint32_t xb = x * (uint32_t)UWS;
    const uint32_t lxd = (uint32_t)(int32_t)d.x_dir, lyd = (uint32_t)(int32_t)d.y_dir;
    bool ended = false;
    for (int k = 0; k < 16; k++) {
        float area, height;
        dda_increment_carried(d, p0, CARRY || finite, area, height);
        const uint32_t px = xb + ((ly << 3) | lx);
        if (fx) {
            atomicAdd(reinterpret_cast<double*>(&s.acc_a[px]), fx_of(area));
            atomicAdd(reinterpret_cast<double*>(&s.acc_h[px]), fx_of(height));
        } else {
            atomicAdd(reinterpret_cast<float*>(&s.acc_a[px]), area);
            atomicAdd(reinterpret_cast<float*>(&s.acc_h[px]), height);
        }
        // dda_step, with the in-tile pixel moved alongside
        float c;
        if (d.row_t1 < d.col_t1) {
            c = d.row_t1;
            d.row_t1 = fminf(d.row_t1 + d.y_step, 1.0f);
            d.y = wrap16(d.y + d.y_dir);
            ly += lyd;
        } else {
            c = d.col_t1;
            d.col_t1 = fminf(d.col_t1 + d.x_step, 1.0f);
            d.x = wrap16(d.x + d.x_dir);
            lx += lxd;
        }
        d.t0 = c;
        if (c == 1.0f || (lx | ly) > 7u) {
            ended = true;
            break;
        }
    }
    if (!ended) atomicOr(&cnt->error, (uint32_t)ERR_REPLAY);
}

// Replays a run (x: its tile within the window) into the window's LDS sums
// (fixed-point or float atomics, arrival order).
template <bool EMIT>
__device__ inline void rt_replay_run(RowLds<EMIT>& s, uint32_t x, const RunRec& rr, const float4& L, bool fx,
                                     Counters* cnt) {
    Dda d;
    const V2 last = mk(L.x, L.y), point = mk(L.z, L.w);
    if (rr.line >> 31) {
        dda_init(d, last, point);
    } else {
        dda_consts(d, last, point);
        d.x = (int16_t)(rr.xy & 0xffffu);
        d.y = (int16_t)(rr.xy >> 16);
        d.row_t1 = rr.row_t1;
        d.col_t1 = rr.col_t1;
        d.t0 = 0.0f;
        dda_step(d);
    }
    const bool finite = line_finite(last, point);
    const V2 p0 = dda_point(d, d.t0);
    // a wave-uniform choice: the start-point recomputation only where a line needs it
    if (__ballot(!finite) == 0) rt_replay_steps<EMIT, true>(s, d, p0, finite, x, fx, cnt);
    else rt_replay_steps<EMIT, false>(s, d, p0, finite, x, fx, cnt);
}

// RT_U entries per thread of a window (entries e0 + u * RT_NT), with their
// run records and lines, loaded level by level (each level's loads in flight
// together).  The emit pass loads the next window's first batch while the
// current window's ranks, carries and quantisation run.
struct RtBatch {
    uint32_t kk[RT_U], vv[RT_U];
    RunRec rr[RT_U];
    float4 L[RT_U];
};
__device__ inline void rt_load_keys(RtBatch& b, const RowTileArgs& a, uint32_t e0, uint32_t ee) {
#pragma unroll
    for (int u = 0; u < RT_U; u++) {
        const uint32_t e = e0 + (uint32_t)(u * RT_NT);
        b.kk[u] = e < ee ? a.keys[e] : 0u;
        b.vv[u] = e < ee ? a.vals[e] : 0x80000000u;
    }
}
// An entry's value: bit 31 a crossing (its sign in the low byte); else a
// run record's index with RT_REC set, or the line whose first run it is.
__device__ inline void rt_load_recs(RtBatch& b, const RowTileArgs& a) {
#pragma unroll
    for (int u = 0; u < RT_U; u++) {
        b.rr[u].line = (b.vv[u] & (RT_REC - 1u)) | 0x80000000u;  // a line's first run: replayed from its start
        b.rr[u].xy = 0;
        b.rr[u].row_t1 = b.rr[u].col_t1 = 0.0f;
        if ((b.vv[u] >> 30) == 1u) b.rr[u] = a.recs[b.vv[u] & (RT_REC - 1u)];
    }
}
__device__ inline void rt_load_lines(RtBatch& b, const RowTileArgs& a) {
#pragma unroll
    for (int u = 0; u < RT_U; u++)
        b.L[u] = !(b.vv[u] >> 31) ? a.lines[b.rr[u].line & 0x7fffffffu] : make_float4(0, 0, 0, 0);
}

// One batch of a window's entries: occupancy (into the thread's bits, reduced
// per wave after the window's pass: every run of a wave would otherwise hit the
// one occupancy word) and crossing sums from the keys, and (emit pass) the
// runs replayed.
static_assert(UW == 32, "one occupancy word per window");
template <bool EMIT>
__device__ inline void rt_entries(RowLds<EMIT>& s, const RtBatch& b, uint32_t e0, uint32_t ee, uint32_t xmask,
                                  int32_t x0, bool fx, Counters* cnt, uint32_t& occ_bits) {
#pragma unroll
    for (int u = 0; u < RT_U; u++) {
        if (e0 + (uint32_t)(u * RT_NT) >= ee) break;
        const uint32_t x = (b.kk[u] & xmask) - (uint32_t)x0;
        if (b.vv[u] >> 31) {
            const int sg = (int)(int8_t)(b.vv[u] & 0xffu);
            if (sg) atomicAdd(&s.tsign[x], sg);
        } else {
            occ_bits |= 1u << x;
            if constexpr (EMIT) rt_replay_run(s, x, b.rr[u], b.L[u], fx, cnt);
        }
    }
}

template <bool EMIT>
__global__ __launch_bounds__(RT_NT) void k_row_tiles(RowTileArgs a) {
    __shared__ __align__(16) RowLds<EMIT> s;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint32_t xmask = (1u << a.xbits) - 1u;
    const uint64_t lt = lanemask_lt();
    if constexpr (EMIT) {
        // The window sums start at zero: cleared here once, and after each
        // window the quantisation clears the slots it read (runs only add into
        // their own, occupied, tiles).
        constexpr int n16 = (int)sizeof(s.acc_a) / 16;  // 16 bytes at a time
        static_assert(sizeof(s.acc_a) % 16 == 0, "16-byte clears");
        float4* za = reinterpret_cast<float4*>(s.acc_a);
        float4* zh = reinterpret_cast<float4*>(s.acc_h);
        for (int k = tid; k < n16; k += RT_NT) {
            za[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            zh[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        __syncthreads();
    }
    const int64_t grid_n = gridDim.x;
    for (int64_t row = xcd_first_row(); row < a.n_rows; row += grid_n) {
        const PathGrid G = a.grid[a.row_path[row]];
        const int32_t r = (int32_t)(row - (int64_t)G.row_base);
        if (r >= G.h) continue;  // the empty-path row
        const int32_t ty = G.ty0 + r;
        const uint32_t nwin = ((uint32_t)G.w + UW - 1) / UW;
        const uint32_t b0 = a.row_wbase[row];
        uint32_t row_tiles = 0, row_spans = 0;
        if (EMIT) {
            const uint64_t off = a.row_off[row];
            row_tiles = (uint32_t)(off >> 32);
            row_spans = (uint32_t)off;
        }
        // wave 0's running state along the row
        int32_t winding = a.row_wb[row];
        bool have_prev = false;
        int32_t prev_x = 0, prev_w = 0;
        uint32_t nt = 0, ns = 0;
        float carry0 = 0.0f, carry1 = 0.0f;  // wave w: the height carries of pixel rows RT_ROWS * w + j
        // the first window holding entries (a window with none has no run
        // and no crossing)
        uint32_t wi = 0, eb = a.bbeg[b0], ee = a.bbeg[b0 + 1];
        while (ee == eb && ++wi < nwin) {
            eb = ee;
            ee = a.bbeg[b0 + wi + 1];
        }
        RtBatch nb;       // the first batch of window wi, when pf
        bool pf = false;
        while (wi < nwin) {
            // the next window holding entries
            uint32_t wn = wi + 1, nbeg = ee, nend = ee;
            while (wn < nwin) {
                nend = a.bbeg[b0 + wn + 1];
                if (nend != nbeg) break;
                wn++;
            }
            const bool has_next = EMIT && wn < nwin;
            const int32_t x0 = (int32_t)(wi * UW);
            const bool fx = G.w < 8192 && ee - eb < RT_FX_MAX_ENTRIES;  // fixed-point sums (see RowLds)
            if (tid < UW) s.tsign[tid] = 0;
            if (tid < UW / 32) s.occ[tid] = 0;
            __syncthreads();
            // one pass over the window's entries, a batch at a time
            uint32_t e0 = eb + (uint32_t)tid, occ_bits = 0;
            if (pf) {
                rt_entries(s, nb, e0, ee, xmask, x0, fx, a.cnt, occ_bits);
                e0 += RT_U * RT_NT;
            }
            for (; e0 < ee; e0 += RT_U * RT_NT) {
                RtBatch b;
                rt_load_keys(b, a, e0, ee);
                if constexpr (EMIT) {
                    rt_load_recs(b, a);
                    rt_load_lines(b, a);
                }
                rt_entries(s, b, e0, ee, xmask, x0, fx, a.cnt, occ_bits);
            }
            for (int o = 1; o < WAVE; o <<= 1) occ_bits |= __shfl_xor(occ_bits, o, WAVE);
            if (lane == 0 && occ_bits) atomicOr(&s.occ[0], occ_bits);
            __syncthreads();
            if (has_next) rt_load_keys(nb, a, nbeg + (uint32_t)tid, nend);
            if (wid == 0) {
                // lane = tile x0 + lane: ranks, the winding after each tile,
                // and the spans between consecutive tiles (rasterizer.rs:252-264)
                const bool occ = lane < UW && ((s.occ[lane >> 5] >> (lane & 31)) & 1u);
                const int32_t incl = wave_incl_scan(lane < UW ? s.tsign[lane] : 0, OpAdd());
                const int32_t w_at = winding + incl;
                const uint64_t om = __ballot(occ);
                const uint64_t below = om & lt;
                int32_t px = prev_x, pw = prev_w;
                bool hp = have_prev;
                const int pl = below ? 63 - __clzll((long long)below) : 0;
                const int32_t w_pl = __shfl(w_at, pl, WAVE);
                if (below) {
                    px = x0 + pl;
                    pw = w_pl;
                    hp = true;
                }
                const int32_t gx = x0 + lane;
                const bool span = occ && hp && gx > px + 1 && pw != 0;
                const uint64_t sm = __ballot(span);
                if (EMIT && span) {
                    const uint32_t si = row_spans + ns + (uint32_t)__popcll(sm & lt);
                    a.span_xy[2 * si] = (int16_t)((G.tx0 + px + 1) * TILE);
                    a.span_xy[2 * si + 1] = (int16_t)(ty * TILE);
                    a.span_w[si] = (uint16_t)((uint16_t)(gx - px - 1) * (uint16_t)TILE);
                }
                if (EMIT && occ) s.tlist[__popcll(om & lt)] = (uint8_t)lane;
                if (om) {
                    const int last = 63 - __clzll((long long)om);
                    prev_x = x0 + last;
                    prev_w = __shfl(w_at, last, WAVE);
                    have_prev = true;
                }
                winding = __shfl(w_at, WAVE - 1, WAVE);
                if (lane == 0) {
                    s.nt = nt;
                    s.nocc = (uint32_t)__popcll(om);
                }
                nt += (uint32_t)__popcll(om);
                ns += (uint32_t)__popcll(sm);
            }
            if constexpr (EMIT) {
                __syncthreads();
                if (has_next) rt_load_recs(nb, a);
                const uint32_t nocc = s.nocc, nt_w = s.nt;
                // The height carry: wave w takes RT_ROWS pixel rows from RT_ROWS * w,
                // lane k the window's k-th occupied tile, which folds its row's 8
                // heights into an in-tile prefix and a total; a wave scan over
                // the tiles' totals gives each tile its incoming carry, and the
                // quantise adds the two.  (Regrouping the row's f32 sums this way
                // moves a pixel's carry by a few ulps: within the 1-LSB contract.)
#pragma unroll
                for (int j = 0; j < RT_ROWS; j++) {
                    const int prow = RT_ROWS * wid + j;
                    float tot = 0.0f;
                    if ((uint32_t)lane < nocc) {
                        const int base = (int)s.tlist[lane] * UWS + prow * 8;
                        float* fh = reinterpret_cast<float*>(s.acc_h);  // the slots' low words
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const float h = fx ? fx_to_f(reinterpret_cast<const double*>(s.acc_h)[base + k])
                                               : fh[2 * (base + k)];
                            fh[2 * (base + k)] = tot;
                            tot += h;
                        }
                    }
                    const float incl = wave_incl_scan(tot, OpAdd());
                    const float up = __shfl_up(incl, 1, WAVE);
                    const float excl = lane ? up : 0.0f;
                    float& carry = j ? carry1 : carry0;
                    if ((uint32_t)lane < nocc) s.rtot[lane * 8 + prow] = carry + excl;
                    carry += __shfl(incl, WAVE - 1, WAVE);
                }
                __syncthreads();
                if (has_next) rt_load_lines(nb, a);
                for (uint32_t item = (uint32_t)tid; item < nocc * 16u; item += RT_NT) {
                    // item -> (the occupied tile of rank item / 16, 4 pixels)
                    const uint32_t rk = item >> 4, quad = item & 15u;
                    const int t = s.tlist[rk];
                    const int base = t * UWS + (int)quad * 4;
                    uint32_t word = 0;
                    const uint64_t ti = (uint64_t)row_tiles + nt_w + rk;
                    const float cin = s.rtot[rk * 8 + (quad >> 1)];  // pixel row quad / 2
                    const float* fh = reinterpret_cast<const float*>(s.acc_h);
                    const float* fa = reinterpret_cast<const float*>(s.acc_a);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float av = fx ? fx_to_f(reinterpret_cast<const double*>(s.acc_a)[base + k])
                                            : fa[2 * (base + k)];
                        const float v = (cin + fh[2 * (base + k)]) + av;
                        word |= (uint32_t)fminf(fabsf(v) * 256.0f, 255.0f) << (8 * k);
                        if (a.tile_cov) a.tile_cov[ti * 64 + quad * 4 + (uint32_t)k] = fminf(fabsf(v), 1.0f);
                        s.acc_a[base + k] = 0;  // cleared for the next window
                        s.acc_h[base + k] = 0;
                    }
                    reinterpret_cast<uint32_t*>(a.tile_alpha + ti * 64)[quad] = word;
                    if (quad == 0) {
                        a.tile_xy[2 * ti] = (int16_t)((G.tx0 + x0 + t) * TILE);
                        a.tile_xy[2 * ti + 1] = (int16_t)(ty * TILE);
                    }
                }
            }
            __syncthreads();
            pf = has_next;
            wi = wn;
            eb = nbeg;
            ee = nend;
        }
        if (!EMIT && tid == 0) a.row_cnt[row] = ((uint64_t)nt << 32) | ns;
    }
}

// Per-path output ranges and the empty-path tile.
__global__ void k_path_out(const PathGrid* grid, int64_t n_paths, const uint64_t* row_off, const uint64_t* total,
                           int64_t* tstart, int32_t* tcount, int64_t* sstart, int32_t* scount, int16_t* tile_xy,
                           uint8_t* tile_alpha, float* tile_cov, int64_t tile_base, int64_t span_base) {
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    PathGrid g = grid[p];
    uint64_t b = row_off[g.row_base];
    uint64_t e = (p + 1 < n_paths) ? row_off[grid[p + 1].row_base] : *total;
    uint64_t fb = row_off[g.row_base + g.h];
    tstart[p] = tile_base + (int64_t)(b >> 32);
    tcount[p] = (int32_t)((e >> 32) - (b >> 32));
    sstart[p] = span_base + (int64_t)(uint32_t)b;
    scount[p] = (int32_t)((uint32_t)e - (uint32_t)b);
    if ((e >> 32) - (fb >> 32) == 1) {  // the empty-path tile
        uint64_t ti = fb >> 32;
        tile_xy[2 * ti] = 0;
        tile_xy[2 * ti + 1] = 0;
        uint4* dst = reinterpret_cast<uint4*>(tile_alpha + ti * 64);
        for (int k = 0; k < 4; k++) dst[k] = make_uint4(0, 0, 0, 0);
        if (tile_cov)
            for (int k = 0; k < 64; k++) tile_cov[ti * 64 + k] = 0.0f;
    }
}

// ------------------------------------------------------------- host driver

#define HIPCHK(x) HIPCHK_RET(x)

static inline uint32_t grid_for(int64_t n, int bs) { return (uint32_t)((n + bs - 1) / bs); }

struct HostTotals {
    uint32_t lines;
    uint32_t error;
    uint32_t max_w;
    uint32_t rows;
    uint64_t dda;     // (runs << 32) | entries
    uint32_t groups;
    uint32_t pad;
    uint64_t out;     // (tiles << 32) | spans
    uint64_t entries;
};

// Internal status of one pass: the input is over the per-pass limits.
constexpr int ST_SPLIT = -100;

// Folds a pass's error bits into the call's word: the emit kernels report a
// replay overrun (ERR_REPLAY) after the host's last read of the pass's
// counters, which the next pass clears.
__global__ void k_fold_error(const Counters* cnt, uint32_t* call_error) {
    if (threadIdx.x == 0 && cnt->error) atomicOr(call_error, cnt->error);
}

struct StoreBucketStart {
    uint32_t* beg;
    uint32_t* cur;
    __device__ void operator()(int64_t i, uint32_t excl, uint32_t) const {
        beg[i] = excl;
        cur[i] = excl;
    }
};

static int run_plain_pass(Context* ctx, const SegInput& in, OutputDesc& out, hipStream_t s, bool unordered) {
    const int64_t V = in.n_verbs;
    const int64_t NP = in.n_paths;
    if (NP <= 0) {
        out.n_tiles = out.n_spans = 0;
        return 0;
    }
    if (V >= (1ll << 31) || in.n_points >= (1ll << 32)) {
        ctx->split_ratio = std::max((double)V / (double)(1ll << 31), (double)in.n_points / (double)(1ll << 32));
        return ST_SPLIT;
    }
    HIPCHK(ctx->counters.ensure(256));
    HIPCHK(ctx->totals.ensure(256));
    if (!ctx->host_pinned) HIPCHK(hipHostMalloc(&ctx->host_pinned, 4096));
    Counters* cnt = ctx->counters.as<Counters>();
    uint64_t* tot = ctx->totals.as<uint64_t>();  // [0] lines, [1] rows, [2] dda, [3] groups, [4] out
    HostTotals* ht = reinterpret_cast<HostTotals*>(ctx->host_pinned);
    fill<uint32_t>(reinterpret_cast<uint32_t*>(cnt), 64, 0u, s);
    fill<uint64_t>(tot, 32, 0ull, s);

    HIPCHK(ensure_scan_capacity(ctx, std::max<int64_t>({V, NP, 1})));

    // ---- (a) flatten
    HIPCHK(ctx->cmd_scan.ensure(std::max<int64_t>(V, 1) * sizeof(CmdAcc)));
    HIPCHK(ctx->cmd_lines.ensure(std::max<int64_t>(V, 1) * 4));
    HIPCHK(ctx->cmd_line_off.ensure(std::max<int64_t>(V, 1) * 4));
    HIPCHK(ctx->path_box.ensure(NP * sizeof(PathBox)));
    HIPCHK(ctx->path_grid.ensure(NP * sizeof(PathGrid)));
    HIPCHK(ctx->path_rows.ensure(NP * 4));
    CmdAcc* cscan = ctx->cmd_scan.as<CmdAcc>();
    PathBox* box = ctx->path_box.as<PathBox>();
    {
        PathBox init;
        init.minx = init.miny = INT32_MAX;
        init.maxx = init.maxy = INT32_MIN;
        init.flags = 0;
        init.pad[0] = init.pad[1] = init.pad[2] = 0;
        fill<PathBox>(box, NP, init, s);
    }
    HIPCHK(launch_scan<CmdAcc>(ctx, V, LoadCmd{in.verbs}, StoreCmd{cscan}, OpCmd(), CmdAcc{0, 0, -1, 0},
                               (CmdAcc*)nullptr, s));
    FlattenArgs fa;
    fa.in = in;
    fa.scan = cscan;
    fa.cmd_lines = ctx->cmd_lines.as<uint32_t>();
    fa.cmd_line_off = ctx->cmd_line_off.as<uint32_t>();
    fa.box = box;
    fa.cnt = cnt;
    fa.lines = nullptr;
    fa.line_path = nullptr;
    if (V > 0) {
        hipLaunchKernelGGL(k_flatten<false>, dim3(grid_for(V, 256)), dim3(256), 0, s, fa);
        HIPCHK(launch_scan<uint32_t>(ctx, V, LoadU32{fa.cmd_lines}, StoreU32{ctx->cmd_line_off.as<uint32_t>()},
                                     OpAdd(), 0u, reinterpret_cast<uint32_t*>(&tot[0]), s));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&ht->lines, &tot[0], 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->error, &cnt->error, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ht->error & ERR_INPUT) return -1;
    if (ht->error & ERR_FLATTEN) return -5;
    const int64_t L = ht->lines;
    if (L > ctx->max_lines) {
        ctx->split_ratio = (double)L / (double)ctx->max_lines;
        return ST_SPLIT;
    }

    HIPCHK(ctx->lines.ensure(std::max<int64_t>(L, 1) * 16));
    HIPCHK(ctx->line_path.ensure(std::max<int64_t>(L, 1) * 4));
    HIPCHK(ctx->line_cnt.ensure(std::max<int64_t>(L, 1) * 8));
    HIPCHK(ensure_scan_capacity(ctx, std::max<int64_t>(L, 1)));
    fa.lines = ctx->lines.as<float4>();
    fa.line_path = ctx->line_path.as<uint32_t>();
    if (V > 0) hipLaunchKernelGGL(k_flatten<true>, dim3(grid_for(V, 256)), dim3(256), 0, s, fa);
    PathGrid* grid = ctx->path_grid.as<PathGrid>();
    hipLaunchKernelGGL(k_path_grid, dim3(grid_for(NP, 256)), dim3(256), 0, s, box, grid,
                       ctx->path_rows.as<uint32_t>(), NP, cnt);
    HIPCHK(launch_scan<uint32_t>(ctx, NP, LoadU32{ctx->path_rows.as<uint32_t>()}, StoreRowBase{grid}, OpAdd(), 0u,
                                 reinterpret_cast<uint32_t*>(&tot[1]), s));
    DdaArgs da;
    da.lines = fa.lines;
    da.line_path = fa.line_path;
    da.n_lines = L;
    da.line_cnt = ctx->line_cnt.as<uint64_t>();
    da.grid = grid;
    da.cnt = cnt;
    da.xbits = 0;
    da.keys = da.vals = nullptr;
    da.recs = nullptr;
    da.unordered = unordered;
    const int64_t nblk = (int64_t)grid_for(std::max<int64_t>(L, 1), 256);
    HIPCHK(ctx->dda_blk.ensure((size_t)nblk * 16));
    da.blk_entries = ctx->dda_blk.as<uint64_t>();
    da.blk_cnt = da.blk_entries + nblk;
    if (L > 0) {
        hipLaunchKernelGGL(k_dda<false>, dim3(grid_for(L, 256)), dim3(256), 0, s, da);
        hipLaunchKernelGGL(k_sum_u64, dim3(std::min<uint32_t>(grid_for(nblk, 256), 256)), dim3(256), 0, s,
                           da.blk_entries, nblk, &cnt->entries);
        HIPCHK(launch_scan<uint64_t>(ctx, nblk, LoadU64{da.blk_cnt}, StoreU64{da.blk_cnt}, OpAdd(), 0ull, &tot[2], s));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&ht->rows, &tot[1], 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->dda, &tot[2], 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->max_w, &cnt->max_w, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->error, &cnt->error, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->entries, &cnt->entries, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ht->error & ERR_FLATTEN) return -5;
    // entries >= runs (every run record is also a sort entry), so below the
    // limit both 32-bit halves of the packed total are exact
    if (ht->entries > (uint64_t)ctx->max_entries) {
        ctx->split_ratio = (double)ht->entries / (double)ctx->max_entries;
        return ST_SPLIT;
    }
    const int64_t R = (int64_t)(uint32_t)ht->dda;
    const int64_t RUNS = (int64_t)(ht->dda >> 32);
    const int64_t ROWS = ht->rows;
    const uint32_t xbits = std::max<uint32_t>(1, bits_for(std::max<uint32_t>(ht->max_w, 1)));
    const uint32_t rbits = bits_for((uint64_t)ROWS);
    if (xbits + rbits > 32) {
        ctx->split_ratio = 2.0;
        return ST_SPLIT;
    }
    const uint32_t kbits = xbits + rbits;

    // ---- (b) binning + sort
    HIPCHK(ctx->keys[0].ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->keys[1].ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->vals[0].ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->vals[1].ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->records.ensure(std::max<int64_t>(RUNS, 1) * sizeof(RunRec)));
    const uint32_t rs_blocks = (uint32_t)std::max<int64_t>(1, (R + RS_TILE - 1) / RS_TILE);
    HIPCHK(ctx->rs_counts.ensure((size_t)rs_blocks * 256 * 4));
    HIPCHK(ensure_scan_capacity(ctx, std::max<int64_t>({R, (int64_t)rs_blocks * 256, ROWS, 1})));
    da.xbits = xbits;
    da.keys = ctx->keys[0].as<uint32_t>();
    da.vals = ctx->vals[0].as<uint32_t>();
    da.recs = ctx->records.as<RunRec>();
    if (L > 0) hipLaunchKernelGGL(k_dda<true>, dim3(grid_for(L, 256)), dim3(256), 0, s, da);
    if (unordered) {
        // ---- window buckets instead of the sort
        uint32_t* keys0 = ctx->keys[0].as<uint32_t>();
        uint32_t* vals0 = ctx->vals[0].as<uint32_t>();
        HIPCHK(ctx->u_pwbase.ensure((size_t)NP * 4));
        HIPCHK(ctx->u_rwbase.ensure((size_t)ROWS * 4));
        HIPCHK(ctx->row_path.ensure(ROWS * 4));
        HIPCHK(ctx->row_sign.ensure(ROWS * 4));
        HIPCHK(ctx->row_wb.ensure(ROWS * 4));
        HIPCHK(ctx->row_cnt.ensure(ROWS * 8));
        HIPCHK(ctx->row_off.ensure(ROWS * 8));
        uint32_t* pwbase = ctx->u_pwbase.as<uint32_t>();
        uint32_t* rwbase = ctx->u_rwbase.as<uint32_t>();
        uint32_t* nb_dev = reinterpret_cast<uint32_t*>(&tot[6]);
        HIPCHK(launch_scan<uint32_t>(ctx, NP, LoadNWin{grid}, StoreU32{pwbase}, OpAdd(), 0u, nb_dev, s));
        hipLaunchKernelGGL(k_rows_setup, dim3(grid_for(NP, 256)), dim3(256), 0, s, grid, NP, pwbase,
                           ctx->row_path.as<uint32_t>(), rwbase);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&ht->groups, nb_dev, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const int64_t NB = ht->groups;
        HIPCHK(ctx->u_bcnt.ensure((size_t)std::max<int64_t>(NB, 1) * 4));
        HIPCHK(ctx->u_bbeg.ensure((size_t)(NB + 1) * 4));
        HIPCHK(ctx->u_bcur.ensure((size_t)std::max<int64_t>(NB, 1) * 4));
        HIPCHK(ensure_scan_capacity(ctx, std::max<int64_t>({NB, ROWS, 1})));
        uint32_t* bcnt = ctx->u_bcnt.as<uint32_t>();
        uint32_t* bbeg = ctx->u_bbeg.as<uint32_t>();
        uint32_t* bcur = ctx->u_bcur.as<uint32_t>();
        fill<uint32_t>(bcnt, std::max<int64_t>(NB, 1), 0u, s);
        fill<int32_t>(ctx->row_sign.as<int32_t>(), ROWS, 0, s);
        const uint32_t bk_blocks = grid_for(R, BK_NT * BK_ITEMS);
        if (R > 0)
            hipLaunchKernelGGL(k_bucket_hist, dim3(bk_blocks), dim3(BK_NT), 0, s, keys0, vals0, R, xbits, rwbase,
                               bcnt, ctx->row_sign.as<int32_t>());
        if (NB > 0)
            HIPCHK(launch_scan<uint32_t>(ctx, NB, LoadU32{bcnt}, StoreBucketStart{bbeg, bcur}, OpAdd(), 0u, bbeg + NB,
                                         s));
        else
            fill<uint32_t>(bbeg, 1, 0u, s);
        if (R > 0)
            hipLaunchKernelGGL(k_bucket_scatter, dim3(bk_blocks), dim3(BK_NT), 0, s, keys0, vals0, R, xbits, rwbase,
                               bcur, ctx->keys[1].as<uint32_t>(), ctx->vals[1].as<uint32_t>());
        hipLaunchKernelGGL(k_path_rows_u, dim3(grid_for(NP, 256)), dim3(256), 0, s, grid, NP, pwbase, bbeg,
                           ctx->row_sign.as<int32_t>(), ctx->row_wb.as<int32_t>(), ctx->row_cnt.as<uint64_t>());
        RowTileArgs ra;
        ra.grid = grid;
        ra.row_path = ctx->row_path.as<uint32_t>();
        ra.row_wbase = rwbase;
        ra.row_wb = ctx->row_wb.as<int32_t>();
        ra.n_rows = ROWS;
        ra.bbeg = bbeg;
        ra.keys = ctx->keys[1].as<uint32_t>();
        ra.vals = ctx->vals[1].as<uint32_t>();
        ra.xbits = xbits;
        ra.recs = ctx->records.as<RunRec>();
        ra.lines = ctx->lines.as<float4>();
        ra.row_cnt = ctx->row_cnt.as<uint64_t>();
        ra.row_off = ctx->row_off.as<uint64_t>();
        ra.tile_xy = out.tile_xy;
        ra.tile_alpha = out.tile_alpha;
        ra.tile_cov = out.tile_cov;
        ra.span_xy = out.span_xy;
        ra.span_w = out.span_w;
        ra.cnt = cnt;
        const uint32_t row_blocks = row_grid(ROWS, 1 << 16);
        hipLaunchKernelGGL(k_row_tiles<false>, dim3(row_blocks), dim3(RT_NT), 0, s, ra);
        HIPCHK(launch_scan<uint64_t>(ctx, ROWS, LoadU64{ra.row_cnt}, StoreU64{ctx->row_off.as<uint64_t>()}, OpAdd(),
                                     0ull, &tot[4], s));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&ht->out, &tot[4], 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&ht->error, &cnt->error, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (ht->error & (ERR_BBOX | ERR_REPLAY)) return -3;
        const int64_t T = (int64_t)(ht->out >> 32);
        const int64_t S = (int64_t)(uint32_t)ht->out;
        out.n_tiles = T;
        out.n_spans = S;
        if (T > out.tile_capacity || S > out.span_capacity) return -4;
        if (ctx->timing) {
            // pass statistics beside the kernel times: lines, sort entries,
            // runs, tile rows, windows and the windows holding entries
            std::vector<uint32_t> bc((size_t)NB);
            if (NB > 0) HIPCHK(hipMemcpyAsync(bc.data(), bcnt, (size_t)NB * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            int64_t used = 0;
            for (uint32_t c : bc) used += c != 0;
            const std::pair<const char*, int64_t> st[] = {{"pass.lines", L},   {"pass.entries", R},
                                                         {"pass.runs", RUNS},  {"pass.rows", ROWS},
                                                         {"pass.windows", NB}, {"pass.windows_used", used}};
            for (const auto& kv : st) {
                auto& t = ctx->timings[kv.first];
                t.first += (double)kv.second;
                t.second += 1;
            }
            if (!ctx->ev[0]) HIPCHK(hipEventCreate(&ctx->ev[0]));
            if (!ctx->ev[1]) HIPCHK(hipEventCreate(&ctx->ev[1]));
            HIPCHK(hipEventRecord(ctx->ev[0], s));
        }
        hipLaunchKernelGGL(k_row_tiles<true>, dim3(row_blocks), dim3(RT_NT), 0, s, ra);
        if (ctx->timing) HIPCHK(hipEventRecord(ctx->ev[1], s));
        hipLaunchKernelGGL(k_path_out, dim3(grid_for(NP, 256)), dim3(256), 0, s, grid, NP, ra.row_off,
                           (const uint64_t*)&tot[4], out.path_tile_start, out.path_tile_count, out.path_span_start,
                           out.path_span_count, out.tile_xy, out.tile_alpha, out.tile_cov, out.tile_base, out.span_base);
        hipLaunchKernelGGL(k_fold_error, dim3(1), dim3(64), 0, s, cnt, ctx->call_error.as<uint32_t>());
        HIPCHK(hipGetLastError());
        if (ctx->timing) {
            HIPCHK(hipEventSynchronize(ctx->ev[1]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            auto& t = ctx->timings["k_row_tiles"];
            t.first += ms;
            t.second += 1;
        }
        return 0;
    }
    int cur = 0;
    const int passes = (int)((kbits + 7) / 8);
    uint32_t* rsc = ctx->rs_counts.as<uint32_t>();
    for (int pass = 0; pass < passes && R > 0; pass++) {
        int shift = pass * 8;
        hipLaunchKernelGGL(k_rs_count, dim3(rs_blocks), dim3(RS_NT), 0, s, ctx->keys[cur].as<uint32_t>(), R, shift,
                           rsc, rs_blocks);
        HIPCHK(launch_scan<uint32_t>(ctx, (int64_t)rs_blocks * 256, LoadU32{rsc}, StoreU32{rsc}, OpAdd(), 0u,
                                     (uint32_t*)nullptr, s));
        // values are the entry payloads (run index, or crossing sign with bit 31)
        hipLaunchKernelGGL(k_rs_scatter, dim3(rs_blocks), dim3(RS_NT), 0, s, ctx->keys[cur].as<uint32_t>(),
                           ctx->vals[cur].as<uint32_t>(), R, shift, rsc, rs_blocks,
                           ctx->keys[cur ^ 1].as<uint32_t>(), ctx->vals[cur ^ 1].as<uint32_t>());
        cur ^= 1;
    }
    const uint32_t* skeys = ctx->keys[cur].as<uint32_t>();
    const uint32_t* svals = ctx->vals[cur].as<uint32_t>();

    // ---- groups (one per distinct key) and tile rows
    HIPCHK(ctx->grp_begin.ensure((std::max<int64_t>(R, 1) + 1) * 4));
    HIPCHK(ctx->grp_key.ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->grp_hasrun.ensure(std::max<int64_t>(R, 1) * 4));
    HIPCHK(ctx->grp_sign.ensure(std::max<int64_t>(R, 1) * 4));
    fill<uint32_t>(ctx->grp_hasrun.as<uint32_t>(), R, 0u, s);
    fill<int32_t>(ctx->grp_sign.as<int32_t>(), R, 0, s);
    if (R > 0) {
        HIPCHK(launch_scan<uint32_t>(ctx, R, LoadGroupStart{skeys},
                                     StoreGroup{skeys, svals, ctx->grp_begin.as<uint32_t>(),
                                                ctx->grp_key.as<uint32_t>(), ctx->grp_hasrun.as<uint32_t>(),
                                                ctx->grp_sign.as<int32_t>(), R},
                                     OpAdd(), 0u, reinterpret_cast<uint32_t*>(&tot[3]), s));
    }
    HIPCHK(ctx->row_gbegin.ensure(ROWS * 4));
    HIPCHK(ctx->row_gend.ensure(ROWS * 4));
    HIPCHK(ctx->row_sign.ensure(ROWS * 4));
    HIPCHK(ctx->row_wb.ensure(ROWS * 4));
    HIPCHK(ctx->row_path.ensure(ROWS * 4));
    HIPCHK(ctx->row_cnt.ensure(ROWS * 8));
    HIPCHK(ctx->row_off.ensure(ROWS * 8));
    fill<uint32_t>(ctx->row_gbegin.as<uint32_t>(), ROWS, 0u, s);
    fill<uint32_t>(ctx->row_gend.as<uint32_t>(), ROWS, 0u, s);
    fill<int32_t>(ctx->row_sign.as<int32_t>(), ROWS, 0, s);
    RowArgs ra;
    ra.gkey = ctx->grp_key.as<uint32_t>();
    ra.gsign = ctx->grp_sign.as<int32_t>();
    ra.n_groups = reinterpret_cast<const uint32_t*>(&tot[3]);
    ra.xbits = xbits;
    ra.row_gbegin = ctx->row_gbegin.as<uint32_t>();
    ra.row_gend = ctx->row_gend.as<uint32_t>();
    ra.row_sign = ctx->row_sign.as<int32_t>();
    if (R > 0)
        hipLaunchKernelGGL(k_row_table, dim3(std::min<uint32_t>(grid_for(R, 256), 16384)), dim3(256), 0, s, ra, R);
    hipLaunchKernelGGL(k_path_rows, dim3(grid_for(NP, 64)), dim3(64), 0, s, grid, NP, ra.row_gbegin, ra.row_gend,
                       ra.row_sign, ctx->row_wb.as<int32_t>(), ctx->row_path.as<uint32_t>(),
                       ctx->row_cnt.as<uint64_t>());
    TileArgs ta;
    ta.grid = grid;
    ta.row_path = ctx->row_path.as<uint32_t>();
    ta.row_gbegin = ra.row_gbegin;
    ta.row_gend = ra.row_gend;
    ta.row_wb = ctx->row_wb.as<int32_t>();
    ta.n_rows = ROWS;
    ta.gbegin = ctx->grp_begin.as<uint32_t>();
    ta.gkey = ra.gkey;
    ta.ghasrun = ctx->grp_hasrun.as<uint32_t>();
    ta.gsign = ra.gsign;
    ta.xbits = xbits;
    ta.vals = svals;
    ta.recs = ctx->records.as<RunRec>();
    ta.lines = ctx->lines.as<float4>();
    ta.row_cnt = ctx->row_cnt.as<uint64_t>();
    ta.row_off = ctx->row_off.as<uint64_t>();
    ta.tile_xy = out.tile_xy;
    ta.tile_alpha = out.tile_alpha;
    ta.tile_cov = out.tile_cov;
    ta.span_xy = out.span_xy;
    ta.span_w = out.span_w;
    ta.cnt = cnt;
    const uint32_t tile_blocks = (uint32_t)std::min<int64_t>((ROWS + ROWS_PER_WAVE - 1) / ROWS_PER_WAVE, 1 << 20);
    hipLaunchKernelGGL(k_tile_counts, dim3(tile_blocks), dim3(WAVE), 0, s, ta);
    HIPCHK(launch_scan<uint64_t>(ctx, ROWS, LoadU64{ta.row_cnt}, StoreU64{ctx->row_off.as<uint64_t>()}, OpAdd(),
                                 0ull, &tot[4], s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&ht->out, &tot[4], 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ht->error, &cnt->error, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ht->error & (ERR_BBOX | ERR_REPLAY)) return -3;
    const int64_t T = (int64_t)(ht->out >> 32);
    const int64_t S = (int64_t)(uint32_t)ht->out;
    out.n_tiles = T;
    out.n_spans = S;
    if (T > out.tile_capacity || S > out.span_capacity) return -4;

    // ---- (c)(d)(e)
    if (ctx->timing) {
        if (!ctx->ev[0]) HIPCHK(hipEventCreate(&ctx->ev[0]));
        if (!ctx->ev[1]) HIPCHK(hipEventCreate(&ctx->ev[1]));
        HIPCHK(hipEventRecord(ctx->ev[0], s));
    }
    hipLaunchKernelGGL(k_tile_rows, dim3(row_grid(ROWS, 1 << 20)), dim3(WAVE), 0, s, ta);
    if (ctx->timing) HIPCHK(hipEventRecord(ctx->ev[1], s));
    hipLaunchKernelGGL(k_path_out, dim3(grid_for(NP, 256)), dim3(256), 0, s, grid, NP, ta.row_off,
                       (const uint64_t*)&tot[4], out.path_tile_start, out.path_tile_count, out.path_span_start,
                       out.path_span_count, out.tile_xy, out.tile_alpha, out.tile_cov, out.tile_base, out.span_base);
    hipLaunchKernelGGL(k_fold_error, dim3(1), dim3(64), 0, s, cnt, ctx->call_error.as<uint32_t>());
    HIPCHK(hipGetLastError());
    if (ctx->timing) {
        HIPCHK(hipEventSynchronize(ctx->ev[1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        auto& t = ctx->timings["k_tile_rows"];
        t.first += ms;
        t.second += 1;
    }
    return 0;
}

// Offsets of paths [p0, p0 + n] rebased to the range's first verb, point and
// weight, packed as three arrays of n + 1.
__global__ void k_rebase_offsets(const int64_t* voff, const int64_t* poff, const int64_t* woff, int64_t p0, int64_t n,
                                 int64_t* dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > n) return;
    dst[i] = voff[p0 + i] - voff[p0];
    dst[n + 1 + i] = poff[p0 + i] - poff[p0];
    dst[2 * (n + 1) + i] = woff[p0 + i] - woff[p0];
}

// The emit-phase error bits of a call (folded in by k_fold_error after each
// pass), read back once the call's work is queued: a replay overrun makes the
// call fail with -3 instead of returning truncated alpha.
static int check_call_error(Context* ctx, int st, hipStream_t s) {
    if (st != 0 && st != -4) return st;
    uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(ctx->host_pinned) + 2048);
    HIPCHK(hipMemcpyAsync(h, ctx->call_error.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (*h & (ERR_BBOX | ERR_REPLAY)) return -3;
    return st;
}

int run_plain(Context* ctx, const SegInput& in, OutputDesc& out, hipStream_t s, bool unordered) {
    HIPCHK(ctx->call_error.ensure(256));
    if (!ctx->host_pinned) HIPCHK(hipHostMalloc(&ctx->host_pinned, 4096));
    fill<uint32_t>(ctx->call_error.as<uint32_t>(), 1, 0u, s);
    int st = run_plain_pass(ctx, in, out, s, unordered);
    if (st != ST_SPLIT) return check_call_error(ctx, st, s);
    // Paths are independent Rasterizers: rasterise path ranges one after
    // another, each range's tiles and spans following the previous range's.
    if (in.path_seg_off || in.seg_path || in.n_paths < 2) return -7;
    const int64_t NP = in.n_paths;
    std::vector<int64_t> voff(NP + 1), poff(NP + 1), woff(NP + 1);
    HIPCHK(hipMemcpyAsync(voff.data(), in.seg_verb_off, (NP + 1) * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(poff.data(), in.seg_point_off, (NP + 1) * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(woff.data(), in.seg_weight_off, (NP + 1) * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(ctx->split_off.ensure((size_t)3 * (NP + 1) * 8));
    int64_t* roff = ctx->split_off.as<int64_t>();

    // ranges still to do, the next one at the back; a range over the limits is
    // cut into pieces of about equal verb counts, sized from how far the
    // failed pass was over
    std::vector<std::pair<int64_t, int64_t>> todo;
    auto cut = [&](int64_t p0, int64_t p1) {
        const int k = (int)std::min<double>(std::ceil(std::max(2.0, 1.25 * ctx->split_ratio)), (double)(p1 - p0));
        std::vector<int64_t> cuts{p0};
        const int64_t v0 = voff[p0], vr = voff[p1] - v0;
        for (int i = 1; i < k; i++) {
            const int64_t target = v0 + (int64_t)((double)vr * i / k);
            int64_t c = std::lower_bound(voff.begin() + p0, voff.begin() + p1, target) - voff.begin();
            if (c > cuts.back() && c < p1) cuts.push_back(c);
        }
        if (cuts.size() == 1) cuts.push_back(p0 + (p1 - p0) / 2);
        cuts.push_back(p1);
        for (size_t i = cuts.size() - 1; i > 0; i--) todo.push_back({cuts[i - 1], cuts[i]});
    };
    cut(0, NP);
    int64_t T = 0, S = 0;
    bool room = true;
    while (!todo.empty()) {
        const int64_t p0 = todo.back().first, p1 = todo.back().second;
        todo.pop_back();
        const int64_t n = p1 - p0;
        hipLaunchKernelGGL(k_rebase_offsets, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, in.seg_verb_off,
                           in.seg_point_off, in.seg_weight_off, p0, n, roff);
        HIPCHK(hipGetLastError());
        SegInput si;
        si.n_segs = n;
        si.n_paths = n;
        si.verbs = in.verbs + voff[p0];
        si.points = in.points + 2 * poff[p0];
        si.weights = in.weights ? in.weights + woff[p0] : nullptr;
        si.seg_verb_off = roff;
        si.seg_point_off = roff + (n + 1);
        si.seg_weight_off = roff + 2 * (n + 1);
        si.seg_xf = in.seg_xf + p0;
        si.seg_raw = in.seg_raw ? in.seg_raw + p0 : nullptr;
        si.n_verbs = voff[p1] - voff[p0];
        si.n_points = poff[p1] - poff[p0];
        si.n_weights = woff[p1] - woff[p0];
        OutputDesc od;
        od.tile_xy = out.tile_xy + (room ? 2 * T : 0);
        od.tile_alpha = out.tile_alpha + (room ? 64 * T : 0);
        od.tile_cov = out.tile_cov ? out.tile_cov + (room ? 64 * T : 0) : nullptr;
        od.tile_capacity = room ? out.tile_capacity - T : 0;
        od.span_xy = out.span_xy + (room ? 2 * S : 0);
        od.span_w = out.span_w + (room ? S : 0);
        od.span_capacity = room ? out.span_capacity - S : 0;
        od.path_tile_start = out.path_tile_start + p0;
        od.path_tile_count = out.path_tile_count + p0;
        od.path_span_start = out.path_span_start + p0;
        od.path_span_count = out.path_span_count + p0;
        od.tile_base = out.tile_base + T;
        od.span_base = out.span_base + S;
        st = run_plain_pass(ctx, si, od, s, unordered);
        if (st == ST_SPLIT) {
            if (n < 2) return -7;
            cut(p0, p1);
            continue;
        }
        if (st == -4) room = false;
        else if (st) return st;
        T += od.n_tiles;
        S += od.n_spans;
    }
    out.n_tiles = T;
    out.n_spans = S;
    return check_call_error(ctx, (room && T <= out.tile_capacity && S <= out.span_capacity) ? 0 : -4, s);
}

}  // namespace ochre
