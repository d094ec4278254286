// C ABI (include/ochre.h) and the host-side Rasterizer.
//
// The host Rasterizer records commands exactly as the reference's
// Rasterizer receives them (fill: one segment per call with its transform;
// move_to / line_to / command: untransformed segments) and defers all
// arithmetic to the device pipeline at finish().  Nothing is rasterised on
// the CPU: without a usable GPU, ochre_rasterizer_new returns NULL and the
// batch entry points return OCHRE_ERR_NO_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../../include/ochre.h"
#include "pipeline.hpp"

using namespace ochre;

struct ochre_context {
    Context c;
    std::mutex mu;
};

static_assert(sizeof(ochre_transform) == sizeof(Xform), "transform layout");

namespace {

bool device_usable(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return false;
    return true;
}

std::mutex g_default_mu;
ochre_context* g_default = nullptr;

ochre_context* default_context() {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) {
        if (!device_usable(0)) return nullptr;
        ochre_context* c = new (std::nothrow) ochre_context();
        if (!c) return nullptr;
        c->c.device = 0;
        g_default = c;
    }
    return g_default;
}

}  // namespace

extern "C" {

const char* ochre_build_info(void) {
    return "ochre-mi355x: gfx950 HIP kernels (flatten, DDA binning, radix sort, tile accumulation, row scan, "
           "compaction)";
}

int ochre_context_create(int device, ochre_context** out) {
    if (!out) return OCHRE_ERR_INVALID_ARG;
    *out = nullptr;
    if (!device_usable(device)) return OCHRE_ERR_NO_DEVICE;
    ochre_context* c = new (std::nothrow) ochre_context();
    if (!c) return OCHRE_ERR_OUT_OF_MEMORY;
    c->c.device = device;
    // the allocation-free entry point's scratch address on this device
    // (ochre_rasterize_batched_noalloc fails with OCHRE_ERR_DEVICE without it)
    if (hipSetDevice(device) == hipSuccess) (void)noalloc_init_device(device);
    *out = c;
    return OCHRE_OK;
}

void ochre_context_destroy(ochre_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    delete ctx;
}

int ochre_context_set_timing(ochre_context* ctx, int enabled) {
    if (!ctx) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.timing = enabled != 0;
    ctx->c.timings.clear();
    return OCHRE_OK;
}

int ochre_context_timing(ochre_context* ctx, const char* kernel, double* total_ms, int64_t* launches) {
    if (!ctx || !kernel) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->c.timings.find(kernel);
    if (total_ms) *total_ms = it == ctx->c.timings.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->c.timings.end() ? 0 : it->second.second;
    return OCHRE_OK;
}

int ochre_rasterize_plain(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out, void* stream) {
    if (!ctx || !in || !out) return OCHRE_ERR_INVALID_ARG;
    if (in->n_paths < 0 || in->n_verbs < 0 || in->n_points < 0 || in->n_weights < 0) return OCHRE_ERR_INVALID_ARG;
    if (in->n_paths > 0 && (!in->verb_off || !in->point_off || !in->weight_off || !in->transforms))
        return OCHRE_ERR_INVALID_ARG;
    if (in->n_verbs > 0 && (!in->verbs)) return OCHRE_ERR_INVALID_ARG;
    if (in->n_points > 0 && !in->points) return OCHRE_ERR_INVALID_ARG;
    if (in->n_weights > 0 && !in->weights) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    SegInput si;
    si.n_segs = in->n_paths;
    si.n_paths = in->n_paths;
    si.verbs = in->verbs;
    si.points = in->points;
    si.weights = in->weights;
    si.seg_verb_off = in->verb_off;
    si.seg_point_off = in->point_off;
    si.seg_weight_off = in->weight_off;
    si.seg_xf = reinterpret_cast<const Xform*>(in->transforms);
    si.n_verbs = in->n_verbs;
    si.n_points = in->n_points;
    si.n_weights = in->n_weights;
    OutputDesc od;
    od.tile_xy = out->tile_xy;
    od.tile_alpha = out->tile_alpha;
    od.tile_cov = out->tile_coverage;
    od.tile_capacity = out->tile_capacity;
    od.span_xy = out->span_xy;
    od.span_w = out->span_width;
    od.span_capacity = out->span_capacity;
    od.path_tile_start = out->path_tile_start;
    od.path_tile_count = out->path_tile_count;
    od.path_span_start = out->path_span_start;
    od.path_span_count = out->path_span_count;
    int st = run_plain(&ctx->c, si, od, reinterpret_cast<hipStream_t>(stream));
    out->n_tiles = od.n_tiles;
    out->n_spans = od.n_spans;
    return st;
}

static int to_seg_input(const ochre_path_batch* in, SegInput* si) {
    if (in->n_paths < 0 || in->n_verbs < 0 || in->n_points < 0 || in->n_weights < 0) return OCHRE_ERR_INVALID_ARG;
    if (in->n_paths > 0 && (!in->verb_off || !in->point_off || !in->weight_off || !in->transforms))
        return OCHRE_ERR_INVALID_ARG;
    if ((in->n_verbs > 0 && !in->verbs) || (in->n_points > 0 && !in->points) || (in->n_weights > 0 && !in->weights))
        return OCHRE_ERR_INVALID_ARG;
    si->n_segs = in->n_paths;
    si->n_paths = in->n_paths;
    si->verbs = in->verbs;
    si->points = in->points;
    si->weights = in->weights;
    si->seg_verb_off = in->verb_off;
    si->seg_point_off = in->point_off;
    si->seg_weight_off = in->weight_off;
    si->seg_xf = reinterpret_cast<const Xform*>(in->transforms);
    si->n_verbs = in->n_verbs;
    si->n_points = in->n_points;
    si->n_weights = in->n_weights;
    return OCHRE_OK;
}

static OutputDesc to_output(const ochre_batch_output* out) {
    OutputDesc od;
    od.tile_xy = out->tile_xy;
    od.tile_alpha = out->tile_alpha;
    od.tile_cov = out->tile_coverage;
    od.tile_capacity = out->tile_capacity;
    od.span_xy = out->span_xy;
    od.span_w = out->span_width;
    od.span_capacity = out->span_capacity;
    od.path_tile_start = out->path_tile_start;
    od.path_tile_count = out->path_tile_count;
    od.path_span_start = out->path_span_start;
    od.path_span_count = out->path_span_count;
    return od;
}

int ochre_rasterize_batched(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out, void* stream) {
    if (!ctx || !in || !out) return OCHRE_ERR_INVALID_ARG;
    SegInput si;
    int st = to_seg_input(in, &si);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    OutputDesc od = to_output(out);
    st = run_batched(&ctx->c, si, od, reinterpret_cast<hipStream_t>(stream), ctx->c.batched_waves_per_cu,
                     ctx->c.batched_gran, ctx->c.batched_tc, ctx->c.batched_graph);
    out->n_tiles = od.n_tiles;
    out->n_spans = od.n_spans;
    return st;
}

int ochre_rasterize_batched_noalloc(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                                    void* stream) {
    if (!ctx || !in || !out) return OCHRE_ERR_INVALID_ARG;
    SegInput si;
    int st = to_seg_input(in, &si);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    OutputDesc od = to_output(out);
    st = run_batched_noalloc(si, od, reinterpret_cast<hipStream_t>(stream));
    out->n_tiles = od.n_tiles;
    out->n_spans = od.n_spans;
    return st;
}

static int stroke_batch(ochre_context* ctx, const ochre_path_batch* in, const float* widths,
                        const uint8_t* stroke_flags, ochre_batch_output* out, void* stream, bool batched) {
    if (!ctx || !in || !out || (in->n_paths > 0 && !widths)) return OCHRE_ERR_INVALID_ARG;
    SegInput si;
    int st = to_seg_input(in, &si);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    OutputDesc od = to_output(out);
    st = run_stroke(&ctx->c, si, widths, stroke_flags, od, reinterpret_cast<hipStream_t>(stream), batched);
    out->n_tiles = od.n_tiles;
    out->n_spans = od.n_spans;
    return st;
}

int ochre_stroke_plain(ochre_context* ctx, const ochre_path_batch* in, const float* widths, ochre_batch_output* out,
                       void* stream) {
    return stroke_batch(ctx, in, widths, nullptr, out, stream, false);
}

int ochre_stroke_batched(ochre_context* ctx, const ochre_path_batch* in, const float* widths,
                         ochre_batch_output* out, void* stream) {
    return stroke_batch(ctx, in, widths, nullptr, out, stream, true);
}

int ochre_draw_plain(ochre_context* ctx, const ochre_path_batch* in, const uint8_t* stroke_flags,
                     const float* widths, ochre_batch_output* out, void* stream) {
    if (in && in->n_paths > 0 && !stroke_flags) return OCHRE_ERR_INVALID_ARG;
    return stroke_batch(ctx, in, widths, stroke_flags, out, stream, false);
}

int ochre_draw_batched(ochre_context* ctx, const ochre_path_batch* in, const uint8_t* stroke_flags,
                       const float* widths, ochre_batch_output* out, void* stream) {
    if (in && in->n_paths > 0 && !stroke_flags) return OCHRE_ERR_INVALID_ARG;
    return stroke_batch(ctx, in, widths, stroke_flags, out, stream, true);
}

int ochre_build_atlas(ochre_context* ctx, const ochre_batch_output* in, int64_t n_paths, const uint8_t* colors,
                      uint8_t* atlas, int atlas_size, ochre_atlas_vertex* vertices, int64_t vertex_capacity,
                      uint32_t* indices, int64_t index_capacity, int64_t* n_vertices, int64_t* n_indices,
                      void* stream) {
    if (!ctx || !in || !n_vertices || !n_indices) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    OutputDesc od = to_output(in);
    return run_atlas(&ctx->c, od, n_paths, colors, atlas, atlas_size, vertices, vertex_capacity, indices,
                     index_capacity, n_vertices, n_indices, reinterpret_cast<hipStream_t>(stream));
}

int ochre_context_set_batched_params(ochre_context* ctx, int waves_per_cu, int paths_per_pop, int tiles_per_chunk,
                                     int use_graph) {
    if (!ctx || waves_per_cu < 1 || waves_per_cu > 32 || paths_per_pop < 1 || paths_per_pop > 1024 ||
        (tiles_per_chunk != 16 && tiles_per_chunk != 32))
        return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.batched_waves_per_cu = waves_per_cu;
    ctx->c.batched_gran = paths_per_pop;
    ctx->c.batched_tc = tiles_per_chunk;
    ctx->c.batched_graph = use_graph != 0;
    return OCHRE_OK;
}

int ochre_context_get_batched_params(ochre_context* ctx, int* waves_per_cu, int* paths_per_pop, int* tiles_per_chunk,
                                     int* use_graph) {
    if (!ctx) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (waves_per_cu) *waves_per_cu = ctx->c.batched_waves_per_cu;
    if (paths_per_pop) *paths_per_pop = ctx->c.batched_gran;
    if (tiles_per_chunk) *tiles_per_chunk = ctx->c.batched_tc;
    if (use_graph) *use_graph = ctx->c.batched_graph ? 1 : 0;
    return OCHRE_OK;
}

int ochre_context_set_dense_params(ochre_context* ctx, int enabled, int groups_per_cu, int waves_per_group) {
    if (!ctx || groups_per_cu < 1 || groups_per_cu > 8 ||
        (waves_per_group != 2 && waves_per_group != 4 && waves_per_group != 8))
        return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.dense = enabled != 0;
    ctx->c.dense_groups_per_cu = groups_per_cu;
    ctx->c.dense_waves = waves_per_group;
    return OCHRE_OK;
}

int ochre_context_set_dense_small(ochre_context* ctx, int groups_per_cu) {
    if (!ctx || groups_per_cu < 0 || groups_per_cu > 9) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.dense_small_groups = groups_per_cu;
    return OCHRE_OK;
}

int ochre_context_set_pass_limits(ochre_context* ctx, int64_t max_lines, int64_t max_entries) {
    if (!ctx || max_lines < 0 || max_entries < 0) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.max_lines = (max_lines == 0) ? PLAIN_MAX_LINES : std::min<int64_t>(max_lines, PLAIN_MAX_LINES);
    ctx->c.max_entries = (max_entries == 0) ? PLAIN_MAX_ENTRIES : std::min<int64_t>(max_entries, PLAIN_MAX_ENTRIES);
    return OCHRE_OK;
}

int ochre_context_set_dense_large(ochre_context* ctx, int enabled) {
    if (!ctx || enabled < 0 || enabled > 1) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->c.dense_large = enabled;
    return OCHRE_OK;
}

int ochre_context_get_dense_large(ochre_context* ctx, int* enabled) {
    if (!ctx || !enabled) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *enabled = ctx->c.dense_large;
    return OCHRE_OK;
}

int ochre_context_get_dense_small(ochre_context* ctx, int* groups_per_cu) {
    if (!ctx || !groups_per_cu) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *groups_per_cu = ctx->c.dense_small_groups;
    return OCHRE_OK;
}

int ochre_context_get_dense_params(ochre_context* ctx, int* enabled, int* groups_per_cu, int* waves_per_group) {
    if (!ctx) return OCHRE_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (enabled) *enabled = ctx->c.dense ? 1 : 0;
    if (groups_per_cu) *groups_per_cu = ctx->c.dense_groups_per_cu;
    if (waves_per_group) *waves_per_group = ctx->c.dense_waves;
    return OCHRE_OK;
}

int ochre_coverage_adjoint(ochre_context* ctx, const ochre_path_batch* in, const ochre_coverage_target* target,
                           float* loss, float* d_points, float* d_transforms, void* stream) {
    if (!ctx || !in || !target || !loss || !d_points || !d_transforms) return OCHRE_ERR_INVALID_ARG;
    if (in->n_paths > 0 && (!target->tile_start || !target->tile_count || !target->span_start || !target->span_count))
        return OCHRE_ERR_INVALID_ARG;
    SegInput si;
    int st = to_seg_input(in, &si);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    AdjTarget t;
    t.tile_start = target->tile_start;
    t.tile_count = target->tile_count;
    t.tile_xy = target->tile_xy;
    t.tile_cov = target->tile_coverage;
    t.span_start = target->span_start;
    t.span_count = target->span_count;
    t.span_xy = target->span_xy;
    t.span_w = target->span_width;
    // The adjoint's own launch shape, not the batched call's tuned one (which
    // is tuned for the ordered kernel behind the dense ones and may leave it a
    // wave per CU): on 4,096 16-cubic glyphs 32-tile chunks, a path per pop and
    // 8 or more waves per CU measured 1.58-1.61 ms, 16-tile chunks 2.37-2.42,
    // one wave per CU 6.6-9.8 ms.
    return run_adjoint(&ctx->c, si, t, loss, d_points, d_transforms, reinterpret_cast<hipStream_t>(stream), 8, 1,
                       32);
}

int ochre_autotune_batched(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out, void* stream,
                           int reps, ochre_batched_tuning* results, int max_results, int* n_results) {
    if (!ctx || !in || !out || reps < 1) return OCHRE_ERR_INVALID_ARG;
    SegInput si;
    int st = to_seg_input(in, &si);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    static const int waves[] = {2, 3, 4, 6, 8, 12};
    static const int grans[] = {1, 2, 4, 8};
    static const int tcs[] = {16, 32};
    static const int dgroups[] = {1, 2, 3, 4};
    static const int dwaves[] = {2, 4, 8};
    ochre_batched_tuning best{ctx->c.batched_waves_per_cu, ctx->c.batched_gran, ctx->c.batched_tc, 1e30f, 0, 0, 0, 0};
    int best_dg = ctx->c.dense_groups_per_cu, best_dw = ctx->c.dense_waves;
    int k = 0;
    auto measure = [&](int w, int g, int tc, float* ms_out) -> int {
        OutputDesc od = to_output(out);
        // warm-up (also captures the graph for this configuration)
        int rc = run_batched(&ctx->c, si, od, s, w, g, tc, true);
        if (rc) return rc;
        std::vector<double> t;
        for (int r = 0; r < reps; r++) {
            auto t0 = std::chrono::steady_clock::now();
            rc = run_batched(&ctx->c, si, od, s, w, g, tc, true);
            if (rc) return rc;
            t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        std::sort(t.begin(), t.end());
        *ms_out = (float)t[t.size() / 2];
        return 0;
    };
    if (ctx->c.dense) {
        // the ordered kernel only sees the paths the dense one hands on:
        // tune the dense launch shape, with waves/CU and tiles per chunk
        // encoded as (groups per CU, waves per group) in the results
        for (int dg : dgroups)
            for (int dw : dwaves) {
                if (dw == 8 && dg > 2) continue;  // 16 waves of 128 registers fill a CU
                ctx->c.dense_groups_per_cu = dg;
                ctx->c.dense_waves = dw;
                float ms = 0;
                st = measure(best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, &ms);
                if (st) return st;
                if (results && k < max_results)
                    results[k] = ochre_batched_tuning{best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, ms,
                                                      dg, dw, ctx->c.dense_small_groups, ctx->c.dense_large};
                k++;
                if (ms < best.ms) {
                    best.ms = ms;
                    best_dg = dg;
                    best_dw = dw;
                }
            }
        ctx->c.dense_groups_per_cu = best_dg;
        ctx->c.dense_waves = best_dw;
        // the small-grid instance in front: off or 4-8 workgroups per CU
        int best_small = ctx->c.dense_small_groups;
        for (int sg : {0, 4, 6, 8}) {
            ctx->c.dense_small_groups = sg;
            float ms = 0;
            st = measure(best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, &ms);
            if (st) return st;
            if (results && k < max_results)
                results[k] = ochre_batched_tuning{best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, ms,
                                                  best_dg, best_dw, sg, ctx->c.dense_large};
            k++;
            if (ms < best.ms) {
                best.ms = ms;
                best_small = sg;
            }
        }
        ctx->c.dense_small_groups = best_small;
        // the 256-tile instance behind: off or on
        int best_large = ctx->c.dense_large;
        for (int lg : {0, 1}) {
            ctx->c.dense_large = lg;
            float ms = 0;
            st = measure(best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, &ms);
            if (st) return st;
            if (results && k < max_results)
                results[k] = ochre_batched_tuning{best.waves_per_cu, best.paths_per_pop, best.tiles_per_chunk, ms,
                                                  best_dg, best_dw, best_small, lg};
            k++;
            if (ms < best.ms) {
                best.ms = ms;
                best_large = lg;
            }
        }
        ctx->c.dense_large = best_large;
        // then the persistent grid of the ordered kernel behind them, which
        // only works through the paths the dense kernel hands on
        for (int w : {1, 2, 4, 8}) {
            float ms = 0;
            st = measure(w, best.paths_per_pop, best.tiles_per_chunk, &ms);
            if (st) return st;
            if (results && k < max_results)
                results[k] = ochre_batched_tuning{w, best.paths_per_pop, best.tiles_per_chunk, ms, best_dg, best_dw,
                                                  ctx->c.dense_small_groups, ctx->c.dense_large};
            k++;
            if (ms < best.ms) {
                best.ms = ms;
                best.waves_per_cu = w;
            }
        }
    } else {
        for (int w : waves)
            for (int g : grans)
                for (int tc : tcs) {
                    float ms = 0;
                    st = measure(w, g, tc, &ms);
                    if (st) return st;
                    if (results && k < max_results) results[k] = ochre_batched_tuning{w, g, tc, ms, 0, 0, 0, 0};
                    k++;
                    if (ms < best.ms) best = ochre_batched_tuning{w, g, tc, ms, 0, 0, 0, 0};
                }
    }
    if (n_results) *n_results = k;
    ctx->c.batched_waves_per_cu = best.waves_per_cu;
    ctx->c.batched_gran = best.paths_per_pop;
    ctx->c.batched_tc = best.tiles_per_chunk;
    ctx->c.batched_graph = true;
    return OCHRE_OK;
}

}  // extern "C"

// ------------------------------------------------------- autotune cache
//
// Tuned parameters keyed by the device (its name and compute unit count), the
// dense kernels' setting and a coarse signature of the batch: the
// power-of-two buckets of its path count and of its mean commands and points
// per path (half-octave steps).  One table per process; a cache file holds
// one "key = values" line per entry and is read once per path.
namespace {

struct TuneEntry {
    int waves_per_cu, gran, tc, dense_groups, dense_waves, dense_small, dense_large;
};

std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune;
std::set<std::string> g_tune_loaded;

int bucket(double v, double per_octave) { return v > 0 ? (int)std::floor(std::log2(v) * per_octave) : -1; }

std::string tune_key(const Context& c, const ochre_path_batch* in) {
    hipDeviceProp_t prop;
    std::string name = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, c.device) == hipSuccess) {
        name = prop.gcnArchName;
        cus = prop.multiProcessorCount;
    }
    for (char& ch : name)
        if (ch == ' ' || ch == '=') ch = '_';
    const double np = (double)std::max<int64_t>(in->n_paths, 1);
    char buf[256];
    snprintf(buf, sizeof buf, "%s/cu%d/dense%d/paths%d/verbs%d/points%d", name.c_str(), cus, c.dense ? 1 : 0,
             bucket((double)in->n_paths, 1.0), bucket((double)in->n_verbs / np, 2.0),
             bucket((double)in->n_points / np, 2.0));
    return buf;
}

void load_tune_file(const char* path) {
    if (!path || g_tune_loaded.count(path)) return;
    g_tune_loaded.insert(path);
    FILE* f = fopen(path, "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char key[320];
        TuneEntry e;
        if (sscanf(line, "%319s = %d %d %d %d %d %d %d", key, &e.waves_per_cu, &e.gran, &e.tc, &e.dense_groups,
                   &e.dense_waves, &e.dense_small, &e.dense_large) == 8)
            g_tune[key] = e;
    }
    fclose(f);
}

}  // namespace

extern "C" {

int ochre_autotune_batched_cached(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                                  void* stream, int reps, const char* cache_path, int* tuned) {
    if (!ctx || !in || !out || reps < 1) return OCHRE_ERR_INVALID_ARG;
    if (tuned) *tuned = 0;
    std::string key;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        key = tune_key(ctx->c, in);
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        load_tune_file(cache_path);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            const TuneEntry& e = it->second;
            std::lock_guard<std::mutex> lk2(ctx->mu);
            ctx->c.batched_waves_per_cu = e.waves_per_cu;
            ctx->c.batched_gran = e.gran;
            ctx->c.batched_tc = e.tc;
            ctx->c.batched_graph = true;
            if (ctx->c.dense) {
                ctx->c.dense_groups_per_cu = e.dense_groups;
                ctx->c.dense_waves = e.dense_waves;
                ctx->c.dense_small_groups = e.dense_small;
                ctx->c.dense_large = e.dense_large;
            }
            return OCHRE_OK;
        }
    }
    int st = ochre_autotune_batched(ctx, in, out, stream, reps, nullptr, 0, nullptr);
    if (st) return st;
    TuneEntry e;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        e = TuneEntry{ctx->c.batched_waves_per_cu, ctx->c.batched_gran, ctx->c.batched_tc,
                      ctx->c.dense_groups_per_cu, ctx->c.dense_waves, ctx->c.dense_small_groups, ctx->c.dense_large};
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = e;
    if (cache_path) {
        FILE* f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%s = %d %d %d %d %d %d %d\n", key.c_str(), e.waves_per_cu, e.gran, e.tc, e.dense_groups,
                    e.dense_waves, e.dense_small, e.dense_large);
            fclose(f);
        }
    }
    if (tuned) *tuned = 1;
    return OCHRE_OK;
}

}  // extern "C"

// --------------------------------------------------------------- Rasterizer

struct ochre_rasterizer {
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<float> weights;
    std::vector<int64_t> seg_verb_off{0}, seg_point_off{0}, seg_weight_off{0};
    std::vector<Xform> seg_xf;
    std::vector<uint8_t> seg_raw;
    // stroke(): the path-space input of stroke k; segment i with
    // seg_stroke[i] == k is empty until finish() puts k's outline there.
    std::vector<int32_t> seg_stroke;
    std::vector<uint8_t> s_verbs;
    std::vector<float> s_points, s_weights, s_width;
    std::vector<int64_t> s_voff{0}, s_poff{0}, s_woff{0};
    ochre_context* ctx = nullptr;

    void open_segment(bool raw, const Xform& xf) {
        seg_verb_off.push_back((int64_t)verbs.size());
        seg_point_off.push_back((int64_t)points.size() / 2);
        seg_weight_off.push_back((int64_t)weights.size());
        seg_xf.push_back(xf);
        seg_raw.push_back(raw ? 1 : 0);
        seg_stroke.push_back(-1);
    }
    // Segment ends are the next segment's starts; the trailing entries are
    // kept equal to the current sizes.
    void close_segment() {
        seg_verb_off.back() = (int64_t)verbs.size();
        seg_point_off.back() = (int64_t)points.size() / 2;
        seg_weight_off.back() = (int64_t)weights.size();
    }
    void raw_command(uint8_t verb, const float* p, int np, float w) {
        if (seg_raw.empty() || !seg_raw.back()) {
            Xform id{1, 0, 0, 1, 0, 0};
            // seg_*_off holds n_segs + 1 entries: the last one is the end.
            seg_xf.push_back(id);
            seg_raw.push_back(1);
            seg_stroke.push_back(-1);
            seg_verb_off.push_back((int64_t)verbs.size());
            seg_point_off.push_back((int64_t)points.size() / 2);
            seg_weight_off.push_back((int64_t)weights.size());
        }
        verbs.push_back(verb);
        for (int k = 0; k < 2 * np; k++) points.push_back(p[k]);
        if (verb == OCHRE_VERB_CONIC) weights.push_back(w);
        close_segment();
    }
};

namespace {

int verb_npts(uint8_t v) { return ochre::verb_points(v); }

struct DevPath {
    DBuf tile_xy, tile_alpha, span_xy, span_w, pts, ptc, pss, psc;
    // the path's arrays, packed into one pinned host buffer and sent with one copy
    void* stage_host = nullptr;
    size_t stage_cap = 0;
    DBuf stage_dev;
    // a single fill's output and the kernel's control block (at the tail), in
    // pinned host memory that the kernel writes through its device mapping
    void* arena_host = nullptr;
    void* arena_dev = nullptr;
    size_t arena_host_cap = 0;
    int64_t fast_tiles = 256, fast_spans = 64;  // the arena's capacities
    ~DevPath() {
        if (stage_host) (void)hipHostFree(stage_host);
        if (arena_host) (void)hipHostFree(arena_host);
    }
};

// Packs host arrays at 256-byte aligned offsets of dp's pinned staging buffer
// and copies them to the device in one transfer; dev[i] receives array i's
// device address.
struct HostPiece {
    const void* p;
    size_t bytes;
};
static int upload_packed(DevPath* dp, const HostPiece* pc, int n, void** dev, hipStream_t s) {
    size_t off[16];
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        off[i] = total;
        total += (std::max<size_t>(pc[i].bytes, 16) + 255) & ~(size_t)255;
    }
    if (total > dp->stage_cap) {
        if (dp->stage_host) (void)hipHostFree(dp->stage_host);
        dp->stage_host = nullptr;
        dp->stage_cap = 0;
        const size_t cap = std::max<size_t>(total, 1 << 16);
        if (hipHostMalloc(&dp->stage_host, cap) != hipSuccess) return OCHRE_ERR_OUT_OF_MEMORY;
        dp->stage_cap = cap;
    }
    hipError_t e = dp->stage_dev.ensure(total);
    if (e != hipSuccess) return map_hip_error(e);
    for (int i = 0; i < n; i++) {
        if (pc[i].bytes) memcpy(static_cast<uint8_t*>(dp->stage_host) + off[i], pc[i].p, pc[i].bytes);
        dev[i] = static_cast<uint8_t*>(dp->stage_dev.p) + off[i];
    }
    e = hipMemcpyAsync(dp->stage_dev.p, dp->stage_host, total, hipMemcpyHostToDevice, s);
    return e == hipSuccess ? 0 : map_hip_error(e);
}


// Runs the recorded strokes through the device expansion (flatten + stroke)
// and splices each outline into its placeholder segment.  The outlines come
// back to the host because the segment layout is rebuilt there; all of their
// arithmetic happens on the device.
int expand_strokes(ochre_context* ctx, ochre_rasterizer* r, DevPath* dp, hipStream_t s) {
    const int64_t K = (int64_t)r->s_width.size();
    std::vector<Xform> xf((size_t)K, Xform{1, 0, 0, 1, 0, 0});
    // the recorded strokes in one packed upload
    const HostPiece pc[8] = {{r->s_verbs.data(), r->s_verbs.size()},
                             {r->s_points.data(), r->s_points.size() * sizeof(float)},
                             {r->s_weights.data(), r->s_weights.size() * sizeof(float)},
                             {r->s_voff.data(), r->s_voff.size() * sizeof(int64_t)},
                             {r->s_poff.data(), r->s_poff.size() * sizeof(int64_t)},
                             {r->s_woff.data(), r->s_woff.size() * sizeof(int64_t)},
                             {r->s_width.data(), r->s_width.size() * sizeof(float)},
                             {xf.data(), xf.size() * sizeof(Xform)}};
    void* dv[8];
    int st = upload_packed(dp, pc, 8, dv, s);
    if (st) return st;
    SegInput si;
    si.n_segs = K;
    si.n_paths = K;
    si.verbs = static_cast<const uint8_t*>(dv[0]);
    si.points = static_cast<const float*>(dv[1]);
    si.weights = static_cast<const float*>(dv[2]);
    si.seg_verb_off = static_cast<const int64_t*>(dv[3]);
    si.seg_point_off = static_cast<const int64_t*>(dv[4]);
    si.seg_weight_off = static_cast<const int64_t*>(dv[5]);
    si.seg_xf = static_cast<const Xform*>(dv[7]);
    si.n_verbs = (int64_t)r->s_verbs.size();
    si.n_points = (int64_t)r->s_points.size() / 2;
    si.n_weights = (int64_t)r->s_weights.size();
    StrokeOutline o;
    if ((st = stroke_expand(&ctx->c, si, static_cast<const float*>(dv[6]), nullptr, o, s))) return st;
    std::vector<int64_t> ov((size_t)K + 1), op((size_t)K + 1);
    std::vector<uint8_t> overbs((size_t)o.in.n_verbs);
    std::vector<float> opts((size_t)o.in.n_points * 2);
    hipError_t e = hipMemcpyAsync(ov.data(), o.in.seg_verb_off, ov.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(op.data(), o.in.seg_point_off, op.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !overbs.empty())
        e = hipMemcpyAsync(overbs.data(), o.in.verbs, overbs.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !opts.empty())
        e = hipMemcpyAsync(opts.data(), o.in.points, opts.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return map_hip_error(e);
    // rebuild the segment layout with every outline in place
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<int64_t> voff{0}, poff{0}, woff{0};
    const size_t nseg = r->seg_xf.size();
    for (size_t i = 0; i < nseg; i++) {
        const int32_t k = r->seg_stroke[i];
        if (k < 0) {
            verbs.insert(verbs.end(), r->verbs.begin() + r->seg_verb_off[i], r->verbs.begin() + r->seg_verb_off[i + 1]);
            points.insert(points.end(), r->points.begin() + 2 * r->seg_point_off[i],
                          r->points.begin() + 2 * r->seg_point_off[i + 1]);
        } else {
            verbs.insert(verbs.end(), overbs.begin() + ov[k], overbs.begin() + ov[k + 1]);
            points.insert(points.end(), opts.begin() + 2 * op[k], opts.begin() + 2 * op[k + 1]);
        }
        voff.push_back((int64_t)verbs.size());
        poff.push_back((int64_t)points.size() / 2);
        woff.push_back(r->seg_weight_off[i + 1]);  // outlines add no weights
    }
    r->verbs.swap(verbs);
    r->points.swap(points);
    r->seg_verb_off.swap(voff);
    r->seg_point_off.swap(poff);
    r->seg_weight_off.swap(woff);
    std::fill(r->seg_stroke.begin(), r->seg_stroke.end(), -1);
    return 0;
}

}  // namespace

extern "C" {

ochre_rasterizer* ochre_rasterizer_new(void) {
    ochre_context* ctx = default_context();
    if (!ctx) return nullptr;
    ochre_rasterizer* r = new (std::nothrow) ochre_rasterizer();
    if (!r) return nullptr;
    r->ctx = ctx;
    return r;
}

void ochre_rasterizer_free(ochre_rasterizer* r) { delete r; }

int ochre_rasterizer_move_to(ochre_rasterizer* r, float x, float y) {
    if (!r) return OCHRE_ERR_INVALID_ARG;
    float p[2] = {x, y};
    r->raw_command(OCHRE_VERB_MOVE, p, 1, 0.0f);
    return OCHRE_OK;
}

int ochre_rasterizer_line_to(ochre_rasterizer* r, float x, float y) {
    if (!r) return OCHRE_ERR_INVALID_ARG;
    float p[2] = {x, y};
    r->raw_command(OCHRE_VERB_LINE, p, 1, 0.0f);
    return OCHRE_OK;
}

int ochre_rasterizer_command(ochre_rasterizer* r, uint8_t verb, const float* points, float weight) {
    if (!r) return OCHRE_ERR_INVALID_ARG;
    int np = verb_npts(verb);
    if (np < 0 || (np > 0 && !points)) return OCHRE_ERR_INVALID_ARG;
    r->raw_command(verb, points, np, weight);
    return OCHRE_OK;
}

}  // extern "C"

static int check_path(const uint8_t* verbs, size_t n_verbs, const float* points, size_t n_points,
                      const float* weights, size_t n_weights) {
    if ((n_verbs && !verbs) || (n_points && !points) || (n_weights && !weights)) return OCHRE_ERR_INVALID_ARG;
    size_t np = 0, nw = 0;
    for (size_t i = 0; i < n_verbs; i++) {
        int k = verb_npts(verbs[i]);
        if (k < 0) return OCHRE_ERR_INVALID_ARG;
        np += (size_t)k;
        nw += verbs[i] == OCHRE_VERB_CONIC;
    }
    return np == n_points && nw == n_weights ? OCHRE_OK : OCHRE_ERR_INVALID_ARG;
}

extern "C" {

int ochre_rasterizer_stroke(ochre_rasterizer* r, const uint8_t* verbs, size_t n_verbs, const float* points,
                            size_t n_points, const float* weights, size_t n_weights, float width,
                            const ochre_transform* transform) {
    if (!r || !transform) return OCHRE_ERR_INVALID_ARG;
    int st = check_path(verbs, n_verbs, points, n_points, weights, n_weights);
    if (st) return st;
    Xform xf;
    memcpy(&xf, transform, sizeof(xf));
    r->s_verbs.insert(r->s_verbs.end(), verbs, verbs + n_verbs);
    r->s_points.insert(r->s_points.end(), points, points + 2 * n_points);
    r->s_weights.insert(r->s_weights.end(), weights, weights + n_weights);
    r->s_voff.push_back((int64_t)r->s_verbs.size());
    r->s_poff.push_back((int64_t)r->s_points.size() / 2);
    r->s_woff.push_back((int64_t)r->s_weights.size());
    r->s_width.push_back(width);
    r->seg_xf.push_back(xf);
    r->seg_raw.push_back(0);
    r->seg_stroke.push_back((int32_t)r->s_width.size() - 1);
    r->seg_verb_off.push_back((int64_t)r->verbs.size());
    r->seg_point_off.push_back((int64_t)r->points.size() / 2);
    r->seg_weight_off.push_back((int64_t)r->weights.size());
    return OCHRE_OK;
}

int ochre_rasterizer_fill(ochre_rasterizer* r, const uint8_t* verbs, size_t n_verbs, const float* points,
                          size_t n_points, const float* weights, size_t n_weights,
                          const ochre_transform* transform) {
    if (!r || !transform) return OCHRE_ERR_INVALID_ARG;
    int st = check_path(verbs, n_verbs, points, n_points, weights, n_weights);
    if (st) return st;
    Xform xf;
    memcpy(&xf, transform, sizeof(xf));
    // open a transformed segment
    r->seg_xf.push_back(xf);
    r->seg_raw.push_back(0);
    r->seg_stroke.push_back(-1);
    r->seg_verb_off.push_back((int64_t)r->verbs.size());
    r->seg_point_off.push_back((int64_t)r->points.size() / 2);
    r->seg_weight_off.push_back((int64_t)r->weights.size());
    r->verbs.insert(r->verbs.end(), verbs, verbs + n_verbs);
    r->points.insert(r->points.end(), points, points + 2 * n_points);
    r->weights.insert(r->weights.end(), weights, weights + n_weights);
    r->close_segment();
    return OCHRE_OK;
}

// The TileBuilder replay in the reference's call order: a span follows the
// tile to its left (rasterizer.rs:241,263).
static void replay(const ochre_tile_builder* builder, const int16_t* txy, const uint8_t* ta, int64_t nt,
                   const int16_t* sxy, const uint16_t* sw, int64_t ns) {
    if (!builder) return;
    int64_t k = 0;
    for (int64_t t = 0; t < nt; t++) {
        const int16_t x = txy[2 * t], y = txy[2 * t + 1];
        if (builder->tile) builder->tile(builder->user, x, y, ta + (size_t)t * 64);
        while (k < ns && sxy[2 * k + 1] == y && sxy[2 * k] == (int16_t)(x + 8)) {
            if (builder->span) builder->span(builder->user, sxy[2 * k], sxy[2 * k + 1], sw[k]);
            k++;
        }
    }
}

// A Rasterizer holding one fill: one upload (already queued), one launch
// whose output and counts land in mapped pinned host memory, one
// synchronisation, then the replay.  Returns 1 when the fill needs the general
// path (a path over k_single's limits, or more output than the arena holds:
// its capacities grow for the next call).
static int finish_one_fill(DevPath* dp, const SegInput& si, hipStream_t s, const ochre_tile_builder* builder) {
    const int64_t tc = dp->fast_tiles, sc = dp->fast_spans;
    const size_t o_txy = (size_t)tc * 64, o_sxy = o_txy + (size_t)tc * 4, o_sw = o_sxy + (size_t)sc * 4;
    const size_t o_rng = (o_sw + (size_t)sc * 2 + 255) & ~(size_t)255, total = o_rng + 256;
    if (dp->arena_host_cap < total + SMALL_CTL_BYTES) {
        if (dp->arena_host) (void)hipHostFree(dp->arena_host);
        dp->arena_host = dp->arena_dev = nullptr;
        dp->arena_host_cap = 0;
        if (hipHostMalloc(&dp->arena_host, total + SMALL_CTL_BYTES, hipHostMallocMapped) != hipSuccess)
            return OCHRE_ERR_OUT_OF_MEMORY;
        hipError_t e = hipHostGetDevicePointer(&dp->arena_dev, dp->arena_host, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(dp->arena_host);
            dp->arena_host = dp->arena_dev = nullptr;
            return map_hip_error(e);
        }
        dp->arena_host_cap = total + SMALL_CTL_BYTES;
    }
    uint8_t* a = static_cast<uint8_t*>(dp->arena_dev);
    OutputDesc od;
    od.tile_alpha = a;
    od.tile_xy = reinterpret_cast<int16_t*>(a + o_txy);
    od.tile_capacity = tc;
    od.span_xy = reinterpret_cast<int16_t*>(a + o_sxy);
    od.span_w = reinterpret_cast<uint16_t*>(a + o_sw);
    od.span_capacity = sc;
    od.path_tile_start = reinterpret_cast<int64_t*>(a + o_rng);
    od.path_tile_count = reinterpret_cast<int32_t*>(a + o_rng + 8);
    od.path_span_start = reinterpret_cast<int64_t*>(a + o_rng + 16);
    od.path_span_count = reinterpret_cast<int32_t*>(a + o_rng + 24);
    int st = launch_single_fill(si, od, a + total, s);
    if (st) return st;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return map_hip_error(e);
    const uint8_t* h = static_cast<const uint8_t*>(dp->arena_host);
    const SmallBatchResult res = small_batch_result(h + total);
    if (res.status == OCHRE_ERR_CAPACITY || res.tiles > tc || res.spans > sc) {
        dp->fast_tiles = std::max<int64_t>(tc, res.tiles + res.tiles / 4 + 64);
        dp->fast_spans = std::max<int64_t>(sc, res.spans + res.spans / 4 + 64);
        return 1;
    }
    if (res.status) return res.status;
    if (res.deferred) return 1;
    replay(builder, reinterpret_cast<const int16_t*>(h + o_txy), h, res.tiles,
           reinterpret_cast<const int16_t*>(h + o_sxy), reinterpret_cast<const uint16_t*>(h + o_sw), res.spans);
    return 0;
}

int ochre_rasterizer_finish(ochre_rasterizer* r, const ochre_tile_builder* builder) {
    if (!r) return OCHRE_ERR_INVALID_ARG;
    ochre_context* ctx = r->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    static DevPath* dp = nullptr;  // guarded by the default context's mutex
    if (!dp) dp = new (std::nothrow) DevPath();
    int st = 0;
    do {
        if (!dp) {
            st = OCHRE_ERR_OUT_OF_MEMORY;
            break;
        }
        if (hipSetDevice(ctx->c.device) != hipSuccess) {
            st = OCHRE_ERR_DEVICE;
            break;
        }
        hipStream_t s = nullptr;
        if (!r->s_width.empty() && (st = expand_strokes(ctx, r, dp, s))) break;
        // a path with no segments is one empty segment
        if (r->seg_xf.empty()) {
            Xform id{1, 0, 0, 1, 0, 0};
            r->seg_xf.push_back(id);
            r->seg_raw.push_back(1);
            r->seg_stroke.push_back(-1);
            r->seg_verb_off.push_back(0);
            r->seg_point_off.push_back(0);
            r->seg_weight_off.push_back(0);
        }
        const int64_t nseg = (int64_t)r->seg_xf.size();
        std::vector<int32_t> seg_path((size_t)nseg, 0);
        std::vector<int64_t> path_seg{0, nseg};
        const HostPiece pc[10] = {{r->verbs.data(), r->verbs.size()},
                                  {r->points.data(), r->points.size() * sizeof(float)},
                                  {r->weights.data(), r->weights.size() * sizeof(float)},
                                  {r->seg_verb_off.data(), r->seg_verb_off.size() * sizeof(int64_t)},
                                  {r->seg_point_off.data(), r->seg_point_off.size() * sizeof(int64_t)},
                                  {r->seg_weight_off.data(), r->seg_weight_off.size() * sizeof(int64_t)},
                                  {r->seg_xf.data(), r->seg_xf.size() * sizeof(Xform)},
                                  {r->seg_raw.data(), r->seg_raw.size()},
                                  {seg_path.data(), seg_path.size() * sizeof(int32_t)},
                                  {path_seg.data(), path_seg.size() * sizeof(int64_t)}};
        void* dv[10];
        if ((st = upload_packed(dp, pc, 10, dv, s))) break;
        SegInput si;
        si.n_segs = nseg;
        si.n_paths = 1;
        si.verbs = static_cast<const uint8_t*>(dv[0]);
        si.points = static_cast<const float*>(dv[1]);
        si.weights = static_cast<const float*>(dv[2]);
        si.seg_verb_off = static_cast<const int64_t*>(dv[3]);
        si.seg_point_off = static_cast<const int64_t*>(dv[4]);
        si.seg_weight_off = static_cast<const int64_t*>(dv[5]);
        si.seg_xf = static_cast<const Xform*>(dv[6]);
        si.seg_raw = static_cast<const uint8_t*>(dv[7]);
        si.seg_path = static_cast<const int32_t*>(dv[8]);
        si.path_seg_off = static_cast<const int64_t*>(dv[9]);
        si.n_verbs = (int64_t)r->verbs.size();
        si.n_points = (int64_t)r->points.size() / 2;
        si.n_weights = (int64_t)r->weights.size();
        OutputDesc od;
        // one fill, or one stroke (its outline, expanded above, is filled)
        const bool one_fill = nseg == 1 && !r->seg_raw[0];
        if (one_fill && !ctx->c.timing) {
            SegInput s1 = si;
            s1.seg_path = nullptr;
            s1.path_seg_off = nullptr;
            st = finish_one_fill(dp, s1, s, builder);
            if (st != 1) break;  // done (or an error); 1: the general path below
            st = 0;
        }
        int64_t tcap = (int64_t)(dp->tile_alpha.bytes / 64), scap = (int64_t)(dp->span_w.bytes / 2);
        for (int attempt = 0; attempt < 2; attempt++) {
            hipError_t e;
            if ((e = dp->tile_xy.ensure(std::max<int64_t>(tcap, 64) * 4)) != hipSuccess ||
                (e = dp->tile_alpha.ensure(std::max<int64_t>(tcap, 64) * 64)) != hipSuccess ||
                (e = dp->span_xy.ensure(std::max<int64_t>(scap, 64) * 4)) != hipSuccess ||
                (e = dp->span_w.ensure(std::max<int64_t>(scap, 64) * 2)) != hipSuccess ||
                (e = dp->pts.ensure(8)) != hipSuccess || (e = dp->ptc.ensure(8)) != hipSuccess ||
                (e = dp->pss.ensure(8)) != hipSuccess || (e = dp->psc.ensure(8)) != hipSuccess) {
                st = map_hip_error(e);
                break;
            }
            od.tile_xy = dp->tile_xy.as<int16_t>();
            od.tile_alpha = dp->tile_alpha.as<uint8_t>();
            od.tile_capacity = (int64_t)(dp->tile_alpha.bytes / 64);
            od.span_xy = dp->span_xy.as<int16_t>();
            od.span_w = dp->span_w.as<uint16_t>();
            od.span_capacity = (int64_t)(dp->span_w.bytes / 2);
            od.path_tile_start = dp->pts.as<int64_t>();
            od.path_tile_count = dp->ptc.as<int32_t>();
            od.path_span_start = dp->pss.as<int64_t>();
            od.path_span_count = dp->psc.as<int32_t>();
            if (one_fill) {
                // one fill (the reference's demo makes a Rasterizer per fill):
                // the ordered batched kernel, bit-exact like the plain pipeline,
                // in one launch and one host synchronisation; a path over its
                // limits goes to the plain pipeline within the call
                SegInput s1 = si;
                s1.seg_path = nullptr;
                s1.path_seg_off = nullptr;
                const bool dense = ctx->c.dense;
                ctx->c.dense = false;
                st = run_batched(&ctx->c, s1, od, s, ctx->c.batched_waves_per_cu, ctx->c.batched_gran,
                                 ctx->c.batched_tc, false);
                ctx->c.dense = dense;
            } else {
                st = run_plain(&ctx->c, si, od, s);
            }
            if (st != OCHRE_ERR_CAPACITY) break;
            tcap = od.n_tiles;
            scap = od.n_spans;
        }
        if (st) break;
        std::vector<int16_t> txy((size_t)od.n_tiles * 2), sxy((size_t)od.n_spans * 2);
        std::vector<uint8_t> ta((size_t)od.n_tiles * 64);
        std::vector<uint16_t> sw((size_t)od.n_spans);
        hipError_t e = hipSuccess;
        if (od.n_tiles) {
            if ((e = hipMemcpyAsync(txy.data(), od.tile_xy, txy.size() * 2, hipMemcpyDeviceToHost, s)) ||
                (e = hipMemcpyAsync(ta.data(), od.tile_alpha, ta.size(), hipMemcpyDeviceToHost, s))) {
            }
        }
        if (e == hipSuccess && od.n_spans) {
            if ((e = hipMemcpyAsync(sxy.data(), od.span_xy, sxy.size() * 2, hipMemcpyDeviceToHost, s)) ||
                (e = hipMemcpyAsync(sw.data(), od.span_w, sw.size() * 2, hipMemcpyDeviceToHost, s))) {
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            st = map_hip_error(e);
            break;
        }
        replay(builder, txy.data(), ta.data(), od.n_tiles, sxy.data(), sw.data(), od.n_spans);
    } while (0);
    delete r;
    return st;
}

// ------------------------------------------------------- finish_many
//
// Rasterizers holding exactly one fill or one stroke (or nothing) -- the
// demo's pattern, one Rasterizer per fill or stroke finishing into one
// Builder (examples/svg.rs:140-156) -- are rasterised together: one upload,
// one draw-list call through the ordered (bit-exact) kernels, one read-back,
// then each one's TileBuilder calls in order.
static bool single_draw(const ochre_rasterizer* r) {
    if (r->seg_xf.empty()) return true;  // empty: one zero tile
    if (r->seg_xf.size() != 1 || r->seg_raw[0]) return false;
    return r->s_width.empty() || (r->s_width.size() == 1 && r->seg_stroke[0] == 0);
}

static int finish_draws(ochre_rasterizer* const* rs, size_t m, const ochre_tile_builder* builder) {
    ochre_context* ctx = rs[0]->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    static DevPath* dp = nullptr;  // guarded by the default context's mutex
    if (!dp) dp = new (std::nothrow) DevPath();
    if (!dp) return OCHRE_ERR_OUT_OF_MEMORY;
    if (hipSetDevice(ctx->c.device) != hipSuccess) return OCHRE_ERR_DEVICE;
    hipStream_t s = nullptr;
    std::vector<uint8_t> verbs, flags(m, 0);
    std::vector<float> points, weights, widths(m, 0.0f);
    std::vector<int64_t> voff{0}, poff{0}, woff{0};
    std::vector<Xform> xf(m);
    bool any_stroke = false;
    for (size_t k = 0; k < m; k++) {
        const ochre_rasterizer* r = rs[k];
        xf[k] = r->seg_xf.empty() ? Xform{1, 0, 0, 1, 0, 0} : r->seg_xf[0];
        const bool stroke = !r->s_width.empty();
        const std::vector<uint8_t>& v = stroke ? r->s_verbs : r->verbs;
        const std::vector<float>& pt = stroke ? r->s_points : r->points;
        const std::vector<float>& w = stroke ? r->s_weights : r->weights;
        verbs.insert(verbs.end(), v.begin(), v.end());
        points.insert(points.end(), pt.begin(), pt.end());
        weights.insert(weights.end(), w.begin(), w.end());
        voff.push_back((int64_t)verbs.size());
        poff.push_back((int64_t)points.size() / 2);
        woff.push_back((int64_t)weights.size());
        if (stroke) {
            flags[k] = 1;
            widths[k] = r->s_width[0];
            any_stroke = true;
        }
    }
    const HostPiece pc[9] = {{verbs.data(), verbs.size()},
                             {points.data(), points.size() * sizeof(float)},
                             {weights.data(), weights.size() * sizeof(float)},
                             {voff.data(), voff.size() * sizeof(int64_t)},
                             {poff.data(), poff.size() * sizeof(int64_t)},
                             {woff.data(), woff.size() * sizeof(int64_t)},
                             {xf.data(), xf.size() * sizeof(Xform)},
                             {flags.data(), flags.size()},
                             {widths.data(), widths.size() * sizeof(float)}};
    void* dv[9];
    int st = upload_packed(dp, pc, 9, dv, s);
    if (st) return st;
    SegInput si;
    si.n_segs = (int64_t)m;
    si.n_paths = (int64_t)m;
    si.verbs = static_cast<const uint8_t*>(dv[0]);
    si.points = static_cast<const float*>(dv[1]);
    si.weights = static_cast<const float*>(dv[2]);
    si.seg_verb_off = static_cast<const int64_t*>(dv[3]);
    si.seg_point_off = static_cast<const int64_t*>(dv[4]);
    si.seg_weight_off = static_cast<const int64_t*>(dv[5]);
    si.seg_xf = static_cast<const Xform*>(dv[6]);
    si.n_verbs = (int64_t)verbs.size();
    si.n_points = (int64_t)points.size() / 2;
    si.n_weights = (int64_t)weights.size();
    OutputDesc od;
    int64_t tcap = std::max<int64_t>((int64_t)(dp->tile_alpha.bytes / 64), 64 * (int64_t)m);
    int64_t scap = std::max<int64_t>((int64_t)(dp->span_w.bytes / 2), 8 * (int64_t)m);
    const bool dense = ctx->c.dense;
    ctx->c.dense = false;  // the ordered kernels: bit-exact, as a single finish is
    for (int attempt = 0; attempt < 2; attempt++) {
        hipError_t e;
        if ((e = dp->tile_xy.ensure(tcap * 4)) != hipSuccess || (e = dp->tile_alpha.ensure(tcap * 64)) != hipSuccess ||
            (e = dp->span_xy.ensure(std::max<int64_t>(scap, 1) * 4)) != hipSuccess ||
            (e = dp->span_w.ensure(std::max<int64_t>(scap, 1) * 2)) != hipSuccess ||
            (e = dp->pts.ensure(m * 8)) != hipSuccess || (e = dp->ptc.ensure(m * 4)) != hipSuccess ||
            (e = dp->pss.ensure(m * 8)) != hipSuccess || (e = dp->psc.ensure(m * 4)) != hipSuccess) {
            st = map_hip_error(e);
            break;
        }
        od.tile_xy = dp->tile_xy.as<int16_t>();
        od.tile_alpha = dp->tile_alpha.as<uint8_t>();
        od.tile_capacity = (int64_t)(dp->tile_alpha.bytes / 64);
        od.span_xy = dp->span_xy.as<int16_t>();
        od.span_w = dp->span_w.as<uint16_t>();
        od.span_capacity = (int64_t)(dp->span_w.bytes / 2);
        od.path_tile_start = dp->pts.as<int64_t>();
        od.path_tile_count = dp->ptc.as<int32_t>();
        od.path_span_start = dp->pss.as<int64_t>();
        od.path_span_count = dp->psc.as<int32_t>();
        st = any_stroke ? run_stroke(&ctx->c, si, static_cast<const float*>(dv[8]),
                                     static_cast<const uint8_t*>(dv[7]), od, s, true)
                        : run_batched(&ctx->c, si, od, s, ctx->c.batched_waves_per_cu, ctx->c.batched_gran,
                                      ctx->c.batched_tc, false);
        if (st != OCHRE_ERR_CAPACITY) break;
        tcap = od.n_tiles;
        scap = od.n_spans;
    }
    ctx->c.dense = dense;
    if (st) return st;
    std::vector<int16_t> txy((size_t)od.n_tiles * 2), sxy((size_t)od.n_spans * 2);
    std::vector<uint8_t> ta((size_t)od.n_tiles * 64);
    std::vector<uint16_t> sw((size_t)od.n_spans);
    std::vector<int64_t> ts(m), ss(m);
    std::vector<int32_t> tc(m), sc(m);
    hipError_t e = hipSuccess;
    auto get = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    get(txy.data(), od.tile_xy, txy.size() * 2);
    get(ta.data(), od.tile_alpha, ta.size());
    get(sxy.data(), od.span_xy, sxy.size() * 2);
    get(sw.data(), od.span_w, sw.size() * 2);
    get(ts.data(), od.path_tile_start, m * 8);
    get(tc.data(), od.path_tile_count, m * 4);
    get(ss.data(), od.path_span_start, m * 8);
    get(sc.data(), od.path_span_count, m * 4);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return map_hip_error(e);
    for (size_t k = 0; k < m; k++)
        replay(builder, txy.data() + 2 * ts[k], ta.data() + 64 * ts[k], tc[k], sxy.data() + 2 * ss[k],
               sw.data() + ss[k], sc[k]);
    return OCHRE_OK;
}

int ochre_rasterizer_finish_many(ochre_rasterizer* const* rs, size_t n, const ochre_tile_builder* builder) {
    if (n && !rs) return OCHRE_ERR_INVALID_ARG;
    int first_error = OCHRE_OK;
    size_t i = 0;
    while (i < n) {
        size_t j = i;
        while (j < n && rs[j] && single_draw(rs[j]) && rs[j]->ctx == rs[i]->ctx) j++;
        int st;
        if (j - i >= 2) {
            st = finish_draws(rs + i, j - i, builder);
            for (size_t k = i; k < j; k++) delete rs[k];
            i = j;
        } else {
            st = rs[i] ? ochre_rasterizer_finish(rs[i], builder) : OCHRE_ERR_INVALID_ARG;
            i++;
        }
        if (st && first_error == OCHRE_OK) first_error = st;
    }
    return first_error;
}

}  // extern "C"
