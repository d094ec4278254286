// Internal interfaces of the MI355X rasterisation pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "scan.hpp"

namespace ochre {

// Device input.  A path is one Rasterizer; it consists of one or more
// segments, each a run of commands filled with one transform (one segment
// per Rasterizer::fill call; `raw` segments come from move_to / line_to /
// command and are not transformed).  The batch API maps path i to segment i.
struct SegInput {
    int64_t n_segs = 0;
    int64_t n_paths = 0;
    const uint8_t* verbs = nullptr;
    const float* points = nullptr;    // (x, y) pairs
    const float* weights = nullptr;
    const int64_t* seg_verb_off = nullptr;   // [n_segs + 1]
    const int64_t* seg_point_off = nullptr;  // [n_segs + 1]
    const int64_t* seg_weight_off = nullptr; // [n_segs + 1]
    const Xform* seg_xf = nullptr;           // [n_segs]
    const uint8_t* seg_raw = nullptr;        // [n_segs] or null (none raw)
    const int32_t* seg_path = nullptr;       // [n_segs] or null (segment i -> path i)
    const int64_t* path_seg_off = nullptr;   // [n_paths + 1] or null (identity)
    int64_t n_verbs = 0, n_points = 0, n_weights = 0;
};

struct OutputDesc {
    int16_t* tile_xy = nullptr;
    uint8_t* tile_alpha = nullptr;
    float* tile_cov = nullptr;  // optional f32 pre-quantisation coverage (batched pipeline)
    int64_t tile_capacity = 0;
    int16_t* span_xy = nullptr;
    uint16_t* span_w = nullptr;
    int64_t span_capacity = 0;
    int64_t* path_tile_start = nullptr;
    int32_t* path_tile_count = nullptr;
    int64_t* path_span_start = nullptr;
    int32_t* path_span_count = nullptr;
    int64_t n_tiles = 0, n_spans = 0;
    // added to the per-path starts written (the output arrays above already
    // point at tile_base / span_base of the caller's arrays)
    int64_t tile_base = 0, span_base = 0;
};

// Per-path tile grid: keys are ((row_base + ty - ty0) << xbits) | (tx - tx0).
struct PathGrid {
    int32_t tx0, ty0, w, h;
    uint32_t row_base;
    int32_t pad[3];
};

// Per-path bounding box of everything the DDA can visit (pixel coords).
struct PathBox {
    int32_t minx, miny, maxx, maxy;
    uint32_t flags;  // bit0: full i16 range needed, bit1: has lines
    uint32_t pad[3];
};

// Device counters shared by the kernels (one 256-byte block).
struct Counters {
    uint32_t error;        // OR of error bits
    uint32_t max_w;        // max grid width in tiles
    unsigned long long entries;  // sort entries of all lines (64-bit: a pass over 2^32 is detected)
    uint32_t pad[60];
};

enum ErrorBits : uint32_t {
    ERR_FLATTEN = 1u,
    ERR_BBOX = 2u,       // a run fell outside its path grid (internal invariant)
    ERR_INPUT = 4u,      // verbs/points/weights inconsistent
    ERR_REPLAY = 8u,     // a run replay exceeded its step bound (internal invariant)
    ERR_CAPACITY = 16u,  // a fixed-capacity buffer overflowed (batched path)
};

// Growable device buffer.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need);
    ~DBuf();
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

// Device buffers of the stroke expansion (flattened polylines, outlines).
struct StrokeBuffers {
    DBuf cnt, fv, fp, ov, op, ow, err, fverbs, fpoints, overbs, opoints, oweights;
    void* host = nullptr;  // pinned: sizes and error bits read back
    ~StrokeBuffers();
};

// Device buffers of the atlas builder (per-path tile and span bases).
struct AtlasState {
    DBuf tile_base, span_base;
    void* host = nullptr;  // pinned: totals read back
    ~AtlasState();
};

// The plain five-kernel pipeline.  One pass holds at most PLAIN_MAX_LINES
// flattened lines and PLAIN_MAX_ENTRIES sort entries (about 66 GB of
// intermediates at both limits); a batch of independent paths above either is
// split into path ranges that are rasterised one after another into the same
// output (same result: paths are independent).  A single path above them, or
// a segment input (several fills per path), returns -7.
constexpr int64_t PLAIN_MAX_LINES = 1ll << 29;
constexpr int64_t PLAIN_MAX_ENTRIES = 1ll << 30;

struct Context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    // tile sums of the device-wide scans
    DBuf scan_partials;
    // pipeline scratch
    DBuf cmd_scan, cmd_seg, cmd_lines, cmd_line_off;
    DBuf lines, line_path, line_cnt, line_off, dda_blk;
    DBuf path_box, path_grid, path_rows;
    DBuf counters, totals;
    DBuf call_error;  // emit-phase error bits of the current plain call (all passes)
    DBuf keys[2], vals[2], records, rs_counts;
    DBuf grp_begin, grp_key, grp_hasrun, grp_sign;
    DBuf row_gbegin, row_gend, row_sign, row_wb, row_path, row_cnt, row_off;
    DBuf split_off;  // rebased per-path offsets of a path range (run_plain's split)
    DBuf u_pwbase, u_rwbase, u_bcnt, u_bbeg, u_bcur;  // unordered tile stage: window buckets
    double split_ratio = 2.0;  // how far the last pass that had to split was over its limits
    int64_t max_lines = PLAIN_MAX_LINES, max_entries = PLAIN_MAX_ENTRIES;  // per-pass limits (lowered by tests)
    // host staging
    void* host_pinned = nullptr;
    // timing
    bool timing = false;
    std::map<std::string, std::pair<double, int64_t>> timings;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // batched-path state lives in batched.hip
    void* batched = nullptr;
    int batched_waves_per_cu = 6;
    int batched_gran = 2;
    int batched_tc = 32;
    bool batched_graph = true;
    // dense single-walk kernel in front of the ordered one: on/off, resident
    // workgroups per CU, waves per workgroup (2 or 4)
    bool dense = true;
    int dense_groups_per_cu = 4;
    int dense_waves = 4;
    int dense_small_groups = 0;  // small-grid instance in front (0: off)
    int dense_large = 1;         // 256-tile instance behind (1 workgroup of 16 waves per CU; 0: off)
    StrokeBuffers stroke;
    AtlasState atlas;
    ~Context();
};

// The plain five-kernel pipeline; see the per-pass limits above.
// `unordered` (used for the paths the batched pipeline hands on, under the
// dense kernels' 1-LSB contract) replaces the sort and the ordered tile pass
// by window buckets and LDS float-atomic sums.
int run_plain(Context* ctx, const SegInput& in, OutputDesc& out, hipStream_t stream, bool unordered = false);
// Persistent work-queue pipeline (one wave per path); paths that do not fit a
// wave are deferred to run_plain.  waves_per_cu sets the persistent grid,
// gran the paths taken per queue pop; use_graph replays (reset, kernel) from
// a captured HIP graph.
int run_batched(Context* ctx, const SegInput& in, OutputDesc& out, hipStream_t stream, int waves_per_cu, int gran,
                int tiles_per_chunk, bool use_graph);
// One path by k_single with no host round trip: the output goes where `out`
// points and the control block (SMALL_CTL_BYTES, reset by the kernel) to
// `ctl`, both device-visible (pinned host memory mapped into the device's
// address space, for the drop-in finish); after synchronising,
// small_batch_result gives the counts, the status of run_batched and whether
// the path was over the kernel's limits (its output is not written: rerun
// through run_batched).
constexpr size_t SMALL_CTL_BYTES = 256;
struct SmallBatchResult {
    int64_t tiles, spans;
    uint32_t deferred;
    int status;
};
int launch_single_fill(const SegInput& in, const OutputDesc& out, void* ctl, hipStream_t stream);
SmallBatchResult small_batch_result(const void* host_ctl);

// The dense kernels alone (the 64-tile instance, then the 256-tile one over
// what it hands on) with caller-provided device memory: `ctl` of
// DENSE_CTL_BYTES (cleared by a kernel at the start) and `ids` of n_paths
// entries.  finish_dense_chain reads the control block back after the launch:
// the statuses of run_batched, and -7 when a path was over the 256-tile
// instance (its output range is not written).  Neither allocates.
constexpr size_t DENSE_CTL_BYTES = 256;
int launch_dense_chain(const SegInput& in, const OutputDesc& out, uint8_t* ctl, uint32_t* ids, hipStream_t stream);
int finish_dense_chain(const uint8_t* ctl, OutputDesc& out, hipStream_t stream);
// The allocation-free batched call (noalloc.hip): the dense chain with its
// control block and handed-on list in a fixed device array.  noalloc_init_device
// looks the array's address up on the current device (once per device).
int noalloc_init_device(int device);
int run_batched_noalloc(const SegInput& in, OutputDesc& out, hipStream_t stream);

// Target render of the coverage adjoint (device pointers, per path ranges).
struct AdjTarget {
    const int64_t* tile_start;
    const int32_t* tile_count;
    const int16_t* tile_xy;
    const float* tile_cov;
    const int64_t* span_start;
    const int32_t* span_count;
    const int16_t* span_xy;
    const uint16_t* span_w;
};
// Loss and gradients of the f32 coverage (one wave per path).
int run_adjoint(Context* ctx, const SegInput& in, const AdjTarget& tgt, float* loss, float* d_points, float* d_xf,
                hipStream_t stream, int waves_per_cu, int gran, int tiles_per_chunk);

// Outline of a stroked batch: a fill batch (one segment per path, the input's
// transforms) whose buffers belong to the context until its next stroke call.
struct StrokeOutline {
    SegInput in;
};
// Rasterizer::stroke for every path of a batch: path-space flattening and
// stroke expansion on the device (widths: [n_paths], device memory).  With
// stroke_flags (device, [n_paths]) the batch is a draw list: a path whose
// flag is 0 is a fill and passes through unchanged.
int stroke_expand(Context* ctx, const SegInput& in, const float* widths, const uint8_t* stroke_flags,
                  StrokeOutline& out, hipStream_t stream);
// stroke_expand, then the plain or batched fill of the outlines.
int run_stroke(Context* ctx, const SegInput& in, const float* widths, const uint8_t* stroke_flags, OutputDesc& out,
               hipStream_t stream, bool batched);

}  // namespace ochre
struct ochre_atlas_vertex;
namespace ochre {
// The demo's atlas Builder (examples/svg.rs:15-88) over a batch's device
// output, paths in index order; colours [n_paths][4] device memory.
int run_atlas(Context* ctx, const OutputDesc& in, int64_t n_paths, const uint8_t* colors, uint8_t* atlas,
              int atlas_size, ochre_atlas_vertex* verts, int64_t vcap, uint32_t* idx, int64_t icap, int64_t* n_verts,
              int64_t* n_idx, hipStream_t stream);

#define HIPCHK_RET(x)                                      \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return ::ochre::map_hip_error(e_); \
    } while (0)

int map_hip_error(hipError_t e);

}  // namespace ochre
