/*
 * ochre.h -- C ABI of the MI355X (gfx950) implementation of ochre's
 * rasterisation path: Rasterizer::fill(&[PathCmd], Transform) ->
 * Rasterizer::finish(&mut impl TileBuilder) emitting 8x8 u8 alpha tiles and
 * solid n x 8 spans.
 *
 * Reference interface each entry point replaces (glowcoil/ochre v0.2.0):
 *   ochre_rasterizer_new        Rasterizer::new        src/rasterizer.rs:50-58
 *   ochre_rasterizer_move_to    Rasterizer::move_to    src/rasterizer.rs:61-69
 *   ochre_rasterizer_line_to    Rasterizer::line_to    src/rasterizer.rs:72-140
 *   ochre_rasterizer_command    Rasterizer::command    src/rasterizer.rs:145-157
 *   ochre_rasterizer_fill       Rasterizer::fill       src/rasterizer.rs:161-165
 *   ochre_rasterizer_stroke     Rasterizer::stroke     src/rasterizer.rs:169-171
 *   ochre_rasterizer_finish     Rasterizer::finish     src/rasterizer.rs:180-268
 *   ochre_tile_builder          trait TileBuilder      src/rasterizer.rs:12-22
 *   ochre_transform             Transform / Mat2x2     src/geom.rs:131-133,203-207
 *   OCHRE_VERB_*                enum PathCmd           src/path.rs:5-12
 * The batch entry points (ochre_rasterize_plain / ochre_rasterize_batched)
 * run many independent Rasterizers (one per path, as examples/svg.rs:143-145
 * does) on device-resident inputs; they have no single reference
 * counterpart and are the throughput interface.
 *
 * Paths are passed structure-of-arrays: one verb byte per PathCmd, the
 * command points as consecutive f32 (x, y) pairs (Move/Line 1, Quadratic 2,
 * Cubic 3, Conic 2, Close 0), and one weight per Conic.
 *
 * Errors: every entry point returns an int status (0 = OCHRE_OK, negative
 * on failure); nothing unwinds across the ABI.  The reference API is
 * infallible; failures here are argument, memory or device errors.
 */
#ifndef OCHRE_H
#define OCHRE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCHRE_TILE_SIZE 8

enum ochre_verb {
    OCHRE_VERB_MOVE = 0,
    OCHRE_VERB_LINE = 1,
    OCHRE_VERB_QUADRATIC = 2,
    OCHRE_VERB_CUBIC = 3,
    OCHRE_VERB_CONIC = 4,
    OCHRE_VERB_CLOSE = 5
};

enum ochre_status {
    OCHRE_OK = 0,
    OCHRE_ERR_INVALID_ARG = -1,
    OCHRE_ERR_OUT_OF_MEMORY = -2,
    OCHRE_ERR_DEVICE = -3,
    OCHRE_ERR_CAPACITY = -4,      /* output arrays too small; required sizes reported */
    OCHRE_ERR_FLATTEN_LIMIT = -5, /* a curve needed 65536 or more lines, a conic more than 32
                                     levels of subdivision, or a line more than 2^20 DDA steps
                                     (inputs on which the reference does not terminate) */
    OCHRE_ERR_NO_DEVICE = -6,
    OCHRE_ERR_TOO_LARGE = -7      /* one path exceeds the per-pass limits (internal 32-bit indexing) */
};

/* Transform (src/geom.rs:203-207): m is the row-major Mat2x2
 * (m[0] m[1]; m[2] m[3]), applied as (m0*x + m1*y + ox, m2*x + m3*y + oy). */
typedef struct ochre_transform {
    float m[4];
    float offset[2];
} ochre_transform;

/* TileBuilder (src/rasterizer.rs:12-22).  `alpha` is 64 bytes, row-major
 * (alpha[row * 8 + col]), valid only during the call.  Calls arrive in the
 * reference's order: row-major by (tile y, tile x), each span right after the
 * tile to its left. */
typedef struct ochre_tile_builder {
    void* user;
    void (*tile)(void* user, int16_t x, int16_t y, const uint8_t* alpha);
    void (*span)(void* user, int16_t x, int16_t y, uint16_t width);
} ochre_tile_builder;

/* ------------------------------------------------------------ Rasterizer */

typedef struct ochre_rasterizer ochre_rasterizer;

/* Rasterizer::new.  Returns NULL on allocation failure or when no GPU is
 * usable (the library has no CPU fallback). */
ochre_rasterizer* ochre_rasterizer_new(void);
int ochre_rasterizer_move_to(ochre_rasterizer* r, float x, float y);
int ochre_rasterizer_line_to(ochre_rasterizer* r, float x, float y);
/* `points` holds the command's points (verb-dependent count); `weight` is
 * read for conics only.  No transform is applied (as Rasterizer::command). */
int ochre_rasterizer_command(ochre_rasterizer* r, uint8_t verb, const float* points, float weight);
int ochre_rasterizer_fill(ochre_rasterizer* r, const uint8_t* verbs, size_t n_verbs, const float* points,
                          size_t n_points, const float* weights, size_t n_weights,
                          const ochre_transform* transform);
/* Rasterizer::stroke: the path is flattened and stroked (butt caps, miter
 * joins; path.rs:114-274) in path space with the given width, and the outline
 * is filled with the transform.  Both expansions run on the GPU at finish. */
int ochre_rasterizer_stroke(ochre_rasterizer* r, const uint8_t* verbs, size_t n_verbs, const float* points,
                            size_t n_points, const float* weights, size_t n_weights, float width,
                            const ochre_transform* transform);
/* Rasterizer::finish: rasterises on the GPU, then calls the builder.
 * Consumes the rasterizer (the handle is freed, whatever the status). */
int ochre_rasterizer_finish(ochre_rasterizer* r, const ochre_tile_builder* builder);
/* Finishes n rasterizers into one builder, in order: the calls builder
 * receives are those of ochre_rasterizer_finish on each in turn (the demo's
 * loop, one Rasterizer per fill or stroke finishing into one Builder,
 * examples/svg.rs:140-156).  Consecutive rasterizers that each hold one fill,
 * one stroke or nothing are rasterised together in one device call (the
 * ordered, bit-exact kernels); others are finished one at a time.  Consumes
 * every handle; returns the first error (the remaining rasterizers are still
 * finished).  No single reference counterpart. */
int ochre_rasterizer_finish_many(ochre_rasterizer* const* rasterizers, size_t n, const ochre_tile_builder* builder);
/* Frees a rasterizer that will not be finished. */
void ochre_rasterizer_free(ochre_rasterizer* r);

/* -------------------------------------------------------- batch interface */

/* Device-resident input: n_paths paths, each filled with its own transform.
 * All pointers are device pointers; *_off arrays have n_paths + 1 entries
 * (offsets in elements: verbs, points (pairs), weights). */
typedef struct ochre_path_batch {
    int64_t n_paths;
    const uint8_t* verbs;
    const int64_t* verb_off;
    const float* points;
    const int64_t* point_off;
    const float* weights;
    const int64_t* weight_off;
    const ochre_transform* transforms;
    int64_t n_verbs;   /* = verb_off[n_paths] (host copy) */
    int64_t n_points;  /* = point_off[n_paths] */
    int64_t n_weights; /* = weight_off[n_paths] */
} ochre_path_batch;

/* Device-resident output, caller-owned.  Tiles of path p occupy
 * [path_tile_start[p], path_tile_start[p] + path_tile_count[p]) in row-major
 * (y, x) order (likewise spans); tile_alpha is 64 bytes per tile.  On
 * OCHRE_ERR_CAPACITY nothing is written except n_tiles / n_spans, which
 * hold the required sizes. */
typedef struct ochre_batch_output {
    int16_t* tile_xy;        /* [tile_capacity][2] */
    uint8_t* tile_alpha;     /* [tile_capacity][64] */
    int64_t tile_capacity;
    int16_t* span_xy;        /* [span_capacity][2] */
    uint16_t* span_width;    /* [span_capacity] */
    int64_t span_capacity;
    int64_t* path_tile_start; /* [n_paths] */
    int32_t* path_tile_count; /* [n_paths] */
    int64_t* path_span_start; /* [n_paths] */
    int32_t* path_span_count; /* [n_paths] */
    int64_t n_tiles;         /* out */
    int64_t n_spans;         /* out */
    float* tile_coverage;    /* optional: [tile_capacity][64] pre-quantisation coverage
                                min(|accum + area|, 1) (rasterizer.rs:235 before * 256), written for
                                every tile (handed-on paths included) when not NULL */
} ochre_batch_output;

typedef struct ochre_context ochre_context;

/* One context per device and thread; it owns scratch memory that grows on
 * demand and is reused across calls. */
int ochre_context_create(int device, ochre_context** out);
void ochre_context_destroy(ochre_context* ctx);

/* The five-kernel pipeline: (a) flatten, (b) DDA binning + device radix
 * sort, (c)-(e) per-tile accumulation, row scan and compaction.  Paths are
 * output contiguously in path order.  `stream` is a hipStream_t (NULL =
 * default stream).  Synchronises the stream to size its scratch and, at the
 * end, to read the call's error flags, so the outputs are complete when it
 * returns.  An internal invariant broken on the device (a run replay past its
 * step bound, a run outside its path's grid) returns OCHRE_ERR_DEVICE. */
int ochre_rasterize_plain(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                          void* stream);

/* The batched pipeline: one persistent work-queue launch (one wave64 per
 * path; the launch and its queue reset are replayed from a captured HIP
 * graph), for batches of many small paths.  Paths whose tile grid exceeds
 * 32 x 32 tiles or which flatten to more than 8192 lines are handed to the
 * plain pipeline within the same call.  Tiles of a path are contiguous and in
 * row-major order, but paths are laid out in the order they finish. */
int ochre_rasterize_batched(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                            void* stream);

/* ochre_rasterize_batched for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  It runs the dense kernels alone --
 * the 64-tile instance over every path, then the 256-tile instance over the
 * paths it hands on -- with their queues, counters and the list of handed-on
 * paths in a fixed device array of 4 MiB + 256 B per device, held until the
 * process ends; it calls no allocator (the caller passes the inputs and the
 * outputs, as for ochre_rasterize_batched) and synchronises the stream once to
 * read the counters back.  Limits: at most 2^20 paths per call
 * (OCHRE_ERR_TOO_LARGE before anything runs); a path over the 256-tile
 * instance (a tile grid of more than 256 tiles or 32 on a side, more than
 * 128 lines over 8 pixel steps, a conic weight that is not positive and
 * finite) makes the call return OCHRE_ERR_TOO_LARGE after the other paths are
 * written, with that path's range unwritten (ochre_rasterize_batched handles
 * it).  Every call on a device shares that device's array, which keeps the
 * last call's contents; each call clears the counters it uses, and calls that
 * overlap on one device race.  Results equal ochre_rasterize_batched with the
 * dense kernels on: tile and span sets exact, alpha within 1 LSB. */
int ochre_rasterize_batched_noalloc(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                                    void* stream);

/* Rasterizer::stroke for every path of a batch (widths: [n_paths], device
 * memory): flattening and stroke expansion on the device (one wave per path),
 * then the plain or batched fill of the outlines with the paths' transforms.
 * Returns OCHRE_ERR_INVALID_ARG for verbs that do not match their points or
 * weights and OCHRE_ERR_FLATTEN_LIMIT for a curve the reference would not
 * finish flattening. */
int ochre_stroke_plain(ochre_context* ctx, const ochre_path_batch* in, const float* widths, ochre_batch_output* out,
                       void* stream);
int ochre_stroke_batched(ochre_context* ctx, const ochre_path_batch* in, const float* widths,
                         ochre_batch_output* out, void* stream);

/* Device-side tile consumer: the demo's atlas Builder (examples/svg.rs:15-88)
 * applied to a batch's device output, as if one Rasterizer per path finished
 * into one Builder in path order (svg.rs:140-156).  Writes the R8 atlas
 * (atlas_size^2 bytes, atlas_size a multiple of 8: slot 0 solid, tile g of
 * the batch in slot g + 1, row-major slots of 8x8, unused slots cleared),
 * 4 vertices and 6 indices per TileBuilder call, in call order (tiles
 * textured from their slot, spans at uv 0), colour colors[p] (4 bytes per
 * path, device memory).  `in` is an ochre_batch_output filled by any of the
 * rasterize/stroke entry points for these n_paths.  OCHRE_ERR_CAPACITY when
 * the tiles do not fit the atlas or the vertex/index arrays are too small
 * (*n_vertices / *n_indices report the sizes needed).  Synchronises the
 * stream once to size the output. */
typedef struct ochre_atlas_vertex {
    int16_t pos[2];  /* svg.rs:15-20 Vertex */
    uint16_t uv[2];
    uint8_t col[4];
} ochre_atlas_vertex;
int ochre_build_atlas(ochre_context* ctx, const ochre_batch_output* in, int64_t n_paths, const uint8_t* colors,
                      uint8_t* atlas, int atlas_size, ochre_atlas_vertex* vertices, int64_t vertex_capacity,
                      uint32_t* indices, int64_t index_capacity, int64_t* n_vertices, int64_t* n_indices,
                      void* stream);

/* Tunables of the batched pipeline (the autotuner sets them): persistent
 * waves per compute unit (1-32), paths taken per queue pop (1-1024), tiles
 * accumulated per pass in LDS (16 or 32: fewer tiles, more waves per CU),
 * and whether to replay the launch from a HIP graph. */
int ochre_context_set_batched_params(ochre_context* ctx, int waves_per_cu, int paths_per_pop, int tiles_per_chunk,
                                     int use_graph);
int ochre_context_get_batched_params(ochre_context* ctx, int* waves_per_cu, int* paths_per_pop, int* tiles_per_chunk,
                                     int* use_graph);

/* The dense single-walk kernel (on by default) runs in front of the ordered
 * one and takes every path whose tile grid (line bounding box + a 2 px
 * margin) has at most 64 tiles: one DDA walk adds each increment to its
 * pixel's sums with LDS float atomics, so f32 sums are taken in arrival
 * order and an alpha can differ from the ordered result by 1 LSB (tile and
 * span sets are exact).  Off: every path goes through the ordered kernel,
 * whose sums follow emission order.  groups_per_cu 1-8 resident workgroups,
 * waves_per_group 2, 4 or 8 (waves sharing one path; with 8, at most 2
 * workgroups per CU are resident). */
int ochre_context_set_dense_params(ochre_context* ctx, int enabled, int groups_per_cu, int waves_per_group);
int ochre_context_get_dense_params(ochre_context* ctx, int* enabled, int* groups_per_cu, int* waves_per_group);
/* A small-grid dense instance (tile grids of at most 32 tiles, 2 waves and
 * 16 KB of LDS per workgroup, so twice as many paths in flight per CU) runs in
 * front of the 64-tile one and hands it the larger paths; groups_per_cu 0
 * (off) to 9, default 0 (on the 10,000-glyph run one 64-tile instance is
 * faster; the autotuner tries both). */
int ochre_context_set_dense_small(ochre_context* ctx, int groups_per_cu);
int ochre_context_get_dense_small(ochre_context* ctx, int* groups_per_cu);
/* A 256-tile dense instance (one workgroup of 16 waves and 130 KB of LDS per
 * CU) behind the 64-tile one, over the paths it hands on (grids of 65-256
 * tiles go there instead of to the ordered kernel); enabled 0 or 1, default
 * 1.  The autotuner tries both. */
int ochre_context_set_dense_large(ochre_context* ctx, int enabled);
int ochre_context_get_dense_large(ochre_context* ctx, int* enabled);

/* Per-pass limits of the plain pipeline: one pass holds at most max_lines
 * flattened lines and max_entries sort entries (defaults 2^29 and 2^30, about
 * 66 GB of intermediates at both).  A batch over either is rasterised as
 * consecutive path ranges into the same output, with the same result; a
 * single path over them returns OCHRE_ERR_TOO_LARGE.  0 restores a default;
 * values above the defaults are clamped to them.  Tests lower the limits to
 * exercise the split on small batches.  No reference counterpart (the
 * reference holds every increment of one path in host memory). */
int ochre_context_set_pass_limits(ochre_context* ctx, int64_t max_lines, int64_t max_entries);

/* Coverage adjoint.  The loss of path p is the sum, over the pixels of the
 * union of its tiles and its target's tiles, of (cov - target)^2, where a
 * tile pixel's cov is min(|accum + area|, 1) and a pixel outside a render's
 * tiles has coverage 1 inside one of its spans and 0 elsewhere.  Writes the
 * f32 loss per path, dL/d(points) (path space, pairs, same layout as the
 * input points) and dL/d(transform) (m00 m01 m10 m11 ox oy per path).  The
 * target is a render in the ochre_batch_output layout (tile_coverage set).
 * A path within one wave's limits (a 32 x 32 tile grid, 256 tiles, 1024
 * commands, 4096 points, 8192 lines) is differentiated by one wave; a larger
 * one after the batch, alone on the device, with its lines recomputed from the
 * commands and its tile grid in device memory.  OCHRE_ERR_TOO_LARGE only for a
 * path over 2^24 lines or a grid over 2^21 tiles, or with coordinates beyond
 * the 16-bit pixel range.  No reference counterpart (the reference has no
 * derivatives). */
typedef struct ochre_coverage_target {
    const int64_t* tile_start;
    const int32_t* tile_count;
    const int16_t* tile_xy;
    const float* tile_coverage;
    const int64_t* span_start;
    const int32_t* span_count;
    const int16_t* span_xy;
    const uint16_t* span_width;
} ochre_coverage_target;
int ochre_coverage_adjoint(ochre_context* ctx, const ochre_path_batch* in, const ochre_coverage_target* target,
                           float* loss, float* d_points, float* d_transforms, void* stream);

/* Autotuner: with the dense kernel on, times every (workgroups per CU
 * {1,2,3,4}, waves per workgroup {2,4,8}) of the 64-tile dense launch, then
 * the small-grid instance {off, 4, 6, 8 workgroups per CU}, the 256-tile
 * instance {off, on}, then waves per CU {1,2,4,8} of the ordered kernel
 * behind them; with it off,
 * every combination of waves per CU {2,3,4,6,8,12}, paths per pop {1,2,4,8}
 * and tiles per chunk {16,32} of the ordered launch.  Each configuration is
 * timed on the caller's batch (`reps` calls, median); the fastest is kept in
 * the context and every measurement reported in `results` (up to
 * max_results entries). */
typedef struct ochre_batched_tuning {
    int waves_per_cu;
    int paths_per_pop;
    int tiles_per_chunk;
    float ms;
    int dense_groups_per_cu; /* 0: the dense kernel was off */
    int dense_waves_per_group;
    int dense_small_groups_per_cu; /* the small-grid dense instance (0: off) */
    int dense_large;               /* the 256-tile dense instance (0: off) */
} ochre_batched_tuning;
int ochre_autotune_batched(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out, void* stream,
                           int reps, ochre_batched_tuning* results, int max_results, int* n_results);

/* The autotuner with a cache.  The key is the device (its architecture name
 * and compute unit count), whether the dense kernels are on, and a coarse
 * signature of the batch (power-of-two buckets of its path count and
 * half-octave buckets of its mean commands and points per path).  A key
 * found in the process's table -- which first reads `cache_path`, when not
 * NULL, once per path -- sets the stored parameters in the context without
 * timing anything (*tuned = 0); otherwise the batch is tuned as by
 * ochre_autotune_batched (reps calls per configuration), the choice stored,
 * and, with a cache path, appended to that file as one text line (*tuned = 1).
 * The cache is the library's: every caller of the C ABI shares it. */
int ochre_autotune_batched_cached(ochre_context* ctx, const ochre_path_batch* in, ochre_batch_output* out,
                                  void* stream, int reps, const char* cache_path, int* tuned);

/* Kernel-time accounting for measurement: when enabled, the pipeline
 * records HIP events around its dominant kernel; ochre_context_timing
 * returns the summed milliseconds and launch count since the last reset.
 * The unordered tile stage of the plain pipeline also records pass
 * statistics under "pass.lines", "pass.entries", "pass.runs", "pass.rows",
 * "pass.windows" and "pass.windows_used" (the sum over passes in total_ms,
 * the pass count in launches). */
int ochre_context_set_timing(ochre_context* ctx, int enabled);
int ochre_context_timing(ochre_context* ctx, const char* kernel, double* total_ms, int64_t* launches);

/* A draw list: path p is its own Rasterizer, filled (stroke_flags[p] == 0,
 * Rasterizer::fill) or stroked with widths[p] (stroke_flags[p] == 1,
 * Rasterizer::stroke), in path order -- the sequence of Rasterizers the
 * demo's SVG walk finishes into one Builder (examples/svg.rs:140-156).
 * stroke_flags and widths are device arrays of n_paths.  Errors as for
 * ochre_stroke_*.  With every flag 0 the result equals ochre_rasterize_*;
 * with every flag 1 it equals ochre_stroke_*. */
int ochre_draw_plain(ochre_context* ctx, const ochre_path_batch* in, const uint8_t* stroke_flags,
                     const float* widths, ochre_batch_output* out, void* stream);
int ochre_draw_batched(ochre_context* ctx, const ochre_path_batch* in, const uint8_t* stroke_flags,
                       const float* widths, ochre_batch_output* out, void* stream);

/* ------------------------------------------------------------- SVG input */

/* SVG ingestion (host): parses a document into the draw calls the demo makes
 * from it (examples/svg.rs:103-164, which reads the file with usvg 0.13):
 * for every path element (and rect, circle, ellipse, line, polyline,
 * polygon) in document order, a fill draw when its fill is a solid colour,
 * then a stroke draw when its stroke is a solid colour of positive width.
 * Path data is converted as usvg does (absolute f64 MoveTo / LineTo /
 * CurveTo / ClosePath, quadratics and arcs as cubics) and cast to f32; a
 * draw's transform is the product of its element's and its ancestors'
 * transforms up to the nearest group usvg keeps (opacity, clip-path, mask or
 * filter), as the demo reads node.transform() (svg.rs:107-111).  The viewBox
 * is reported, not applied (the demo does not apply it).  Gradients,
 * patterns, text, images, <use> and CSS style sheets are not drawn.
 * Returns OCHRE_ERR_INVALID_ARG when the text is not an XML document whose
 * root element is <svg>. */
typedef struct ochre_svg_document ochre_svg_document;
typedef struct ochre_svg_draw {
    int64_t path;              /* index of the draw's path (its verb and point ranges) */
    int32_t stroke;            /* 0: Rasterizer::fill, 1: Rasterizer::stroke with `width` */
    float width;               /* stroke width in user units */
    uint8_t color[4];          /* rgb and opacity: the Builder's colour for the call (svg.rs:141,150) */
    ochre_transform transform; /* applied to the path's points */
} ochre_svg_draw;
typedef struct ochre_svg_summary {
    int64_t n_paths, n_verbs, n_points, n_draws;
    int64_t n_skipped;  /* elements not drawn (not graphics, unsupported, display="none") */
    float view_box[4];  /* min-x, min-y, width, height (0 when absent) */
    float width, height;
} ochre_svg_summary;
int ochre_svg_parse(const char* text, size_t len, ochre_svg_document** out);
int ochre_svg_info(const ochre_svg_document* doc, ochre_svg_summary* info);
/* Paths of the document, structure-of-arrays as in ochre_path_batch (host
 * memory owned by the document; offsets have n_paths + 1 entries). */
const uint8_t* ochre_svg_verbs(const ochre_svg_document* doc);
const int64_t* ochre_svg_verb_offsets(const ochre_svg_document* doc);
const float* ochre_svg_points(const ochre_svg_document* doc);
const int64_t* ochre_svg_point_offsets(const ochre_svg_document* doc);
const ochre_svg_draw* ochre_svg_draws(const ochre_svg_document* doc);
void ochre_svg_free(ochre_svg_document* doc);

/* Version / build information (NUL-terminated, static storage). */
const char* ochre_build_info(void);

#ifdef __cplusplus
}
#endif

#endif
