/*
 * ochre_oracle.h -- CPU oracle for the ochre rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential, plain-C restatement of the
 * reference rasteriser (glowcoil/ochre, Rust) used as the checker for the HIP
 * implementation in ochre_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  It is never part of the product
 * path.
 *
 * Every function cites the reference file:line it restates.  Numerics follow
 * Rust's f32 semantics: no FMA contraction (built with -ffp-contract=off),
 * IEEE division and sqrt, NaN-ignoring min/max, saturating f32->i16 casts,
 * wrapping i16 arithmetic.
 *
 * Parity pinning (the reference ships no tests or fixtures, and its Rust
 * toolchain is absent here, so it cannot be built or run):
 *   - flatten: closed-form segment counts (README path: 23 + 47 lines),
 *     end point exactness, degenerate curves, f64 chord-error bound;
 *   - DDA / coverage: f64 brute-force polygon clipping per pixel on tiny
 *     inputs, axis-aligned rectangles (255 inside / 0 outside), half-pixel
 *     edges (128), quarter corners (64), shoelace-area conservation;
 *   - spans: hand-derived square rings and winding +-1/+-2 cases;
 *   - quirks: empty path -> one zero tile at (0,0); Close is a no-op;
 *     repeated fill() accumulates.
 *   The tie order of the reference's unstable bin sort (rasterizer.rs:209)
 *   cannot be pinned without running the Rust code: this oracle sums
 *   same-tile bins in emission order (a stable sort), which is one of the
 *   orders the reference may produce; alpha parity is therefore stated as
 *   +-1 LSB against the reference and bit-exact against this oracle.
 *   Conic flattening (path.rs:75-104) and stroking (path.rs:152-274) are
 *   pinned by closed forms and invariants only (see tests/).
 */
#ifndef OCHRE_ORACLE_H
#define OCHRE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PathCmd discriminants (path.rs:5-12). */
enum {
    ORC_MOVE = 0,
    ORC_LINE = 1,
    ORC_QUAD = 2,
    ORC_CUBIC = 3,
    ORC_CONIC = 4,
    ORC_CLOSE = 5
};

/* Error flags (the reference has none; these mark inputs on which the
 * reference would loop forever or overflow its stack). */
enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,
    ORC_ERR_FLATTEN_LIMIT = 2
};

typedef struct orc_rasterizer orc_rasterizer;

typedef void (*orc_tile_fn)(void* user, int16_t x, int16_t y, const uint8_t* alpha64);
typedef void (*orc_span_fn)(void* user, int16_t x, int16_t y, uint16_t width);

/* Rasterizer::new (rasterizer.rs:50-58). */
orc_rasterizer* orc_new(void);
void orc_free(orc_rasterizer* r);
/* Rasterizer::move_to / line_to (rasterizer.rs:61-140). */
void orc_move_to(orc_rasterizer* r, float x, float y);
void orc_line_to(orc_rasterizer* r, float x, float y);
/* Rasterizer::command (rasterizer.rs:145-157); pts holds the command's
 * points (x,y pairs), weight is used by conics only. */
int orc_command(orc_rasterizer* r, int verb, const float* pts, float weight);
/* Rasterizer::fill (rasterizer.rs:161-165); xf = {m0,m1,m2,m3,ox,oy}. */
int orc_fill(orc_rasterizer* r, const uint8_t* verbs, size_t n_verbs,
             const float* points, size_t n_points,
             const float* weights, size_t n_weights, const float* xf);
/* Rasterizer::stroke (rasterizer.rs:169-171). */
int orc_stroke(orc_rasterizer* r, const uint8_t* verbs, size_t n_verbs,
               const float* points, size_t n_points,
               const float* weights, size_t n_weights, float width, const float* xf);
/* Rasterizer::finish (rasterizer.rs:180-268).  Does not free r. */
int orc_finish(orc_rasterizer* r, void* user, orc_tile_fn tile, orc_span_fn span);
/* Diagnostics: counts accumulated so far (lines with point != last,
 * per-pixel increments, tile-row crossings). */
void orc_stats(const orc_rasterizer* r, int64_t* lines, int64_t* increments, int64_t* tile_increments);

/* Free-standing path helpers (path.rs:114-274).  Output buffers are grown
 * by the callee; release with orc_path_free. */
typedef struct {
    uint8_t* verbs;
    float* points;
    float* weights;
    size_t n_verbs, n_points, n_weights;
    size_t cap_verbs, cap_points, cap_weights;
} orc_path;
int orc_flatten_path(const uint8_t* verbs, size_t n_verbs, const float* points, size_t n_points,
                     const float* weights, size_t n_weights, float tolerance, orc_path* out);
int orc_stroke_path(const uint8_t* verbs, size_t n_verbs, const float* points, size_t n_points,
                    float width, orc_path* out);
void orc_path_free(orc_path* p);

/* The demo's atlas Builder (examples/svg.rs:15-88) fed by one Rasterizer
 * per path in path order (svg.rs:140-156), colours [n_paths][4].  Clears the
 * atlas (atlas_size^2 bytes), marks slot (0, 0) solid, then packs tiles from
 * slot 1 on.  `verts` holds 12-byte vertices {i16 pos[2]; u16 uv[2];
 * u8 col[4]}.  Returns the or of the per-path statuses, plus 0x100 when the
 * atlas or the vertex/index capacity is exceeded (*n_verts / *n_idx still
 * report the required counts). */
int orc_atlas_batch(int64_t n_paths, const uint8_t* verbs, const int64_t* verb_off, const float* points,
                    const int64_t* point_off, const float* weights, const int64_t* weight_off, const float* xforms,
                    const float* widths, const uint8_t* stroke_flags, const uint8_t* colors, int atlas_size,
                    uint8_t* atlas, void* verts, size_t vcap, uint32_t* idx, size_t icap, size_t* n_verts,
                    size_t* n_idx);

/* Batch runner: one Rasterizer per path (as examples/svg.rs:143-145 does),
 * fill (or stroke when widths != NULL) with the path's transform, finish.
 * With widths and stroke_flags both set the batch is a draw list: path p is
 * stroked when stroke_flags[p] != 0 and filled otherwise (the demo's
 * sequence of fills and strokes, svg.rs:140-156). */
typedef struct orc_result orc_result;
orc_result* orc_run_batch(int64_t n_paths,
                          const uint8_t* verbs, const int64_t* verb_off,
                          const float* points, const int64_t* point_off,
                          const float* weights, const int64_t* weight_off,
                          const float* xforms, const float* widths, const uint8_t* stroke_flags);
int orc_result_status(const orc_result* res);
void orc_result_sizes(const orc_result* res, int64_t* n_tiles, int64_t* n_spans);
void orc_result_copy(const orc_result* res, int16_t* tile_xy, uint8_t* tile_alpha,
                     int16_t* span_xy, uint16_t* span_w,
                     int64_t* path_tile_off, int64_t* path_span_off);
void orc_result_stats(const orc_result* res, int64_t* lines, int64_t* increments, int64_t* tile_increments);
void orc_result_free(orc_result* res);

#ifdef __cplusplus
}
#endif

#endif
