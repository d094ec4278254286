/*
 * Test infrastructure: the demo's tile consumer (examples/svg.rs:15-88), the
 * `Builder` TileBuilder that packs every 8x8 alpha tile into one R8 atlas and
 * emits a textured quad per tile and an untextured quad per span, restated
 * sequentially in C.  A batch is consumed as examples/svg.rs:143-156 does:
 * one Rasterizer per path, in path order, all finishing into the one Builder
 * with the path's colour set first.  Only the tests load this.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ochre_oracle.h"

typedef struct {
    int16_t pos[2];
    uint16_t uv[2];
    uint8_t col[4];
} orc_vertex;

typedef struct {
    uint8_t* atlas;
    int atlas_size;
    orc_vertex* verts;
    size_t vcap, nv;
    uint32_t* idx;
    size_t icap, ni;
    uint16_t next_row, next_col;  /* svg.rs:27-28, Builder::new: row 0, col 1 */
    uint8_t color[4];
    int overflow;
} builder;

static void push_quad(builder* b, int16_t x0, int16_t y0, int16_t x1, int16_t y1, uint16_t u1, uint16_t v1,
                      uint16_t u2, uint16_t v2) {
    if (b->nv + 4 > b->vcap || b->ni + 6 > b->icap) {
        b->overflow = 1;
        b->nv += 4;
        b->ni += 6;
        return;
    }
    uint32_t base = (uint32_t)b->nv;
    orc_vertex q[4] = {
        {{x0, y0}, {u1, v1}, {0}}, {{x1, y0}, {u2, v1}, {0}}, {{x1, y1}, {u2, v2}, {0}}, {{x0, y1}, {u1, v2}, {0}}};
    for (int k = 0; k < 4; k++) {
        memcpy(q[k].col, b->color, 4);
        b->verts[b->nv++] = q[k];
    }
    const uint32_t ix[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
    for (int k = 0; k < 6; k++) b->idx[b->ni++] = ix[k];
}

/* Builder::tile (svg.rs:52-78) */
static void on_tile(void* user, int16_t x, int16_t y, const uint8_t* data) {
    builder* b = (builder*)user;
    const uint16_t u1 = (uint16_t)(b->next_col * 8), u2 = (uint16_t)((b->next_col + 1) * 8);
    const uint16_t v1 = (uint16_t)(b->next_row * 8), v2 = (uint16_t)((b->next_row + 1) * 8);
    push_quad(b, x, y, (int16_t)(x + 8), (int16_t)(y + 8), u1, v1, u2, v2);
    const int slots = b->atlas_size / 8;
    if (b->next_row < slots) {
        for (int row = 0; row < 8; row++)
            memcpy(b->atlas + ((size_t)b->next_row * 8 + row) * (size_t)b->atlas_size + (size_t)b->next_col * 8,
                   data + row * 8, 8);
    } else {
        b->overflow = 1;  /* the reference indexes past its atlas and panics */
    }
    b->next_col++;
    if (b->next_col == slots) {
        b->next_col = 0;
        b->next_row++;
    }
}

/* Builder::span (svg.rs:80-87) */
static void on_span(void* user, int16_t x, int16_t y, uint16_t width) {
    builder* b = (builder*)user;
    push_quad(b, x, y, (int16_t)(x + (int16_t)width), (int16_t)(y + 8), 0, 0, 0, 0);
}

int orc_atlas_batch(int64_t n_paths, const uint8_t* verbs, const int64_t* verb_off, const float* points,
                    const int64_t* point_off, const float* weights, const int64_t* weight_off, const float* xforms,
                    const float* widths, const uint8_t* stroke_flags, const uint8_t* colors, int atlas_size,
                    uint8_t* atlas, void* verts,
                    size_t vcap, uint32_t* idx, size_t icap, size_t* n_verts, size_t* n_idx) {
    builder b;
    memset(&b, 0, sizeof(b));
    b.atlas = atlas;
    b.atlas_size = atlas_size;
    b.verts = (orc_vertex*)verts;
    b.vcap = vcap;
    b.idx = idx;
    b.icap = icap;
    b.next_row = 0;
    b.next_col = 1;
    /* Builder::new (svg.rs:33-48): a cleared atlas whose slot (0, 0) is solid */
    memset(atlas, 0, (size_t)atlas_size * (size_t)atlas_size);
    for (int row = 0; row < 8; row++) memset(atlas + (size_t)row * (size_t)atlas_size, 255, 8);
    int st = 0;
    for (int64_t p = 0; p < n_paths; p++) {
        memcpy(b.color, colors + 4 * p, 4);
        orc_rasterizer* r = orc_new();
        const uint8_t* v = verbs + verb_off[p];
        size_t nv = (size_t)(verb_off[p + 1] - verb_off[p]);
        const float* pts = points + 2 * point_off[p];
        size_t np = (size_t)(point_off[p + 1] - point_off[p]);
        const float* w = weights ? weights + weight_off[p] : NULL;
        size_t nw = (size_t)(weight_off[p + 1] - weight_off[p]);
        if (widths && (!stroke_flags || stroke_flags[p]))
            st |= orc_stroke(r, v, nv, pts, np, w, nw, widths[p], xforms + 6 * p);
        else st |= orc_fill(r, v, nv, pts, np, w, nw, xforms + 6 * p);
        st |= orc_finish(r, &b, on_tile, on_span);
        orc_free(r);
    }
    *n_verts = b.nv;
    *n_idx = b.ni;
    if (b.overflow) st |= 0x100;
    return st;
}
