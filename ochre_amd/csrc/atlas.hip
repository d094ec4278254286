// Device-side tile consumer: the demo's Builder (examples/svg.rs:15-88) run
// on the rasteriser's device output, so tiles never cross PCIe and no
// per-tile callback runs.  The reference feeds one Builder from one
// Rasterizer per path, in path order (svg.rs:140-156); every TileBuilder call
// appends one quad (4 vertices, 6 indices) and every tile also takes the next
// 8x8 atlas slot (slot 0 is the solid one).  All of that is a function of a
// call's position in the global call sequence, which here is computed
// instead of replayed:
//   tile g (in path order)  -> atlas slot g + 1
//   call k (tile or span)    -> vertices 4k .. 4k + 3, indices 6k .. 6k + 5
// and a path's calls are its tiles and spans merged by (y, x): finish emits
// tiles in row-major order with each span right after the tile on its left
// (rasterizer.rs:241,263), so a tile's call index within its path is its
// tile index plus the number of the path's spans before it in (y, x) order,
// and symmetrically for a span.  One thread per tile and one per span.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ochre.h"
#include "pipeline.hpp"
#include "scan.hpp"

namespace ochre {

namespace {

struct AtlasIn {
    int64_t n_paths;
    const int16_t* txy;
    const uint8_t* ta;
    const int16_t* sxy;
    const uint16_t* sw;
    const int64_t* pts;
    const int32_t* ptc;
    const int64_t* pss;
    const int32_t* psc;
    const uint8_t* colors;    // [n_paths][4]
    const int64_t* tile_base; // [n_paths + 1]: tiles of earlier paths
    const int64_t* span_base; // [n_paths + 1]
};

struct AtlasOut {
    uint8_t* atlas;
    int32_t atlas_size;
    ochre_atlas_vertex* verts;
    uint32_t* idx;
};

// Last path p with base[p] <= g (base is non-decreasing, base[0] = 0).
__device__ inline int64_t owner(const int64_t* base, int64_t n, int64_t g) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Number of entries of xy[0 .. n) (row-major sorted) before (y, x).
__device__ inline int64_t count_before(const int16_t* xy, int64_t n, int16_t y, int16_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int16_t my = xy[2 * mid + 1], mx = xy[2 * mid];
        if (my < y || (my == y && mx < x)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ inline void put_quad(const AtlasOut& o, int64_t call, int16_t x0, int16_t y0, int16_t x1, int16_t y1,
                                uint16_t u1, uint16_t v1, uint16_t u2, uint16_t v2, uint32_t col) {
    ochre_atlas_vertex* v = o.verts + 4 * call;
    const uint32_t c = col;
    ochre_atlas_vertex q[4];
    q[0].pos[0] = x0; q[0].pos[1] = y0; q[0].uv[0] = u1; q[0].uv[1] = v1;
    q[1].pos[0] = x1; q[1].pos[1] = y0; q[1].uv[0] = u2; q[1].uv[1] = v1;
    q[2].pos[0] = x1; q[2].pos[1] = y1; q[2].uv[0] = u2; q[2].uv[1] = v2;
    q[3].pos[0] = x0; q[3].pos[1] = y1; q[3].uv[0] = u1; q[3].uv[1] = v2;
    for (int k = 0; k < 4; k++) {
        __builtin_memcpy(q[k].col, &c, 4);
        v[k] = q[k];
    }
    const uint32_t b = (uint32_t)(4 * call);
    uint32_t* ix = o.idx + 6 * call;
    ix[0] = b;
    ix[1] = b + 1;
    ix[2] = b + 2;
    ix[3] = b;
    ix[4] = b + 2;
    ix[5] = b + 3;
}

__global__ void k_atlas_tiles(AtlasIn in, AtlasOut o, int64_t n_tiles) {
    const int32_t slots = o.atlas_size / 8;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n_tiles;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = owner(in.tile_base, in.n_paths, g);
        const int64_t i = g - in.tile_base[p];
        const int64_t t = in.pts[p] + i;
        const int16_t x = in.txy[2 * t], y = in.txy[2 * t + 1];
        const int64_t call = in.tile_base[p] + in.span_base[p] + i +
                             count_before(in.sxy + 2 * in.pss[p], in.psc[p], y, x);
        const int64_t slot = g + 1;
        const uint16_t col = (uint16_t)(slot % slots), row = (uint16_t)(slot / slots);
        uint32_t c;
        __builtin_memcpy(&c, in.colors + 4 * p, 4);
        put_quad(o, call, x, y, (int16_t)(x + 8), (int16_t)(y + 8), (uint16_t)(col * 8), (uint16_t)(row * 8),
                 (uint16_t)((col + 1) * 8), (uint16_t)((row + 1) * 8), c);
        const uint2* src = reinterpret_cast<const uint2*>(in.ta + t * 64);
        uint8_t* dst = o.atlas + (size_t)row * 8 * (size_t)o.atlas_size + (size_t)col * 8;
#pragma unroll
        for (int r = 0; r < 8; r++) *reinterpret_cast<uint2*>(dst + (size_t)r * (size_t)o.atlas_size) = src[r];
    }
}

__global__ void k_atlas_spans(AtlasIn in, AtlasOut o, int64_t n_spans) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < n_spans;
         h += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = owner(in.span_base, in.n_paths, h);
        const int64_t j = h - in.span_base[p];
        const int64_t s = in.pss[p] + j;
        const int16_t x = in.sxy[2 * s], y = in.sxy[2 * s + 1];
        const uint16_t w = in.sw[s];
        const int64_t call = in.tile_base[p] + in.span_base[p] + j +
                             count_before(in.txy + 2 * in.pts[p], in.ptc[p], y, x);
        uint32_t c;
        __builtin_memcpy(&c, in.colors + 4 * p, 4);
        put_quad(o, call, x, y, (int16_t)(x + (int16_t)w), (int16_t)(y + 8), 0, 0, 0, 0, c);
    }
}

// Builder::new (svg.rs:33-48) for the slots the batch leaves unused: slot 0
// solid, slots n_tiles + 1 .. zero.  One thread per (slot, pixel row).
__global__ void k_atlas_clear(AtlasOut o, int64_t n_tiles) {
    const int64_t slots = o.atlas_size / 8;
    const int64_t total = slots * slots;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total * 8;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t slot = k >> 3, r = k & 7;
        if (slot != 0 && slot <= n_tiles) continue;
        const uint32_t v = slot == 0 ? 0xffffffffu : 0u;
        uint8_t* dst = o.atlas + ((slot / slots) * 8 + r) * (size_t)o.atlas_size + (slot % slots) * 8;
        *reinterpret_cast<uint2*>(dst) = make_uint2(v, v);
    }
}

struct LoadI32 {
    const int32_t* p;
    int64_t n;
    __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct StoreOff {
    int64_t* p;
    __device__ void operator()(int64_t i, int64_t excl, int64_t) const { p[i] = excl; }
};

hipError_t count_offsets(Context* ctx, const int32_t* cnt, int64_t* off, int64_t n, hipStream_t s) {
    if (n <= SCAN_TILE) {
        hipLaunchKernelGGL((k_scan_single<int64_t, OpAdd, LoadI32, StoreOff>), dim3(1), dim3(SCAN_NT), 0, s,
                           LoadI32{cnt, n}, StoreOff{off}, n, OpAdd(), (int64_t)0, off + n);
        return hipGetLastError();
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipError_t e = ctx->scan_partials.ensure((size_t)nb * 16);
    if (e != hipSuccess) return e;
    int64_t* partials = ctx->scan_partials.as<int64_t>();
    hipLaunchKernelGGL((k_scan_reduce<int64_t, OpAdd, LoadI32>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s,
                       LoadI32{cnt, n}, n, partials, OpAdd(), (int64_t)0);
    hipLaunchKernelGGL((k_scan_partials<int64_t, OpAdd>), dim3(1), dim3(SCAN_NT), 0, s, partials, nb, OpAdd(),
                       (int64_t)0, off + n);
    hipLaunchKernelGGL((k_scan_down<int64_t, OpAdd, LoadI32, StoreOff>), dim3((uint32_t)nb), dim3(SCAN_NT), 0, s,
                       LoadI32{cnt, n}, StoreOff{off}, n, (const int64_t*)partials, OpAdd(), (int64_t)0);
    return hipGetLastError();
}

uint32_t grid_for(int64_t n) { return (uint32_t)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

}  // namespace

int run_atlas(Context* ctx, const OutputDesc& in, int64_t n_paths, const uint8_t* colors, uint8_t* atlas,
              int atlas_size, ochre_atlas_vertex* verts, int64_t vcap, uint32_t* idx, int64_t icap, int64_t* n_verts,
              int64_t* n_idx, hipStream_t s) {
    *n_verts = *n_idx = 0;
    if (atlas_size < 8 || atlas_size % 8 != 0 || atlas_size > 65536 || n_paths < 0) return -1;
    if (n_paths > 0 && (!colors || !in.path_tile_start || !in.path_tile_count || !in.path_span_start ||
                        !in.path_span_count))
        return -1;
    AtlasState& A = ctx->atlas;
    HIPCHK_RET(A.tile_base.ensure((size_t)(n_paths + 1) * 8));
    HIPCHK_RET(A.span_base.ensure((size_t)(n_paths + 1) * 8));
    if (!A.host) HIPCHK_RET(hipHostMalloc(&A.host, 64));
    int64_t* h = reinterpret_cast<int64_t*>(A.host);
    int64_t* tb = A.tile_base.as<int64_t>();
    int64_t* sb = A.span_base.as<int64_t>();
    int64_t T = 0, S = 0;
    if (n_paths > 0) {
        HIPCHK_RET(count_offsets(ctx, in.path_tile_count, tb, n_paths, s));
        HIPCHK_RET(count_offsets(ctx, in.path_span_count, sb, n_paths, s));
        HIPCHK_RET(hipMemcpyAsync(h, tb + n_paths, 8, hipMemcpyDeviceToHost, s));
        HIPCHK_RET(hipMemcpyAsync(h + 1, sb + n_paths, 8, hipMemcpyDeviceToHost, s));
        HIPCHK_RET(hipStreamSynchronize(s));
        T = h[0];
        S = h[1];
    }
    *n_verts = 4 * (T + S);
    *n_idx = 6 * (T + S);
    const int64_t slots = (int64_t)(atlas_size / 8) * (atlas_size / 8);
    if (T + 1 > slots || *n_verts > vcap || *n_idx > icap || *n_verts > (int64_t)UINT32_MAX) return -4;
    if (!atlas || (T + S > 0 && (!verts || !idx))) return -1;
    AtlasIn ai{n_paths, in.tile_xy, in.tile_alpha, in.span_xy, in.span_w, in.path_tile_start, in.path_tile_count,
               in.path_span_start, in.path_span_count, colors, tb, sb};
    AtlasOut ao{atlas, atlas_size, verts, idx};
    hipLaunchKernelGGL(k_atlas_clear, dim3(grid_for(slots * 8)), dim3(256), 0, s, ao, T);
    if (T > 0) hipLaunchKernelGGL(k_atlas_tiles, dim3(grid_for(T)), dim3(256), 0, s, ai, ao, T);
    if (S > 0) hipLaunchKernelGGL(k_atlas_spans, dim3(grid_for(S)), dim3(256), 0, s, ai, ao, S);
    HIPCHK_RET(hipGetLastError());
    return 0;
}

}  // namespace ochre
