// Stable LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits, for
// wave64.  Each pass: per-block digit counts -> device scan (digit-major) ->
// scatter with stable block-local ranks computed by per-wave ballot matching.
// Stability is what makes equal tile keys keep the emission order that the
// per-pixel f32 sums depend on.
#pragma once

#include "scan.hpp"

namespace ochre {

constexpr int RS_NT = 256;                 // threads per block
constexpr int RS_ITEMS = 16;               // keys per thread
constexpr int RS_TILE = RS_NT * RS_ITEMS;  // 4096 keys per block
constexpr int RS_WAVES = RS_NT / WAVE;
constexpr int RS_WAVE_ITEMS = RS_TILE / RS_WAVES;  // 1024

__device__ inline uint64_t lanemask_lt() {
    const int lane = lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// counts[digit * nblocks + block]
__global__ __launch_bounds__(RS_NT) void k_rs_count(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                    uint32_t* __restrict__ counts, uint32_t nblocks) {
    __shared__ uint32_t hist[RS_WAVES][256];
    const int tid = threadIdx.x;
    const int wid = tid / WAVE;
    for (int i = tid; i < RS_WAVES * 256; i += RS_NT) (&hist[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        int64_t i = base + (int64_t)j * RS_NT + tid;
        if (i < n) {
            uint32_t d = (keys[i] >> shift) & 255u;
            atomicAdd(&hist[wid][d], 1u);
        }
    }
    __syncthreads();
    for (int d = tid; d < 256; d += RS_NT) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) s += hist[w][d];
        counts[(int64_t)d * nblocks + blockIdx.x] = s;
    }
}

// Scatter.  offsets[digit * nblocks + block] = exclusive global start.
// vals_in == nullptr means the identity permutation (value = index).
// The block's keys are first placed in LDS in digit order (block-local stable
// positions: digit start + the wave's offset in the digit + the rank within
// the wave), then written out in that order, so that consecutive threads
// store consecutive addresses of one digit's run; the values follow through
// the same positions.
static_assert(RS_NT == 256, "the scatter takes one digit per thread");

struct RsScatterLds {
    uint32_t cnt[RS_WAVES][256];  // per wave and digit: count, then the wave's start within the digit
    uint32_t dstart[256];         // the digit's first position in the block
    uint32_t goff[256];           // the digit's global start for this block
    uint32_t buf[RS_TILE];        // keys, then values, in digit order
    uint8_t dig[RS_TILE];         // the digit at each position
};

__global__ __launch_bounds__(RS_NT, 4) void k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                      const uint32_t* __restrict__ vals_in, int64_t n, int shift,
                                                      const uint32_t* __restrict__ offsets, uint32_t nblocks,
                                                      uint32_t* __restrict__ keys_out,
                                                      uint32_t* __restrict__ vals_out) {
    __shared__ RsScatterLds s;
    const int tid = threadIdx.x;
    const int wid = tid / WAVE;
    const int lane = lane_id();
    for (int i = tid; i < RS_WAVES * 256; i += RS_NT) (&s.cnt[0][0])[i] = 0;
    for (int d = tid; d < 256; d += RS_NT) s.goff[d] = offsets[(int64_t)d * nblocks + blockIdx.x];
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    // Items of wave w: base + w*1024 + j*64 + lane, so (j, lane) order is index order.
    const int64_t bbase = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = bbase + (int64_t)wid * RS_WAVE_ITEMS;
    const int nb = (int)(n - bbase < (int64_t)RS_TILE ? n - bbase : (int64_t)RS_TILE);  // items in this block
    uint32_t key[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        int64_t i = wbase + (int64_t)j * WAVE + lane;
        bool valid = i < n;
        uint32_t k = valid ? keys_in[i] : 0u;
        key[j] = k;
        uint32_t d = (k >> shift) & 255u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t prior = valid ? s.cnt[wid][d] : 0u;
        rank[j] = prior + (uint32_t)__popcll(match & lt);
        // The lowest lane of each match group advances the wave-private counter.
        if (valid && (match & lt) == 0) s.cnt[wid][d] = prior + (uint32_t)__popcll(match);
    }
    __syncthreads();
    // per digit: the waves' starts within the digit, and the digit's total
    uint32_t tot = 0;
    {
        const int d = tid;  // RS_NT == 256: one thread per digit
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            const uint32_t c = s.cnt[w][d];
            s.cnt[w][d] = tot;
            tot += c;
        }
    }
    // the digits' starts in the block: a block scan of the totals
    {
        const uint32_t incl = wave_incl_scan(tot, OpAdd());
        if (lane == WAVE - 1) s.dstart[wid] = incl;  // wave totals (dstart reused below)
        __syncthreads();
        uint32_t before = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++)
            if (w < wid) before += s.dstart[w];
        __syncthreads();
        s.dstart[tid] = before + incl - tot;
    }
    __syncthreads();
    // rank[j] becomes the item's position in the block
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        int64_t i = wbase + (int64_t)j * WAVE + lane;
        if (i < n) {
            const uint32_t d = (key[j] >> shift) & 255u;
            rank[j] += s.dstart[d] + s.cnt[wid][d];
            s.buf[rank[j]] = key[j];
            s.dig[rank[j]] = (uint8_t)d;
        }
    }
    __syncthreads();
    for (int i = tid; i < nb; i += RS_NT) {
        const uint32_t d = s.dig[i];
        keys_out[s.goff[d] + (uint32_t)i - s.dstart[d]] = s.buf[i];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        int64_t i = wbase + (int64_t)j * WAVE + lane;
        if (i < n) s.buf[rank[j]] = vals_in ? vals_in[i] : (uint32_t)i;
    }
    __syncthreads();
    for (int i = tid; i < nb; i += RS_NT) {
        const uint32_t d = s.dig[i];
        vals_out[s.goff[d] + (uint32_t)i - s.dstart[d]] = s.buf[i];
    }
}

}  // namespace ochre
