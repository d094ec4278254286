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
__global__ __launch_bounds__(RS_NT) void k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                      const uint32_t* __restrict__ vals_in, int64_t n, int shift,
                                                      const uint32_t* __restrict__ offsets, uint32_t nblocks,
                                                      uint32_t* __restrict__ keys_out,
                                                      uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t cnt[RS_WAVES][256];
    __shared__ uint32_t wave_off[RS_WAVES][256];
    const int tid = threadIdx.x;
    const int wid = tid / WAVE;
    const int lane = lane_id();
    for (int i = tid; i < RS_WAVES * 256; i += RS_NT) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    // Items of wave w: base + w*1024 + j*64 + lane, so (j, lane) order is index order.
    const int64_t wbase = (int64_t)blockIdx.x * RS_TILE + (int64_t)wid * RS_WAVE_ITEMS;
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
        uint32_t prior = valid ? cnt[wid][d] : 0u;
        rank[j] = prior + (uint32_t)__popcll(match & lt);
        // The lowest lane of each match group advances the wave-private counter.
        if (valid && (match & lt) == 0) cnt[wid][d] = prior + (uint32_t)__popcll(match);
    }
    __syncthreads();
    for (int d = tid; d < 256; d += RS_NT) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            wave_off[w][d] = s;
            s += cnt[w][d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        int64_t i = wbase + (int64_t)j * WAVE + lane;
        if (i < n) {
            uint32_t d = (key[j] >> shift) & 255u;
            uint32_t pos = offsets[(int64_t)d * nblocks + blockIdx.x] + wave_off[wid][d] + rank[j];
            keys_out[pos] = key[j];
            vals_out[pos] = vals_in ? vals_in[i] : (uint32_t)i;
        }
    }
}

}  // namespace ochre
