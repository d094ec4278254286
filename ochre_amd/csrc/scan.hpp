// Wave/block scan helpers for wave64 and a device-wide exclusive scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <type_traits>

namespace ochre {

constexpr int WAVE = 64;

__device__ inline int lane_id() { return (int)__lane_id(); }

template <typename T>
__device__ inline T shfl_up_any(T v, int d) {
    static_assert(sizeof(T) % 4 == 0, "scan type must be a multiple of 4 bytes");
    constexpr int N = sizeof(T) / 4;
    uint32_t w[N];
    memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < N; i++) w[i] = __shfl_up(w[i], d, WAVE);
    T r;
    memcpy(&r, w, sizeof(T));
    return r;
}

template <typename T>
__device__ inline T shfl_any(T v, int src) {
    constexpr int N = sizeof(T) / 4;
    uint32_t w[N];
    memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < N; i++) w[i] = __shfl(w[i], src, WAVE);
    T r;
    memcpy(&r, w, sizeof(T));
    return r;
}

// Inclusive scan across the 64 lanes of a wave (all lanes must be active).
template <typename T, typename Op>
__device__ inline T wave_incl_scan(T v, Op op) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T o = shfl_up_any(v, d);
        if (lane >= d) v = op(o, v);
    }
    return v;
}

// Exclusive block scan for blockDim.x == NT (multiple of 64).  Returns the
// exclusive prefix of this thread; *total receives the block aggregate.
// `smem` must hold NT/64 values of T.  Contains __syncthreads().
template <int NT, typename T, typename Op>
__device__ inline T block_excl_scan(T v, Op op, T identity, T* smem, T* total) {
    constexpr int NW = NT / WAVE;
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    T incl = wave_incl_scan(v, op);
    if (lane == WAVE - 1) smem[wid] = incl;
    __syncthreads();
    T wave_prefix = identity;
    T tot = identity;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        T s = smem[w];
        if (w < wid) wave_prefix = op(wave_prefix, s);
        tot = op(tot, s);
    }
    __syncthreads();
    T excl = shfl_up_any(incl, 1);
    if (lane == 0) excl = identity;
    *total = tot;
    return op(wave_prefix, excl);
}

// Device-wide exclusive scan as reduce-then-scan: (1) every block reduces its
// tile, (2) one block scans the tile sums, (3) every block rescans its tile
// from its prefix.  All three launches are fully parallel; on a multi-XCD
// part a single-pass look-back chain would poll flags across dies (one
// Infinity Fabric round trip per link of the chain).
constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

// A store functor with a `chunk_tag` type member takes a thread's whole run
// of SCAN_ITEMS consecutive items at once (chunk(base, n, excl, v)), so that
// it can aggregate across them; others take one item at a time.
template <typename S, typename = void>
struct IsChunkStore : std::false_type {};
template <typename S>
struct IsChunkStore<S, std::void_t<typename S::chunk_tag>> : std::true_type {};

template <typename T, typename Op, typename Store>
__device__ inline void tile_store(const Store& store, int64_t base, int64_t n, T r, const T* v, Op op) {
    if constexpr (IsChunkStore<Store>::value) {
        store.chunk(base, n, r, v);
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) {
            int64_t i = base + j;
            if (i < n) store(i, r, v[j]);
            r = op(r, v[j]);
        }
    }
}

template <typename T, typename Op, typename Load>
__device__ inline T tile_load_reduce(Load load, int64_t base, int64_t n, Op op, T identity, T* v) {
    T acc = identity;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        int64_t i = base + j;
        v[j] = i < n ? load(i) : identity;
        acc = op(acc, v[j]);
    }
    return acc;
}

template <typename T, typename Op, typename Load>
__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(Load load, int64_t n, T* partials, Op op, T identity) {
    __shared__ T s_red[SCAN_NT / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T acc = tile_load_reduce<T>(load, base, n, op, identity, v);
    T total;
    (void)block_excl_scan<SCAN_NT>(acc, op, identity, s_red, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// One block: exclusive scan of partials[0, nb) in place; *total = sum.
template <typename T, typename Op>
__global__ __launch_bounds__(SCAN_NT) void k_scan_partials(T* partials, int64_t nb, Op op, T identity, T* total) {
    __shared__ T s_red[SCAN_NT / WAVE];
    T carry = identity;
    for (int64_t b0 = 0; b0 < nb; b0 += SCAN_TILE) {
        const int64_t base = b0 + (int64_t)threadIdx.x * SCAN_ITEMS;
        T v[SCAN_ITEMS];
        T acc = identity;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) {
            int64_t i = base + j;
            v[j] = i < nb ? partials[i] : identity;
            acc = op(acc, v[j]);
        }
        T chunk;
        T r = op(carry, block_excl_scan<SCAN_NT>(acc, op, identity, s_red, &chunk));
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) {
            int64_t i = base + j;
            if (i < nb) partials[i] = r;
            r = op(r, v[j]);
        }
        carry = op(carry, chunk);
    }
    if (threadIdx.x == 0 && total != nullptr) *total = carry;
}

template <typename T, typename Op, typename Load, typename Store>
__global__ __launch_bounds__(SCAN_NT) void k_scan_down(Load load, Store store, int64_t n, const T* partials, Op op,
                                                       T identity) {
    __shared__ T s_red[SCAN_NT / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T acc = tile_load_reduce<T>(load, base, n, op, identity, v);
    T total;
    T r = op(partials ? partials[blockIdx.x] : identity, block_excl_scan<SCAN_NT>(acc, op, identity, s_red, &total));
    tile_store(store, base, n, r, v, op);
}

// Single-tile scan that also writes the total.
template <typename T, typename Op, typename Load, typename Store>
__global__ __launch_bounds__(SCAN_NT) void k_scan_single(Load load, Store store, int64_t n, Op op, T identity,
                                                         T* total_out) {
    __shared__ T s_red[SCAN_NT / WAVE];
    const int64_t base = (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T acc = tile_load_reduce<T>(load, base, n, op, identity, v);
    T total;
    T r = block_excl_scan<SCAN_NT>(acc, op, identity, s_red, &total);
    tile_store(store, base, n, r, v, op);
    if (threadIdx.x == 0 && total_out != nullptr) *total_out = total;
}

struct OpAdd {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a + b; }
};


}  // namespace ochre
