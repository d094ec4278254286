// Device-wide single-pass exclusive scan (decoupled look-back) and wave/block
// scan helpers for wave64.
//
// Look-back state is tagged with an epoch read from device memory at kernel
// start; the last block to finish bumps the epoch and resets the block
// dispenser, so the state never needs clearing (no memset nodes in captured
// graphs, no host round trip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

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

// Look-back scratch for one scan call site.  Sized for max_blocks blocks.
template <typename T>
struct ScanState {
    uint32_t* flags;     // per block: (epoch << 2) | status, status 1 = aggregate, 2 = inclusive
    T* aggregate;
    T* inclusive;
    uint32_t* ctl;       // [0] block dispenser, [1] finished blocks, [2] epoch
};

template <typename T>
__device__ inline void store_release(uint32_t* flag, uint32_t v) {
    __threadfence();
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

// Exclusive scan: store(i, exclusive_prefix_i, value_i) for i in [0, n).
// Load: T load(int64_t i) for i < n.  The final total is written to *total
// if total != nullptr.
template <typename T, typename Op, typename Load, typename Store>
__global__ __launch_bounds__(SCAN_NT) void k_scan(Load load, Store store, int64_t n, ScanState<T> st,
                                                  Op op, T identity, T* total) {
    __shared__ uint32_t s_bid, s_epoch;
    __shared__ T s_red[SCAN_NT / WAVE];
    __shared__ T s_prefix;
    __shared__ int s_last;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_bid = atomicAdd(&st.ctl[0], 1u);
        s_epoch = __hip_atomic_load(&st.ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t bid = s_bid;
    const uint32_t epoch = s_epoch;
    const int64_t base = (int64_t)bid * SCAN_TILE + (int64_t)tid * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T acc = identity;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        int64_t i = base + j;
        v[j] = i < n ? load(i) : identity;
        acc = op(acc, v[j]);
    }
    T block_total;
    T excl = block_excl_scan<SCAN_NT>(acc, op, identity, s_red, &block_total);
    if (tid == 0) {
        T prefix = identity;
        const uint32_t tag = epoch << 2;
        if (bid == 0) {
            st.inclusive[0] = block_total;
            store_release<T>(&st.flags[0], tag | 2u);
        } else {
            st.aggregate[bid] = block_total;
            store_release<T>(&st.flags[bid], tag | 1u);
            int64_t p = (int64_t)bid - 1;
            for (;;) {
                uint32_t f = __hip_atomic_load(&st.flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((f & ~3u) != tag || (f & 3u) == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                __threadfence();
                if ((f & 3u) == 2u) {
                    prefix = op(st.inclusive[p], prefix);
                    break;
                }
                prefix = op(st.aggregate[p], prefix);
                p--;
            }
            st.inclusive[bid] = op(prefix, block_total);
            store_release<T>(&st.flags[bid], tag | 2u);
        }
        s_prefix = prefix;
    }
    __syncthreads();
    T r = op(s_prefix, excl);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        int64_t i = base + j;
        if (i < n) store(i, r, v[j]);
        r = op(r, v[j]);
    }
    const uint32_t nblocks = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    if (total != nullptr && bid == nblocks - 1 && tid == SCAN_NT - 1) *total = r;
    // The last block to finish retires this epoch.
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        uint32_t done = atomicAdd(&st.ctl[1], 1u);
        s_last = (done == nblocks - 1);
    }
    __syncthreads();
    if (s_last && tid == 0) {
        st.ctl[0] = 0;
        st.ctl[1] = 0;
        __hip_atomic_store(&st.ctl[2], epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
    }
}

struct OpAdd {
    template <typename T>
    __device__ T operator()(T a, T b) const { return a + b; }
};

template <typename T>
inline uint32_t scan_blocks(int64_t n) {
    return (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
}

}  // namespace ochre
