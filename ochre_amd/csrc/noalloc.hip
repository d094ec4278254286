// The batched call for callers that cannot allocate device memory: code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc.
//
// It runs the dense kernels alone (the 64-tile instance over every path, then
// the 256-tile instance over the paths it hands on; dense.inc) with their
// control block (queues, counters, error bits) and the list of handed-on
// paths in `scratch`, a fixed file-scope device array of NOALLOC_SCRATCH
// bytes (256 B + 4 MiB) that each device holds until the process ends.  Its
// address on a device is looked up once, when a context is created on that
// device (noalloc_init_device), and kept per device; host code passes that
// address, never the array's name, to the kernels and to the copies.
//
// Every call on a device shares that device's array: the control block is
// cleared by a kernel at the start of each call (not by a memset), so a call
// starts from zero whatever the last call left, and calls that overlap on one
// device race.  A batch over NOALLOC_MAX_PATHS paths is refused with
// OCHRE_ERR_TOO_LARGE before anything runs; a path over the 256-tile instance
// (a tile grid over 256 tiles or 32 on a side, more long lines than its queue
// holds, a conic weight that is not positive and finite) makes the call
// return OCHRE_ERR_TOO_LARGE after the other paths are written (the batched
// call handles those paths).
#include <hip/hip_runtime.h>

#include <mutex>

#include "pipeline.hpp"

namespace ochre {

constexpr int64_t NOALLOC_MAX_PATHS = (int64_t)1 << 20;
constexpr size_t NOALLOC_SCRATCH = DENSE_CTL_BYTES + (size_t)NOALLOC_MAX_PATHS * 4;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t* g_ptr[NOALLOC_MAX_DEVICES];  // the array's address on each device
}  // namespace

// Looks up the array's address on `dev`, which must be the current device.
// Called when a context is created on the device; idempotent.
int noalloc_init_device(int dev) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return -1;
    if (g_ptr[dev]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return -3;
    g_ptr[dev] = reinterpret_cast<uint8_t*>(p);
    return 0;
}

int run_batched_noalloc(const SegInput& in, OutputDesc& out, hipStream_t s) {
    out.n_tiles = out.n_spans = 0;
    if (in.n_paths <= 0) return 0;
    if (in.n_paths > NOALLOC_MAX_PATHS) return -7;
    int dev = 0;
    HIPCHK_RET(hipGetDevice(&dev));
    uint8_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev >= 0 && dev < NOALLOC_MAX_DEVICES) base = g_ptr[dev];
    }
    if (!base) return -3;  // no context was created on this device
    uint8_t* ctl = base;
    uint32_t* ids = reinterpret_cast<uint32_t*>(base + DENSE_CTL_BYTES);
    int st = launch_dense_chain(in, out, ctl, ids, s);
    if (st) return st;
    return finish_dense_chain(ctl, out, s);
}

}  // namespace ochre
