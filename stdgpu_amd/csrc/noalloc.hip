// device_range() for callers that cannot allocate: code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc.  The tile counts of the range kernels live in a file-scope
// device array, one copy per GPU (each GPU holds it until the process ends), instead of the table's
// scratch, and the entry count is written to caller device memory instead of being copied to the
// host, so a call allocates nothing and never waits for the stream.
//
// The array's host-side name is not a device address: its address on each device is looked up once
// with hipGetSymbolAddress, with that device current, on the first call there, and only that address
// is handed to kernels.  Every call on one GPU shares the array; calls that overlap on a GPU (two
// streams, two threads) race on it.  Nothing in it needs to start at zero: each call's count kernel
// writes every tile count it later reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "htab_internal.hpp"

using namespace htab_host;

// tile counts of one range (one per kRangeTile = 4096 slots) plus the grand total: 256 KiB per GPU
static __device__ __align__(256) int64_t scratch[HTAB_NOALLOC_RANGE_TILES + 1];

namespace
{

constexpr int kMaxDevices = 64;

std::mutex g_mu;
int64_t* g_scratch[kMaxDevices] = {};

int
scratch_of_current_device(int64_t** out)
{
    int dev = 0;
    HTAB_HIP_TRY(hipGetDevice(&dev), "htab_range_noalloc: current device");
    if (dev < 0 || dev >= kMaxDevices) return fail(HTAB_E_NODEV, "htab_range_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev])
    {
        void* p = nullptr;
        HTAB_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "htab_range_noalloc: scratch address");
        if (!p) return fail(HTAB_E_HIP, "htab_range_noalloc: scratch address");
        g_scratch[dev] = static_cast<int64_t*>(p);
    }
    *out = g_scratch[dev];
    return HTAB_OK;
}

} // namespace

extern "C" {

int
htab_range_noalloc(htab* h, int64_t* idx_out, int64_t cap, int64_t* n_out_dev, void* stream)
{
    if (!h || !n_out_dev) return fail(HTAB_E_ARG, "htab_range_noalloc: null argument");
    if (cap < 0) return fail(HTAB_E_ARG, "htab_range_noalloc: negative capacity");
    if (cap > 0 && !idx_out) return fail(HTAB_E_ARG, "htab_range_noalloc: null index array");
    const int64_t total = h->info.physical_slots + 2;
    const int64_t tiles = (total + static_cast<int64_t>(htab_dev::kRangeTile) - 1) / static_cast<int64_t>(htab_dev::kRangeTile);
    if (tiles > HTAB_NOALLOC_RANGE_TILES)
        return fail(HTAB_E_LIMIT, "htab_range_noalloc: table has more slots than HTAB_NOALLOC_RANGE_TILES * 4096");
    int64_t* sc = nullptr;
    const int rc = scratch_of_current_device(&sc);
    if (rc != HTAB_OK) return rc;
    htab_call c;
    c.op = call_op::range;
    c.h = h;
    c.idx_out = idx_out;
    c.cap = cap;
    c.scratch_ext = sc;
    c.count_dev = n_out_dev;
    c.s = static_cast<hipStream_t>(stream);
    return dispatch(h->kind, c);
}

} // extern "C"
