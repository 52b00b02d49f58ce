// Free-list deque<uint32_t> with bulk push_back / pop_front.
//
// Replaces the reference's per-element path (src/stdgpu/impl/deque_detail.cuh:202-260 push_back,
// 390-448 pop_front: atomic size +-1, atomicInc of begin/end modulo capacity, then a spin on a
// per-slot lock bit and an occupancy-bit flip).  One bulk call only pushes or only pops, so it
// reserves all of its ring positions at once: a one-thread kernel reads the size and the end (push)
// or begin (pop) counter, takes min(n, room) positions and advances both, then a streaming kernel
// moves element i to or from ring position (counter + i) mod capacity.  No per-element atomic, lock
// or occupancy bit remains, and the result is the one of a sequential loop of calls.
//
// Each slot also keeps a sequence word (2t: free for turn t; 2t + 1: holds the element of ring
// position p, t = p / capacity), which the bulk kernels keep current so that user kernels can push
// and pop per thread between bulk calls (include/stdgpu_amd/deque_device.hpp).  size = end - begin.
#include <hip/hip_runtime.h>

#include <new>

#include <stdgpu_amd/htab.h>

namespace
{

// control block: int64 words, each count on a 128 B line of its own
constexpr int kReady = 0;     // elements published by their push and not yet reserved by a pop
constexpr int kBegin = 16;    // uint64 front counter
constexpr int kEnd = 32;      // uint64 back counter
constexpr int kResBase = 48;  // reservation of the current bulk call: first counter value
constexpr int kResCount = 49; // ... and the number of elements it covers
constexpr int kFree = 64;     // slots released by their pop and not yet reserved by a push
constexpr size_t kDequeCtlBytes = 1024;

__device__ inline uint32_t
seq_full(uint64_t p, uint64_t cap)
{
    return static_cast<uint32_t>(2 * (p / cap) + 1);
}

__device__ inline uint32_t
seq_free(uint64_t p, uint64_t cap)
{
    return static_cast<uint32_t>(2 * (p / cap));
}

// the clamped size end - begin (deque_detail.cuh:482-512), then the call's reservation
__global__ void
k_deque_reserve(int64_t* ctl, int64_t cap, int64_t n, int push)
{
    uint64_t* begin = reinterpret_cast<uint64_t*>(ctl + kBegin);
    uint64_t* end = reinterpret_cast<uint64_t*>(ctl + kEnd);
    int64_t s = static_cast<int64_t>(*end - *begin);
    s = s < 0 ? 0 : (s > cap ? cap : s);
    const int64_t room = push ? cap - s : s;
    const int64_t k = n < room ? n : room;
    uint64_t* counter = push ? end : begin;
    ctl[kResBase] = static_cast<int64_t>(*counter);
    ctl[kResCount] = k;
    *counter += static_cast<uint64_t>(k);
    ctl[kReady] += push ? k : -k;
    ctl[kFree] += push ? -k : k;
}

__global__ __launch_bounds__(256) void
k_deque_push(uint32_t* __restrict__ data, uint32_t* __restrict__ seq, const int64_t* __restrict__ ctl, int64_t cap,
             const uint32_t* __restrict__ v, int64_t n, uint8_t* __restrict__ pushed)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = i < ctl[kResCount];
    if (ok)
    {
        const uint64_t p = static_cast<uint64_t>(ctl[kResBase]) + static_cast<uint64_t>(i);
        const uint64_t slot = p % static_cast<uint64_t>(cap);
        data[slot] = v[i];
        seq[slot] = seq_full(p, static_cast<uint64_t>(cap));
    }
    if (pushed) pushed[i] = ok ? 1 : 0;
}

__global__ __launch_bounds__(256) void
k_deque_pop(const uint32_t* __restrict__ data, uint32_t* __restrict__ seq, const int64_t* __restrict__ ctl, int64_t cap,
            uint32_t* __restrict__ out, int64_t n, uint8_t* __restrict__ popped)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = i < ctl[kResCount];
    uint32_t v = 0xFFFFFFFFU;
    if (ok)
    {
        const uint64_t p = static_cast<uint64_t>(ctl[kResBase]) + static_cast<uint64_t>(i);
        const uint64_t slot = p % static_cast<uint64_t>(cap);
        v = data[slot];
        seq[slot] = seq_free(p + static_cast<uint64_t>(cap), static_cast<uint64_t>(cap));
    }
    out[i] = v;
    if (popped) popped[i] = ok ? 1 : 0;
}

// fill_iota: slot i holds value i as ring position i; clear: every slot free for turn 0.  Both reset the
// counters (begin = 0, end = capacity or 0).
__global__ __launch_bounds__(256) void
k_deque_reset(uint32_t* __restrict__ data, uint32_t* __restrict__ seq, int64_t* ctl, int64_t cap, int iota)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < static_cast<uint64_t>(cap);
         i += stride)
    {
        if (iota) data[i] = static_cast<uint32_t>(i);
        seq[i] = iota ? 1U : 0U;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        ctl[kBegin] = 0;
        ctl[kEnd] = iota ? cap : 0;
        ctl[kReady] = iota ? cap : 0;
        ctl[kFree] = iota ? 0 : cap;
    }
}

__global__ __launch_bounds__(256) void
k_deque_copy(const uint32_t* __restrict__ data, int64_t cap, uint64_t begin, int64_t n, uint32_t* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = data[(begin + static_cast<uint64_t>(i)) % static_cast<uint64_t>(cap)];
}

bool
aligned4(const void* p)
{
    return p == nullptr || reinterpret_cast<uintptr_t>(p) % 4 == 0;
}

int
dq_status()
{
    return hipGetLastError() == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

} // namespace

struct htab_deque
{
    uint32_t* data = nullptr;
    uint32_t* seq = nullptr;
    int64_t* ctl = nullptr;
    int64_t capacity = 0;
};

namespace
{

int
deque_reset(htab_deque* q, int iota, hipStream_t s)
{
    int64_t grid = (q->capacity + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_deque_reset, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->seq, q->ctl,
                       q->capacity, iota);
    return dq_status();
}

} // namespace

extern "C" {

int
htab_deque_create(int64_t capacity, htab_deque** out)
{
    if (!out || capacity <= 0 || capacity > (int64_t(1) << 32)) return HTAB_E_ARG;
    *out = nullptr;
    htab_deque* q = new (std::nothrow) htab_deque();
    if (!q) return HTAB_E_OOM;
    void* mem = nullptr;
    if (hipMalloc(&mem, kDequeCtlBytes + 2 * static_cast<size_t>(capacity) * sizeof(uint32_t)) != hipSuccess)
    {
        delete q;
        return HTAB_E_OOM;
    }
    q->ctl = static_cast<int64_t*>(mem);
    q->data = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + kDequeCtlBytes);
    q->seq = q->data + capacity;
    q->capacity = capacity;
    // empty, every slot free for turn 0 (the reference creates an empty deque, deque_detail.cuh:40-80)
    if (hipMemset(mem, 0, kDequeCtlBytes) != hipSuccess || deque_reset(q, 0, nullptr) != HTAB_OK ||
        hipStreamSynchronize(nullptr) != hipSuccess)
    {
        (void)hipFree(mem);
        delete q;
        return HTAB_E_HIP;
    }
    *out = q;
    return HTAB_OK;
}

int
htab_deque_destroy(htab_deque* q)
{
    if (!q) return HTAB_E_ARG;
    const hipError_t e = hipFree(q->ctl);
    delete q;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_deque_push_back(htab_deque* q, const uint32_t* values, int64_t n, uint8_t* pushed, void* stream)
{
    if (!q || n < 0 || (n > 0 && !values) || !aligned4(values)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_deque_reserve, dim3(1), dim3(1), 0, s, q->ctl, q->capacity, n, 1);
    hipLaunchKernelGGL(k_deque_push, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->seq, q->ctl,
                       q->capacity, values, n, pushed);
    return dq_status();
}

int
htab_deque_pop_front(htab_deque* q, uint32_t* out, int64_t n, uint8_t* popped, void* stream)
{
    if (!q || n < 0 || (n > 0 && !out) || !aligned4(out)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_deque_reserve, dim3(1), dim3(1), 0, s, q->ctl, q->capacity, n, 0);
    hipLaunchKernelGGL(k_deque_pop, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->seq, q->ctl,
                       q->capacity, out, n, popped);
    return dq_status();
}

int
htab_deque_fill_iota(htab_deque* q, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return deque_reset(q, 1, static_cast<hipStream_t>(stream));
}

int
htab_deque_clear(htab_deque* q, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return deque_reset(q, 0, static_cast<hipStream_t>(stream));
}

int
htab_deque_get_view(const htab_deque* q, htab_deque_view* out)
{
    if (!q || !out) return HTAB_E_ARG;
    out->data = q->data;
    out->seq = q->seq;
    out->begin = reinterpret_cast<uint64_t*>(q->ctl + kBegin);
    out->end = reinterpret_cast<uint64_t*>(q->ctl + kEnd);
    out->ready = q->ctl + kReady;
    out->free_slots = q->ctl + kFree;
    out->capacity = q->capacity;
    return HTAB_OK;
}

static int
deque_read_ctl(htab_deque* q, int64_t* size, uint64_t* begin, hipStream_t s)
{
    int64_t ctl[kEnd + 1];
    if (hipMemcpyAsync(ctl, q->ctl, sizeof(ctl), hipMemcpyDeviceToHost, s) != hipSuccess) return HTAB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    int64_t n = static_cast<int64_t>(static_cast<uint64_t>(ctl[kEnd]) - static_cast<uint64_t>(ctl[kBegin]));
    // the reference's size() clamps to [0, capacity] (deque_detail.cuh:482-512)
    n = n < 0 ? 0 : (n > q->capacity ? q->capacity : n);
    *size = n;
    *begin = static_cast<uint64_t>(ctl[kBegin]);
    return HTAB_OK;
}

int
htab_deque_size(htab_deque* q, int64_t* out, void* stream)
{
    if (!q || !out) return HTAB_E_ARG;
    uint64_t begin = 0;
    return deque_read_ctl(q, out, &begin, static_cast<hipStream_t>(stream));
}

int
htab_deque_contents(htab_deque* q, uint32_t* out, int64_t* n_out, void* stream)
{
    if (!q || !out || !n_out) return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t n = 0;
    uint64_t begin = 0;
    int rc = deque_read_ctl(q, &n, &begin, s);
    if (rc) return rc;
    if (n > 0)
    {
        hipLaunchKernelGGL(k_deque_copy, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, q->data,
                           q->capacity, begin, n, out);
        if (hipGetLastError() != hipSuccess) return HTAB_E_HIP;
        if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    }
    *n_out = n;
    return HTAB_OK;
}

} // extern "C"
