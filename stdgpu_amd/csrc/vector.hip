// vector<uint32_t> as a stack (the reference's excess free-list container) with bulk push_back /
// pop_back.
//
// Replaces the reference's per-element path (src/stdgpu/impl/vector_detail.cuh:181-240 push_back,
// 242-302 pop_back: an atomic ++size / --size per element, then a spin on a per-slot lock bit and an
// occupancy-bit flip, run one lane at a time by warp_convergent_execute).  One bulk call only pushes
// or only pops, so it reserves all of its positions at once: a one-thread kernel reads the size,
// takes min(n, room) positions and writes the new size, then a streaming kernel moves element i to or
// from position base + i (push) or top - 1 - i (pop).  No per-element atomic, lock or occupancy bit
// remains, and the result is the one of a sequential loop of push_back / pop_back calls.
#include <hip/hip_runtime.h>

#include <new>

#include <stdgpu_amd/htab.h>

namespace
{

constexpr size_t kVectorCtlBytes = 256;
constexpr int kSize = 0;    // int64 size word
constexpr int kResBase = 16; // reservation of the current bulk call: first position (push) or top (pop)
constexpr int kResCount = 17; // ... and the number of elements it covers

// the clamped size (vector_detail.cuh:454-512), then the call's reservation
__global__ void
k_vector_reserve(int64_t* ctl, int64_t cap, int64_t n, int push)
{
    int64_t s = ctl[kSize];
    s = s < 0 ? 0 : (s > cap ? cap : s);
    const int64_t room = push ? cap - s : s;
    const int64_t k = n < room ? n : room;
    ctl[kSize] = push ? s + k : s - k;
    ctl[kResBase] = s;
    ctl[kResCount] = k;
}

__global__ __launch_bounds__(256) void
k_vector_push(uint32_t* __restrict__ data, const int64_t* __restrict__ ctl, const uint32_t* __restrict__ v, int64_t n,
              uint8_t* __restrict__ pushed)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = i < ctl[kResCount];
    if (ok) data[ctl[kResBase] + i] = v[i];
    if (pushed) pushed[i] = ok ? 1 : 0;
}

__global__ __launch_bounds__(256) void
k_vector_pop(const uint32_t* __restrict__ data, const int64_t* __restrict__ ctl, uint32_t* __restrict__ out, int64_t n,
             uint8_t* __restrict__ popped)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = i < ctl[kResCount];
    if (out) out[i] = ok ? data[ctl[kResBase] - 1 - i] : 0xFFFFFFFFU;
    if (popped) popped[i] = ok ? 1 : 0;
}

__global__ void
k_vector_clear(int64_t* ctl)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[kSize] = 0;
}

// vector_clear_iota (vector_detail.cuh:361-368): data[i] = first + i, size = capacity
__global__ __launch_bounds__(256) void
k_vector_iota(uint32_t* __restrict__ data, int64_t* ctl, int64_t cap, uint32_t first)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < static_cast<uint64_t>(cap);
         i += stride)
        data[i] = first + static_cast<uint32_t>(i);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[kSize] = cap;
}

bool
aligned4(const void* p)
{
    return p == nullptr || reinterpret_cast<uintptr_t>(p) % 4 == 0;
}

int
vec_status()
{
    return hipGetLastError() == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

} // namespace

struct htab_vector
{
    uint32_t* data = nullptr;
    int64_t* ctl = nullptr; // size word and the reservation of the current bulk call
    int64_t capacity = 0;
};

namespace
{

int
vec_size(const htab_vector* q, int64_t* out, hipStream_t s)
{
    int64_t n = 0;
    if (hipMemcpyAsync(&n, q->ctl + kSize, sizeof(n), hipMemcpyDeviceToHost, s) != hipSuccess) return HTAB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    // the reference's size() clamps to [0, capacity] (vector_detail.cuh:454-512)
    *out = n < 0 ? 0 : (n > q->capacity ? q->capacity : n);
    return HTAB_OK;
}

} // namespace

extern "C" {

int
htab_vector_create(int64_t capacity, htab_vector** out)
{
    if (!out || capacity <= 0 || capacity > (int64_t(1) << 32)) return HTAB_E_ARG;
    *out = nullptr;
    htab_vector* q = new (std::nothrow) htab_vector();
    if (!q) return HTAB_E_OOM;
    void* mem = nullptr;
    if (hipMalloc(&mem, kVectorCtlBytes + static_cast<size_t>(capacity) * sizeof(uint32_t)) != hipSuccess)
    {
        delete q;
        return HTAB_E_OOM;
    }
    q->ctl = static_cast<int64_t*>(mem);
    q->data = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + kVectorCtlBytes);
    q->capacity = capacity;
    if (hipMemset(mem, 0, kVectorCtlBytes) != hipSuccess)
    {
        (void)hipFree(mem);
        delete q;
        return HTAB_E_HIP;
    }
    *out = q;
    return HTAB_OK;
}

int
htab_vector_destroy(htab_vector* q)
{
    if (!q) return HTAB_E_ARG;
    const hipError_t e = hipFree(q->ctl);
    delete q;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_vector_push_back(htab_vector* q, const uint32_t* values, int64_t n, uint8_t* pushed, void* stream)
{
    if (!q || n < 0 || (n > 0 && !values) || !aligned4(values)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_vector_reserve, dim3(1), dim3(1), 0, s, q->ctl, q->capacity, n, 1);
    hipLaunchKernelGGL(k_vector_push, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->ctl, values, n,
                       pushed);
    return vec_status();
}

int
htab_vector_pop_back(htab_vector* q, uint32_t* out, int64_t n, uint8_t* popped, void* stream)
{
    if (!q || n < 0 || !aligned4(out)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_vector_reserve, dim3(1), dim3(1), 0, s, q->ctl, q->capacity, n, 0);
    hipLaunchKernelGGL(k_vector_pop, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->ctl, out, n,
                       popped);
    return vec_status();
}

int
htab_vector_clear(htab_vector* q, void* stream)
{
    if (!q) return HTAB_E_ARG;
    hipLaunchKernelGGL(k_vector_clear, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), q->ctl);
    return vec_status();
}

int
htab_vector_get_view(htab_vector* q, htab_vector_view* out)
{
    if (!q || !out) return HTAB_E_ARG;
    out->data = q->data;
    out->size = q->ctl + kSize;
    out->capacity = q->capacity;
    return HTAB_OK;
}

int
htab_vector_fill_iota(htab_vector* q, uint32_t first, void* stream)
{
    if (!q) return HTAB_E_ARG;
    int64_t grid = (q->capacity + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_vector_iota, dim3(static_cast<uint32_t>(grid)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), q->data, q->ctl, q->capacity, first);
    return vec_status();
}

int
htab_vector_size(htab_vector* q, int64_t* out, void* stream)
{
    if (!q || !out) return HTAB_E_ARG;
    return vec_size(q, out, static_cast<hipStream_t>(stream));
}

int
htab_vector_contents(htab_vector* q, uint32_t* out, int64_t* n_out, void* stream)
{
    if (!q || !out || !n_out) return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t n = 0;
    const int rc = vec_size(q, &n, s);
    if (rc) return rc;
    if (n > 0 && hipMemcpyAsync(out, q->data, static_cast<size_t>(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s) !=
                         hipSuccess)
        return HTAB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    *n_out = n;
    return HTAB_OK;
}

} // extern "C"
