// vector<T> and deque<T> of any element size: one ring of slots behind both C-ABI families
// (htab_vector_*, htab_deque_*), with bulk calls for host callers and a view for the per-thread calls
// of user kernels (include/stdgpu_amd/ring_device.hpp).
//
// Replaces the reference's per-element paths (src/stdgpu/impl/vector_detail.cuh:181-302 and
// deque_detail.cuh:202-448: an atomic size change per element, an atomicInc / atomicDec of begin or
// end modulo the capacity, then a spin on a per-slot lock bit and an occupancy-bit flip).  One bulk
// call only pushes or only pops, at one end, so it reserves all of its positions at once: a
// one-thread kernel reads begin and end, takes min(n, room) positions, moves the counter and the
// ready / free counts, then a streaming kernel copies element i to or from its position and steps
// the slot's state (FREE -> FULL, FULL -> FREE).  No per-element atomic remains, and the result is
// the one of a sequential loop of the reference's calls: push_back puts element i at end + i,
// pop_back takes end - 1 - i, push_front puts it at begin - 1 - i, pop_front takes begin + i.
//
// A vector is the ring with begin fixed at 0 (its positions are the indices of data()); a deque
// starts both counters at capacity * 2^30, so push_front can move begin down for as long as a
// program runs while position p stays in slot p mod capacity.
#include <hip/hip_runtime.h>

#include <new>

#include <stdgpu_amd/htab.h>
#include <stdgpu_amd/ring_device.hpp>

namespace
{

// control block: int64 words, each count on a 128 B line of its own
constexpr int kReady = 0;     // elements published by their push and not yet reserved by a pop
constexpr int kBegin = 16;    // uint64 front position
constexpr int kEnd = 32;      // uint64 back position
constexpr int kResBase = 48;  // reservation of the current bulk call: the counter before the call
constexpr int kResCount = 49; // ... and the number of elements it covers
constexpr int kFree = 64;     // slots released by their pop and not yet reserved by a push
constexpr int kBad = 80;      // valid(): violations found
constexpr size_t kRingCtlBytes = 1024;

enum ring_op
{
    op_push_back = 0,
    op_pop_back = 1,
    op_push_front = 2,
    op_pop_front = 3
};

__host__ __device__ inline bool
is_push(int op)
{
    return op == op_push_back || op == op_push_front;
}

// the clamped size end - begin (deque_detail.cuh:482-512), then the call's reservation
__global__ void
k_ring_reserve(int64_t* ctl, int64_t cap, int64_t n, int op)
{
    uint64_t* begin = reinterpret_cast<uint64_t*>(ctl + kBegin);
    uint64_t* end = reinterpret_cast<uint64_t*>(ctl + kEnd);
    int64_t s = static_cast<int64_t>(*end - *begin);
    s = s < 0 ? 0 : (s > cap ? cap : s);
    const bool push = is_push(op);
    const int64_t room = push ? cap - s : s;
    const int64_t k = n < room ? n : room;
    const bool back = op == op_push_back || op == op_pop_back;
    uint64_t* counter = back ? end : begin;
    ctl[kResBase] = static_cast<int64_t>(*counter);
    ctl[kResCount] = k;
    // push_back and pop_front move their counter up, pop_back and push_front down
    const bool up = op == op_push_back || op == op_pop_front;
    *counter = up ? *counter + static_cast<uint64_t>(k) : *counter - static_cast<uint64_t>(k);
    ctl[kReady] += push ? k : -k;
    ctl[kFree] += push ? -k : k;
}

// copies one element of eb bytes (4-byte words when eb and both addresses allow it)
__device__ inline void
copy_elem(char* __restrict__ dst, const char* __restrict__ src, int64_t eb)
{
    if ((eb & 3) == 0 && (reinterpret_cast<uintptr_t>(dst) & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 3) == 0)
    {
        for (int64_t w = 0; w < eb; w += 4)
            *reinterpret_cast<uint32_t*>(dst + w) = *reinterpret_cast<const uint32_t*>(src + w);
    }
    else
    {
        for (int64_t b = 0; b < eb; ++b) dst[b] = src[b];
    }
}

// element i of a bulk call: its position, and a copy to or from the caller's array.  The call runs
// alone on the ring, so each slot's state steps by plain stores (FREE -> FULL for a push, FULL -> FREE
// for a pop: +2, as a claim and a release would).
__global__ __launch_bounds__(256) void
k_ring_move(char* __restrict__ data, uint32_t* __restrict__ state, const int64_t* __restrict__ ctl, int64_t cap,
            int64_t eb, int op, char* __restrict__ io, int64_t n, uint8_t* __restrict__ ok_out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool ok = i < ctl[kResCount];
    if (ok)
    {
        const uint64_t base = static_cast<uint64_t>(ctl[kResBase]);
        const bool up = op == op_push_back || op == op_pop_front;
        const uint64_t p = up ? base + static_cast<uint64_t>(i) : base - 1 - static_cast<uint64_t>(i);
        const uint64_t slot = p % static_cast<uint64_t>(cap);
        char* elem = data + slot * static_cast<uint64_t>(eb);
        if (is_push(op))
            copy_elem(elem, io + i * eb, eb);
        else if (io)
            copy_elem(io + i * eb, elem, eb);
        state[slot] += 2;
    }
    else if (!is_push(op) && io)
    {
        // a pop that found the container empty: all-ones bytes (0xFFFFFFFF for 4-byte elements)
        for (int64_t b = 0; b < eb; ++b) io[i * eb + b] = static_cast<char>(0xFF);
    }
    if (ok_out) ok_out[i] = ok ? 1 : 0;
}

// clear (every slot free, no element) or fill_iota (4-byte elements: element i = first + i at
// position begin + i, full); both reset the counters
__global__ __launch_bounds__(256) void
k_ring_reset(char* __restrict__ data, uint32_t* __restrict__ state, int64_t* ctl, int64_t cap, uint64_t base, int iota,
             uint32_t first)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < static_cast<uint64_t>(cap);
         i += stride)
    {
        if (iota) reinterpret_cast<uint32_t*>(data)[(base + i) % static_cast<uint64_t>(cap)] = first + static_cast<uint32_t>(i);
        state[i] = iota ? htab_dev::kSlotFull : htab_dev::kSlotFree;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        ctl[kBegin] = static_cast<int64_t>(base);
        ctl[kEnd] = static_cast<int64_t>(base + (iota ? static_cast<uint64_t>(cap) : 0));
        ctl[kReady] = iota ? cap : 0;
        ctl[kFree] = iota ? 0 : cap;
        ctl[kBad] = 0;
    }
}

__global__ __launch_bounds__(256) void
k_ring_copy(const char* __restrict__ data, int64_t cap, int64_t eb, uint64_t begin, int64_t n, char* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) copy_elem(out + i * eb, data + ((begin + static_cast<uint64_t>(i)) % static_cast<uint64_t>(cap)) * eb, eb);
}

// valid() (deque_detail.cuh:514-573, vector_detail.cuh:514-577): begin <= end <= begin + capacity, the
// slots of the positions [begin, end) full and every other slot free, ready and free equal to the
// element and free-slot counts.  Meant for a ring no kernel is changing.
__global__ __launch_bounds__(256) void
k_ring_valid(const uint32_t* __restrict__ state, int64_t* ctl, int64_t cap)
{
    const uint64_t b = static_cast<uint64_t>(ctl[kBegin]);
    const uint64_t e = static_cast<uint64_t>(ctl[kEnd]);
    const uint64_t n = e - b;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long bad = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        if (e < b || n > static_cast<uint64_t>(cap)) ++bad;
        if (ctl[kReady] != static_cast<int64_t>(n) || ctl[kFree] != cap - static_cast<int64_t>(n)) ++bad;
    }
    if (n <= static_cast<uint64_t>(cap))
    {
        const uint64_t first = b % static_cast<uint64_t>(cap);
        for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < static_cast<uint64_t>(cap);
             s += stride)
        {
            // slot s holds position b + ((s - first) mod cap), an element when that offset is below n
            const uint64_t off = (s + static_cast<uint64_t>(cap) - first) % static_cast<uint64_t>(cap);
            const uint32_t want = off < n ? htab_dev::kSlotFull : htab_dev::kSlotFree;
            if ((state[s] & 3U) != want) ++bad;
        }
    }
    if (bad) atomicAdd(reinterpret_cast<unsigned long long*>(ctl + kBad), bad);
}

int
ring_status()
{
    return hipGetLastError() == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

bool
aligned_to(const void* p, int64_t eb)
{
    const int64_t a = (eb % 8 == 0) ? 8 : (eb % 4 == 0) ? 4 : (eb % 2 == 0) ? 2 : 1;
    return p == nullptr || reinterpret_cast<uintptr_t>(p) % static_cast<uintptr_t>(a) == 0;
}

} // namespace

struct htab_ring
{
    char* data = nullptr;
    uint32_t* state = nullptr;
    int64_t* ctl = nullptr;
    int64_t capacity = 0;
    int64_t elem_bytes = 4;
    uint64_t base = 0; // begin and end after clear(): 0 for a vector, capacity * 2^30 for a deque
};

struct htab_vector
{
    htab_ring r;
};
struct htab_deque
{
    htab_ring r;
};

namespace
{

int
ring_reset(htab_ring* q, int iota, uint32_t first, hipStream_t s)
{
    int64_t grid = (q->capacity + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_ring_reset, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->state, q->ctl,
                       q->capacity, q->base, iota, first);
    return ring_status();
}

int
ring_create(int64_t capacity, int64_t elem_bytes, bool deque, htab_ring* q)
{
    if (capacity <= 0 || capacity > (int64_t(1) << 32) || elem_bytes < 1 || elem_bytes > 256) return HTAB_E_ARG;
    const size_t data_bytes = ((static_cast<size_t>(capacity) * static_cast<size_t>(elem_bytes) + 255) / 256) * 256;
    void* mem = nullptr;
    if (hipMalloc(&mem, kRingCtlBytes + data_bytes + static_cast<size_t>(capacity) * sizeof(uint32_t)) != hipSuccess)
        return HTAB_E_OOM;
    q->ctl = static_cast<int64_t*>(mem);
    q->data = static_cast<char*>(mem) + kRingCtlBytes;
    q->state = reinterpret_cast<uint32_t*>(q->data + data_bytes);
    q->capacity = capacity;
    q->elem_bytes = elem_bytes;
    q->base = deque ? static_cast<uint64_t>(capacity) << 30 : 0;
    // empty, every slot free (the reference creates empty containers, vector_detail.cuh:40-80,
    // deque_detail.cuh:40-80)
    if (hipMemset(mem, 0, kRingCtlBytes) != hipSuccess || ring_reset(q, 0, 0, nullptr) != HTAB_OK ||
        hipStreamSynchronize(nullptr) != hipSuccess)
    {
        (void)hipFree(mem);
        return HTAB_E_HIP;
    }
    return HTAB_OK;
}

int
ring_call(htab_ring* q, int op, const void* in, void* out, int64_t n, uint8_t* ok, void* stream)
{
    if (n < 0) return HTAB_E_ARG;
    if (is_push(op) && n > 0 && !in) return HTAB_E_ARG;
    if (!aligned_to(in, q->elem_bytes) || !aligned_to(out, q->elem_bytes)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_ring_reserve, dim3(1), dim3(1), 0, s, q->ctl, q->capacity, n, op);
    char* io = is_push(op) ? const_cast<char*>(static_cast<const char*>(in)) : static_cast<char*>(out);
    hipLaunchKernelGGL(k_ring_move, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->data, q->state, q->ctl,
                       q->capacity, q->elem_bytes, op, io, n, ok);
    return ring_status();
}

void
ring_view(const htab_ring* q, htab_ring_view* out)
{
    out->data = q->data;
    out->state = q->state;
    out->begin = reinterpret_cast<uint64_t*>(q->ctl + kBegin);
    out->end = reinterpret_cast<uint64_t*>(q->ctl + kEnd);
    out->ready = q->ctl + kReady;
    out->free_slots = q->ctl + kFree;
    out->capacity = q->capacity;
    out->elem_bytes = q->elem_bytes;
}

int
ring_read_ctl(const htab_ring* q, int64_t* size, uint64_t* begin, hipStream_t s)
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
ring_contents(const htab_ring* q, void* out, int64_t* n_out, hipStream_t s)
{
    if (!out || !n_out || !aligned_to(out, q->elem_bytes)) return HTAB_E_ARG;
    int64_t n = 0;
    uint64_t begin = 0;
    int rc = ring_read_ctl(q, &n, &begin, s);
    if (rc) return rc;
    if (n > 0)
    {
        hipLaunchKernelGGL(k_ring_copy, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, q->data,
                           q->capacity, q->elem_bytes, begin, n, static_cast<char*>(out));
        if (hipGetLastError() != hipSuccess) return HTAB_E_HIP;
        if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    }
    *n_out = n;
    return HTAB_OK;
}

int
ring_valid(htab_ring* q, int32_t* out, hipStream_t s)
{
    if (!out) return HTAB_E_ARG;
    unsigned long long zero = 0, bad = 0;
    if (hipMemcpyAsync(q->ctl + kBad, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess) return HTAB_E_HIP;
    int64_t grid = (q->capacity + 255) / 256;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_ring_valid, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, q->state, q->ctl, q->capacity);
    if (hipGetLastError() != hipSuccess) return HTAB_E_HIP;
    if (hipMemcpyAsync(&bad, q->ctl + kBad, sizeof(bad), hipMemcpyDeviceToHost, s) != hipSuccess) return HTAB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    *out = bad == 0 ? 1 : 0;
    return HTAB_OK;
}

} // namespace

extern "C" {

// ------------------------------------------------------------------ vector
int
htab_vector_create_sized(int64_t capacity, int64_t elem_bytes, htab_vector** out)
{
    if (!out) return HTAB_E_ARG;
    *out = nullptr;
    htab_vector* v = new (std::nothrow) htab_vector();
    if (!v) return HTAB_E_OOM;
    const int rc = ring_create(capacity, elem_bytes, false, &v->r);
    if (rc != HTAB_OK)
    {
        delete v;
        return rc;
    }
    *out = v;
    return HTAB_OK;
}

int
htab_vector_create(int64_t capacity, htab_vector** out)
{
    return htab_vector_create_sized(capacity, 4, out);
}

int
htab_vector_destroy(htab_vector* v)
{
    if (!v) return HTAB_E_ARG;
    const hipError_t e = hipFree(v->r.ctl);
    delete v;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_vector_push_back(htab_vector* v, const void* values, int64_t n, uint8_t* pushed, void* stream)
{
    if (!v) return HTAB_E_ARG;
    return ring_call(&v->r, op_push_back, values, nullptr, n, pushed, stream);
}

int
htab_vector_pop_back(htab_vector* v, void* out, int64_t n, uint8_t* popped, void* stream)
{
    if (!v) return HTAB_E_ARG;
    return ring_call(&v->r, op_pop_back, nullptr, out, n, popped, stream);
}

int
htab_vector_clear(htab_vector* v, void* stream)
{
    if (!v) return HTAB_E_ARG;
    return ring_reset(&v->r, 0, 0, static_cast<hipStream_t>(stream));
}

int
htab_vector_get_view(htab_vector* v, htab_vector_view* out)
{
    if (!v || !out) return HTAB_E_ARG;
    ring_view(&v->r, out);
    return HTAB_OK;
}

int
htab_vector_fill_iota(htab_vector* v, uint32_t first, void* stream)
{
    if (!v || v->r.elem_bytes != 4) return HTAB_E_ARG;
    return ring_reset(&v->r, 1, first, static_cast<hipStream_t>(stream));
}

int
htab_vector_size(htab_vector* v, int64_t* out, void* stream)
{
    if (!v || !out) return HTAB_E_ARG;
    uint64_t begin = 0;
    return ring_read_ctl(&v->r, out, &begin, static_cast<hipStream_t>(stream));
}

int
htab_vector_contents(htab_vector* v, void* out, int64_t* n_out, void* stream)
{
    if (!v) return HTAB_E_ARG;
    return ring_contents(&v->r, out, n_out, static_cast<hipStream_t>(stream));
}

int
htab_vector_valid(htab_vector* v, int32_t* out, void* stream)
{
    if (!v) return HTAB_E_ARG;
    return ring_valid(&v->r, out, static_cast<hipStream_t>(stream));
}

// ------------------------------------------------------------------ deque
int
htab_deque_create_sized(int64_t capacity, int64_t elem_bytes, htab_deque** out)
{
    if (!out) return HTAB_E_ARG;
    *out = nullptr;
    htab_deque* q = new (std::nothrow) htab_deque();
    if (!q) return HTAB_E_OOM;
    const int rc = ring_create(capacity, elem_bytes, true, &q->r);
    if (rc != HTAB_OK)
    {
        delete q;
        return rc;
    }
    *out = q;
    return HTAB_OK;
}

int
htab_deque_create(int64_t capacity, htab_deque** out)
{
    return htab_deque_create_sized(capacity, 4, out);
}

int
htab_deque_destroy(htab_deque* q)
{
    if (!q) return HTAB_E_ARG;
    const hipError_t e = hipFree(q->r.ctl);
    delete q;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_deque_push_back(htab_deque* q, const void* values, int64_t n, uint8_t* pushed, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_call(&q->r, op_push_back, values, nullptr, n, pushed, stream);
}

int
htab_deque_push_front(htab_deque* q, const void* values, int64_t n, uint8_t* pushed, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_call(&q->r, op_push_front, values, nullptr, n, pushed, stream);
}

int
htab_deque_pop_front(htab_deque* q, void* out, int64_t n, uint8_t* popped, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_call(&q->r, op_pop_front, nullptr, out, n, popped, stream);
}

int
htab_deque_pop_back(htab_deque* q, void* out, int64_t n, uint8_t* popped, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_call(&q->r, op_pop_back, nullptr, out, n, popped, stream);
}

int
htab_deque_fill_iota(htab_deque* q, void* stream)
{
    if (!q || q->r.elem_bytes != 4) return HTAB_E_ARG;
    return ring_reset(&q->r, 1, 0, static_cast<hipStream_t>(stream));
}

int
htab_deque_clear(htab_deque* q, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_reset(&q->r, 0, 0, static_cast<hipStream_t>(stream));
}

int
htab_deque_get_view(const htab_deque* q, htab_deque_view* out)
{
    if (!q || !out) return HTAB_E_ARG;
    ring_view(&q->r, out);
    return HTAB_OK;
}

int
htab_deque_size(htab_deque* q, int64_t* out, void* stream)
{
    if (!q || !out) return HTAB_E_ARG;
    uint64_t begin = 0;
    return ring_read_ctl(&q->r, out, &begin, static_cast<hipStream_t>(stream));
}

int
htab_deque_contents(htab_deque* q, void* out, int64_t* n_out, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_contents(&q->r, out, n_out, static_cast<hipStream_t>(stream));
}

int
htab_deque_valid(htab_deque* q, int32_t* out, void* stream)
{
    if (!q) return HTAB_E_ARG;
    return ring_valid(&q->r, out, static_cast<hipStream_t>(stream));
}

} // extern "C"
