// bitset over 32-bit blocks with bulk per-bit updates and tests (src/stdgpu/bitset.cuh,
// impl/bitset_detail.cuh).  set / reset / flip of bit n is one atomic or / and / xor on its block and
// returns the bit's previous value (bitset_detail.cuh:47-112, 265-329); set() / reset() fill every
// block and flip() inverts every block (:247-319); count() sums the population counts of the blocks
// with the last block masked to the bits below size() (:131-160), here with a workgroup reduction and
// one atomic per workgroup.
#include <hip/hip_runtime.h>

#include <new>

#include <stdgpu_amd/htab.h>

#include "wave.hpp"

using namespace htab_dev;

namespace
{

__global__ __launch_bounds__(256) void
k_bitset_update(uint32_t* __restrict__ blocks, int64_t nbits, const int64_t* __restrict__ pos, int64_t n, int op,
                uint8_t* __restrict__ old_out, uint8_t* __restrict__ in_range)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = pos[i];
    const bool ok = p >= 0 && p < nbits;
    uint32_t old = 0;
    if (ok)
    {
        uint32_t* w = blocks + (p >> 5);
        const uint32_t bit = 1U << (p & 31);
        if (op == HTAB_BIT_SET)
            old = __hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else if (op == HTAB_BIT_RESET)
            old = __hip_atomic_fetch_and(w, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            old = __hip_atomic_fetch_xor(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        old = (old & bit) ? 1U : 0U;
    }
    if (old_out) old_out[i] = static_cast<uint8_t>(old);
    if (in_range) in_range[i] = ok ? 1 : 0;
}

__global__ __launch_bounds__(256) void
k_bitset_test(const uint32_t* __restrict__ blocks, int64_t nbits, const int64_t* __restrict__ pos, int64_t n,
              uint8_t* __restrict__ out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = pos[i];
    out[i] = (p >= 0 && p < nbits && ((blocks[p >> 5] >> (p & 31)) & 1U)) ? 1 : 0;
}

// op 0: every block = 0, op 1: every block = all ones, op 2: every block inverted
__global__ __launch_bounds__(256) void
k_bitset_all(uint32_t* __restrict__ blocks, int64_t nblocks, int op)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < nblocks; b += stride)
        blocks[b] = op == 0 ? 0U : (op == 1 ? ~0U : ~blocks[b]);
}

__global__ __launch_bounds__(256) void
k_bitset_count(const uint32_t* __restrict__ blocks, int64_t nbits, unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long s_wave[4];
    const int64_t nblocks = (nbits + 31) >> 5;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    unsigned long long c = 0;
    for (int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < nblocks; b += stride)
    {
        const int64_t rem = nbits - (b << 5);
        const uint32_t mask = rem >= 32 ? ~0U : ((1U << rem) - 1U);
        c += static_cast<unsigned long long>(__builtin_popcount(blocks[b] & mask));
    }
    const int lane = lane_id();
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += shfl64(c, lane ^ off);
    if (lane == 0) s_wave[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const unsigned long long sum = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        if (sum) atomicAdd(total, sum);
    }
}

__global__ void
k_zero_counter(unsigned long long* c)
{
    *c = 0;
}

bool
aligned(const void* p, uintptr_t bytes)
{
    return p == nullptr || reinterpret_cast<uintptr_t>(p) % bytes == 0;
}

int
bs_status()
{
    return hipGetLastError() == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

constexpr size_t kBitsetCtlBytes = 256;

} // namespace

struct htab_bitset
{
    uint32_t* blocks = nullptr;
    unsigned long long* counter = nullptr;
    int64_t nbits = 0;
};

extern "C" {

int
htab_bitset_create(int64_t nbits, htab_bitset** out)
{
    if (!out || nbits <= 0 || nbits > (int64_t(1) << 40)) return HTAB_E_ARG;
    *out = nullptr;
    htab_bitset* b = new (std::nothrow) htab_bitset();
    if (!b) return HTAB_E_OOM;
    const size_t nblocks = static_cast<size_t>((nbits + 31) >> 5);
    void* mem = nullptr;
    if (hipMalloc(&mem, kBitsetCtlBytes + nblocks * sizeof(uint32_t)) != hipSuccess)
    {
        delete b;
        return HTAB_E_OOM;
    }
    b->counter = static_cast<unsigned long long*>(mem);
    b->blocks = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + kBitsetCtlBytes);
    b->nbits = nbits;
    // every bit starts reset (bitset_detail.cuh:190-204)
    if (hipMemset(mem, 0, kBitsetCtlBytes + nblocks * sizeof(uint32_t)) != hipSuccess)
    {
        (void)hipFree(mem);
        delete b;
        return HTAB_E_HIP;
    }
    *out = b;
    return HTAB_OK;
}

int
htab_bitset_destroy(htab_bitset* b)
{
    if (!b) return HTAB_E_ARG;
    const hipError_t e = hipFree(b->counter);
    delete b;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_bitset_update(htab_bitset* b, int op, const int64_t* positions, int64_t n, uint8_t* old_out, uint8_t* in_range,
                   void* stream)
{
    if (!b || n < 0 || (n > 0 && !positions) || op < HTAB_BIT_SET || op > HTAB_BIT_FLIP) return HTAB_E_ARG;
    if (!aligned(positions, 8)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipLaunchKernelGGL(k_bitset_update, dim3(static_cast<uint32_t>(grid)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), b->blocks, b->nbits, positions, n, op, old_out, in_range);
    return bs_status();
}

int
htab_bitset_test(htab_bitset* b, const int64_t* positions, int64_t n, uint8_t* out, void* stream)
{
    if (!b || n < 0 || (n > 0 && (!positions || !out))) return HTAB_E_ARG;
    if (!aligned(positions, 8)) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
    hipLaunchKernelGGL(k_bitset_test, dim3(static_cast<uint32_t>(grid)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), b->blocks, b->nbits, positions, n, out);
    return bs_status();
}

int
htab_bitset_all_bits(htab_bitset* b, int op, void* stream)
{
    if (!b || op < HTAB_BIT_SET || op > HTAB_BIT_FLIP) return HTAB_E_ARG;
    const int64_t nblocks = (b->nbits + 31) >> 5;
    int64_t grid = (nblocks + 255) / 256;
    if (grid > 65536) grid = 65536;
    const int kop = op == HTAB_BIT_SET ? 1 : (op == HTAB_BIT_RESET ? 0 : 2);
    hipLaunchKernelGGL(k_bitset_all, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       b->blocks, nblocks, kop);
    return bs_status();
}

int
htab_bitset_get_view(const htab_bitset* b, htab_bitset_view* out)
{
    if (!b || !out) return HTAB_E_ARG;
    out->blocks = b->blocks;
    out->nbits = b->nbits;
    return HTAB_OK;
}

int
htab_bitset_count(htab_bitset* b, int64_t* out, void* stream)
{
    if (!b || !out) return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nblocks = (b->nbits + 31) >> 5;
    int64_t grid = (nblocks + 255) / 256;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, b->counter);
    hipLaunchKernelGGL(k_bitset_count, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, b->blocks, b->nbits,
                       b->counter);
    if (hipGetLastError() != hipSuccess) return HTAB_E_HIP;
    unsigned long long c = 0;
    if (hipMemcpyAsync(&c, b->counter, sizeof(c), hipMemcpyDeviceToHost, s) != hipSuccess) return HTAB_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return HTAB_E_HIP;
    *out = static_cast<int64_t>(c);
    return HTAB_OK;
}

} // extern "C"
