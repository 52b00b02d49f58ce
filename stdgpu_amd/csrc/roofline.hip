// Random-gather and random-atomic microbenchmarks: the measured ceiling the hash-table kernels
// are compared against on the same chip.  Built into libstdgpu_amd_roofline.so (measurement only,
// not part of the container API).
//
//   gather: `n` random, aligned reads of `line` bytes (64 or 128) from a buffer of `bytes`
//           bytes.  Like a bucket probe, `line/16` lanes read one line together with 16-byte
//           loads, and every lane keeps 8 independent lines in flight.
//   cas   : `n` random 64-bit compare-and-swap operations whose old value is consumed.
//   add   : `n` random 64-bit atomic adds without return.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace
{

__device__ inline uint64_t
mixr(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

template <int G>
__global__ __launch_bounds__(256) void
k_gather(const uint4* __restrict__ buf, uint64_t nlines, uint64_t n, uint64_t seed, uint32_t* __restrict__ sink)
{
    constexpr int R = 8;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int li = lane & (G - 1);
    const uint64_t group = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
    const uint64_t ngroups = static_cast<uint64_t>(gridDim.x) * blockDim.x / G;
    uint32_t acc = 0;
    for (uint64_t base = group * R; base < n; base += ngroups * R)
    {
        uint4 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            const uint64_t i = base + r;
            const uint64_t line = ((mixr(i ^ seed) >> 32) * nlines) >> 32;
            v[r] = (i < n) ? buf[line * G + li] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) acc += v[r].x ^ v[r].y ^ v[r].z ^ v[r].w;
    }
    if (acc == 0x9E3779B9U) sink[0] = acc; // keeps the loads alive; practically never stores
}

__global__ __launch_bounds__(256) void
k_cas(unsigned long long* __restrict__ buf, uint64_t nwords, uint64_t n, uint64_t seed, uint32_t* __restrict__ sink)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    uint64_t acc = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t w = ((mixr(i ^ seed) >> 32) * nwords) >> 32;
        acc += atomicCAS(buf + w, 0ULL, static_cast<unsigned long long>(i + 1));
    }
    if (acc == 0x9E3779B97F4A7C15ULL) sink[0] = static_cast<uint32_t>(acc);
}

__global__ __launch_bounds__(256) void
k_add(unsigned long long* __restrict__ buf, uint64_t nwords, uint64_t n, uint64_t seed)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const uint64_t w = ((mixr(i ^ seed) >> 32) * nwords) >> 32;
        atomicAdd(buf + w, 1ULL);
    }
}

__global__ __launch_bounds__(256) void
k_fill(uint4* __restrict__ buf, uint64_t n16)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
        buf[i] = make_uint4(static_cast<uint32_t>(i), 1, 2, 3);
}

// plain streaming copy of known size: the calibration kernel for FETCH_SIZE / WRITE_SIZE
__global__ __launch_bounds__(256) void
k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

} // namespace

extern "C" {

// All functions: `buf` is caller-owned device memory of `bytes` bytes; returns 0 or a HIP error.
// Times are taken by the caller with events on `stream`.
int
roofline_fill(void* buf, uint64_t bytes, void* stream)
{
    hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, static_cast<hipStream_t>(stream), static_cast<uint4*>(buf),
                       bytes / 16);
    return static_cast<int>(hipGetLastError());
}

int
roofline_gather(const void* buf, uint64_t bytes, int line, uint64_t n, uint64_t seed, uint32_t* sink, void* stream)
{
    const uint64_t nlines = bytes / static_cast<uint64_t>(line);
    if (nlines == 0 || nlines > (1ULL << 32)) return static_cast<int>(hipErrorInvalidValue);
    const dim3 grid(16384), block(256);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (line == 64)
        hipLaunchKernelGGL(k_gather<4>, grid, block, 0, s, static_cast<const uint4*>(buf), nlines, n, seed, sink);
    else if (line == 128)
        hipLaunchKernelGGL(k_gather<8>, grid, block, 0, s, static_cast<const uint4*>(buf), nlines, n, seed, sink);
    else if (line == 256)
        hipLaunchKernelGGL(k_gather<16>, grid, block, 0, s, static_cast<const uint4*>(buf), nlines, n, seed, sink);
    else
        return static_cast<int>(hipErrorInvalidValue);
    return static_cast<int>(hipGetLastError());
}

int
roofline_cas(void* buf, uint64_t bytes, uint64_t n, uint64_t seed, uint32_t* sink, void* stream)
{
    const uint64_t nwords = bytes / 8;
    if (nwords == 0 || nwords > (1ULL << 32)) return static_cast<int>(hipErrorInvalidValue);
    hipLaunchKernelGGL(k_cas, dim3(16384), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<unsigned long long*>(buf), nwords, n, seed, sink);
    return static_cast<int>(hipGetLastError());
}

int
roofline_add(void* buf, uint64_t bytes, uint64_t n, uint64_t seed, void* stream)
{
    const uint64_t nwords = bytes / 8;
    if (nwords == 0 || nwords > (1ULL << 32)) return static_cast<int>(hipErrorInvalidValue);
    hipLaunchKernelGGL(k_add, dim3(16384), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<unsigned long long*>(buf), nwords, n, seed);
    return static_cast<int>(hipGetLastError());
}

int
roofline_copy(const void* src, void* dst, uint64_t bytes, void* stream)
{
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), bytes / 16);
    return static_cast<int>(hipGetLastError());
}

} // extern "C"
