// Wavefront (wave64) helpers for gfx950: lane id, ballot and lane shuffles.
#ifndef STDGPU_AMD_WAVE_HPP
#define STDGPU_AMD_WAVE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace htab_dev
{

// ------------------------------------------------------------------ wave helpers
__device__ inline int
lane_id()
{
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0U, __builtin_amdgcn_mbcnt_lo(~0U, 0U)));
}

__device__ inline uint64_t
ballot(bool p)
{
    return __builtin_amdgcn_ballot_w64(p);
}

__device__ inline uint32_t
shfl32(uint32_t v, int src)
{
    return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(v)));
}

__device__ inline uint64_t
shfl64(uint64_t v, int src)
{
    const uint32_t lo = shfl32(static_cast<uint32_t>(v), src);
    const uint32_t hi = shfl32(static_cast<uint32_t>(v >> 32), src);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <typename T>
__device__ inline T
shfl(T v, int src)
{
    if constexpr (sizeof(T) == 8)
        return static_cast<T>(shfl64(static_cast<uint64_t>(v), src));
    else
        return static_cast<T>(shfl32(static_cast<uint32_t>(v), src));
}

__device__ inline int
ctz64(uint64_t x)
{
    return __builtin_ctzll(x);
}

} // namespace htab_dev

#endif // STDGPU_AMD_WAVE_HPP
