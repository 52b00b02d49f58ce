// Device-side push_back / pop_back on vector<uint32_t> for user kernels, called per thread as the
// reference's callers call them (src/stdgpu/vector.cuh:219-239; vector_detail.cuh:181-240 push_back,
// 242-302 pop_back).  gfx950 only.
//
// The calling lanes of a wavefront take their positions with one atomic on the size word: a push
// reserves k positions from the top (fetch_add), a pop takes the k topmost (fetch_sub).  Lanes whose
// position falls outside [0, capacity) report failure, as the reference's calls do on a full or empty
// vector, and their share of the size change is given back with one more atomic.  A call never waits,
// so it cannot stall on a lane of its own wavefront; the price is that one kernel pushes or pops but
// does not do both on one vector (a pop could take a position whose push has not written it yet).
// deque<uint32_t> (deque_device.hpp) is the container for pushes and pops mixed in one kernel.
#ifndef STDGPU_AMD_VECTOR_DEVICE_HPP
#define STDGPU_AMD_VECTOR_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <stdgpu_amd/htab.h>

namespace htab_dev
{

struct vector_ref
{
    uint32_t* data = nullptr;
    int64_t* size_word = nullptr;
    int64_t capacity = 0;

    __host__ __device__ static vector_ref from_view(const htab_vector_view& v)
    {
        vector_ref r;
        r.data = v.data;
        r.size_word = v.size;
        r.capacity = v.capacity;
        return r;
    }

    // size() clamped to [0, capacity] as the reference's (vector_detail.cuh:454-512)
    __device__ int64_t size() const
    {
        const int64_t n = __hip_atomic_load(size_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return n < 0 ? 0 : (n > capacity ? capacity : n);
    }

    __device__ bool push_back(uint32_t v) const
    {
        int64_t pos = 0;
        if (!take(1, &pos)) return false;
        data[pos] = v;
        return true;
    }

    __device__ bool pop_back(uint32_t* out) const
    {
        int64_t pos = 0;
        if (!take(-1, &pos)) return false;
        *out = data[pos];
        return true;
    }

private:
    // dir +1: positions old, old + 1, ... for the calling lanes in lane order; -1: old - 1, old - 2, ...
    __device__ bool take(int64_t dir, int64_t* pos) const
    {
        const uint64_t lanes = __builtin_amdgcn_ballot_w64(true); // the lanes making this call
        const int lane = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0U, __builtin_amdgcn_mbcnt_lo(~0U, 0U)));
        const int leader = __builtin_ctzll(lanes);
        const int64_t k = __builtin_popcountll(lanes);
        const int64_t rank = __builtin_popcountll(lanes & ((uint64_t(1) << lane) - 1));
        int64_t old = 0;
        if (lane == leader) old = __hip_atomic_fetch_add(size_word, dir * k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        old = static_cast<int64_t>(
                static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(old), leader))) |
                (static_cast<uint64_t>(static_cast<uint32_t>(
                         __builtin_amdgcn_readlane(static_cast<uint32_t>(static_cast<uint64_t>(old) >> 32), leader)))
                 << 32));
        // this lane's position; every lane checks it against [0, capacity) itself, so no access leaves
        // the array even when a size change is still to be given back
        *pos = (dir > 0) ? old + rank : old - 1 - rank;
        const bool ok = *pos >= 0 && *pos < capacity;
        const int64_t good = __builtin_popcountll(__builtin_amdgcn_ballot_w64(ok));
        if (good < k && lane == leader)
            __hip_atomic_fetch_add(size_word, -dir * (k - good), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return ok;
    }
};

} // namespace htab_dev

#endif // STDGPU_AMD_VECTOR_DEVICE_HPP
