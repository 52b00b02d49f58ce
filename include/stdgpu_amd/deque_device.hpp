// Device-side operations on the free-list deque<uint32_t> for user kernels: push_back and pop_front
// called per thread, as the reference's voxel pipelines call them (src/stdgpu/deque.cuh:199-239;
// deque_detail.cuh:202-260 push_back, 390-448 pop_front).  gfx950 only.
//
// Ring position p (a 64-bit counter value) lives in slot p mod capacity.  seq[slot] holds the slot's
// state for turn t = p / capacity: 2t when it is free for the push of position p, 2t + 1 when it holds
// position p's element, and 2(t + 1) once that element is popped (free for position p + capacity).
// begin and end are the front and back counters; size = end - begin.
//
// Two counts gate the calls: ready (elements whose push has published them) and free (slots whose pop
// has released them).  The calling lanes of a wavefront reserve from ready (pop) or free (push) with one
// fetch_sub, give back what was not there with one fetch_add, and take that many consecutive ring
// positions with one fetch_add on begin or end, in lane order; lanes past the reservation get false,
// as the reference's calls do on an empty or full deque (it gates them the same way with its size
// counter, deque_detail.cuh:202-260, 390-448).  A reserved position may still be in the hands of the
// other call (its push still writing, or its previous pop still reading): the lane waits for that
// slot's seq word, as the reference waits on its per-slot lock, then copies, publishes, and the
// wavefront credits the opposite count once.  The wait ends: the count it reserved from only counts
// calls that have taken their positions, a call that has taken a position finishes it with no wait
// other than this same one for an earlier position, and within one wavefront a whole call runs in one
// branch, so a lane never waits on a diverged lane of its own wavefront.  A transient under-count (a
// give-back not yet done) can only make a concurrent call report empty or full early, as in the
// reference.
#ifndef STDGPU_AMD_DEQUE_DEVICE_HPP
#define STDGPU_AMD_DEQUE_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <stdgpu_amd/htab.h>

namespace htab_dev
{

struct deque_ref
{
    uint32_t* data = nullptr;
    uint32_t* seq = nullptr;
    uint64_t* begin = nullptr;
    uint64_t* end = nullptr;
    int64_t* ready = nullptr;
    int64_t* free_slots = nullptr;
    uint64_t capacity = 0;

    __host__ __device__ static deque_ref from_view(const htab_deque_view& v)
    {
        deque_ref d;
        d.data = v.data;
        d.seq = v.seq;
        d.begin = v.begin;
        d.end = v.end;
        d.ready = v.ready;
        d.free_slots = v.free_slots;
        d.capacity = static_cast<uint64_t>(v.capacity);
        return d;
    }

    __device__ uint32_t seq_free(uint64_t p) const { return static_cast<uint32_t>(2 * (p / capacity)); }
    __device__ uint32_t seq_full(uint64_t p) const { return static_cast<uint32_t>(2 * (p / capacity) + 1); }

    __device__ uint64_t load_counter(const uint64_t* c) const
    {
        return __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // size() clamped to [0, capacity] as the reference's (deque_detail.cuh:482-512)
    __device__ int64_t size() const
    {
        const uint64_t b = load_counter(begin);
        const uint64_t e = load_counter(end);
        const int64_t n = static_cast<int64_t>(e - b);
        return n < 0 ? 0 : (n > static_cast<int64_t>(capacity) ? static_cast<int64_t>(capacity) : n);
    }

    // the element at index n from the front (no synchronisation, as the reference's operator[])
    __device__ uint32_t* at(int64_t n) const
    {
        return data + (load_counter(begin) + static_cast<uint64_t>(n)) % capacity;
    }

    __device__ bool push_back(uint32_t v) const
    {
        const wave_call w = begin_call();
        uint64_t pos = 0;
        const uint32_t run = reserve(w, free_slots, end, &pos);
        const bool mine = w.rank < run;
        if (mine)
        {
            const uint64_t slot = pos % capacity;
            wait_seq(slot, seq_free(pos));
            data[slot] = v;
            __hip_atomic_store(seq + slot, seq_full(pos), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (w.lane == w.leader && run > 0)
            __hip_atomic_fetch_add(ready, static_cast<int64_t>(run), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return mine;
    }

    __device__ bool pop_front(uint32_t* out) const
    {
        const wave_call w = begin_call();
        uint64_t pos = 0;
        const uint32_t run = reserve(w, ready, begin, &pos);
        const bool mine = w.rank < run;
        if (mine)
        {
            const uint64_t slot = pos % capacity;
            wait_seq(slot, seq_full(pos));
            *out = data[slot];
            __hip_atomic_store(seq + slot, seq_free(pos + capacity), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (w.lane == w.leader && run > 0)
            __hip_atomic_fetch_add(free_slots, static_cast<int64_t>(run), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        return mine;
    }

private:
    struct wave_call
    {
        uint64_t lanes;
        int lane;
        int leader;
        uint32_t k;
        uint32_t rank;
    };

    // the lanes of the wavefront that make this call, and this lane's rank among them
    __device__ static wave_call begin_call()
    {
        wave_call w;
        w.lanes = __builtin_amdgcn_ballot_w64(true); // the lanes making this call (a convergent vote)
        w.lane = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0U, __builtin_amdgcn_mbcnt_lo(~0U, 0U)));
        w.leader = __builtin_ctzll(w.lanes);
        w.k = static_cast<uint32_t>(__builtin_popcountll(w.lanes));
        w.rank = static_cast<uint32_t>(__builtin_popcountll(w.lanes & ((uint64_t(1) << w.lane) - 1)));
        return w;
    }

    // reserves up to k units from count, gives back the shortfall, and takes that many positions from
    // counter; returns the number taken (uniform) and this lane's position in *pos
    __device__ uint32_t reserve(const wave_call& w, int64_t* count, uint64_t* counter, uint64_t* pos) const
    {
        uint64_t base = 0;
        uint32_t run = 0;
        if (w.lane == w.leader)
        {
            const int64_t k = static_cast<int64_t>(w.k);
            const int64_t old = __hip_atomic_fetch_add(count, -k, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            const int64_t got = old <= 0 ? 0 : (old < k ? old : k);
            if (got < k) __hip_atomic_fetch_add(count, k - got, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (got > 0)
                base = __hip_atomic_fetch_add(counter, static_cast<uint64_t>(got), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            run = static_cast<uint32_t>(got);
        }
        run = __builtin_amdgcn_readlane(run, w.leader);
        base = static_cast<uint64_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(base), w.leader)) |
               (static_cast<uint64_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(base >> 32), w.leader)) << 32);
        *pos = base + w.rank;
        return run;
    }

    // waits until the other call on this slot is done (acquire: a pop then reads the element its push
    // released, a push writes after the previous pop of its slot released it)
    __device__ void wait_seq(uint64_t slot, uint32_t want) const
    {
        while (__hip_atomic_load(seq + slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want)
            __builtin_amdgcn_s_sleep(1);
    }
};

} // namespace htab_dev

#endif // STDGPU_AMD_DEQUE_DEVICE_HPP
