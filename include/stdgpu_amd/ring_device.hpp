// Per-thread push and pop at either end of a ring of slots, called from user kernels as the
// reference's callers call them: the device side of stdgpu::deque<T> (src/stdgpu/deque.cuh:199-239;
// deque_detail.cuh:202-448 push_back / pop_back / push_front / pop_front) and of stdgpu::vector<T>
// (src/stdgpu/vector.cuh:219-239; vector_detail.cuh:181-302 push_back / pop_back), a ring that only
// grows and shrinks at its back and never wraps.  gfx950 only.
//
// Positions are 64-bit counters; position p lives in slot p mod capacity, and the elements are the
// positions [begin, end).  Two counts gate the calls: ready (elements whose push has published them)
// and free_slots (slots whose pop has released them).  The calling lanes of a wavefront reserve from
// one count with one fetch_sub (the shortfall is given back with one fetch_add), take that many
// consecutive positions with one more atomic on end (back) or begin (front), in lane order, and the
// lanes past the reservation report false, as the reference's calls do on a full or empty container.
// So no lane ever takes a position outside [0, capacity) of a vector, and the positions [begin, end)
// never span more than capacity slots.
//
// Each slot carries a state word whose low two bits are its phase: FREE, WRITING, FULL, READING.  A
// push claims its slot by compare-and-swap FREE -> WRITING, writes the element and releases FULL; a
// pop claims FULL -> READING, reads and releases FREE, each in one step of its retry loop (the word
// counts up by one per step, so every state is a new value).  The claim is the reference's per-slot lock with its occupancy check
// (deque_detail.cuh:221-248): pushes and pops of both ends may be mixed in one kernel, and when a
// push and a pop meet at one slot, whichever comes second waits for the first.  A reserved position
// is always one that a push of this or an earlier call has taken (a pop) or that a pop has given up
// (a push), so every wait ends as long as the call it waits for runs: on one slot the calls that have
// taken positions alternate push, pop, push, ... in the order they took them, so one of them can always
// claim; a call holds no claim while it waits; a call only waits inside itself, after taking its
// positions; and the lanes of one wavefront that make one call run it together, so a lane never waits
// on a diverged lane of its own wavefront that has not taken its position yet.
// Under concurrent calls of both ends, which element a pop receives depends on their order, as in the
// reference; the multiset of elements is always kept.
#ifndef STDGPU_AMD_RING_DEVICE_HPP
#define STDGPU_AMD_RING_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <stdgpu_amd/htab.h>

namespace htab_dev
{

constexpr uint32_t kSlotFree = 0;
constexpr uint32_t kSlotWriting = 1;
constexpr uint32_t kSlotFull = 2;
constexpr uint32_t kSlotReading = 3;

struct ring_ref
{
    char* data = nullptr;
    uint32_t* state = nullptr;
    uint64_t* begin = nullptr;
    uint64_t* end = nullptr;
    int64_t* ready = nullptr;
    int64_t* free_slots = nullptr;
    uint64_t capacity = 0;
    uint64_t elem_bytes = 0;

    __host__ __device__ static ring_ref from_view(const htab_ring_view& v)
    {
        ring_ref r;
        r.data = static_cast<char*>(v.data);
        r.state = v.state;
        r.begin = v.begin;
        r.end = v.end;
        r.ready = v.ready;
        r.free_slots = v.free_slots;
        r.capacity = static_cast<uint64_t>(v.capacity);
        r.elem_bytes = static_cast<uint64_t>(v.elem_bytes);
        return r;
    }

    __device__ static uint64_t load_counter(const uint64_t* c)
    {
        return __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // size() clamped to [0, capacity] as the reference's (deque_detail.cuh:482-512, vector_detail.cuh:454-512)
    __device__ int64_t size() const
    {
        const int64_t n = static_cast<int64_t>(load_counter(end) - load_counter(begin));
        return n < 0 ? 0 : (n > static_cast<int64_t>(capacity) ? static_cast<int64_t>(capacity) : n);
    }

    // the element at index n from the front (no synchronisation, as the reference's operator[])
    template <typename T>
    __device__ T* at(int64_t n) const
    {
        return element<T>((load_counter(begin) + static_cast<uint64_t>(n)) % capacity);
    }

    template <typename T>
    __device__ T* element(uint64_t slot) const
    {
        return reinterpret_cast<T*>(data + slot * elem_bytes);
    }

    template <typename T>
    __device__ bool push_back(const T& v) const
    {
        return push<T>(true, v);
    }
    template <typename T>
    __device__ bool push_front(const T& v) const
    {
        return push<T>(false, v);
    }
    template <typename T>
    __device__ bool pop_back(T* out) const
    {
        return pop<T>(true, out);
    }
    template <typename T>
    __device__ bool pop_front(T* out) const
    {
        return pop<T>(false, out);
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

    // the builtin returns a signed int: each half goes through uint32_t, or a low half of 2^31 and up
    // would sign-extend into the high half (a deque's positions below its start have such low halves)
    __device__ static uint64_t readlane64(uint64_t v, int lane)
    {
        const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), lane));
        const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), lane));
        return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
    }

    // Reserves up to k units from count and gives back the shortfall, then takes that many positions:
    // grow = true moves the counter outwards (push_back: end up, push_front: begin down), false inwards.
    // Returns the number taken (uniform) and this lane's position in *pos.
    __device__ uint32_t reserve(const wave_call& w, int64_t* count, uint64_t* counter, bool up, uint64_t* pos) const
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
            {
                const uint64_t g = static_cast<uint64_t>(got);
                base = __hip_atomic_fetch_add(counter, up ? g : 0 - g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            run = static_cast<uint32_t>(got);
        }
        run = __builtin_amdgcn_readlane(run, w.leader);
        base = readlane64(base, w.leader);
        // moving up, the lanes take base, base + 1, ...; moving down, base - 1, base - 2, ...
        *pos = up ? base + w.rank : base - 1 - w.rank;
        return run;
    }

    // The lanes with `mine` set claim their `slot` in phase `want` (FREE for a push, FULL for a pop),
    // run `body` (the copy) and release the next phase, all in one iteration of the retry loop, waiting
    // while the call that owns the slot finishes.  A lane never holds a claimed slot while it waits:
    // a claim held until the other lanes of its wavefront are done would make a lane of another
    // wavefront that needs this slot wait on them, and two wavefronts could wait on each other for
    // ever.  So the loop runs while any lane of the wavefront still waits (a uniform exit on a
    // ballot): with a per-lane exit, the compiler places the copy and the release on the exit path,
    // after the loop, where they run only once every lane of the wavefront has claimed.  Acquire on
    // the claim: a pop reads what its push released, a push writes after the previous pop of the slot
    // has read; release on the store.
    template <typename Body>
    __device__ void exchange(bool mine, uint64_t slot, uint32_t want, const Body& body) const
    {
        uint32_t* w = state + slot;
        bool waiting = mine;
        while (__builtin_amdgcn_ballot_w64(waiting) != 0)
        {
            if (waiting)
            {
                uint32_t s = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((s & 3U) == want &&
                    __hip_atomic_compare_exchange_strong(w, &s, s + 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT))
                {
                    body();
                    __hip_atomic_store(w, s + 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    waiting = false;
                }
            }
            if (__builtin_amdgcn_ballot_w64(waiting) != 0) __builtin_amdgcn_s_sleep(1);
        }
    }

    template <typename T>
    __device__ bool push(bool back, const T& v) const
    {
        const wave_call w = begin_call();
        uint64_t pos = 0;
        const uint32_t run = reserve(w, free_slots, back ? end : begin, back, &pos);
        const bool mine = w.rank < run;
        const uint64_t slot = mine ? pos % capacity : 0;
        exchange(mine, slot, kSlotFree, [&]() { *element<T>(slot) = v; });
        if (w.lane == w.leader && run > 0)
            __hip_atomic_fetch_add(ready, static_cast<int64_t>(run), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return mine;
    }

    template <typename T>
    __device__ bool pop(bool back, T* out) const
    {
        const wave_call w = begin_call();
        uint64_t pos = 0;
        const uint32_t run = reserve(w, ready, back ? end : begin, !back, &pos);
        const bool mine = w.rank < run;
        const uint64_t slot = mine ? pos % capacity : 0;
        exchange(mine, slot, kSlotFull, [&]() { *out = *element<T>(slot); });
        if (w.lane == w.leader && run > 0)
            __hip_atomic_fetch_add(free_slots, static_cast<int64_t>(run), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return mine;
    }
};

} // namespace htab_dev

#endif // STDGPU_AMD_RING_DEVICE_HPP
