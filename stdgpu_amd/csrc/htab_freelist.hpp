// Free-list calls between a map and a deque<M> of slot indices: the reference's voxel-hashing caller
// pattern (pop a slot from the free list for each new block and insert (block, slot); erase the blocks
// that left the region of interest and push their slots back) as two bulk calls that need no host
// synchronisation and no compaction of keys on the host side:
//
//   * take: after a bulk insert wrote its per-element flags, every key the insert added receives the
//     next value from the front of the deque (deque::pop_front, deque_detail.cuh:390-448);
//   * evict: one pass over the slot array erases every entry whose key, read as a packed 3-D block
//     coordinate, lies farther than a radius from a point, and pushes its value onto the back of the
//     deque (deque::push_back, deque_detail.cuh:204-260).
//
// Both are bulk calls: like the deque's own bulk calls they run alone on the deque, so positions are
// handed out by rank, not claimed slot by slot.  A one-thread kernel snapshots the end of the deque the
// call works on and how many elements (take) or free slots (evict) it may use; the main kernel ranks
// the calling threads of each workgroup with one atomic on a counter in the table's control block;
// a one-thread kernel moves the deque's counters by what was used.  The evict pass is an erase-only
// launch on the table and follows k_erase's rules (an erased slot becomes EMPTY when its bucket holds
// an EMPTY slot or no key runs through it, otherwise TOMB; a key stored past its home bucket lowers
// the overflow counts of the buckets it ran through and its away count in the home bucket).
#ifndef STDGPU_AMD_HTAB_FREELIST_HPP
#define STDGPU_AMD_HTAB_FREELIST_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <stdgpu_amd/htab.h>
#include <stdgpu_amd/htab_device.hpp>

#include "wave.hpp"

namespace htab_dev
{

constexpr int kFreelistThreads = 1024;

// words of the table's control block used by one free-list call: [0] ranks handed out,
// [1] the deque position rank 0 maps to, [2] how many ranks the deque can serve (a template, as the
// other kernels of headers included by several translation units)
template <int = 0>
__global__ void
k_freelist_begin(htab_ring_view r, int64_t* w, int push)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t have = push ? *r.free_slots : *r.ready;
    w[0] = 0;
    w[1] = static_cast<int64_t>(push ? *r.end : *r.begin);
    w[2] = have < 0 ? 0 : have;
}

// Ranking: each workgroup owns a contiguous run of at most kFreelistTiles tiles of kFreelistThreads
// elements.  Pass 1 reads all of them at once (the tile loop is unrolled, so the loads of every tile
// are in flight together) and records which elements of each tile want a rank (one mask per
// wavefront per tile, in LDS); thread 0 turns the masks into prefix counts and takes the workgroup's
// run with one atomic on the call's counter; pass 2 works on each wanting element with the rank its
// position gives it.  So a call makes at most one counter atomic per workgroup, however sparse or
// dense the wanting elements are.
constexpr int kFreelistTiles = 8;
constexpr int kFreelistWaves = kFreelistThreads / 64;
// both kernels are built for 8 wavefronts per SIMD (two workgroups per CU): their work is latency of
// scattered loads, so residency is what hides it

struct rank_lds
{
    uint64_t mask[kFreelistTiles][kFreelistWaves];
    int64_t prefix[kFreelistTiles][kFreelistWaves];
    int64_t base;
    int64_t total;
};

// the first tile of this workgroup and its tile count, for `tiles` tiles in all
__device__ inline void
block_tiles(uint64_t tiles, uint64_t* t0, uint32_t* nt)
{
    const uint64_t per = (tiles + gridDim.x - 1) / gridDim.x;
    *t0 = static_cast<uint64_t>(blockIdx.x) * per;
    const uint64_t t1 = *t0 + per < tiles ? *t0 + per : tiles;
    *nt = *t0 < t1 ? static_cast<uint32_t>(t1 - *t0) : 0U;
}

// between the passes: prefix counts and the workgroup's base (every thread calls it)
__device__ inline void
rank_settle(rank_lds& L, uint32_t ntiles, int64_t* counter)
{
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int64_t run = 0;
        for (uint32_t t = 0; t < ntiles; ++t)
            for (int v = 0; v < kFreelistWaves; ++v)
            {
                L.prefix[t][v] = run;
                run += __builtin_popcountll(L.mask[t][v]);
            }
        L.total = run;
        L.base = run ? __hip_atomic_fetch_add(counter, run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    }
    __syncthreads();
}

// Dealing a wavefront's sparse work: the elements marked in the wave-uniform per-tile masks m are
// numbered in tile order, then lane order, and element c goes to lane c % 64 in round c / 64, so a
// wavefront with a few marked elements in each of its tiles runs their dependent memory accesses in
// one round instead of one round per tile.  dealt() gives this lane's element of round rd: its tile
// and the lane it belongs to.
__device__ inline bool
dealt(const uint64_t (&m)[kFreelistTiles], uint32_t rd, uint32_t* tile, uint32_t* src)
{
    uint32_t c = rd * 64U + static_cast<uint32_t>(lane_id());
#pragma unroll
    for (int u = 0; u < kFreelistTiles; ++u)
    {
        const uint32_t cnt = static_cast<uint32_t>(__builtin_popcountll(m[u]));
        if (c < cnt)
        {
            uint64_t mm = m[u];
            for (uint32_t i = 0; i < c; ++i) mm &= mm - 1; // the c-th marked lane of tile u
            *tile = static_cast<uint32_t>(u);
            *src = static_cast<uint32_t>(__builtin_ctzll(mm));
            return true;
        }
        c -= cnt;
    }
    return false;
}

// ---------------------------------------------------------------- take
// flags[i] == 1: keys[i] gets the value at deque position w[1] + rank (2 when the deque ran out
// first; 3 when keys[i] is not in the table, nothing popped for it).  The finds of a wavefront's
// flagged elements are dealt one per lane; the lane that finds an element also assigns it.
template <typename K, typename M>
__global__ __launch_bounds__(kFreelistThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_take_values(table_ref<K, M> t, const K* __restrict__ keys, uint8_t* __restrict__ flags, int64_t n,
              htab_ring_view r, int64_t* w)
{
    using value_type = typename slot_traits<K, M>::value_type;
    __shared__ rank_lds L;
    const uint64_t pos0 = static_cast<uint64_t>(w[1]);
    const int64_t avail = w[2];
    const uint64_t cap = static_cast<uint64_t>(r.capacity);
    const M* data = static_cast<const M*>(r.data);
    uint64_t t0;
    uint32_t nt;
    block_tiles((static_cast<uint64_t>(n) + kFreelistThreads - 1) / kFreelistThreads, &t0, &nt);
    const int wave = static_cast<int>(threadIdx.x) / 64;
    const int64_t i0 = static_cast<int64_t>(t0 * kFreelistThreads + threadIdx.x);
    const int64_t w0 = static_cast<int64_t>(t0 * kFreelistThreads) + wave * 64; // lane 0 of this wavefront, tile 0
    // the elements the insert added, one ballot per tile
    uint64_t m[kFreelistTiles];
#pragma unroll
    for (int u = 0; u < kFreelistTiles; ++u)
    {
        const int64_t i = i0 + static_cast<int64_t>(u) * kFreelistThreads;
        m[u] = ballot(static_cast<uint32_t>(u) < nt && i < n && flags[i] == 1);
    }
    if (lane_id() == 0)
    {
#pragma unroll
        for (int u = 0; u < kFreelistTiles; ++u) L.mask[u][wave] = 0;
    }
    // their finds, dealt one per lane; a found element is marked for ranking (an LDS or into its
    // tile's mask; the zeroing above precedes it in this wavefront's order) and kept as one word per
    // round: its slot << 11 | tile << 6 | lane (kNone: nothing found this round)
    constexpr uint64_t kNone = ~uint64_t(0);
    uint64_t found[kFreelistTiles];
#pragma unroll
    for (int rd = 0; rd < kFreelistTiles; ++rd)
    {
        found[rd] = kNone;
        uint32_t u = 0, src = 0;
        if (dealt(m, static_cast<uint32_t>(rd), &u, &src))
        {
            const int64_t i = w0 + static_cast<int64_t>(u) * kFreelistThreads + src;
            value_type* q = t.find(keys[i]);
            if (q == t.end_ptr())
            {
                flags[i] = 3;
            }
            else
            {
                found[rd] = (static_cast<uint64_t>(q - t.slots) << 11) | (u << 6) | src;
                atomicOr(reinterpret_cast<unsigned long long*>(&L.mask[u][wave]), 1ULL << src);
            }
        }
    }
    rank_settle(L, nt, w);
#pragma unroll
    for (int rd = 0; rd < kFreelistTiles; ++rd)
    {
        if (found[rd] == kNone) continue;
        const uint32_t u = static_cast<uint32_t>(found[rd] >> 6) & 31U, src = static_cast<uint32_t>(found[rd]) & 63U;
        const int64_t rank = L.base + L.prefix[u][wave] + __builtin_popcountll(L.mask[u][wave] & ((uint64_t(1) << src) - 1));
        if (rank < avail)
        {
            const uint64_t slot = (pos0 + static_cast<uint64_t>(rank)) % cap;
            t.slots[found[rd] >> 11].second = data[slot];
            r.state[slot] += 2; // full -> free: the call runs alone on the deque
        }
        else
        {
            flags[w0 + static_cast<int64_t>(u) * kFreelistThreads + src] = 2;
        }
    }
}

// pop_front of what k_take_values used
template <int = 0>
__global__ void
k_take_end(htab_ring_view r, const int64_t* w, int64_t* count)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t k = w[0] < w[2] ? w[0] : w[2];
    *r.begin += static_cast<uint64_t>(k);
    *r.ready -= k;
    *r.free_slots += k;
    if (count) *count = k;
}

// ---------------------------------------------------------------- evict
// a packed key's block centre is farther than g.radius from g.center (the same arithmetic as the
// host-side selection: centre = (coordinate + 0.5) * block_size per axis, squared distances summed
// x, y, z in that order, no fused multiply-add)
__device__ inline bool
far_block(const htab_block_region& g, uint64_t key)
{
    const uint64_t m = (uint64_t(1) << g.bits) - 1;
    const int64_t cx = static_cast<int64_t>((key >> (2 * g.bits)) & m) - g.offset;
    const int64_t cy = static_cast<int64_t>((key >> g.bits) & m) - g.offset;
    const int64_t cz = static_cast<int64_t>(key & m) - g.offset;
    const double x = __dmul_rn(__dadd_rn(static_cast<double>(cx), 0.5), g.block_size);
    const double y = __dmul_rn(__dadd_rn(static_cast<double>(cy), 0.5), g.block_size);
    const double z = __dmul_rn(__dadd_rn(static_cast<double>(cz), 0.5), g.block_size);
    const double dx = __dadd_rn(x, -g.center[0]);
    const double dy = __dadd_rn(y, -g.center[1]);
    const double dz = __dadd_rn(z, -g.center[2]);
    const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    return d2 > __dmul_rn(g.radius, g.radius);
}

template <typename K, typename M>
__global__ __launch_bounds__(kFreelistThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_evict_far(table_ref<K, M> t, htab_block_region g, htab_ring_view r, int64_t* w)
{
    __shared__ rank_lds L;
    const uint64_t pos0 = static_cast<uint64_t>(w[1]);
    const int64_t room = w[2];
    const uint64_t cap = static_cast<uint64_t>(r.capacity);
    M* data = static_cast<M*>(r.data);
    const uint64_t ns = t.nslots();
    uint64_t t0;
    uint32_t nt;
    block_tiles((ns + kFreelistThreads - 1) / kFreelistThreads, &t0, &nt);
    const int wave = static_cast<int>(threadIdx.x) / 64;
    const uint64_t s0 = t0 * kFreelistThreads + threadIdx.x;
    const uint64_t w0 = t0 * kFreelistThreads + static_cast<uint64_t>(wave) * 64; // lane 0 of this wavefront, tile 0
    // plain loads: nothing but this launch's erases changes the table meanwhile, and each slot is read
    // by one thread here and erased by at most one
    K k[kFreelistTiles];
#pragma unroll
    for (int u = 0; u < kFreelistTiles; ++u)
    {
        const uint64_t s = s0 + static_cast<uint64_t>(u) * kFreelistThreads;
        k[u] = (static_cast<uint32_t>(u) < nt && s < ns) ? *t.key_ptr(s) : empty_key<K>();
    }
    uint64_t m[kFreelistTiles];
#pragma unroll
    for (int u = 0; u < kFreelistTiles; ++u)
    {
        m[u] = ballot(!table_ref<K, M>::is_reserved(k[u]) && far_block(g, static_cast<uint64_t>(k[u])));
        if (static_cast<uint32_t>(u) < nt && lane_id() == 0) L.mask[u][wave] = m[u];
    }
    rank_settle(L, nt, w);
    if (threadIdx.x == 0)
    {
        // this workgroup's erases: its ranks below the room
        int64_t e = room - L.base;
        e = e < 0 ? 0 : (e > L.total ? L.total : e);
        const uint32_t stripe = t.nstripes > 1 ? (blockIdx.x & (t.nstripes - 1)) : 0U;
        if (e)
            __hip_atomic_fetch_add(t.stripes + static_cast<uint64_t>(stripe) * kStripeStride, -e, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    // the erases of a wavefront's far slots, dealt one per lane
#pragma unroll
    for (int rd = 0; rd < kFreelistTiles; ++rd)
    {
        uint32_t u = 0, src = 0;
        if (!dealt(m, static_cast<uint32_t>(rd), &u, &src)) continue;
        const int64_t rank = L.base + L.prefix[u][wave] + __builtin_popcountll(m[u] & ((uint64_t(1) << src) - 1));
        if (rank >= room) continue;
        const uint64_t s = w0 + static_cast<uint64_t>(u) * kFreelistThreads + src;
        const K key = *t.key_ptr(s);
        const uint64_t b = s / t.W;
        // EMPTY when no key runs through the bucket (an EMPTY slot in it says so too), else TOMB; the
        // bucket's keys and count are loaded together
        const M v = t.slots[s].second;
        bool ends = t.overflow(b) == 0;
        for (uint32_t j0 = 0; j0 < t.W; j0 += table_ref<K, M>::kScanBatch)
        {
            K kk[table_ref<K, M>::kScanBatch];
            t.load_scan(b, 0, j0, kk);
#pragma unroll
            for (uint32_t u2 = 0; u2 < table_ref<K, M>::kScanBatch; ++u2) ends |= kk[u2] == empty_key<K>();
        }
        __hip_atomic_store(t.key_ptr(s), ends ? empty_key<K>() : tomb_key<K>(), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t h = home_bucket(static_cast<uint64_t>(key), t.nbuckets);
        if (h != b)
        {
            t.overflow_add(h, t.distance(h, b), -1);
            if (t.away_counters() != 0)
            {
                const uint32_t f = t.away_slot(key);
                if (t.away(h, f) < kAwaySaturated) t.away_add_fast(h, f, -1);
            }
        }
        const uint64_t slot = (pos0 + static_cast<uint64_t>(rank)) % cap;
        data[slot] = v;
        r.state[slot] += 2; // free -> full: the call runs alone on the deque
    }
}

// the two reserved keys (kept in the side slots) under the same rule, then push_back of what the
// call erased
template <typename K, typename M>
__global__ void
k_evict_end(table_ref<K, M> t, htab_block_region g, htab_ring_view r, int64_t* w, int64_t* count)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint64_t cap = static_cast<uint64_t>(r.capacity);
    M* data = static_cast<M*>(r.data);
    const K side[2] = { empty_key<K>(), tomb_key<K>() };
    for (int j = 0; j < 2; ++j)
    {
        const K k = side[j];
        const uint32_t bit = table_ref<K, M>::side_bit(k);
        if (!(load_agent(t.flags) & bit) || !far_block(g, static_cast<uint64_t>(k)) || w[0] >= w[2]) continue;
        __hip_atomic_fetch_and(t.flags, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t slot = (static_cast<uint64_t>(w[1]) + static_cast<uint64_t>(w[0])) % cap;
        data[slot] = t.slots[t.side_index(k)].second;
        r.state[slot] += 2;
        w[0] += 1;
        __hip_atomic_fetch_add(t.stripes, static_cast<int64_t>(-1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int64_t k = w[0] < w[2] ? w[0] : w[2];
    *r.end += static_cast<uint64_t>(k);
    *r.ready += k;
    *r.free_slots -= k;
    if (count) *count = k;
}

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_FREELIST_HPP
