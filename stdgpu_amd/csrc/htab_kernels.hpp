// Bulk kernels of the hashed containers: wavefront-cooperative probing on gfx950.
//
// A wavefront owns 64 consecutive input elements, one per lane.  Lanes are split into
// groups of G lanes; a group probes one bucket (16*G bytes) with a single 16-byte load per
// lane, so one vector-memory instruction probes 64/G buckets.  A group works through the G
// elements owned by its own lanes:
//   1. first probes: the home buckets of R of the group's elements (R = probe batch, a tuning
//      knob) are loaded back to back, so a wave keeps 64*R/G bucket lines in flight;
//   2. each element is resolved against its home bucket with __ballot votes: a key match,
//      an EMPTY slot or ovf == 0 (end of the probe sequence) or, for insert, a CAS on the
//      first free slot elected with ctz over the ballot mask (the R claims are issued back to
//      back, then their results read);
//   3. elements the home bucket does not resolve (overflow, lost CAS) continue one group
//      at a time along the probe sequence, from the next bucket when the home bucket's
//      findings are kept (find, erase, and inserts that did not lose a claim);
//   4. results return to the owning lane and leave in one coalesced store; size changes are
//      summed per workgroup into one atomic on one of 32 counter stripes.
//
// Reference call stacks these kernels replace (paths relative to the stdgpu source tree):
//   find   : unordered_base_detail.cuh:653-706 via a for_each_index functor
//   insert : unordered_base_detail.cuh:470-495, 708-809, 967-1017 (lock bits + excess stack)
//   erase  : unordered_base_detail.cuh:497-522, 811-956, 1019-1066
#ifndef STDGPU_AMD_HTAB_KERNELS_HPP
#define STDGPU_AMD_HTAB_KERNELS_HPP

#include <stdgpu_amd/htab_device.hpp>

#include <type_traits>

#include "wave.hpp"

namespace htab_dev
{

// ------------------------------------------------------------------ 16-byte lane chunk of a bucket
// A lane's chunk holds SPL = 16/sizeof(value_type) slots; (lo, hi) are its two 64-bit halves.
template <typename K, typename M>
struct chunk
{
    using traits = slot_traits<K, M>;
    using value_type = typename traits::value_type;
    static constexpr int SB = static_cast<int>(sizeof(value_type));
    static constexpr int SPL = 16 / SB;
    static_assert(SB == 4 || SB == 8 || SB == 16, "value_type must be 4, 8 or 16 bytes");

    uint64_t lo, hi;

    __device__ K key(int j) const
    {
        if constexpr (SB == 16)
            return static_cast<K>(lo);
        else if constexpr (SB == 8)
            return static_cast<K>(j == 0 ? lo : hi);
        else
        {
            const uint64_t w = (j < 2) ? lo : hi;
            return static_cast<K>((j & 1) ? (w >> 32) : w);
        }
    }

    // mapped value of slot j (maps only)
    __device__ uint64_t mapped(int j) const
    {
        if constexpr (SB == 16)
            return hi;
        else if constexpr (SB == 8)
            return (j == 0 ? lo : hi) >> (8 * sizeof(K));
        else
            return 0;
    }

    // plain (L1-cached) load: read-only phases
    __device__ static chunk load_plain(const value_type* bucket, int li)
    {
        const uint4 q = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(bucket) + 16 * li);
        chunk c;
        c.lo = static_cast<uint64_t>(q.x) | (static_cast<uint64_t>(q.y) << 32);
        c.hi = static_cast<uint64_t>(q.z) | (static_cast<uint64_t>(q.w) << 32);
        return c;
    }

    // coherent load of the key fields: phases that write the table
    __device__ static chunk load_keys(const value_type* bucket, int li)
    {
        const uint64_t* p = reinterpret_cast<const uint64_t*>(reinterpret_cast<const char*>(bucket) + 16 * li);
        chunk c;
        c.lo = load_agent(p);
        if constexpr (SB == 16)
            c.hi = 0;
        else
            c.hi = load_agent(p + 1);
        return c;
    }

    // in-bucket ovf count (first slot's padding), valid on the lane holding slot 0 of a plain load
    __device__ uint32_t ovf_word() const
    {
        static_assert(SB == 16, "the in-bucket count lives in 16-byte slots");
        return static_cast<uint32_t>(hi >> 32);
    }

    // bitmask over the chunk's slots j for which pred(key) holds
    template <typename Pred>
    __device__ uint32_t mask(Pred pred) const
    {
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < SPL; ++j)
            if (pred(key(j))) m |= 1U << j;
        return m;
    }
};

// The away counter of key k in the bucket whose chunks the group holds (in-bucket layouts with G > 1
// slots per bucket: lane j of the group holds slot j, whose padding word carries counters 2(j-1) and
// 2j-1).  Called with k uniform over the group.
template <typename K, typename M, int G>
__device__ __forceinline__ uint32_t
group_away(const chunk<K, M>& q, K k, int gbase)
{
    static_assert(G > 1, "a one-slot bucket has no away counters");
    const uint32_t f = away_index(static_cast<uint64_t>(k), 2U * (G - 1));
    const uint32_t w = shfl32(q.ovf_word(), gbase + 1 + static_cast<int>(f / 2));
    return (w >> (16 * (f & 1))) & 0xFFFFU;
}

template <typename K, typename M>
__device__ inline void
store_mapped(typename slot_traits<K, M>::value_type* p, uint64_t v)
{
    if constexpr (slot_traits<K, M>::is_map) p->second = static_cast<M>(v);
}

// per-workgroup sum of `delta` (one value per thread) -> one atomic on a counter stripe.
// Returns the stripe's previous value (only meaningful to callers that need the exact count).
__device__ inline int64_t
block_count_add(int64_t* stripes, uint32_t nstripes, int64_t delta, int64_t* s_wave, int64_t* s_old,
                int64_t* wave_prefix_out)
{
    const int lane = lane_id();
    const int wid = static_cast<int>(threadIdx.x) >> 6;
    const int nw = static_cast<int>(blockDim.x) >> 6;
    // wave-level sum via DPP-free shuffles (delta is 0/1 or 0/-1 in practice)
    int64_t v = delta;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += static_cast<int64_t>(shfl64(static_cast<uint64_t>(v), lane ^ off));
    if (lane == 0) s_wave[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int64_t total = 0;
        for (int w = 0; w < nw; ++w)
        {
            const int64_t c = s_wave[w];
            s_wave[w] = total; // exclusive prefix
            total += c;
        }
        int64_t old = 0;
        if (total != 0)
        {
            const uint32_t s = (nstripes == 1) ? 0U : (static_cast<uint32_t>(blockIdx.x) & (nstripes - 1));
            old = __hip_atomic_fetch_add(stripes + static_cast<uint64_t>(s) * kStripeStride, total, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        }
        *s_old = old;
    }
    __syncthreads();
    *wave_prefix_out = s_wave[wid];
    return *s_old;
}

__device__ inline int64_t
block_exclusive_scan(int64_t v, int64_t* s_tmp, int64_t* total)
{
    // LDS-staged scan: wave-level inclusive scan with shuffles, then wave totals in LDS
    const int lane = lane_id();
    const int wid = static_cast<int>(threadIdx.x) >> 6;
    const int nw = static_cast<int>(blockDim.x) >> 6;
    int64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const int64_t y = static_cast<int64_t>(shfl64(static_cast<uint64_t>(x), lane >= off ? lane - off : lane));
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int64_t run = 0;
        for (int w = 0; w < nw; ++w)
        {
            const int64_t c = s_tmp[w];
            s_tmp[w] = run;
            run += c;
        }
        s_tmp[nw] = run;
    }
    __syncthreads();
    const int64_t res = s_tmp[wid] + x - v;
    *total = s_tmp[nw];
    __syncthreads();
    return res;
}

// ================================================================== find
template <typename K, typename M, int G, int R>
__device__ __forceinline__ void
find_body(const table_ref<K, M>& t, const K* __restrict__ keys, const int64_t i, const bool valid,
          uint8_t* __restrict__ hits, void* __restrict__ vals_out)
{
    using C = chunk<K, M>;
    using value_type = typename C::value_type;
    constexpr int SPL = C::SPL;
    const int lane = lane_id();
    const int gbase = lane & ~(G - 1);
    const int li = lane & (G - 1);
    const uint64_t gmask = ((G == 64) ? ~0ULL : ((1ULL << G) - 1)) << gbase;
    const uint32_t W = static_cast<uint32_t>(G * SPL);
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    constexpr bool kAway = kInline && G > 1;

    const K mykey = valid ? keys[i] : empty_key<K>();
    bool myhit = false;
    uint64_t myval = 0;
    const bool reserved = valid && table_ref<K, M>::is_reserved(mykey);
    if (reserved)
    {
        if (__hip_atomic_load(t.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & table_ref<K, M>::side_bit(mykey))
        {
            myhit = true;
            if constexpr (slot_traits<K, M>::is_map) myval = static_cast<uint64_t>(t.slots[t.side_index(mykey)].second);
        }
    }
    const bool probe = valid && !reserved;
    uint32_t pend = 0; // group-uniform: elements not resolved by their home bucket

#pragma unroll
    for (int rb = 0; rb < G; rb += R)
    {
        C q[R];
        K kr[R];
        bool act[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            kr[r] = shfl(mykey, gbase + rb + r);
            act[r] = shfl32(probe ? 1U : 0U, gbase + rb + r) != 0;
            if (act[r])
            {
                const uint64_t b = home_bucket(static_cast<uint64_t>(kr[r]), t.nbuckets);
                q[r] = C::load_plain(t.slots + b * W, li);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            if (act[r])
            {
                const K k = kr[r];
                const uint32_t m = q[r].mask([k](K x) { return x == k; });
                const uint32_t e = q[r].mask([](K x) { return x == empty_key<K>(); });
                const uint64_t bm = ballot(m != 0) & gmask;
                const uint64_t be = ballot(e != 0) & gmask;
                if (bm)
                {
                    const int src = ctz64(bm);
                    const uint64_t cand = m ? q[r].mapped(__builtin_ctz(m)) : 0;
                    const uint64_t v = shfl64(cand, src);
                    if (li == rb + r)
                    {
                        myhit = true;
                        myval = v;
                    }
                }
                else if (!be)
                {
                    // no key runs past a home bucket with ovf == 0 (read with the bucket when in-bucket),
                    // and none of this bucket's keys with k's fingerprint when its away counter is 0
                    if constexpr (kInline)
                    {
                        bool more = shfl32(li == 0 ? q[r].ovf_word() : 0U, gbase) != 0;
                        if constexpr (kAway)
                        {
                            if (more) more = group_away<K, M, G>(q[r], k, gbase) != 0;
                        }
                        if (more) pend |= 1U << (rb + r);
                    }
                    else
                    {
                        pend |= 1U << (rb + r);
                    }
                }
            }
        }
    }

    // overflowed elements: walk the probe sequence one group element at a time.  With the count
    // in the bucket, bucket b's own count decides whether the walk goes on past it; with the count
    // array, bucket b is loaded together with ovf of the bucket before it (p), both in flight at once.
    int cur = -1;
    K ck = 0;
    uint64_t b = 0, p = 0, steps = 0;
    while (true)
    {
        if (cur < 0 && pend)
        {
            cur = __builtin_ctz(pend);
            pend &= pend - 1;
            ck = shfl(mykey, gbase + cur);
            p = home_bucket(static_cast<uint64_t>(ck), t.nbuckets);
            b = next_bucket(p, t.nbuckets);
            steps = 1;
        }
        const bool active = cur >= 0;
        if (ballot(active) == 0) break;
        if (active)
        {
            const C q = C::load_plain(t.slots + b * W, li);
            uint32_t ov = 0;
            if constexpr (kInline)
                ov = shfl32(li == 0 ? q.ovf_word() : 0U, gbase); // bucket b's own count
            else
                ov = shfl32(li == 0 ? t.ovf[p] : 0U, gbase); // count of the bucket before b
            const K k = ck;
            const uint32_t m = q.mask([k](K x) { return x == k; });
            const uint32_t e = q.mask([](K x) { return x == empty_key<K>(); });
            const uint64_t bm = ballot(m != 0) & gmask;
            const uint64_t be = ballot(e != 0) & gmask;
            if (!kInline && ov == 0)
            {
                cur = -1; // no key runs past p: absent
            }
            else if (bm)
            {
                const int src = ctz64(bm);
                const uint64_t cand = m ? q.mapped(__builtin_ctz(m)) : 0;
                const uint64_t v = shfl64(cand, src);
                if (li == cur)
                {
                    myhit = true;
                    myval = v;
                }
                cur = -1;
            }
            else if (be || (kInline && ov == 0) || ++steps >= t.nbuckets)
            {
                cur = -1;
            }
            else
            {
                p = b;
                b = next_bucket(b, t.nbuckets);
            }
        }
    }

    if (valid)
    {
        if (hits) hits[i] = myhit ? 1 : 0;
        if constexpr (slot_traits<K, M>::is_map)
        {
            if (vals_out && myhit) static_cast<M*>(vals_out)[i] = static_cast<M>(myval);
        }
    }
}

// ================================================================== insert
// Elements come either as AoS value_type (`values`) or as split arrays (`keys`, `vals`).
template <typename K, typename M>
__device__ inline void
fetch_element(const void* values, const void* keys, const void* vals, int64_t i, K* k, uint64_t* v)
{
    using value_type = typename slot_traits<K, M>::value_type;
    if (values)
    {
        const value_type e = static_cast<const value_type*>(values)[i];
        *k = slot_traits<K, M>::key(e);
        if constexpr (slot_traits<K, M>::is_map)
            *v = static_cast<uint64_t>(e.second);
        else
            *v = 0;
    }
    else
    {
        *k = static_cast<const K*>(keys)[i];
        if constexpr (slot_traits<K, M>::is_map)
            *v = vals ? static_cast<uint64_t>(static_cast<const M*>(vals)[i]) : 0;
        else
            *v = 0;
    }
}

constexpr uint64_t kNoSlot = ~0ULL;

// What the home bucket showed, kept by an element's own lane when its probe sequence runs on past the
// home bucket, so that the walk starts at the next bucket instead of loading the home bucket again:
// bits 0-7 the index in the bucket of its first free slot in the key's claim order, then flags.
constexpr uint32_t kWalkSet = 1U << 8;     // the walk starts at the home bucket's successor
constexpr uint32_t kWalkFree = 1U << 9;    // the home bucket has a free slot (index in bits 0-7)
constexpr uint32_t kWalkTomb = 1U << 10;   // that slot held TOMB (else EMPTY)
constexpr uint32_t kWalkAbsent = 1U << 11; // no key runs past the home bucket: the key is absent
constexpr uint32_t kWalkAwayHigh = 1U << 12; // the key's away counter there was seen at kAwayFastLimit or above

// The claim order of a key inside a bucket (claim_offset, kOrderedClaimBuckets): htab_device.hpp.

// bits j of a lane's chunk whose bucket slot li*SPL + j lies in [lo, hi)
template <int SPL>
__device__ __forceinline__ uint32_t
slot_range_mask(int li, uint32_t lo, uint32_t hi)
{
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < SPL; ++j)
    {
        const uint32_t x = static_cast<uint32_t>(li * SPL + j);
        if (x >= lo && x < hi) m |= 1U << j;
    }
    return m;
}

// The group's first free slot in the order starting at o, from order position jmin on: returns the
// lane that holds it (-1 if none) and, in *fsel, that lane's candidate bits (the lowest one is it).
// Positions >= jmin are the slots [o + jmin, W) followed by [max(o + jmin - W, 0), o).
template <int SPL>
__device__ __forceinline__ int
first_free_in_order(uint32_t f, uint32_t o, uint32_t jmin, uint32_t W, int li, uint64_t gmask, uint32_t* fsel)
{
    const uint32_t a1 = o + jmin;
    const uint32_t fa = a1 < W ? (f & slot_range_mask<SPL>(li, a1, W)) : 0U;
    const uint64_t ba = ballot(fa != 0) & gmask;
    if (ba)
    {
        *fsel = fa;
        return ctz64(ba);
    }
    if (o == 0)
    {
        *fsel = 0; // plain first-free order: nothing before the start
        return -1;
    }
    const uint32_t fb = f & slot_range_mask<SPL>(li, a1 > W ? a1 - W : 0U, o);
    const uint64_t bb = ballot(fb != 0) & gmask;
    *fsel = fb;
    return bb ? ctz64(bb) : -1;
}

template <typename K, typename M, int G, int R>
__device__ __forceinline__ void
insert_body(const table_ref<K, M>& t, const void* __restrict__ values, const void* __restrict__ keys,
            const void* __restrict__ vals, const int64_t i, const bool valid, uint8_t* __restrict__ inserted,
            const bool striped = false)
{
    using C = chunk<K, M>;
    using value_type = typename C::value_type;
    constexpr int SPL = C::SPL;
    __shared__ int64_t s_wave[16];
    __shared__ int64_t s_old;
    const int lane = lane_id();
    const int gbase = lane & ~(G - 1);
    const int li = lane & (G - 1);
    const uint64_t gmask = ((G == 64) ? ~0ULL : ((1ULL << G) - 1)) << gbase;
    const uint32_t W = static_cast<uint32_t>(G * SPL);
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    constexpr bool kAway = kInline && G > 1;

    K mykey = empty_key<K>();
    uint64_t myv = 0;
    if (valid) fetch_element<K, M>(values, keys, vals, i, &mykey, &myv);
    const bool ordered = t.nbuckets < kOrderedClaimBuckets; // uniform over the launch

    // A table whose limit is below its slots counts its size on one stripe and ranks each insert
    // against the limit, unless the launch was planned as unable to reach it (striped: counts go to
    // all stripes, folded back into the first after the launch).  A table at its exact limit refuses
    // every insert (reference: the !full() loop guard).
    const bool exact = (t.nstripes == 1) && !striped;
    const bool at_limit = exact && (load_agent(t.stripes) >= static_cast<int64_t>(t.limit));

    bool myins = false;
    uint64_t myslot = kNoSlot;
    const bool reserved = valid && table_ref<K, M>::is_reserved(mykey);
    if (reserved && !at_limit)
    {
        const uint32_t bit = table_ref<K, M>::side_bit(mykey);
        const uint32_t old = __hip_atomic_fetch_or(t.flags, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(old & bit))
        {
            myslot = t.side_index(mykey);
            store_mapped<K, M>(t.slots + myslot, myv);
            myins = true;
        }
    }
    const bool probe = valid && !reserved && !at_limit;
    uint32_t pend = 0;
    uint32_t mywalk = 0; // kWalk* for this lane's element

#pragma unroll
    for (int rb = 0; rb < G; rb += R)
    {
        C q[R];
        K kr[R];
        bool act[R];
        // With up to 4 elements in flight the home buckets are recomputed from the keys where they
        // are needed rather than kept in registers across the claims (k_insert<.., 8, 4> then stays
        // within 64 VGPRs, 8 wavefronts per SIMD); with more, the kernel is past that budget anyway
        // and keeping them saves the arithmetic.
        constexpr bool kKeepHome = R > 4;
        uint32_t hbk[kKeepHome ? R : 1];
        const auto home_of = [&](int r) -> uint64_t {
            if constexpr (kKeepHome)
                return hbk[r];
            else
                return home_bucket(static_cast<uint64_t>(kr[r]), t.nbuckets);
        };
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            kr[r] = shfl(mykey, gbase + rb + r);
            act[r] = shfl32(probe ? 1U : 0U, gbase + rb + r) != 0;
            if constexpr (kKeepHome) hbk[r] = static_cast<uint32_t>(home_bucket(static_cast<uint64_t>(kr[r]), t.nbuckets));
            if (act[r])
            {
                // a stale view is harmless here: a slot seen free but taken fails its CAS, and an
                // insert-only launch never frees a slot (the in-bucket count comes with the load)
                q[r] = C::load_plain(t.slots + home_of(r) * W, li);
            }
        }
        // Decide every element's claim first, then issue the R compare-and-swaps together so that
        // they are in flight at once, then act on their results.  Byte r of flp (uniform in the group)
        // is the lane that claims for element r (0xFF: none); that lane keeps the slot's index in the
        // bucket in byte r of sjp and, in bit r of tombp, whether the free slot it saw held TOMB rather
        // than EMPTY.  Packed so that R claims in flight cost a few registers, not 3R.
        using pack_t = std::conditional_t<(R > 4), uint64_t, uint32_t>;
        pack_t flp = ~pack_t(0), sjp = 0;
        uint32_t tombp = 0;
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            if (act[r])
            {
                const K k = kr[r];
                const uint32_t m = q[r].mask([k](K x) { return x == k; });
                const uint32_t e = q[r].mask([](K x) { return x == empty_key<K>(); });
                const uint32_t f = q[r].mask([](K x) { return x == empty_key<K>() || x == tomb_key<K>(); });
                const uint64_t bm = ballot(m != 0) & gmask;
                const uint64_t be = ballot(e != 0) & gmask;
                const uint64_t bf = ballot(f != 0) & gmask;
                bool ends_home = be != 0;
                uint32_t ovh = 1;
                if constexpr (kInline)
                {
                    // no key runs past a home bucket with ovf == 0: a TOMB there is free to take.  The
                    // same holds for k alone when the away counter of its fingerprint is 0 (no key of
                    // this bucket with that fingerprint is stored further on), so ovh then reads 0 too.
                    ovh = shfl32(li == 0 ? q[r].ovf_word() : 0U, gbase);
                    ends_home = ends_home || (bf != 0 && ovh == 0);
                }
                if (bm)
                {
                    // key present: nothing to do (insert never overwrites)
                }
                else if (ends_home)
                {
                    // probe sequence ends in the home bucket: claim its first free slot in the key's order
                    uint32_t fsel = f;
                    int fl = 0;
                    if (ordered)
                        fl = first_free_in_order<SPL>(f, claim_offset(static_cast<uint64_t>(k), W), 0, W, li, gmask,
                                                      &fsel);
                    else
                        fl = ctz64(bf);
                    flp = (flp & ~(pack_t(0xFF) << (8 * r))) | (static_cast<pack_t>(fl) << (8 * r));
                    if (lane == fl)
                    {
                        const int j = __builtin_ctz(fsel);
                        sjp |= static_cast<pack_t>(li * SPL + j) << (8 * r);
                        if (q[r].key(j) == tomb_key<K>()) tombp |= 1U << r;
                    }
                }
                else
                {
                    // the probe sequence runs on past the home bucket: the element's lane keeps what
                    // this bucket showed, and the walk starts at the next bucket
                    uint32_t info = kWalkSet | ((kInline && ovh == 0) ? kWalkAbsent : 0U);
                    if constexpr (kAway)
                    {
                        if (ovh != 0)
                        {
                            const uint32_t aw = group_away<K, M, G>(q[r], k, gbase);
                            if (aw == 0) info |= kWalkAbsent;
                            if (aw >= kAwayFastLimit) info |= kWalkAwayHigh;
                        }
                    }
                    if (bf)
                    {
                        uint32_t fsel = f;
                        int fl = 0;
                        if (ordered)
                            fl = first_free_in_order<SPL>(f, claim_offset(static_cast<uint64_t>(k), W), 0, W, li,
                                                          gmask, &fsel);
                        else
                            fl = ctz64(bf);
                        uint32_t part = 0;
                        if (lane == fl)
                        {
                            const int j = __builtin_ctz(fsel);
                            part = static_cast<uint32_t>(li * SPL + j) | kWalkFree |
                                   ((q[r].key(j) == tomb_key<K>()) ? kWalkTomb : 0U);
                        }
                        info |= shfl32(part, fl);
                        if (info & kWalkAbsent)
                        {
                            // absent past the home bucket and a free slot in it: claim there at once,
                            // with the other elements' claims (as when the probe sequence ends here)
                            flp = (flp & ~(pack_t(0xFF) << (8 * r))) | (static_cast<pack_t>(fl) << (8 * r));
                            if (lane == fl)
                            {
                                sjp |= static_cast<pack_t>(part & 0xFFU) << (8 * r);
                                if (part & kWalkTomb) tombp |= 1U << r;
                            }
                        }
                    }
                    if (!(info & kWalkAbsent) || !(info & kWalkFree))
                    {
                        if (li == rb + r) mywalk = info;
                        pend |= 1U << (rb + r);
                    }
                }
            }
        }
        K got[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            got[r] = 0;
            const int fl = static_cast<int>((flp >> (8 * r)) & 0xFFU);
            if (lane == fl)
            {
                const K o = ((tombp >> r) & 1U) ? tomb_key<K>() : empty_key<K>();
                got[r] = cas_agent(t.key_ptr(home_of(r) * W + static_cast<uint32_t>((sjp >> (8 * r)) & 0xFFU)), o, kr[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            const int fl = static_cast<int>((flp >> (8 * r)) & 0xFFU);
            if (fl != 0xFF)
            {
                // the value travels from the owner lane; the claiming lane stores it
                const uint64_t v = shfl64(myv, gbase + rb + r);
                const uint32_t j = static_cast<uint32_t>((sjp >> (8 * r)) & 0xFFU);
                uint64_t seen = 0;
                if (lane == fl)
                {
                    const K o = ((tombp >> r) & 1U) ? tomb_key<K>() : empty_key<K>();
                    seen = (got[r] == o) ? 1 : ((got[r] == kr[r]) ? 2 : 0);
                    if (seen == 1) store_mapped<K, M>(t.slots + home_of(r) * W + j, v);
                }
                seen = shfl64(seen, fl);
                if (seen == 1)
                {
                    const uint64_t slot = home_of(r) * W + shfl32(j, fl);
                    if (li == rb + r)
                    {
                        myins = true;
                        myslot = slot;
                    }
                }
                else if (seen == 0)
                {
                    pend |= 1U << (rb + r);
                }
            }
        }
    }

    // slow path: the probe sequence with first-free tracking, CAS, and resume after a lost CAS.
    // The key is absent once a bucket without it holds an EMPTY slot or has ovf == 0; it then
    // takes the first free slot (EMPTY or TOMB) seen, or the first one further on.  A key stored
    // past its home bucket raises ovf of every bucket it runs through.
    int cur = -1;
    K ck = 0;
    uint64_t cv = 0;
    uint64_t b = 0, hb = 0, walked = 0, ff = kNoSlot;
    K ffobs = 0;
    uint32_t jmin = 0, co = 0; // order position to resume from; the key's claim offset
    bool absent = false, fresh = false;
    bool away_fast = false; // the key's away counter in its home bucket was seen below kAwayFastLimit
    while (true)
    {
        if (cur < 0 && pend)
        {
            cur = __builtin_ctz(pend);
            pend &= pend - 1;
            ck = shfl(mykey, gbase + cur);
            cv = shfl64(myv, gbase + cur);
            hb = home_bucket(static_cast<uint64_t>(ck), t.nbuckets);
            co = ordered ? claim_offset(static_cast<uint64_t>(ck), W) : 0U;
            b = hb;
            walked = 0;
            ff = kNoSlot;
            jmin = 0;
            absent = false;
            fresh = false;
            // an element sent on by its home bucket (not one that lost a claim there) resumes after it
            const uint32_t info = shfl32(mywalk, gbase + cur);
            away_fast = (info & kWalkSet) && !(info & kWalkAwayHigh);
            if ((info & kWalkSet) && t.nbuckets > 1)
            {
                b = next_bucket(hb, t.nbuckets);
                walked = 1;
                absent = (info & kWalkAbsent) != 0;
                if (info & kWalkFree)
                {
                    ff = hb * W + (info & 0xFFU);
                    ffobs = (info & kWalkTomb) ? tomb_key<K>() : empty_key<K>();
                }
            }
        }
        const bool active = cur >= 0;
        if (ballot(active) == 0) break;
        if (active)
        {
            const K k = ck;
            // an element whose home bucket proved it absent and offered a free slot claims at once
            bool claim = absent && ff != kNoSlot;
            if (!claim)
            {
                // plain (L2-cacheable) loads until the element loses a CAS, coherent loads after that:
                // a stale view only shows taken slots as free or counts as too low, and every claim is
                // checked by its CAS against the value it saw
                C q;
                uint32_t ovl = 0;
                if (fresh)
                {
                    q = C::load_keys(t.slots + b * W, li);
                    ovl = li == 0 ? t.overflow(b) : 0U;
                }
                else
                {
                    q = C::load_plain(t.slots + b * W, li);
                    if constexpr (kInline)
                        ovl = li == 0 ? q.ovf_word() : 0U;
                    else
                        ovl = li == 0 ? t.ovf[b] : 0U;
                }
                const uint32_t ov = shfl32(ovl, gbase);
                // the whole bucket is checked for the key and for EMPTY slots (slots passed after a
                // lost CAS hold other keys for good); free slots count from order position jmin on
                const uint32_t m = q.mask([k](K x) { return x == k; });
                const uint32_t e = q.mask([](K x) { return x == empty_key<K>(); });
                const uint32_t f = q.mask([](K x) { return x == empty_key<K>() || x == tomb_key<K>(); });
                const uint64_t bm = ballot(m != 0) & gmask;
                const uint64_t be = ballot(e != 0) & gmask;
                uint32_t fsel = 0;
                const int fl = first_free_in_order<SPL>(f, co, jmin, W, li, gmask, &fsel);
                if (bm)
                {
                    cur = -1; // present
                }
                else
                {
                    if (ff == kNoSlot && fl >= 0)
                    {
                        uint64_t cand = 0, obs = 0;
                        if (lane == fl)
                        {
                            const int j = __builtin_ctz(fsel);
                            cand = b * W + static_cast<uint64_t>(li * SPL + j);
                            obs = static_cast<uint64_t>(q.key(j));
                        }
                        ff = shfl64(cand, fl);
                        ffobs = static_cast<K>(shfl64(obs, fl));
                    }
                    // no key runs past a bucket with ovf == 0, and a bucket holding an EMPTY slot has
                    // ovf == 0; nor is k stored past its home bucket when its away counter there is 0
                    if (be || ov == 0) absent = true;
                    if constexpr (kAway)
                    {
                        if (!absent && !fresh && b == hb && group_away<K, M, G>(q, k, gbase) == 0) absent = true;
                    }
                    claim = absent && ff != kNoSlot;
                    if (!claim)
                    {
                        b = next_bucket(b, t.nbuckets);
                        jmin = 0;
                        ++walked;
                        // every probe sequence is finite: give up after a full wrap, or early when the
                        // table has reached its limit
                        if (b == hb ||
                            ((walked & 63) == 0 && stripe_sum(t.stripes, t.nstripes) >= static_cast<int64_t>(t.limit)))
                        {
                            cur = -1;
                        }
                    }
                }
            }
            if (claim)
            {
                // claim ff (the group's first lane issues the CAS)
                uint64_t res = 0;
                if (li == 0)
                {
                    const K got = cas_agent(t.key_ptr(ff), ffobs, k);
                    res = (got == ffobs) ? 1 : ((got == k) ? 2 : 0);
                    if (res == 1) store_mapped<K, M>(t.slots + ff, cv);
                }
                res = shfl64(res, gbase);
                if (res == 1)
                {
                    // the key runs through every bucket from its home bucket to ff's
                    const uint64_t fb = ff / W;
                    const uint64_t dist = t.distance(hb, fb);
                    for (uint64_t d = static_cast<uint64_t>(li); d < dist; d += G)
                    {
                        uint64_t x = hb + d;
                        if (x >= t.nbuckets) x -= t.nbuckets;
                        __hip_atomic_fetch_add(t.ovf_ptr(x), 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    // stored past its home bucket: that bucket's away counter of k's fingerprint (a plain
                    // add, not waited for, when the counter was seen well below saturation)
                    if constexpr (kAway)
                    {
                        if (li == 0 && dist != 0)
                        {
                            if (away_fast)
                                t.away_add_fast(hb, t.away_slot(k), 1);
                            else
                                t.away_add(hb, t.away_slot(k), 1);
                        }
                    }
                    if (li == cur)
                    {
                        myins = true;
                        myslot = ff;
                    }
                    cur = -1;
                }
                else if (res == 2)
                {
                    cur = -1;
                }
                else
                {
                    // lost ff to another key: every slot before ff in the key's order holds another key
                    // (an insert-only launch never frees a slot), so the scan resumes right after ff; ff
                    // only moves forward, so this terminates
                    b = ff / W;
                    jmin = ((static_cast<uint32_t>(ff % W) + W - co) & (W - 1)) + 1;
                    ff = kNoSlot;
                    fresh = true;
                    if (jmin == W)
                    {
                        b = next_bucket(b, t.nbuckets);
                        jmin = 0;
                        if (b == hb) cur = -1; // wrapped around: no free slot left
                    }
                }
            }
        }
    }

    int64_t wave_prefix = 0;
    const int64_t old = block_count_add(t.stripes, exact ? 1U : kMaxStripes, myins ? 1 : 0, s_wave, &s_old, &wave_prefix);
    if (exact)
    {
        // exact limit: inserts ranked past it are undone (their slot becomes TOMB)
        const uint64_t bal = ballot(myins);
        const int64_t rank = old + wave_prefix + __builtin_popcountll(bal & ((1ULL << lane) - 1));
        if (myins && rank >= static_cast<int64_t>(t.limit))
        {
            if (myslot >= t.nslots())
            {
                __hip_atomic_fetch_and(t.flags, ~table_ref<K, M>::side_bit(mykey), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            else
            {
                __hip_atomic_store(t.key_ptr(myslot), tomb_key<K>(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t h = home_bucket(static_cast<uint64_t>(mykey), t.nbuckets);
                t.overflow_add(h, t.distance(h, myslot / W), -1);
                if constexpr (kAway)
                {
                    if (myslot / W != h) t.away_add(h, t.away_slot(mykey), -1);
                }
            }
            __hip_atomic_fetch_add(t.stripes, static_cast<int64_t>(-1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            myins = false;
        }
    }
    if (valid && inserted) inserted[i] = myins ? 1 : 0;
}

// ================================================================== erase
// Erase-only launch: a slot is reset to EMPTY when its bucket already holds an EMPTY slot (no
// probe sequence runs through such a bucket, and none can start to while only erases run),
// otherwise to TOMB so probe sequences through the bucket stay intact.  An erased key that was
// stored past its home bucket lowers ovf of the buckets it ran through.
template <typename K, typename M, int G, int R>
__device__ __forceinline__ void
erase_body(const table_ref<K, M>& t, const K* __restrict__ keys, const int64_t i, const bool valid,
           uint8_t* __restrict__ erased, const bool striped = false)
{
    using C = chunk<K, M>;
    constexpr int SPL = C::SPL;
    __shared__ int64_t s_wave[16];
    __shared__ int64_t s_old;
    const int lane = lane_id();
    const int gbase = lane & ~(G - 1);
    const int li = lane & (G - 1);
    const uint64_t gmask = ((G == 64) ? ~0ULL : ((1ULL << G) - 1)) << gbase;
    const uint32_t W = static_cast<uint32_t>(G * SPL);
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    constexpr bool kAway = kInline && G > 1;

    const K mykey = valid ? keys[i] : empty_key<K>();
    bool mydel = false;
    bool mysat = false; // the key's away counter in its home bucket was seen saturated
    const bool reserved = valid && table_ref<K, M>::is_reserved(mykey);
    if (reserved)
    {
        const uint32_t bit = table_ref<K, M>::side_bit(mykey);
        const uint32_t old = __hip_atomic_fetch_and(t.flags, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mydel = (old & bit) != 0;
    }
    const bool probe = valid && !reserved;
    uint32_t pend = 0;

#pragma unroll
    for (int rb = 0; rb < G; rb += R)
    {
        C q[R];
        K kr[R];
        bool act[R];
        uint64_t hb[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            kr[r] = shfl(mykey, gbase + rb + r);
            act[r] = shfl32(probe ? 1U : 0U, gbase + rb + r) != 0;
            hb[r] = 0;
            if (act[r])
            {
                hb[r] = home_bucket(static_cast<uint64_t>(kr[r]), t.nbuckets);
                // a stale view is harmless in an erase-only launch: keys and counts only go away,
                // a key seen but already erased fails its CAS, and a stale count is never too low
                q[r] = C::load_plain(t.slots + hb[r] * W, li);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            if (act[r])
            {
                const K k = kr[r];
                const uint32_t m = q[r].mask([k](K x) { return x == k; });
                const uint32_t e = q[r].mask([](K x) { return x == empty_key<K>(); });
                const uint64_t bm = ballot(m != 0) & gmask;
                const uint64_t be = ballot(e != 0) & gmask;
                // no probe sequence runs through the home bucket: it ends here, and the erased
                // slot may become EMPTY
                bool ends_home = be != 0;
                if constexpr (kInline) ends_home = ends_home || shfl32(li == 0 ? q[r].ovf_word() : 0U, gbase) == 0;
                if (bm)
                {
                    const int ml = ctz64(bm);
                    uint64_t ok = 0;
                    if (lane == ml)
                    {
                        const uint64_t slot = hb[r] * W + static_cast<uint64_t>(li * SPL + __builtin_ctz(m));
                        const K repl = ends_home ? empty_key<K>() : tomb_key<K>();
                        ok = (cas_agent(t.key_ptr(slot), k, repl) == k) ? 1 : 0;
                    }
                    ok = shfl64(ok, ml);
                    if (ok && li == rb + r) mydel = true;
                }
                else if (!ends_home)
                {
                    // k is stored further on only if its away counter in the home bucket says so
                    bool more = true;
                    if constexpr (kAway)
                    {
                        const uint32_t aw = group_away<K, M, G>(q[r], k, gbase);
                        more = aw != 0;
                        if (li == rb + r) mysat = aw >= kAwaySaturated;
                    }
                    if (more) pend |= 1U << (rb + r);
                }
            }
        }
    }

    // overflowed elements: as in find, bucket b is loaded with ovf of the bucket before it; an
    // erased key stops running through the buckets between its home bucket and its own
    int cur = -1;
    K ck = 0;
    uint64_t b = 0, p = 0, hb = 0, steps = 0;
    bool sat = false;
    while (true)
    {
        if (cur < 0 && pend)
        {
            cur = __builtin_ctz(pend);
            pend &= pend - 1;
            ck = shfl(mykey, gbase + cur);
            sat = shfl32(mysat ? 1U : 0U, gbase + cur) != 0;
            hb = home_bucket(static_cast<uint64_t>(ck), t.nbuckets);
            p = hb;
            b = next_bucket(hb, t.nbuckets);
            steps = 1;
        }
        const bool active = cur >= 0;
        if (ballot(active) == 0) break;
        if (active)
        {
            // plain loads: in an erase-only launch a stale view only shows a key that is already gone
            // (its CAS fails), too few EMPTY slots or too high a count (both keep TOMB and walk on)
            const C q = C::load_plain(t.slots + b * W, li);
            // in-bucket layout: bucket b's own count; count array: the count of the bucket before b
            uint32_t ovl = 0;
            if constexpr (kInline)
                ovl = li == 0 ? q.ovf_word() : 0U;
            else
                ovl = li == 0 ? t.ovf[p] : 0U;
            const uint32_t ov = shfl32(ovl, gbase);
            const K k = ck;
            const uint32_t m = q.mask([k](K x) { return x == k; });
            const uint32_t e = q.mask([](K x) { return x == empty_key<K>(); });
            const uint64_t bm = ballot(m != 0) & gmask;
            const uint64_t be = ballot(e != 0) & gmask;
            if (!kInline && ov == 0)
            {
                cur = -1; // no key runs past p: absent
            }
            else if (bm)
            {
                const int ml = ctz64(bm);
                uint64_t ok = 0;
                if (lane == ml)
                {
                    const uint64_t slot = b * W + static_cast<uint64_t>(li * SPL + __builtin_ctz(m));
                    const K repl = (be || (kInline && ov == 0)) ? empty_key<K>() : tomb_key<K>();
                    ok = (cas_agent(t.key_ptr(slot), k, repl) == k) ? 1 : 0;
                }
                ok = shfl64(ok, ml);
                if (ok)
                {
                    const uint64_t dist = t.distance(hb, b);
                    for (uint64_t d = static_cast<uint64_t>(li); d < dist; d += G)
                    {
                        uint64_t x = hb + d;
                        if (x >= t.nbuckets) x -= t.nbuckets;
                        __hip_atomic_fetch_add(t.ovf_ptr(x), static_cast<uint32_t>(-1), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    }
                    // b != hb here: lower the home bucket's away counter with a plain atomic, unless it
                    // was seen saturated (an erase-only launch only lowers counters, so it still is)
                    if constexpr (kAway)
                    {
                        if (li == 0 && !sat) t.away_add_fast(hb, t.away_slot(k), -1);
                    }
                    if (li == cur) mydel = true;
                }
                cur = -1;
            }
            else if (be || (kInline && ov == 0) || ++steps >= t.nbuckets)
            {
                cur = -1;
            }
            else
            {
                p = b;
                b = next_bucket(b, t.nbuckets);
            }
        }
    }

    int64_t wave_prefix = 0;
    // striped: a one-stripe table's launch counts on all stripes (folded back after the launch)
    block_count_add(t.stripes, striped ? kMaxStripes : t.nstripes, mydel ? -1 : 0, s_wave, &s_old, &wave_prefix);
    if (valid && erased) erased[i] = mydel ? 1 : 0;
}

// ================================================================== plain bulk kernels
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(1024) void
k_find(table_ref<K, M> t, const K* __restrict__ keys, int64_t n, uint8_t* __restrict__ hits,
       void* __restrict__ vals_out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    find_body<K, M, G, R>(t, keys, i, i < n, hits, vals_out);
}

// plan: null for tables counted on all stripes; for a one-stripe table, the word k_limit_plan wrote
// (1: this launch cannot reach the limit, count on all stripes without ranking)
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(1024) void
k_insert(table_ref<K, M> t, const void* __restrict__ values, const void* __restrict__ keys,
         const void* __restrict__ vals, int64_t n, uint8_t* __restrict__ inserted, const uint32_t* __restrict__ plan)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool striped = plan != nullptr && *plan != 0;
    insert_body<K, M, G, R>(t, values, keys, vals, i, i < n, inserted, striped);
}

// striped: a one-stripe table counts this launch on all stripes (folded back by k_limit_fold)
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(1024) void
k_erase(table_ref<K, M> t, const K* __restrict__ keys, int64_t n, uint8_t* __restrict__ erased, int striped)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    erase_body<K, M, G, R>(t, keys, i, i < n, erased, striped != 0);
}

// A one-stripe table (limit below its slots) ranks every insert against the limit on one counter,
// which serialises one atomic per workgroup on one address.  Before a bulk insert, this decides
// whether the launch can reach the limit at all (size + n > limit); when it cannot, the launch counts
// on all stripes like any other table.
template <typename K, typename M>
__global__ void
k_limit_plan(table_ref<K, M> t, int64_t n, uint32_t* plan)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *plan = (load_agent(t.stripes) + n <= static_cast<int64_t>(t.limit)) ? 1U : 0U;
}

// after a striped launch on a one-stripe table: stripes 1.. are added into stripe 0 and cleared
template <typename K, typename M>
__global__ void
k_limit_fold(table_ref<K, M> t)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t sum = 0;
    for (uint32_t s = 1; s < kMaxStripes; ++s)
    {
        int64_t* p = t.stripes + static_cast<uint64_t>(s) * kStripeStride;
        sum += *p;
        *p = 0;
    }
    *t.stripes += sum;
}


// ================================================================== batched: many tables, one launch
// A launch works through segments; segment s applies operation op to n elements of table `table`,
// read from / written to its own device arrays (keys, vals, flags = inserted / hit / erased,
// vals_out = mapped values of find hits).  Work is cut into chunks of B elements (the workgroup size) that
// never straddle two segments; chunk_start[s] counts the chunks of segments before s.  A persistent
// grid of B-thread workgroups takes chunks of B elements in a fixed order, workgroup w the chunks
// w, w + grid, w + 2 grid, ..., and exits past the last one.  Chunks are alike in size, so the fixed
// order balances as well as a shared work counter did, without one same-address atomic per chunk
// (those serialise: the counter cost about 9 % of the 4096-table phases).  Segments of one table in
// one launch must share their operation.
constexpr int kOpInsert = 0;
constexpr int kOpFind = 1;
constexpr int kOpErase = 2;

struct batch_segment
{
    const void* keys;
    const void* vals;
    uint8_t* flags;
    void* vals_out;
    int64_t n;
    int32_t table;
    int32_t op;
};

// chunk prefix over segments (one workgroup)
__device__ inline void
plan_chunks(const batch_segment* __restrict__ segs, int32_t nsegs, int chunk, int64_t* __restrict__ chunk_start)
{
    __shared__ int64_t s_tmp[17];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nsegs; base += blockDim.x)
    {
        const int64_t t = base + threadIdx.x;
        int64_t c = 0;
        if (t < nsegs)
        {
            const int64_t cnt = segs[t].n;
            c = cnt > 0 ? (cnt + chunk - 1) / chunk : 0;
        }
        int64_t total = 0;
        const int64_t ex = block_exclusive_scan(c, s_tmp, &total);
        if (t < nsegs) chunk_start[t] = s_carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_start[nsegs] = s_carry;
}

// a template so every translation unit may instantiate it (no duplicate symbols)
template <int Unused = 0>
__global__ __launch_bounds__(1024) void
k_batch_plan(const batch_segment* __restrict__ segs, int32_t nsegs, int chunk, int64_t* __restrict__ chunk_start)
{
    plan_chunks(segs, nsegs, chunk, chunk_start);
}

// segments from a per-table work-offset array: segment t = table t, elements [offsets[t], offsets[t+1])
// of the flattened arrays, operation ops[t] (or op)
// a template so every translation unit may instantiate it (no duplicate symbols)
template <int Unused = 0>
__global__ __launch_bounds__(1024) void
k_batch_plan_offsets(const int64_t* __restrict__ offsets, int32_t ntables, const int8_t* __restrict__ ops, int op,
                     const char* keys, const char* vals, uint8_t* flags, char* vals_out, int key_bytes,
                     int val_bytes, int chunk, batch_segment* __restrict__ segs,
                     int64_t* __restrict__ chunk_start)
{
    for (int32_t t = threadIdx.x; t < ntables; t += blockDim.x)
    {
        const int64_t a = offsets[t];
        batch_segment g;
        g.keys = keys + a * key_bytes;
        g.vals = vals ? vals + a * val_bytes : nullptr;
        g.flags = flags ? flags + a : nullptr;
        g.vals_out = vals_out ? vals_out + a * val_bytes : nullptr;
        g.n = offsets[t + 1] - a;
        g.table = t;
        g.op = ops ? static_cast<int32_t>(ops[t]) : op;
        segs[t] = g;
    }
    __syncthreads();
    plan_chunks(segs, ntables, chunk, chunk_start);
}

template <typename K, typename M, int G, int R, int B>
__global__ __launch_bounds__(B) void
k_batch(const table_ref<K, M>* __restrict__ tables, const batch_segment* __restrict__ segs, int32_t nsegs,
        const int64_t* __restrict__ chunk_start)
{
    __shared__ int s_seg;
    __shared__ int64_t s_begin;
    const int64_t total_chunks = chunk_start[nsegs];
    for (int64_t c = blockIdx.x;; c += gridDim.x)
    {
        if (threadIdx.x == 0)
        {
            int g = -1;
            int64_t b = 0;
            if (c < total_chunks)
            {
                int lo = 0, hi = nsegs; // last s with chunk_start[s] <= c
                while (lo < hi)
                {
                    const int mid = (lo + hi + 1) >> 1;
                    if (chunk_start[mid] <= c)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                g = lo;
                b = (c - chunk_start[g]) * B;
            }
            s_seg = g;
            s_begin = b;
        }
        __syncthreads();
        const int g = s_seg;
        const int64_t b = s_begin;
        __syncthreads();
        if (g < 0) break;
        const batch_segment seg = segs[g];
        const table_ref<K, M> tr = tables[seg.table];
        const int64_t i = b + threadIdx.x;
        const bool valid = i < seg.n;
        if (seg.op == kOpInsert)
            insert_body<K, M, G, R>(tr, nullptr, seg.keys, seg.vals, i, valid, seg.flags);
        else if (seg.op == kOpFind)
            find_body<K, M, G, R>(tr, static_cast<const K*>(seg.keys), i, valid, seg.flags, seg.vals_out);
        else if (seg.op == kOpErase)
            erase_body<K, M, G, R>(tr, static_cast<const K*>(seg.keys), i, valid, seg.flags);
    }
}

// ================================================================== served launch
// One launch of the request server.  Its segments travel by value in the kernel arguments (no
// upload, no planning kernel), one 256-thread workgroup takes one chunk of 256 elements and the
// grid is exactly the launch's chunk count, so no work counter is needed.  k_serve_sizes, queued
// right after on the same stream, stores the sizes of the launch's tables into host-mapped memory
// for the scheduler's admission control (one workgroup; the kernel boundary orders it after every
// size-stripe update of the launch).
constexpr int kServeSegs = 96;
constexpr int kServeBlock = 256;
constexpr int kServeTables = 1024;

struct serve_seg
{
    const void* keys;
    void* vals; // insert: mapped values in (optional); find: mapped values out (optional)
    uint8_t* flags;
    uint32_t n;
    int16_t table;
    int8_t op;
    int8_t pad;
};

struct serve_args
{
    int32_t nsegs;
    uint32_t chunk_start[kServeSegs + 1]; // chunk_start[nsegs] = grid size
    serve_seg segs[kServeSegs];
};

struct serve_tables
{
    int32_t n;
    int16_t table[kServeTables];
};

template <typename K, typename M, int G, int RF, int RI>
__global__ __launch_bounds__(kServeBlock) void
k_serve(const table_ref<K, M>* __restrict__ tables, const serve_args a)
{
    const uint32_t c = blockIdx.x;
    int lo = 0, hi = a.nsegs - 1; // last segment whose first chunk is <= c
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (a.chunk_start[mid] <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    const serve_seg seg = a.segs[lo];
    const table_ref<K, M> tr = tables[seg.table];
    const int64_t i = static_cast<int64_t>(c - a.chunk_start[lo]) * kServeBlock + threadIdx.x;
    const bool valid = i < static_cast<int64_t>(seg.n);
    if (seg.op == kOpInsert)
        insert_body<K, M, G, RI>(tr, nullptr, seg.keys, seg.vals, i, valid, seg.flags);
    else if (seg.op == kOpFind)
        find_body<K, M, G, RF>(tr, static_cast<const K*>(seg.keys), i, valid, seg.flags, seg.vals);
    else
        erase_body<K, M, G, RI>(tr, static_cast<const K*>(seg.keys), i, valid, seg.flags);
}

// 32 lanes per table, one size stripe each, summed with shuffles inside the 32-lane half-wave
template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_serve_sizes(const table_ref<K, M>* __restrict__ tables, const serve_tables a, int64_t* __restrict__ sizes_out)
{
    const int sub = static_cast<int>(threadIdx.x) & 31;
    const int lane = lane_id();
    for (int base = 0; base < a.n; base += 8)
    {
        const int s = base + (static_cast<int>(threadIdx.x) >> 5);
        int64_t v = 0;
        if (s < a.n)
        {
            const table_ref<K, M> t = tables[a.table[s]];
            if (static_cast<uint32_t>(sub) < t.nstripes) v = t.stripes[static_cast<uint64_t>(sub) * kStripeStride];
        }
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) v += static_cast<int64_t>(shfl64(static_cast<uint64_t>(v), lane ^ off));
        if (s < a.n && sub == 0)
            __hip_atomic_store(sizes_out + a.table[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ================================================================== assign
// find(key)->second = value for every present key (one thread per element, native find); a key
// may appear once per launch.  found[i] (optional) = 1 iff keys[i] was present.
template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_assign(table_ref<K, M> t, const K* __restrict__ keys, const M* __restrict__ vals, int64_t n,
         uint8_t* __restrict__ found)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    typename slot_traits<K, M>::value_type* p = t.find(keys[i]);
    const bool hit = p != t.end_ptr();
    if (hit) p->second = vals[i];
    if (found) found[i] = hit ? 1 : 0;
}

// ================================================================== clear / fill
template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_clear(table_ref<K, M> t, uint64_t nchunks)
{
    // every key field becomes EMPTY: the slot area is filled with 0xFF bytes, except the padding
    // word of 16-byte slots that carries the in-bucket ovf count (0); otherwise ovf[] is zeroed
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    uint4* p = reinterpret_cast<uint4*>(t.slots);
    const uint4 fill = make_uint4(~0U, ~0U, ~0U, kInline ? 0U : ~0U);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < nchunks; c += stride) p[c] = fill;
    if constexpr (!kInline)
    {
        for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < t.nbuckets; b += stride)
            t.ovf[b] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kMaxStripes) t.stripes[static_cast<uint64_t>(threadIdx.x) * kStripeStride] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *t.flags = 0;
}

// the side slot of the TOMB key value carries that key, so iterators to it read the right key
template <typename K, typename M>
__global__ void
k_clear_tail(table_ref<K, M> t)
{
    using value_type = typename slot_traits<K, M>::value_type;
    value_type* side = t.slots + t.nslots();
    *reinterpret_cast<K*>(side + 0) = empty_key<K>();
    *reinterpret_cast<K*>(side + 1) = tomb_key<K>();
}

// ================================================================== device_range / contents
// Pass 1: per-workgroup count of stored slots over a contiguous tile of slots.
template <typename K, typename M>
__device__ inline bool
slot_stored(const table_ref<K, M>& t, uint64_t s, uint32_t flags)
{
    const uint64_t ns = t.nslots();
    if (s < ns)
    {
        const K k = *t.key_ptr(s);
        return k != empty_key<K>() && k != tomb_key<K>();
    }
    return (flags & (s == ns ? 1U : 2U)) != 0;
}

constexpr int kRangeThreads = 256;
constexpr int kRangePerThread = 16;
constexpr uint64_t kRangeTile = static_cast<uint64_t>(kRangeThreads) * kRangePerThread;


template <typename K, typename M>
__global__ __launch_bounds__(kRangeThreads) void
k_range_count(table_ref<K, M> t, uint64_t total_slots, int64_t* __restrict__ tile_counts)
{
    __shared__ int64_t s_tmp[17];
    const uint32_t flags = *t.flags;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kRangeTile + static_cast<uint64_t>(threadIdx.x) * kRangePerThread;
    int64_t c = 0;
    for (int j = 0; j < kRangePerThread; ++j)
    {
        const uint64_t s = base + j;
        if (s < total_slots && slot_stored(t, s, flags)) ++c;
    }
    int64_t total = 0;
    block_exclusive_scan(c, s_tmp, &total);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = total;
}

// single-workgroup exclusive scan of the tile counts (in place); grand total at counts[ntiles]
// a template so every translation unit may instantiate it (no duplicate symbols)
template <int Unused = 0>
__global__ __launch_bounds__(1024) void
k_scan_tiles(int64_t* counts, int64_t ntiles)
{
    __shared__ int64_t s_tmp[17];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < ntiles; base += blockDim.x)
    {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < ntiles ? counts[i] : 0;
        int64_t total = 0;
        const int64_t ex = block_exclusive_scan(v, s_tmp, &total);
        if (i < ntiles) counts[i] = s_carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[ntiles] = s_carry;
}

// the grand total of a range into caller device memory (the allocation-free device_range)
template <int Unused = 0>
__global__ __launch_bounds__(64) void
k_copy_count(const int64_t* __restrict__ src, int64_t* __restrict__ dst)
{
    if (threadIdx.x == 0) *dst = *src;
}

template <typename K, typename M>
__global__ __launch_bounds__(kRangeThreads) void
k_range_scatter(table_ref<K, M> t, uint64_t total_slots, const int64_t* __restrict__ tile_offsets,
                int64_t* __restrict__ idx_out, K* __restrict__ keys_out, void* __restrict__ vals_out, int64_t cap)
{
    __shared__ int64_t s_tmp[17];
    const uint32_t flags = *t.flags;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kRangeTile + static_cast<uint64_t>(threadIdx.x) * kRangePerThread;
    uint32_t mask = 0;
    int64_t c = 0;
    for (int j = 0; j < kRangePerThread; ++j)
    {
        const uint64_t s = base + j;
        if (s < total_slots && slot_stored(t, s, flags))
        {
            mask |= 1U << j;
            ++c;
        }
    }
    int64_t total = 0;
    int64_t pos = tile_offsets[blockIdx.x] + block_exclusive_scan(c, s_tmp, &total);
    for (int j = 0; j < kRangePerThread; ++j)
    {
        if (!(mask & (1U << j))) continue;
        const uint64_t s = base + j;
        if (pos < cap)
        {
            if (idx_out) idx_out[pos] = static_cast<int64_t>(s);
            if (keys_out) keys_out[pos] = slot_traits<K, M>::key(t.slots[s]);
            if constexpr (slot_traits<K, M>::is_map)
            {
                if (vals_out) static_cast<M*>(vals_out)[pos] = t.slots[s].second;
            }
        }
        ++pos;
    }
}

// ================================================================== valid()
// Every stored key must be found at its own slot (reachable and unique); bad[0] counts violations,
// bad[1] counts stored entries (compared with the size counter on the host).  Each stored key also
// adds one to want[b] for every bucket b its probe sequence runs through; k_valid_overflow then
// checks ovf against those counts, and that no bucket with ovf > 0 holds an EMPTY slot.
// hashes (optional): the caller's hash of each slot's key, for a table that places keys by those
template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_valid(table_ref<K, M> t, uint64_t total_slots, unsigned long long* bad, uint32_t* want, const uint64_t* hashes)
{
    const uint32_t flags = *t.flags;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t ns = t.nslots();
    unsigned long long nbad = 0, nstored = 0;
    for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < total_slots; s += stride)
    {
        if (!slot_stored(t, s, flags)) continue;
        ++nstored;
        const K k = slot_traits<K, M>::key(t.slots[s]);
        const uint64_t h = t.home_of_hash((hashes && s < ns) ? hashes[s] : static_cast<uint64_t>(k));
        if (t.find_at(k, h) != t.slots + s) ++nbad;
        if (s < ns)
        {
            const uint64_t dist = t.distance(h, s / t.W);
            uint64_t b = h;
            for (uint64_t d = 0; d < dist; ++d)
            {
                atomicAdd(want + b, 1U);
                b = next_bucket(b, t.nbuckets);
            }
        }
    }
    if (nbad) atomicAdd(bad, nbad);
    if (nstored) atomicAdd(bad + 1, nstored);
}

template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_valid_overflow(table_ref<K, M> t, const uint32_t* __restrict__ want, unsigned long long* bad)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long nbad = 0;
    for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < t.nbuckets; b += stride)
    {
        const uint32_t ov = *t.ovf_ptr(b);
        if (ov != want[b]) ++nbad;
        if (ov != 0)
        {
            for (uint32_t j = 0; j < t.W; ++j)
                if (*t.key_ptr(b * t.W + j) == empty_key<K>()) ++nbad;
        }
    }
    if (nbad) atomicAdd(bad, nbad);
}

// Away counters (in-bucket layouts with W > 1): for every home bucket h, the keys homed at h that are
// stored in a later bucket, counted per fingerprint, must equal h's counters, except a saturated one,
// which stands for any count.  Such keys lie in the buckets after h up to the first bucket
// b whose predecessor has ovf == 0 (each of them counts in ovf of every bucket from h to its own).
constexpr uint32_t kMaxAwayCounters = 2 * (16 - 1); // buckets of at most 16 slots

template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_valid_away(table_ref<K, M> t, unsigned long long* bad, const uint64_t* hashes)
{
    const uint32_t nctr = t.away_counters();
    if (nctr == 0 || nctr > kMaxAwayCounters) return;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long nbad = 0;
    for (uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; h < t.nbuckets; h += stride)
    {
        uint32_t cnt[kMaxAwayCounters];
        for (uint32_t f = 0; f < nctr; ++f) cnt[f] = 0;
        uint64_t b = h;
        for (uint64_t step = 1; step < t.nbuckets && t.overflow(b) != 0; ++step)
        {
            b = next_bucket(b, t.nbuckets);
            for (uint32_t j = 0; j < t.W; ++j)
            {
                const K k = *t.key_ptr(b * t.W + j);
                if (k == empty_key<K>() || k == tomb_key<K>()) continue;
                if (t.home_of_hash(hashes ? hashes[b * t.W + j] : static_cast<uint64_t>(k)) != h) continue;
                ++cnt[away_index(static_cast<uint64_t>(k), nctr)];
            }
        }
        for (uint32_t f = 0; f < nctr; ++f)
        {
            const uint32_t have = (*t.away_word(h, f) >> (16 * (f & 1))) & 0xFFFFU;
            if (have < kAwaySaturated && have != cnt[f]) ++nbad;
        }
    }
    if (nbad) atomicAdd(bad, nbad);
}

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_KERNELS_HPP
