// Device-side layout and probing primitives shared by the bulk kernels
// (stdgpu_amd/csrc/htab.hip) and the native per-thread API of the C++ facade
// (include/stdgpu/unordered_{map,set}.cuh).  gfx950 only: wave64, 16-byte lanes.
//
// Layout (replaces the reference's _values/_offsets/_occupied/_locks/excess arrays,
// src/stdgpu/impl/unordered_base.cuh:526-554):
//   * nbuckets buckets of 16*G bytes; a bucket holds W = 16*G/sizeof(value_type) slots
//     of value_type stored AoS, so a probe of one bucket is a single coalesced load by G
//     lanes of a wavefront (16 bytes per lane) and an iterator is still value_type*.
//   * Occupancy lives in the key field: all-ones = EMPTY (never used), all-ones minus one
//     = TOMB (erased).  Those two key values are still storable: they live in two side
//     slots after the buckets, with a presence bit each in `flags`.
//   * Collisions overflow into the next bucket (linear probing over buckets, wrapping).
//   * ovf[b] (one uint32 per bucket) counts the stored keys whose probe sequence runs through
//     bucket b to a later bucket (home bucket <= b < stored bucket, cyclically).  It lives in the
//     padding of the bucket's first slot when value_type has 4 spare bytes, else in an array.  A probe
//     sequence ends at the first bucket that holds the key, holds an EMPTY slot, or has
//     ovf == 0.  A bucket with ovf > 0 never holds an EMPTY slot (an erase there leaves TOMB),
//     so the EMPTY test needs no extra load, and TOMBs never lengthen a probe sequence through
//     buckets that no key overflows.
//   * Home bucket = (hi32(mix64(key)) * nbuckets) >> 32, mix64 = MurmurHash3's fmix64.
#ifndef STDGPU_AMD_HTAB_DEVICE_HPP
#define STDGPU_AMD_HTAB_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <stdgpu_amd/htab.h>

namespace htab_dev
{

constexpr int kWave = 64;
constexpr int kStripeStride = 16; // int64 elements between size stripes (128 B)
constexpr uint32_t kMaxStripes = 32;

// ------------------------------------------------------------------ slot types
template <typename K, typename M>
struct alignas(sizeof(K) > sizeof(M) ? sizeof(K) : sizeof(M)) kv_slot
{
    K first;
    M second;
};

template <typename K, typename M>
struct slot_traits
{
    using key_type = K;
    using mapped_type = M;
    using value_type = kv_slot<K, M>;
    static constexpr bool is_map = true;
    // a bucket's ovf count lives in the padding of its first slot when the slot has 4 spare
    // bytes (unordered_map<u64, u32>: key at 0, value at 8, padding at 12): a probe then reads
    // the count with the bucket, in the same line, and no separate array exists
    static constexpr bool ovf_in_bucket =
            sizeof(value_type) >= sizeof(K) + sizeof(M) + 4 && (sizeof(K) + sizeof(M)) % 4 == 0;
    static constexpr int ovf_offset = static_cast<int>(sizeof(K) + sizeof(M));
    __host__ __device__ static K key(const value_type& v) { return v.first; }
};

template <typename K>
struct slot_traits<K, void>
{
    using key_type = K;
    using mapped_type = void;
    using value_type = K;
    static constexpr bool is_map = false;
    static constexpr bool ovf_in_bucket = false;
    static constexpr int ovf_offset = 0;
    __host__ __device__ static K key(const value_type& v) { return v; }
};

template <typename K>
__host__ __device__ constexpr K
empty_key()
{
    return static_cast<K>(~static_cast<K>(0));
}

template <typename K>
__host__ __device__ constexpr K
tomb_key()
{
    return static_cast<K>(~static_cast<K>(0) - 1);
}

// ------------------------------------------------------------------ hashing
__host__ __device__ inline uint64_t
mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__host__ __device__ inline uint64_t
home_bucket(uint64_t key, uint64_t nbuckets)
{
    const uint64_t h = mix64(key) >> 32;
    return (h * nbuckets) >> 32;
}

__host__ __device__ inline uint64_t
next_bucket(uint64_t b, uint64_t nbuckets)
{
    return (b + 1 == nbuckets) ? 0 : b + 1;
}

// Claim order of a key inside a bucket: slots o, o+1, ..., W-1, 0, ..., o-1, with o taken from the
// key's hash.  Concurrent inserts of different keys into one bucket then aim at different slots,
// while every inserter of one key follows one order over its whole probe sequence, which is what
// the no-duplicate argument needs (a key's claims move forward along that order only).  Tables of
// kOrderedClaimBuckets buckets or more keep o = 0 (plain first-free order): a launch's elements in
// flight (about 2^19) rarely meet in one of so many buckets, and the offset then only costs ALU work.
// The choice depends on the table alone, so all inserters of a key in a launch share it.  The native
// per-thread calls scan a bucket in the same order, so a key that the bulk insert placed is usually
// found with its first load.
constexpr uint64_t kOrderedClaimBuckets = 1ULL << 22;

__host__ __device__ inline uint32_t
claim_offset(uint64_t key, uint32_t W)
{
    return static_cast<uint32_t>(mix64(key)) & (W - 1);
}

// The reference's bucket(): Fibonacci hashing of the identity hash over bucket_count()
// (src/stdgpu/impl/unordered_base_detail.cuh:56-85, functional_detail.h:60-80).
__host__ __device__ inline int64_t
reference_bucket(uint64_t hash, int64_t bucket_count)
{
    int w = 0;
    for (uint64_t x = static_cast<uint64_t>(bucket_count); x != 0; x >>= 1) ++w;
    w -= 1;
    if (w <= 0) return 0;
    const unsigned dropped = static_cast<unsigned>(64 - w);
    const uint64_t improved = hash ^ (hash >> dropped);
    return static_cast<int64_t>((11400714819323198485ULL * improved) >> dropped);
}

// ------------------------------------------------------------------ coherent accessors
// Probes of insert/erase read keys at agent scope (bypassing the non-coherent vector L1),
// so a compare-and-swap that lost a race is followed by a fresh view of the slot.
template <typename T>
__device__ inline T
load_agent(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ inline T
cas_agent(T* p, T expected, T desired)
{
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected; // value seen by the CAS (== the old expected on success)
}

__device__ inline int64_t
stripe_sum(const int64_t* stripes, uint32_t nstripes)
{
    int64_t s = 0;
    for (uint32_t i = 0; i < nstripes; ++i) s += load_agent(stripes + static_cast<uint64_t>(i) * kStripeStride);
    return s;
}

// ------------------------------------------------------------------ typed view
template <typename K, typename M>
struct table_ref
{
    using traits = slot_traits<K, M>;
    using value_type = typename traits::value_type;
    static constexpr int SB = static_cast<int>(sizeof(value_type));

    value_type* slots;
    int64_t* stripes;
    uint32_t* flags;
    uint32_t* ovf; // per bucket: stored keys whose probe sequence runs through it (null when in-bucket)
    uint64_t nbuckets;
    uint64_t limit;
    uint32_t nstripes;
    uint32_t W; // slots per bucket

    __host__ __device__ static table_ref from_view(const htab_view& v)
    {
        table_ref t;
        t.slots = static_cast<value_type*>(v.slots);
        t.stripes = v.stripes;
        t.flags = v.flags;
        t.ovf = v.ovf;
        t.nbuckets = v.nbuckets;
        t.limit = v.limit;
        t.nstripes = v.nstripes;
        t.W = 16U * v.bucket_lanes / static_cast<uint32_t>(SB);
        return t;
    }

    __host__ __device__ uint64_t nslots() const { return nbuckets * W; }
    __host__ __device__ value_type* end_ptr() const { return slots + nslots() + 2; }
    __host__ __device__ static bool is_reserved(K k) { return k == empty_key<K>() || k == tomb_key<K>(); }
    __host__ __device__ uint64_t side_index(K k) const { return nslots() + (k == empty_key<K>() ? 0 : 1); }
    __host__ __device__ static uint32_t side_bit(K k) { return k == empty_key<K>() ? 1U : 2U; }
    __device__ K* key_ptr(uint64_t slot) const { return reinterpret_cast<K*>(slots + slot); }

    __device__ int64_t size() const { return stripe_sum(stripes, nstripes); }

    __host__ __device__ uint32_t* ovf_ptr(uint64_t b) const
    {
        if constexpr (traits::ovf_in_bucket)
            return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(slots + b * W) + traits::ovf_offset);
        else
            return ovf + b;
    }

    __device__ uint32_t overflow(uint64_t b) const { return load_agent(ovf_ptr(b)); }

    // adds d to ovf of the `count` buckets starting at h (cyclic)
    __device__ void overflow_add(uint64_t h, uint64_t count, int32_t d) const
    {
        uint64_t b = h;
        for (uint64_t c = 0; c < count; ++c)
        {
            __hip_atomic_fetch_add(ovf_ptr(b), static_cast<uint32_t>(d), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b = next_bucket(b, nbuckets);
        }
    }

    // buckets from h up to (not including) f, cyclically
    __device__ uint64_t distance(uint64_t h, uint64_t f) const { return f >= h ? f - h : f + nbuckets - h; }

    // ---- native per-thread operations (user kernels; the facade's find/insert/erase) ----

    // where a key's scan of a bucket starts: its claim order (claim_offset); slot j of the scan is
    // (o + j) mod W, W a power of two
    __device__ uint32_t scan_start(K key) const
    {
        return nbuckets < kOrderedClaimBuckets ? claim_offset(static_cast<uint64_t>(key), W) : 0U;
    }

    // find (unordered_base_detail.cuh:653-682): pointer to the entry or end_ptr().
    __device__ value_type* find(K key) const
    {
        if (is_reserved(key))
        {
            return (load_agent(flags) & side_bit(key)) ? slots + side_index(key) : end_ptr();
        }
        uint64_t b = home_bucket(static_cast<uint64_t>(key), nbuckets);
        const uint32_t o = scan_start(key);
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            value_type* base = slots + b * W;
            for (uint32_t j = 0; j < W; ++j)
            {
                const uint32_t x = (o + j) & (W - 1);
                const K k = load_agent(key_ptr(b * W + x));
                if (k == key) return base + x;
                saw_empty |= (k == empty_key<K>());
            }
            if (saw_empty || overflow(b) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return end_ptr();
    }

    // Adds delta (+1 or -1, the same at each call site) to the size for every lane of the wavefront
    // that makes this call at this point, with one atomic for all of them (the size word is the one
    // every successful insert and erase touches).  Returns false for a lane whose +1 would pass the
    // exact limit (tracked only when nstripes == 1): lanes in lane order take what is left below the
    // limit, the rest are taken back with one more atomic.
    __device__ __forceinline__ bool count_add(int64_t delta) const
    {
        const uint64_t lanes = __builtin_amdgcn_ballot_w64(true); // the lanes making this call
        const int lane = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0U, __builtin_amdgcn_mbcnt_lo(~0U, 0U)));
        const int leader = __builtin_ctzll(lanes);
        const int64_t k = __builtin_popcountll(lanes);
        const int64_t rank = __builtin_popcountll(lanes & ((uint64_t(1) << lane) - 1));
        const uint32_t s = nstripes == 1 ? 0U : (static_cast<uint32_t>(blockIdx.x) & (nstripes - 1));
        int64_t old = 0;
        if (lane == leader)
            old = __hip_atomic_fetch_add(stripes + static_cast<uint64_t>(s) * kStripeStride, delta * k, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        old = static_cast<int64_t>(
                static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(old), leader))) |
                (static_cast<uint64_t>(static_cast<uint32_t>(
                         __builtin_amdgcn_readlane(static_cast<uint32_t>(static_cast<uint64_t>(old) >> 32), leader)))
                 << 32));
        if (nstripes == 1 && delta > 0)
        {
            // lanes past the limit: those whose unit lands above it
            int64_t over = old + k - static_cast<int64_t>(limit);
            if (over > 0)
            {
                if (over > k) over = k;
                if (lane == leader) __hip_atomic_fetch_add(stripes, -over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return rank < k - over;
            }
        }
        return true;
    }

    // insert (unordered_base_detail.cuh:708-809, 967-994): never overwrites; claims only EMPTY
    // slots, so it stays duplicate-free when it runs concurrently with erase() in one kernel.
    // Every bucket it walks past without an EMPTY slot gets its ovf raised before the key can
    // appear further on (and lowered again if the key is not inserted), so a concurrent find()
    // that sees the key's slot also sees the way there.
    // Returns {slot pointer, true} on insertion, {end_ptr(), false} otherwise.
    template <typename WriteValue>
    __device__ value_type* insert(K key, const WriteValue& write_value, bool* inserted) const
    {
        *inserted = false;
        if (nstripes == 1 && load_agent(stripes) >= static_cast<int64_t>(limit)) return end_ptr();
        if (is_reserved(key))
        {
            const uint32_t bit = side_bit(key);
            const uint32_t old = __hip_atomic_fetch_or(flags, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old & bit) return end_ptr();
            if (!count_add(1))
            {
                __hip_atomic_fetch_and(flags, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return end_ptr();
            }
            value_type* p = slots + side_index(key);
            write_value(p);
            *inserted = true;
            return p;
        }
        const uint64_t h = home_bucket(static_cast<uint64_t>(key), nbuckets);
        const uint32_t o = scan_start(key);
        uint64_t b = h;
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            for (uint32_t j = 0; j < W; ++j)
            {
                const uint64_t s = b * W + ((o + j) & (W - 1));
                K k = load_agent(key_ptr(s));
                if (k == key)
                {
                    overflow_add(h, step, -1);
                    return end_ptr();
                }
                if (k == empty_key<K>())
                {
                    const K seen = cas_agent(key_ptr(s), empty_key<K>(), key);
                    if (seen == empty_key<K>())
                    {
                        if (!count_add(1))
                        {
                            __hip_atomic_store(key_ptr(s), tomb_key<K>(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            overflow_add(h, step, -1);
                            return end_ptr();
                        }
                        write_value(slots + s);
                        *inserted = true;
                        return slots + s;
                    }
                    if (seen == key)
                    {
                        overflow_add(h, step, -1);
                        return end_ptr();
                    }
                    // lost the slot to another key: keep scanning from the next slot
                }
            }
            // no EMPTY slot left here: the key's probe sequence runs through this bucket
            __hip_atomic_fetch_add(ovf_ptr(b), 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            b = next_bucket(b, nbuckets);
        }
        overflow_add(h, nbuckets, -1);
        return end_ptr();
    }

    // erase (unordered_base_detail.cuh:811-913, 1019-1043): marks the slot TOMB; exactly one of
    // several concurrent erasers of one key succeeds.
    __device__ bool erase(K key) const
    {
        if (is_reserved(key))
        {
            const uint32_t bit = side_bit(key);
            const uint32_t old = __hip_atomic_fetch_and(flags, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!(old & bit)) return false;
            count_add(-1);
            return true;
        }
        const uint64_t h = home_bucket(static_cast<uint64_t>(key), nbuckets);
        const uint32_t o = scan_start(key);
        uint64_t b = h;
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            for (uint32_t j = 0; j < W; ++j)
            {
                const uint64_t s = b * W + ((o + j) & (W - 1));
                const K k = load_agent(key_ptr(s));
                if (k == key)
                {
                    if (cas_agent(key_ptr(s), key, tomb_key<K>()) != key) return false;
                    count_add(-1);
                    overflow_add(h, step, -1);
                    return true;
                }
                saw_empty |= (k == empty_key<K>());
            }
            if (saw_empty || overflow(b) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return false;
    }

    // bucket_size(n) analogue for this layout: stored entries in physical bucket n.
    __device__ int32_t physical_bucket_size(uint64_t n) const
    {
        int32_t c = 0;
        for (uint32_t j = 0; j < W; ++j)
        {
            const K k = load_agent(key_ptr(n * W + j));
            c += (k != empty_key<K>() && k != tomb_key<K>()) ? 1 : 0;
        }
        return c;
    }
};

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_DEVICE_HPP
