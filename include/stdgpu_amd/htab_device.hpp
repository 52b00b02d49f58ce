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
//   * Away counts (in-bucket layouts with W > 1): the padding words of slots 1..W-1 of bucket h hold
//     2(W-1) 16-bit counters; counter f counts the stored keys whose home bucket is h, whose
//     fingerprint (away_index) is f, and that are stored in a later bucket.  A key that is not in its
//     home bucket and whose counter there is 0 is absent, so most probes of absent keys (finds that
//     miss, inserts of new keys into a home bucket with ovf > 0) end at the home bucket although
//     ovf > 0.  A counter at or above kAwaySaturated is never lowered again (it only says "maybe").
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

// counter index of a key among the `nctr` (at most 30) away counters of its home bucket: a 32-bit
// multiplicative hash of the key's two halves folded together, independent of mix64, which places the
// key (home bucket, claim offset), and cheap enough to recompute in registers wherever it is needed
__host__ __device__ inline uint32_t
away_index(uint64_t key, uint32_t nctr)
{
    const uint32_t x = (static_cast<uint32_t>(key) ^ static_cast<uint32_t>(key >> 32)) * 0x9E3779B1U;
    return ((x >> 8) * nctr) >> 24;
}

// A counter at or above kAwaySaturated is never lowered: the keys it stands for are no longer all
// counted.  The bulk insert raises a counter with a plain atomic add (no wait for the result) when
// it saw it below kAwayFastLimit at the start of the launch, else with a compare-and-swap that stops
// at kAwaySaturated; a field can only carry into its neighbour after more than 0xC000 plain adds to
// one counter in one launch, i.e. that many keys of one home bucket and one fingerprint stored away.
constexpr uint32_t kAwaySaturated = 0x8000U;
constexpr uint32_t kAwayFastLimit = 0x4000U;

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

    // The top bit of a bucket's ovf word is the lock that native inserts of keys homed there hold
    // while they claim a slot (see insert()); it is clear whenever no native call runs, so the bulk
    // kernels, valid() and snapshots only ever see counts.
    static constexpr uint32_t kOvfLock = 0x80000000U;
    static constexpr uint32_t kOvfCount = 0x7FFFFFFFU;

    __device__ uint32_t overflow(uint64_t b) const { return load_agent(ovf_ptr(b)) & kOvfCount; }

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

    // ---- away counts (see the layout notes at the top of this file); 0 counters without them
    __host__ __device__ uint32_t away_counters() const { return traits::ovf_in_bucket ? 2U * (W - 1) : 0U; }
    __device__ uint32_t away_slot(K key) const { return away_index(static_cast<uint64_t>(key), away_counters()); }
    __device__ uint32_t* away_word(uint64_t h, uint32_t f) const
    {
        return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(slots + h * W + 1 + f / 2) + traits::ovf_offset);
    }
    // counter f of bucket h (coherent read)
    __device__ uint32_t away(uint64_t h, uint32_t f) const { return (load_agent(away_word(h, f)) >> (16 * (f & 1))) & 0xFFFFU; }
    // counter f of bucket h += d (d = +1 or -1) by compare-and-swap: a raise stops at kAwaySaturated,
    // a saturated counter is not lowered, and no neighbouring counter ever sees a carry or a borrow
    __device__ void away_add(uint64_t h, uint32_t f, int d) const
    {
        uint32_t* w = away_word(h, f);
        const uint32_t sh = 16 * (f & 1);
        uint32_t old = load_agent(w);
        while (((old >> sh) & 0xFFFFU) < kAwaySaturated)
        {
            const uint32_t want = d > 0 ? old + (1U << sh) : old - (1U << sh);
            const uint32_t seen = cas_agent(w, old, want);
            if (seen == old) break;
            old = seen;
        }
    }
    // plain atomic raise / lower of counter f (the bulk kernels, under the rules at kAwaySaturated)
    __device__ void away_add_fast(uint64_t h, uint32_t f, int d) const
    {
        const uint32_t inc = 1U << (16 * (f & 1));
        __hip_atomic_fetch_add(away_word(h, f), d > 0 ? inc : 0U - inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // ---- native per-thread operations (user kernels; the facade's find/insert/erase) ----

    // where a key's scan of a bucket starts: its claim order (claim_offset); slot j of the scan is
    // (o + j) mod W, W a power of two
    __device__ uint32_t scan_start(K key) const
    {
        return nbuckets < kOrderedClaimBuckets ? claim_offset(static_cast<uint64_t>(key), W) : 0U;
    }

    // The keys at scan positions j0 .. j0 + kScanBatch - 1 of bucket b (scan order starting at slot o),
    // loaded back to back so that one memory round trip serves them all; a loop that compared each key
    // before loading the next would wait for every load in turn.  With W below kScanBatch the
    // positions past W repeat slots already loaded.  Four at a time (2^22 native calls at load 0.8
    // on 8-slot buckets): find of present keys 25.9, of absent keys 17.0 G/s, against 19.9 / 17.8 with
    // eight (which loads slots past the key) and 23.8 / 8.0 with one at a time.
    static constexpr uint32_t kScanBatch = 4;
    __device__ void load_scan(uint64_t b, uint32_t o, uint32_t j0, K (&k)[kScanBatch]) const
    {
#pragma unroll
        for (uint32_t u = 0; u < kScanBatch; ++u) k[u] = load_agent(key_ptr(b * W + ((o + j0 + u) & (W - 1))));
    }

    // find (unordered_base_detail.cuh:653-682): pointer to the entry or end_ptr().
    __device__ value_type* find(K key) const { return find_at(key, home_bucket(static_cast<uint64_t>(key), nbuckets)); }

    // home bucket of a key addressed by a caller's hash instead of its bits (home_of_hash(key bits) is
    // the home bucket of the bits, so integer keys hashed by identity land where the bulk kernels put them)
    __host__ __device__ uint64_t home_of_hash(uint64_t hash) const { return home_bucket(hash, nbuckets); }

    // find() with the home bucket given (a table addressed by caller hashes)
    __device__ value_type* find_at(K key, uint64_t home) const
    {
        if (is_reserved(key))
        {
            return (load_agent(flags) & side_bit(key)) ? slots + side_index(key) : end_ptr();
        }
        uint64_t b = home;
        const uint32_t o = scan_start(key);
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            value_type* base = slots + b * W;
            for (uint32_t j0 = 0; j0 < W; j0 += kScanBatch)
            {
                K k[kScanBatch];
                load_scan(b, o, j0, k);
#pragma unroll
                for (uint32_t u = 0; u < kScanBatch; ++u)
                {
                    if (k[u] == key) return base + ((o + j0 + u) & (W - 1));
                    saw_empty |= (k[u] == empty_key<K>());
                }
            }
            if (saw_empty || overflow(b) == 0) break;
            // not in its home bucket and no key of that bucket with its fingerprint stored further on
            if (step == 0 && away_counters() != 0 && away(b, away_slot(key)) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return end_ptr();
    }

    // The stored entry whose key bits satisfy `match`, looked up from home bucket `home` (a transparent
    // lookup: a KeyLike compares with stored keys through the caller's key_equal, so neither the key's
    // bits nor its fingerprint are known, and the probe sequence is read to its end).  At most one
    // stored key may match (the caller's key_equal is consistent with its hasher).
    template <typename Match>
    __device__ value_type* find_match(uint64_t home, const Match& match) const
    {
        const uint32_t f = load_agent(flags);
        if ((f & 1U) && match(empty_key<K>())) return slots + side_index(empty_key<K>());
        if ((f & 2U) && match(tomb_key<K>())) return slots + side_index(tomb_key<K>());
        return find_if(home, [&match](uint64_t, K k) { return match(k); });
    }

    // The stored entry for which present(s, k) holds, where s is an occupied slot and k its slot key,
    // looked up from home bucket `home` to the end of the probe sequence (no away count applies: the
    // caller's key need not be the slot key; the side slots are not searched).
    template <typename Present>
    __device__ value_type* find_if(uint64_t home, const Present& present) const
    {
        uint64_t b = home;
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            for (uint32_t j0 = 0; j0 < W; j0 += kScanBatch)
            {
                K k[kScanBatch];
                load_scan(b, 0, j0, k);
#pragma unroll
                for (uint32_t u = 0; u < kScanBatch; ++u)
                {
                    if (k[u] == empty_key<K>())
                    {
                        saw_empty = true;
                        continue;
                    }
                    const uint64_t s = b * W + ((j0 + u) & (W - 1));
                    if (k[u] != tomb_key<K>() && present(s, k[u])) return slots + s;
                }
            }
            if (saw_empty || overflow(b) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return end_ptr();
    }

    // erase() of the entry for which present(s, k) holds (see find_if and insert_if): the slot
    // becomes TOMB, and the counts along its probe sequence come down as in erase_at
    template <typename Present>
    __device__ bool erase_if(uint64_t h, const Present& present) const
    {
        uint64_t b = h;
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            uint64_t hit = ~uint64_t(0);
            K hk = 0;
            for (uint32_t j0 = 0; j0 < W; j0 += kScanBatch)
            {
                K k[kScanBatch];
                load_scan(b, 0, j0, k);
#pragma unroll
                for (uint32_t u = 0; u < kScanBatch; ++u)
                {
                    if (k[u] == empty_key<K>())
                    {
                        saw_empty = true;
                        continue;
                    }
                    const uint64_t s = b * W + ((j0 + u) & (W - 1));
                    if (hit == ~uint64_t(0) && k[u] != tomb_key<K>() && present(s, k[u]))
                    {
                        hit = s;
                        hk = k[u];
                    }
                }
            }
            if (hit != ~uint64_t(0))
            {
                if (cas_agent(key_ptr(hit), hk, tomb_key<K>()) != hk) return false;
                count_add(-1);
                overflow_add(h, step, -1);
                if (step != 0 && away_counters() != 0) away_add(h, away_slot(hk), -1);
                return true;
            }
            if (saw_empty || overflow(b) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return false;
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

    // Waits until every memory operation this wavefront issued has completed.  The table's words
    // change only through agent-scope atomics, which are performed coherently for the whole device,
    // so once they have completed every later coherent load on any CU sees them; this orders a lock
    // release after the claim it protects without the L2 write-back of an agent-scope release fence.
    __device__ __forceinline__ static void settle()
    {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }

    // insert (unordered_base_detail.cuh:708-809, 967-994): never overwrites; reuses erased (TOMB)
    // slots as well as EMPTY ones, so emplace / erase churn from user kernels never wears a table
    // out.  Returns the slot and *inserted = true on insertion, end_ptr() and false when the key is
    // present or the table is at its limit (try_insert reports a present key with end(),
    // unordered_base_detail.cuh:714, 805).
    //
    // Duplicate-freedom.  A present key is returned by a lock-free find first.  Otherwise the call
    // takes the lock bit of the key's home bucket (the top bit of that bucket's ovf word), so at
    // most one insert of a key runs its claim at a time; erase and find never wait for it.  Under
    // the lock it scans the probe sequence to its end (a bucket that holds an EMPTY slot or has
    // ovf == 0), so a copy of the key anywhere on it is seen, whichever call put it there (native,
    // or a bulk insert before a bulk erase emptied slots in front of it), and claims the first free
    // slot of that scan by CAS, or the next free slot on after a CAS lost to another key.  Before
    // the claim, ovf is raised on every bucket the key runs through, so a find() that sees the key
    // also sees the way there.  EMPTY never appears during a kernel that runs native calls (native
    // erase writes TOMB), so a bucket walked past holds no EMPTY slot and keeps invariant (E).
    //
    // Forward progress.  The lock is tried in a loop whose exit is uniform over the wavefront (a
    // ballot of the lanes still waiting), and a lane that gets the lock finishes its whole claim
    // and releases it in the same iteration, so lanes of one wavefront waiting on each other's
    // lock always make progress.
    template <typename WriteValue>
    __device__ value_type* insert(K key, const WriteValue& write_value, bool* inserted) const
    {
        return insert_at(key, home_bucket(static_cast<uint64_t>(key), nbuckets), write_value, inserted);
    }

    // insert() with the home bucket given (a table addressed by caller hashes)
    template <typename WriteValue>
    __device__ value_type* insert_at(K key, uint64_t h, const WriteValue& write_value, bool* inserted) const
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
        if (find_at(key, h) != end_ptr()) return end_ptr();
        return locked_claim(
                h, [&]() { return find_at(key, h) != end_ptr(); },
                [&]() { return claim_locked(key, key, h, [key](uint64_t, K k) { return k == key; }, true, write_value, inserted); });
    }

    // insert() of an entry whose key the caller compares itself: keys wider than the slot key (the C++
    // facade keeps such a key in the slot's payload record and gives each entry a unique slot key).
    // present(s, k) says whether occupied slot s, holding slot key k, holds the entry's key; `key` is
    // the slot key the new entry gets (neither reserved value), `busy` a value no entry ever has: the
    // free slot is claimed as busy, write_value fills it, and only then is `key` stored (release), so a
    // scan never compares a record before it is written.  Inserts of one home bucket take its lock, so
    // two calls of one key never both claim.
    template <typename Present, typename WriteValue>
    __device__ value_type* insert_if(K key, K busy, uint64_t h, const Present& present, const WriteValue& write_value,
                                     bool* inserted) const
    {
        *inserted = false;
        if (nstripes == 1 && load_agent(stripes) >= static_cast<int64_t>(limit)) return end_ptr();
        if (find_if(h, present) != end_ptr()) return end_ptr();
        return locked_claim(
                h, [&]() { return find_if(h, present) != end_ptr(); },
                [&]() { return claim_locked(key, busy, h, present, false, write_value, inserted); });
    }

    // The home-bucket lock around claim(): the calls of a wavefront that wait on one lock word try it
    // through their lowest lane; found() lets a waiting call leave once another call has inserted its
    // key meanwhile.
    template <typename Found, typename Claim>
    __device__ value_type* locked_claim(uint64_t h, const Found& found, const Claim& claim) const
    {
        uint32_t* lock = ovf_ptr(h);
        const int lane = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0U, __builtin_amdgcn_mbcnt_lo(~0U, 0U)));
        value_type* result = end_ptr();
        bool waiting = true, first = true;
        while (__builtin_amdgcn_ballot_w64(waiting) != 0)
        {
            // A waiting call that can be decided without the lock leaves: the table reached its
            // limit, or another call inserted the key meanwhile (many calls of one key, or many keys
            // into a full tiny table, would otherwise take the lock one after another only to find
            // that out).  Otherwise it tries the lock when it looks free.
            bool ready = false;
            if (waiting)
            {
                if (!first && ((nstripes == 1 && load_agent(stripes) >= static_cast<int64_t>(limit)) || found()))
                    waiting = false;
                else
                    ready = !(load_agent(lock) & kOvfLock);
            }
            first = false;
            // of the lanes of this wavefront that try one lock word, only the lowest does
            const uint64_t cand = __builtin_amdgcn_ballot_w64(ready);
            const int leader = cand ? __builtin_ctzll(cand) : 0;
            const uint64_t la = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(lock));
            const uint64_t leader_lock =
                    static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(la), leader))) |
                    (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(la >> 32), leader)))
                     << 32);
            bool got = false;
            if (ready && (lane == leader || la != leader_lock))
            {
                const uint32_t old = __hip_atomic_fetch_or(lock, kOvfLock, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                got = !(old & kOvfLock);
                if (got)
                {
                    result = claim();
                    settle();
                    __hip_atomic_fetch_and(lock, ~kOvfLock, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    waiting = false;
                }
            }
            // no lane of the wavefront got its lock: back off before looking again
            if (__builtin_amdgcn_ballot_w64(got) == 0) __builtin_amdgcn_s_sleep(4);
        }
        return result;
    }

    // the body of insert() under the home-bucket lock
    // key: the entry's slot key; claim: what the free slot is claimed as (key itself, or a busy value
    // that write_value's record precedes, see insert_if); present(s, k): occupied slot s holds the key;
    // keyed: present is k == key, so the key's away count may end the scan at its home bucket
    template <typename Present, typename WriteValue>
    __device__ value_type* claim_locked(K key, K claim, uint64_t h, const Present& present, bool keyed,
                                        const WriteValue& write_value, bool* inserted) const
    {
        constexpr uint64_t kNone = ~uint64_t(0);
        if (nstripes == 1 && load_agent(stripes) >= static_cast<int64_t>(limit)) return end_ptr();
        const uint32_t o = scan_start(key);
        // 1. scan to the end of the probe sequence: the key, and the first free slot in scan order
        uint64_t b = h, step = 0;
        uint64_t fslot = kNone, fstep = 0;
        uint32_t fj = 0;
        K fseen = empty_key<K>();
        for (; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            for (uint32_t j0 = 0; j0 < W; j0 += kScanBatch)
            {
                K kk[kScanBatch];
                load_scan(b, o, j0, kk);
#pragma unroll
                for (uint32_t u = 0; u < kScanBatch; ++u)
                {
                    const uint32_t j = j0 + u;
                    if (j >= W) break; // positions past W repeat loaded slots (scan order matters here)
                    const K k = kk[u];
                    if (k != empty_key<K>() && k != tomb_key<K>() && present(b * W + ((o + j) & (W - 1)), k))
                        return end_ptr();
                    if (fslot == kNone && (k == empty_key<K>() || k == tomb_key<K>()))
                    {
                        fslot = b * W + ((o + j) & (W - 1));
                        fstep = step;
                        fj = j;
                        fseen = k;
                    }
                    saw_empty |= (k == empty_key<K>());
                }
            }
            if (saw_empty || overflow(b) == 0) break;
            // the away count ends the scan at the home bucket as in find()
            if (keyed && step == 0 && away_counters() != 0 && away(b, away_slot(key)) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        // 2. absent: claim the first free slot from the scan's first one on (the next bucket on
        //    when the scan saw none), raising ovf over the buckets before the one claimed in (and
        //    the key's away count in its home bucket when that is not the one claimed in)
        uint64_t raised = 0; // ovf raised on buckets h .. h + raised - 1
        const bool counts_away = away_counters() != 0;
        const uint32_t af = counts_away ? away_slot(key) : 0U;
        if (fslot == kNone)
        {
            if (step >= nbuckets) return end_ptr(); // every bucket scanned, none free
            fstep = step + 1;
            fj = 0;
            fseen = empty_key<K>();
        }
        uint64_t cstep = fstep; // bucket (steps from h) and order position of the next try
        uint32_t cj = fj;
        bool have = fslot != kNone;
        while (cstep < nbuckets)
        {
            const uint64_t cb = h + cstep >= nbuckets ? h + cstep - nbuckets : h + cstep;
            if (!have)
            {
                // look for the first free slot of bucket cb from order position cj on
                for (; cj < W; ++cj)
                {
                    const uint64_t s = cb * W + ((o + cj) & (W - 1));
                    const K k = load_agent(key_ptr(s));
                    if (k == empty_key<K>() || k == tomb_key<K>())
                    {
                        fslot = s;
                        fseen = k;
                        have = true;
                        break;
                    }
                }
                if (!have)
                {
                    ++cstep;
                    cj = 0;
                    continue;
                }
            }
            if (raised < cstep)
            {
                if (raised == 0 && counts_away) away_add(h, af, 1);
                for (; raised < cstep; ++raised)
                {
                    const uint64_t rb = h + raised >= nbuckets ? h + raised - nbuckets : h + raised;
                    __hip_atomic_fetch_add(ovf_ptr(rb), 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                // the raises complete before the key can be seen: an erase of the key lowers these
                // counts only after it has seen the key, so no count ever dips below zero (into the
                // lock bit)
                settle();
            }
            const K seen = cas_agent(key_ptr(fslot), fseen, claim);
            if (seen == fseen)
            {
                if (!count_add(1))
                {
                    __hip_atomic_store(key_ptr(fslot), tomb_key<K>(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    overflow_add(h, raised, -1);
                    if (raised && counts_away) away_add(h, af, -1);
                    return end_ptr();
                }
                write_value(slots + fslot);
                if (claim != key) __hip_atomic_store(key_ptr(fslot), key, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                *inserted = true;
                return slots + fslot;
            }
            // lost to another key (only keys of other home buckets claim here meanwhile): go on
            have = false;
            ++cj;
        }
        overflow_add(h, raised, -1);
        if (raised && counts_away) away_add(h, af, -1);
        return end_ptr();
    }

    // erase (unordered_base_detail.cuh:811-913, 1019-1043): marks the slot TOMB; exactly one of
    // several concurrent erasers of one key succeeds.
    __device__ bool erase(K key) const { return erase_at(key, home_bucket(static_cast<uint64_t>(key), nbuckets)); }

    // erase() with the home bucket given (a table addressed by caller hashes)
    __device__ bool erase_at(K key, uint64_t h) const
    {
        if (is_reserved(key))
        {
            const uint32_t bit = side_bit(key);
            const uint32_t old = __hip_atomic_fetch_and(flags, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!(old & bit)) return false;
            count_add(-1);
            return true;
        }
        const uint32_t o = scan_start(key);
        uint64_t b = h;
        for (uint64_t step = 0; step < nbuckets; ++step)
        {
            bool saw_empty = false;
            uint64_t hit = ~uint64_t(0);
            for (uint32_t j0 = 0; j0 < W; j0 += kScanBatch)
            {
                K k[kScanBatch];
                load_scan(b, o, j0, k);
#pragma unroll
                for (uint32_t u = 0; u < kScanBatch; ++u)
                {
                    if (k[u] == key) hit = b * W + ((o + j0 + u) & (W - 1));
                    saw_empty |= (k[u] == empty_key<K>());
                }
            }
            if (hit != ~uint64_t(0))
            {
                if (cas_agent(key_ptr(hit), key, tomb_key<K>()) != key) return false;
                count_add(-1);
                overflow_add(h, step, -1);
                if (step != 0 && away_counters() != 0) away_add(h, away_slot(key), -1);
                return true;
            }
            if (saw_empty || overflow(b) == 0) break;
            if (step == 0 && away_counters() != 0 && away(b, away_slot(key)) == 0) break;
            b = next_bucket(b, nbuckets);
        }
        return false;
    }
};

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_DEVICE_HPP
