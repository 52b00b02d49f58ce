// Partitioned bulk insert: a large insert launch without per-element flags, on a table that counts
// its size on all stripes, runs as streaming passes instead of one random compare-and-swap per key.
//
// The bucket array is cut into regions of kRegionBytes (32 KB: 256 buckets of 128 B).  The keys are
// partitioned by the region of their home bucket in two LDS-staged scatter passes (by region / 128
// into super-bins, then by region % 128), each reading and writing the keys once in coalesced runs.
// Every bin has a fixed capacity (its expected share plus slack), so no counting pass precedes a
// scatter; the keys a full bin cannot take (a hot key of a skewed input, or bad luck) spill to a list.
// Then one workgroup per region copies the region into LDS, inserts the region's keys there with LDS
// atomics (the claim protocol of the per-key kernel, so copies of one key in a launch still leave one
// stored key) and writes the region back.  Maps with 8-byte keys claim a slot by a 32-bit
// compare-and-swap on the low word of its value instead of a 64-bit one on its key: gfx950 runs the
// 64-bit LDS compare-and-swap far slower (at 2^26 keys it took 3.2 ms of this kernel's 4.0 ms; the
// 32-bit form and the 64-bit exchange take none of that).  The copy marks every free slot's value word
// unclaimed when the region arrives; the winner of the word stores its key, then its value; a thread
// that loses the word reads the slot again until the winner's key is there.  A stored value may equal
// the mark: a claim that wins the word of a slot already holding a key puts the mark back.  A key whose probe sequence runs past the region's last
// bucket (into a region another workgroup owns meanwhile) is deferred.  The per-key kernel then
// inserts the deferred keys, the spilled keys and the two reserved key values on the whole table.
//
// Why: random 64-bit CAS is performed memory-side on gfx950 (about 18 G/s over a table in HBM,
// 27 G/s inside the 256 MB Infinity Cache, no faster when each XCD keeps to its own slice), so the
// per-key kernel is bound near 18 G keys/s; the partitioned form streams the keys twice and the table
// once.
//
// Correctness of the split: keys of one home bucket land in one region, and a key is inserted in the
// LDS copy only when its probe sequence ends inside the region (a bucket with an EMPTY slot or
// ovf == 0) or it is found there; slots only fill during an insert launch, so copies of one key that
// reach the region either all resolve there (against each other, by LDS CAS) or all defer, and copies
// that spill or defer meet the region's result in the table when the per-key kernel runs.
#ifndef STDGPU_AMD_HTAB_REGION_HPP
#define STDGPU_AMD_HTAB_REGION_HPP

#include "htab_kernels.hpp"

namespace htab_dev
{

constexpr int kRegionBytes = 32768;
constexpr uint32_t kSubBins = 128;   // regions per super-bin (digit of the second pass)
constexpr uint32_t kMaxSuper = 512;  // super-bins (digit of the first pass)
constexpr uint32_t kMaxRegions = kSubBins * kMaxSuper;
// copies of each super-bin: pass-1 chunk c fills copy c % kSuperCopies (a bin of its own, of 1 /
// kSuperCopies of the super-bin's share), so no fill counter takes every chunk's reservation (one
// counter per super-bin took 32,768 same-address atomics at 2^26 keys)
constexpr uint32_t kSuperCopies = 8;
constexpr int kPartThreads = 256;
constexpr int kPartItems = 8;                             // elements per thread of a scatter chunk
constexpr int kPartChunk = kPartThreads * kPartItems;     // 2048
constexpr int kRegionThreads = 256;
// the default insert takes the partitioned path only on tables larger than this (the Infinity Cache)
constexpr uint64_t kRegionDefaultMinBytes = 256ULL << 20;
// insert_regions' answer when the default choice could not get its scratch memory
constexpr int kRegionsDeclined = 1;

template <int G>
__host__ __device__ constexpr uint32_t
region_buckets()
{
    return static_cast<uint32_t>(kRegionBytes / (16 * G));
}

// Device bookkeeping of one partitioned insert.
struct region_ctl
{
    uint32_t super_fill[kMaxSuper * kSuperCopies]; // elements offered to each super-bin copy, copy-major
                                                   // (may pass its capacity)
    uint32_t spilled;               // elements in the spill list
    uint32_t deferred;              // elements the region kernel left (in buffer A once pass 2 is done)
    uint32_t per_key;               // nonzero: the launch's sample found copies of keys, so the per-key
                                    // kernel inserts the whole launch and the partitioned kernels exit
};

struct region_scratch
{
    region_ctl* ctl;
    uint32_t* region_fill; // elements offered to each region (may pass its capacity)
    void* a_keys;          // super-bin d: entries [d * cap1, d * cap1 + min(fill, cap1))
    void* a_vals;
    void* b_keys;          // region r: entries [r * cap2, r * cap2 + min(fill, cap2))
    void* b_vals;
    void* s_keys;          // spill list
    void* s_vals;
    uint32_t cap1, cap2;
};

template <typename K, typename M>
struct region_elem
{
    __device__ static void put(void* keys, void* vals, uint64_t i, K k, uint64_t v)
    {
        static_cast<K*>(keys)[i] = k;
        if constexpr (slot_traits<K, M>::is_map) static_cast<M*>(vals)[i] = static_cast<M>(v);
    }
    __device__ static void get(const void* keys, const void* vals, uint64_t i, K* k, uint64_t* v)
    {
        *k = static_cast<const K*>(keys)[i];
        if constexpr (slot_traits<K, M>::is_map)
            *v = static_cast<uint64_t>(static_cast<const M*>(vals)[i]);
        else
            *v = 0;
    }
};

__device__ __forceinline__ uint32_t
region_of(uint64_t key, uint64_t nbuckets, uint32_t rb)
{
    return static_cast<uint32_t>(home_bucket(key, nbuckets) / rb);
}

// block-wide exclusive scan of n <= 2 * kPartThreads counters in LDS, in place; returns the total
__device__ inline uint32_t
lds_exclusive_scan(uint32_t* a, uint32_t n, uint32_t* s_tmp)
{
    const uint32_t t = threadIdx.x;
    const uint32_t i0 = 2 * t, i1 = 2 * t + 1;
    const uint32_t x0 = i0 < n ? a[i0] : 0U, x1 = i1 < n ? a[i1] : 0U;
    const uint32_t v = x0 + x1;
    const int lane = lane_id();
    const int wid = static_cast<int>(t) >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = shfl32(inc, lane >= off ? lane - off : lane);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_tmp[wid] = inc;
    __syncthreads();
    if (t == 0)
    {
        uint32_t run = 0;
        for (int w = 0; w < kPartThreads / 64; ++w)
        {
            const uint32_t c = s_tmp[w];
            s_tmp[w] = run;
            run += c;
        }
        s_tmp[kPartThreads / 64] = run;
    }
    __syncthreads();
    const uint32_t excl = s_tmp[wid] + inc - v;
    if (i0 < n) a[i0] = excl;
    if (i1 < n) a[i1] = excl + x0;
    const uint32_t total = s_tmp[kPartThreads / 64];
    __syncthreads();
    return total;
}

// ------------------------------------------------------------------ bookkeeping reset (no memset)
template <int = 0>
__global__ __launch_bounds__(256) void
k_region_reset(region_ctl* ctl, uint32_t* region_fill, uint32_t nregion_slots)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t j = i; j < kMaxSuper * kSuperCopies; j += gridDim.x * blockDim.x) ctl->super_fill[j] = 0;
    if (i == 0)
    {
        ctl->spilled = 0;
        ctl->deferred = 0;
        ctl->per_key = 0;
    }
    for (uint32_t r = i; r < nregion_slots; r += gridDim.x * blockDim.x) region_fill[r] = 0;
}

// Shared body of both scatter passes: a chunk's elements are ordered by digit in LDS, each digit's run
// reserves a range of its bin with one atomic, and the runs leave with consecutive threads writing
// consecutive entries; the part of a run past its bin's capacity goes to the spill list.
template <typename K, typename M, int NB>
struct scatter_lds
{
    uint32_t hist[NB];
    uint32_t pre[NB];
    uint32_t gbase[NB];
    uint32_t tmp[kPartThreads / 64 + 1];
    uint32_t spill_base;
    K key[kPartChunk];
    std::conditional_t<slot_traits<K, M>::is_map, std::conditional_t<std::is_void_v<M>, char, M>, char>
            val[slot_traits<K, M>::is_map ? kPartChunk : 1];
};

// digit of a key in a pass: its super-bin (pass 1) or its region within the super-bin (pass 2)
template <bool kFirst>
__device__ __forceinline__ uint32_t
pass_digit(uint64_t key, uint64_t nbuckets, uint32_t rb)
{
    const uint32_t r = region_of(key, nbuckets, rb);
    return kFirst ? r / kSubBins : r % kSubBins;
}

// The thread's elements k / v where ok; bin (bin_base + digit) holds entries [bin * cap, bin * cap +
// cap) of (out_keys, out_vals), and fill[digit] counts what was offered to it.
template <typename K, typename M, bool kFirst>
__device__ __forceinline__ void
scatter_chunk(scatter_lds<K, M, kFirst ? kMaxSuper : kSubBins>& s, const K (&k)[kPartItems],
              const uint64_t (&v)[kPartItems], const bool (&ok)[kPartItems], uint32_t ndigits, uint64_t nbuckets,
              uint32_t rb, uint32_t* fill, uint64_t bin_base, uint32_t cap, void* out_keys, void* out_vals,
              region_ctl* ctl, const region_scratch& sc)
{
    constexpr bool kMap = slot_traits<K, M>::is_map;
    for (uint32_t d = threadIdx.x; d < ndigits; d += kPartThreads) s.hist[d] = 0;
    if (threadIdx.x == 0) s.spill_base = 0;
    __syncthreads();
    uint32_t dr[kPartItems]; // digit << 16 | rank within the chunk's run of that digit
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        dr[j] = ~0U;
        if (ok[j])
        {
            const uint32_t d = pass_digit<kFirst>(static_cast<uint64_t>(k[j]), nbuckets, rb);
            dr[j] = (d << 16) | atomicAdd(&s.hist[d], 1U);
        }
    }
    __syncthreads();
    // reserve each run in its bin; what passes the bin's capacity takes a range of the spill list
    uint32_t spill_mine = 0;
    for (uint32_t d = threadIdx.x; d < ndigits; d += kPartThreads)
    {
        s.pre[d] = s.hist[d];
        if (s.hist[d])
        {
            const uint32_t g = atomicAdd(&fill[d], s.hist[d]);
            s.gbase[d] = g;
            if (g + s.hist[d] > cap) spill_mine += g >= cap ? s.hist[d] : g + s.hist[d] - cap;
        }
    }
    if (spill_mine) atomicAdd(&s.spill_base, spill_mine);
    __syncthreads();
    if (threadIdx.x == 0 && s.spill_base) s.spill_base = atomicAdd(&ctl->spilled, s.spill_base);
    const uint32_t total = lds_exclusive_scan(s.pre, ndigits, s.tmp); // its barriers publish spill_base
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        if (dr[j] != ~0U)
        {
            const uint32_t p = s.pre[dr[j] >> 16] + (dr[j] & 0xFFFFU);
            s.key[p] = k[j];
            if constexpr (kMap) s.val[p] = static_cast<M>(v[j]);
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < total; p += kPartThreads)
    {
        const K kk = s.key[p];
        const uint32_t d = pass_digit<kFirst>(static_cast<uint64_t>(kk), nbuckets, rb);
        const uint32_t pos = s.gbase[d] + (p - s.pre[d]);
        const uint64_t vv = kMap ? static_cast<uint64_t>(s.val[p]) : 0;
        if (pos < cap)
            region_elem<K, M>::put(out_keys, out_vals, (bin_base + d) * cap + pos, kk, vv);
        else
            region_elem<K, M>::put(sc.s_keys, sc.s_vals, atomicAdd(&s.spill_base, 1U), kk, vv);
    }
}

// ------------------------------------------------------------------ pass 1: input -> super-bins
template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part1_scatter(table_ref<K, M> t, const void* __restrict__ values, const void* __restrict__ keys,
                const void* __restrict__ vals, int64_t n, uint32_t rb, uint32_t nsuper, region_scratch sc)
{
    __shared__ scatter_lds<K, M, kMaxSuper> s;
    if (sc.ctl->per_key) return;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kPartChunk;
    K k[kPartItems];
    uint64_t v[kPartItems];
    bool ok[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        const int64_t e = base + j * kPartThreads + threadIdx.x;
        ok[j] = e < n;
        k[j] = empty_key<K>();
        v[j] = 0;
        if (ok[j]) fetch_element<K, M>(values, keys, vals, e, &k[j], &v[j]);
    }
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        if (ok[j] && table_ref<K, M>::is_reserved(k[j]))
        {
            // the two reserved key values go to the per-key kernel through the spill list
            region_elem<K, M>::put(sc.s_keys, sc.s_vals, atomicAdd(&sc.ctl->spilled, 1U), k[j], v[j]);
            ok[j] = false;
        }
    }
    const uint32_t copy = blockIdx.x % kSuperCopies;
    scatter_chunk<K, M, true>(s, k, v, ok, nsuper, t.nbuckets, rb, sc.ctl->super_fill + copy * nsuper,
                              static_cast<uint64_t>(copy) * nsuper, sc.cap1, sc.a_keys,
                              sc.a_vals, sc.ctl, sc);
}

// ------------------------------------------------------------------ pass 2: super-bin -> regions
// workgroup c takes chunk c % chunks_per_super of super-bin c / chunks_per_super
template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part2_scatter(table_ref<K, M> t, uint32_t rb, uint32_t chunks_per_super, uint32_t nsuper, region_scratch sc)
{
    __shared__ scatter_lds<K, M, kSubBins> s;
    if (sc.ctl->per_key) return;
    const uint32_t d = blockIdx.x / chunks_per_super;
    const uint32_t c = blockIdx.x % chunks_per_super;
    const uint32_t f = sc.ctl->super_fill[d];
    const uint32_t count = f < sc.cap1 ? f : sc.cap1;
    const uint32_t lo = c * static_cast<uint32_t>(kPartChunk);
    if (lo >= count) return;
    K k[kPartItems];
    uint64_t v[kPartItems];
    bool ok[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        const uint32_t e = lo + j * kPartThreads + threadIdx.x;
        ok[j] = e < count;
        k[j] = empty_key<K>();
        v[j] = 0;
        if (ok[j]) region_elem<K, M>::get(sc.a_keys, sc.a_vals, static_cast<uint64_t>(d) * sc.cap1 + e, &k[j], &v[j]);
    }
    // bin d is copy d / nsuper of super-bin d % nsuper
    const uint32_t sup = d % nsuper;
    scatter_chunk<K, M, false>(s, k, v, ok, kSubBins, t.nbuckets, rb, sc.region_fill + sup * kSubBins,
                               static_cast<uint64_t>(sup) * kSubBins, sc.cap2, sc.b_keys, sc.b_vals, sc.ctl, sc);
}

// ------------------------------------------------------------------ the region insert
template <typename T>
__device__ __forceinline__ T
lds_load(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T>
__device__ __forceinline__ T
lds_cas(T* p, T expected, T desired)
{
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;
}

// the value word of a free slot in the LDS copy of a region (maps with 8-byte keys), and of a slot
// whose claim is under way
constexpr uint32_t kRegionUnclaimed = 0xFFFFFFFFU;
constexpr uint32_t kRegionClaimed = 0xFFFFFFFEU;

template <typename K, typename M>
struct region_claim
{
    // the claim goes through the low word of the value (offset 8 of the slot)
    static constexpr bool by_value = sizeof(K) == 8 && sizeof(M) >= 4;
};

template <typename K>
struct region_claim<K, void>
{
    static constexpr bool by_value = false;
};

enum class region_result
{
    inserted,
    present,
    deferred
};

// One key into the LDS copy of its region (buckets [0, nb) of the copy are table buckets b0 + ...).
// The protocol of the per-key kernel: scan the probe sequence to its end (a bucket with an EMPTY slot
// or ovf == 0) comparing keys and noting the first free slot in the key's claim order, then claim the
// first free slot from there on by CAS, comparing keys on the way; a CAS lost to another key resumes
// after that slot.  A sequence that leaves the region is deferred before anything is written.
// In LDS, the 16-byte pieces of bucket b are rotated by b (piece q at position (q + b) mod G): lanes
// that read the same slot of different buckets then spread over the banks instead of all meeting in
// the two banks a 128-byte stride maps one slot to.  The copy in and out applies and undoes it.
template <int G>
__device__ __forceinline__ uint32_t
lds_piece(uint32_t b, uint32_t q)
{
    return b * G + ((q + b) & (G - 1));
}

template <typename K, typename M, int G, int W>
__device__ __forceinline__ region_result
region_insert_one(typename slot_traits<K, M>::value_type* lslots, uint32_t* lovf, uint32_t nb, uint32_t lb, uint32_t o,
                  K key, uint64_t val)
{
    using value_type = typename slot_traits<K, M>::value_type;
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    constexpr uint32_t kSeg = W < 16 ? W : 16; // keys of a bucket read together
    constexpr uint32_t SPL = W / G;            // slots per 16-byte piece
    char* lbytes = reinterpret_cast<char*>(lslots);
    auto kp = [&](uint32_t b, uint32_t j) -> K* {
        const uint32_t x = (o + j) & (W - 1);
        return reinterpret_cast<K*>(lbytes + lds_piece<G>(b, x / SPL) * 16 + (x % SPL) * sizeof(value_type));
    };
    auto ovp = [&](uint32_t b) -> uint32_t* {
        if constexpr (kInline)
            return reinterpret_cast<uint32_t*>(lbytes + lds_piece<G>(b, 0) * 16 + slot_traits<K, M>::ovf_offset);
        else
            return lovf + b;
    };
    uint32_t b = lb;
    int32_t fb = -1;
    uint32_t fj = 0;
    while (true)
    {
        bool saw_empty = false;
        const uint32_t ov = lds_load(ovp(b));
        for (uint32_t j0 = 0; j0 < W; j0 += kSeg)
        {
            K kk[kSeg];
#pragma unroll
            for (uint32_t j = 0; j < kSeg; ++j) kk[j] = lds_load(kp(b, j0 + j));
            bool hit = false;
#pragma unroll
            for (uint32_t j = 0; j < kSeg; ++j)
            {
                hit |= kk[j] == key;
                if (fb < 0 && (kk[j] == empty_key<K>() || kk[j] == tomb_key<K>()))
                {
                    fb = static_cast<int32_t>(b);
                    fj = j0 + j;
                }
                saw_empty |= kk[j] == empty_key<K>();
            }
            if (hit) return region_result::present;
        }
        if (saw_empty || ov == 0) break; // absent
        if (b + 1 >= nb) return region_result::deferred; // the sequence runs on into the next region
        ++b;
    }
    // claims: the first free slot from (cb, cj) on in the key's order; after a lost claim the bucket is
    // read again, all its keys in flight together (a walk one slot at a time puts a chain of dependent
    // LDS round trips on every lost claim, and a wavefront waits for its longest chain)
    uint32_t cb = fb >= 0 ? static_cast<uint32_t>(fb) : b + 1;
    uint32_t cj = fb >= 0 ? fj : 0U;
    while (cb < nb)
    {
        int32_t nj = -1;
        K seen = empty_key<K>();
        for (uint32_t j0 = 0; j0 < W; j0 += kSeg)
        {
            K kk[kSeg];
#pragma unroll
            for (uint32_t j = 0; j < kSeg; ++j) kk[j] = lds_load(kp(cb, j0 + j));
            bool hit = false;
#pragma unroll
            for (uint32_t j = 0; j < kSeg; ++j)
            {
                hit |= kk[j] == key;
                if (nj < 0 && j0 + j >= cj && (kk[j] == empty_key<K>() || kk[j] == tomb_key<K>()))
                {
                    nj = static_cast<int32_t>(j0 + j);
                    seen = kk[j];
                }
            }
            if (hit) return region_result::present;
        }
        if (nj < 0)
        {
            ++cb;
            cj = 0;
            continue;
        }
        cj = static_cast<uint32_t>(nj);
        K old;
        bool won;
        if constexpr (region_claim<K, M>::by_value)
        {
            uint32_t* vw = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(kp(cb, cj)) + sizeof(K));
            old = seen;
            won = false;
            if (lds_cas(vw, kRegionUnclaimed, kRegionClaimed) == kRegionUnclaimed)
            {
                old = lds_load(kp(cb, cj));
                if (old == empty_key<K>() || old == tomb_key<K>())
                {
                    __hip_atomic_store(kp(cb, cj), key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    // the key before the value (which may equal the mark): a wavefront's LDS operations
                    // are performed in the order they are issued, so only the compiler must keep it
                    asm volatile("" ::: "memory");
                    old = seen;
                    won = true;
                }
                else
                {
                    // the slot already holds a key whose value's low word is the mark: put it back
                    __hip_atomic_store(vw, kRegionUnclaimed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            // lost the word: old is still a free-slot key until the winner stores its key, and the slot is
            // read again (every claimer stores its key in the same pass as its compare-and-swap)
        }
        else
        {
            old = lds_cas(kp(cb, cj), seen, key);
            won = old == seen;
        }
        if (won)
        {
            store_mapped<K, M>(reinterpret_cast<value_type*>(kp(cb, cj)), val);
            for (uint32_t x = lb; x < cb; ++x) atomicAdd(ovp(x), 1U);
            if constexpr (kInline && W > 1)
            {
                // stored past its home bucket: the home bucket's away counter of the key's fingerprint
                // (slot 1 + f/2's padding word, half f mod 2; it stops at saturation, no carry)
                if (cb != lb)
                {
                    const uint32_t f = away_index(static_cast<uint64_t>(key), 2U * (W - 1));
                    uint32_t* w = reinterpret_cast<uint32_t*>(lbytes + lds_piece<G>(lb, 1 + f / 2) * 16 +
                                                              slot_traits<K, M>::ovf_offset);
                    const uint32_t sh = 16 * (f & 1);
                    uint32_t cur = lds_load(w);
                    while (((cur >> sh) & 0xFFFFU) < kAwaySaturated)
                    {
                        const uint32_t got = lds_cas(w, cur, cur + (1U << sh));
                        if (got == cur) break;
                        cur = got;
                    }
                }
            }
            return region_result::inserted;
        }
        if (old == key) return region_result::present;
        if (old != empty_key<K>() && old != tomb_key<K>()) ++cj;
    }
    return region_result::deferred; // no free slot left in the region
}

// the LDS copy of a region: its buckets, and its ovf counts when they live in an array (exactly
// kRegionBytes for kinds whose counts sit in the buckets, so five workgroups share a CU)
template <typename K, typename M, int G, bool kInline = slot_traits<K, M>::ovf_in_bucket>
struct region_lds
{
    using value_type = typename slot_traits<K, M>::value_type;
    static constexpr uint32_t RB = region_buckets<G>();
    static constexpr uint32_t W = 16 * G / sizeof(value_type);
    value_type slots[RB * W];
    uint32_t ovf[RB];
    __device__ uint32_t* ovf_array() { return ovf; }
};

template <typename K, typename M, int G>
struct region_lds<K, M, G, true>
{
    using value_type = typename slot_traits<K, M>::value_type;
    static constexpr uint32_t RB = region_buckets<G>();
    static constexpr uint32_t W = 16 * G / sizeof(value_type);
    value_type slots[RB * W];
    __device__ uint32_t* ovf_array() { return nullptr; }
};

template <typename K, typename M, int G>
__global__ __launch_bounds__(kRegionThreads) void
k_region_insert(table_ref<K, M> t, region_scratch sc)
{
    using L = region_lds<K, M, G>;
    using value_type = typename slot_traits<K, M>::value_type;
    constexpr uint32_t RB = L::RB;
    constexpr uint32_t W = L::W;
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    constexpr uint32_t kPieces = static_cast<uint32_t>(kRegionBytes / 16 / kRegionThreads); // 16 B per thread each
    __shared__ L s;
    if (sc.ctl->per_key) return;
    const uint32_t r = blockIdx.x;
    const uint32_t f = sc.region_fill[r];
    const uint32_t count = f < sc.cap2 ? f : sc.cap2;
    if (count == 0) return; // no key of this launch lands here: the region is not touched
    const uint64_t kbase = static_cast<uint64_t>(r) * sc.cap2;
    const uint64_t b0 = static_cast<uint64_t>(r) * RB;
    const uint32_t nb = static_cast<uint32_t>((t.nbuckets - b0) < RB ? (t.nbuckets - b0) : RB);
    const uint32_t n16 = nb * W * static_cast<uint32_t>(sizeof(value_type)) / 16;
    // the region and this thread's first keys are loaded together (all in flight at once); the keys
    // then come through a ring of four rounds, each round's load issued four rounds ahead of its use
    // (the region goes straight to LDS: each lane's 16 bytes land at its wavefront's base + lane * 16)
    const uint4* g = reinterpret_cast<const uint4*>(t.slots + b0 * W);
    uint4* l = reinterpret_cast<uint4*>(s.slots);
    const uint32_t wbase = threadIdx.x & ~63U;
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j)
    {
        const uint32_t p = j * kRegionThreads + threadIdx.x; // LDS position of this lane's piece
        const uint32_t b = p / G, q = (p - (p / G) * G - b) & (G - 1);
        if (p < n16) __builtin_amdgcn_global_load_lds(g + b * G + q, l + j * kRegionThreads + wbase, 16, 0, 0);
    }
    K k0 = 0, k1 = 0, k2 = 0, k3 = 0;
    uint64_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    auto fetch = [&](uint32_t q, K* k, uint64_t* v) {
        const uint32_t e = q * kRegionThreads + threadIdx.x;
        if (e < count) region_elem<K, M>::get(sc.b_keys, sc.b_vals, kbase + e, k, v);
    };
    fetch(0, &k0, &v0);
    fetch(1, &k1, &v1);
    fetch(2, &k2, &v2);
    fetch(3, &k3, &v3);
    uint32_t* lovf = s.ovf_array();
    if constexpr (!kInline)
        for (uint32_t i = threadIdx.x; i < nb; i += kRegionThreads) lovf[i] = t.ovf[b0 + i];
    __builtin_amdgcn_s_waitcnt(0); // the direct loads count as vector memory operations
    __syncthreads();
    if constexpr (region_claim<K, M>::by_value)
    {
        // every free slot of the copy starts unclaimed
        for (uint32_t i = threadIdx.x; i < nb * W; i += kRegionThreads)
        {
            const uint32_t b = i / W, x = i - b * W;
            char* sp = reinterpret_cast<char*>(s.slots) + lds_piece<G>(b, x / (W / G)) * 16 + (x % (W / G)) * sizeof(value_type);
            const K kk = *reinterpret_cast<const K*>(sp);
            if (kk == empty_key<K>() || kk == tomb_key<K>()) *reinterpret_cast<uint32_t*>(sp + sizeof(K)) = kRegionUnclaimed;
        }
        __syncthreads();
    }
    const bool ordered = t.nbuckets < kOrderedClaimBuckets;
    int64_t added = 0;
    auto one = [&](uint32_t q, K k, uint64_t v) {
        const uint32_t e = q * kRegionThreads + threadIdx.x;
        if (e >= count) return;
        const uint64_t hb = home_bucket(static_cast<uint64_t>(k), t.nbuckets);
        const uint32_t o = ordered ? claim_offset(static_cast<uint64_t>(k), W) : 0U;
        // every key of the region's bin is homed in the region; one that is not (never, by the
        // partition) would go to the per-key kernel rather than address LDS outside the copy
        const region_result res = (hb >= b0 && hb - b0 < nb)
                                          ? region_insert_one<K, M, G, W>(s.slots, lovf, nb, static_cast<uint32_t>(hb - b0), o, k, v)
                                          : region_result::deferred;
        if (res == region_result::inserted)
            ++added;
        else if (res == region_result::deferred)
            region_elem<K, M>::put(sc.a_keys, sc.a_vals, atomicAdd(&sc.ctl->deferred, 1U), k, v);
    };
    const uint32_t rounds = (count + kRegionThreads - 1) / kRegionThreads;
    for (uint32_t q = 0; q < rounds; q += 4)
    {
        one(q, k0, v0);
        fetch(q + 4, &k0, &v0);
        one(q + 1, k1, v1);
        fetch(q + 5, &k1, &v1);
        one(q + 2, k2, v2);
        fetch(q + 6, &k2, &v2);
        one(q + 3, k3, v3);
        fetch(q + 7, &k3, &v3);
    }
    __syncthreads();
    // LDS -> region
    uint4* gw = reinterpret_cast<uint4*>(t.slots + b0 * W);
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j)
    {
        const uint32_t p = j * kRegionThreads + threadIdx.x;
        const uint32_t b = p / G, q = (p - (p / G) * G - b) & (G - 1);
        if (p < n16) gw[b * G + q] = l[p];
    }
    if constexpr (!kInline)
        for (uint32_t i = threadIdx.x; i < nb; i += kRegionThreads) t.ovf[b0 + i] = lovf[i];
    // one size update per workgroup (the wave sums reuse the LDS copy once it has been written back)
    int64_t x = added;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += static_cast<int64_t>(shfl64(static_cast<uint64_t>(x), lane_id() ^ off));
    __syncthreads();
    int64_t* wave = reinterpret_cast<int64_t*>(s.slots);
    if (lane_id() == 0) wave[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int64_t total = 0;
        for (int w = 0; w < kRegionThreads / 64; ++w) total += wave[w];
        if (total)
            __hip_atomic_fetch_add(t.stripes + static_cast<uint64_t>(r & (t.nstripes - 1)) * kStripeStride, total,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The per-key kernel over a list whose length lives in device memory (the deferred keys, or the
// spilled ones).  A short list (the deferred keys: those whose probe sequences run long) gives each
// lane group one key, so no group walks several long sequences one after another; a long list (keys
// a skewed input spilled, mostly copies of hot keys) takes one key per lane.
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(256) void
k_insert_list(table_ref<K, M> t, const void* __restrict__ keys, const void* __restrict__ vals,
              const uint32_t* __restrict__ count)
{
    const int64_t cnt = static_cast<int64_t>(*count);
    const int64_t nthreads = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const bool sparse = cnt * G <= 4 * nthreads;
    const int64_t per = sparse ? G : 1; // threads per key
    const int64_t stride = nthreads / per;
    const int lane = lane_id();
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * (blockDim.x / per); base < cnt; base += stride)
    {
        const int64_t j = base + static_cast<int64_t>(threadIdx.x) / per;
        const bool mine = (sparse ? (lane & (G - 1)) == 0 : true) && j < cnt;
        insert_body<K, M, G, R>(t, nullptr, keys, vals, j, mine, nullptr, false);
    }
}

// Whether a default partitioned insert goes ahead: kSampleKeys keys spread evenly over the launch go
// into a small LDS table, and when more than one in eight of them repeats one sampled before (a skewed
// input: a few keys hold much of the launch and would overflow their bins) the launch is left to the
// per-key kernel, whose compare-and-swaps on those keys hit in cache.  A uniform input of distinct keys
// shows no repeat.
constexpr uint32_t kSampleKeys = 1024;

template <typename K, typename M>
__global__ __launch_bounds__(256) void
k_region_sample(const void* __restrict__ values, const void* __restrict__ keys, const void* __restrict__ vals,
                int64_t n, region_ctl* ctl)
{
    constexpr uint32_t kSlots = 2 * kSampleKeys;
    __shared__ uint64_t tab[kSlots];
    __shared__ uint32_t repeats;
    constexpr uint64_t kFree = ~static_cast<uint64_t>(0);
    for (uint32_t i = threadIdx.x; i < kSlots; i += blockDim.x) tab[i] = kFree;
    if (threadIdx.x == 0) repeats = 0;
    __syncthreads();
    const uint32_t ns = n < static_cast<int64_t>(kSampleKeys) ? static_cast<uint32_t>(n) : kSampleKeys;
    uint32_t mine = 0;
    for (uint32_t j = threadIdx.x; j < ns; j += blockDim.x)
    {
        // n < 2^31 on this path, so j * n fits
        const int64_t e = static_cast<int64_t>(j) * n / ns;
        K k;
        uint64_t v;
        fetch_element<K, M>(values, keys, vals, e, &k, &v);
        const uint64_t x = static_cast<uint64_t>(k);
        if (x == kFree) continue; // the table's own free mark: not sampled
        uint32_t h = static_cast<uint32_t>(mix64(x)) & (kSlots - 1);
        while (true)
        {
            const uint64_t got = lds_cas(&tab[h], kFree, x);
            if (got == kFree) break;
            if (got == x)
            {
                ++mine;
                break;
            }
            h = (h + 1) & (kSlots - 1);
        }
    }
    if (mine) atomicAdd(&repeats, mine);
    __syncthreads();
    if (threadIdx.x == 0) ctl->per_key = repeats * 8 > ns ? 1U : 0U;
}

// The whole launch by the per-key insert when the sample chose it (else nothing): a grid-stride loop,
// so the launch that finds nothing to do stays small.
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(64) void
k_insert_per_key_if(table_ref<K, M> t, const void* __restrict__ values, const void* __restrict__ keys,
                    const void* __restrict__ vals, int64_t n, const region_ctl* __restrict__ ctl)
{
    if (ctl->per_key == 0) return;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride)
    {
        const int64_t i = base + threadIdx.x;
        insert_body<K, M, G, R>(t, values, keys, vals, i, i < n, nullptr, false);
    }
}

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_REGION_HPP
