// Partitioned bulk insert: a large insert launch without per-element flags, on a table that counts
// its size on all stripes, runs as streaming passes instead of one random compare-and-swap per key.
//
// The bucket array is cut into regions of kRegionBytes (32 KB: 256 buckets of 128 B).  The keys are
// partitioned by the region of their home bucket in two LDS-staged counting passes (by region / 128,
// then by region % 128), each reading and writing the keys once with coalesced runs.  Then one
// workgroup per region copies the region into LDS, inserts the region's keys there with LDS atomics
// (the same claim protocol as the per-key kernel, so copies of one key in the launch still produce
// exactly one stored key), and writes the region back.  A key whose probe sequence runs past the
// region's last bucket (to the next region, which another workgroup owns meanwhile) is deferred, and
// so are the two reserved key values; the per-key kernel inserts the deferred keys afterwards, on
// the whole table.
//
// Why: random 64-bit CAS is performed memory-side on gfx950 (about 18 G/s over a table in HBM,
// 27 G/s inside the 256 MB Infinity Cache, no faster when each XCD keeps to its own slice), so the
// per-key kernel is bound near 18 G keys/s; the partitioned form moves about 8 GB for 2^26 keys into
// a 1.34 GB table at streaming rates.
//
// Correctness of the split: keys of one home bucket land in one region, and a key is inserted in the
// LDS copy only when its probe sequence ends inside the region (a bucket with an EMPTY slot or
// ovf == 0) or it is found there; slots only fill during an insert launch, so copies of one key
// either all resolve inside the region (against each other, by LDS CAS) or all defer.
#ifndef STDGPU_AMD_HTAB_REGION_HPP
#define STDGPU_AMD_HTAB_REGION_HPP

#include "htab_kernels.hpp"

namespace htab_dev
{

constexpr int kRegionBytes = 32768;
constexpr uint32_t kSubBins = 128;   // regions per super-bin (digit of the second pass)
constexpr uint32_t kMaxSuper = 512;  // super-bins (digit of the first pass)
constexpr uint32_t kMaxRegions = kSubBins * kMaxSuper;
constexpr int kPartThreads = 256;
constexpr int kPartItems = 16;                            // elements per thread of a partition chunk
constexpr int kPartChunk = kPartThreads * kPartItems;     // 4096
constexpr int kRegionThreads = 256;

template <int G>
__host__ __device__ constexpr uint32_t
region_buckets()
{
    return static_cast<uint32_t>(kRegionBytes / (16 * G));
}

// Device bookkeeping of one partitioned insert, at the start of the partition scratch.
struct region_ctl
{
    uint32_t super_count[kMaxSuper];
    uint32_t super_start[kMaxSuper + 1];
    uint32_t super_cursor[kMaxSuper];
    uint32_t chunk_start[kMaxSuper + 1]; // second pass: first chunk of each super-bin
    uint32_t reserved;                   // elements holding a reserved key value (top of buffer B)
    uint32_t deferred;                   // elements the region kernel left (bottom of buffer A)
};

// scratch layout: [region_ctl | region_count | region_start | region_cursor | A keys | A vals | B keys | B vals]
struct region_scratch
{
    region_ctl* ctl;
    uint32_t* region_count; // S * kSubBins
    uint32_t* region_start; // S * kSubBins + 1
    uint32_t* region_cursor;
    void* a_keys;
    void* a_vals;
    void* b_keys;
    void* b_vals;
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
    uint32_t v = x0 + x1;
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
k_region_reset(region_ctl* ctl, uint32_t* region_count, uint32_t nregion_slots)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMaxSuper) ctl->super_count[i] = 0;
    if (i == 0)
    {
        ctl->reserved = 0;
        ctl->deferred = 0;
    }
    for (uint32_t r = i; r < nregion_slots; r += gridDim.x * blockDim.x) region_count[r] = 0;
}

// ------------------------------------------------------------------ pass 1: by super-bin
template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part1_count(table_ref<K, M> t, const void* __restrict__ values, const void* __restrict__ keys,
              const void* __restrict__ vals, int64_t n, uint32_t rb, uint32_t nsuper, region_ctl* ctl)
{
    __shared__ uint32_t hist[kMaxSuper];
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads) hist[d] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kPartChunk;
#pragma unroll 4
    for (int j = 0; j < kPartItems; ++j)
    {
        const int64_t e = base + j * kPartThreads + threadIdx.x;
        if (e < n)
        {
            K k;
            uint64_t v;
            fetch_element<K, M>(values, keys, vals, e, &k, &v);
            if (!table_ref<K, M>::is_reserved(k))
                atomicAdd(&hist[region_of(static_cast<uint64_t>(k), t.nbuckets, rb) / kSubBins], 1U);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads)
        if (hist[d]) atomicAdd(&ctl->super_count[d], hist[d]);
}

// one workgroup: super-bin offsets, and the second pass's chunk plan
template <int = 0>
__global__ __launch_bounds__(kPartThreads) void
k_part1_scan(region_ctl* ctl, uint32_t nsuper)
{
    __shared__ uint32_t a[kMaxSuper];
    __shared__ uint32_t c[kMaxSuper];
    __shared__ uint32_t tmp[kPartThreads / 64 + 1];
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads)
    {
        a[d] = ctl->super_count[d];
        c[d] = (ctl->super_count[d] + kPartChunk - 1) / kPartChunk;
    }
    __syncthreads();
    const uint32_t total = lds_exclusive_scan(a, nsuper, tmp);
    const uint32_t chunks = lds_exclusive_scan(c, nsuper, tmp);
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads)
    {
        ctl->super_start[d] = a[d];
        ctl->super_cursor[d] = a[d];
        ctl->chunk_start[d] = c[d];
    }
    if (threadIdx.x == 0)
    {
        ctl->super_start[nsuper] = total;
        ctl->chunk_start[nsuper] = chunks;
    }
}

// Shared body of both scatter passes: the chunk's elements (their digits already known) are ordered
// by digit in LDS, each digit's run gets a range of the output by one atomic per digit, and the runs
// leave with consecutive threads writing consecutive elements.
template <typename K, typename M, int NB>
struct scatter_lds
{
    uint32_t hist[NB];
    uint32_t pre[NB];
    uint32_t gbase[NB];
    uint32_t tmp[kPartThreads / 64 + 1];
    K key[kPartChunk];
    std::conditional_t<slot_traits<K, M>::is_map, std::conditional_t<std::is_void_v<M>, char, M>, char>
            val[slot_traits<K, M>::is_map ? kPartChunk : 1];
};

// ------------------------------------------------------------------ pass 1 scatter
template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part1_scatter(table_ref<K, M> t, const void* __restrict__ values, const void* __restrict__ keys,
                const void* __restrict__ vals, int64_t n, uint32_t rb, uint32_t nsuper, region_ctl* ctl,
                region_scratch sc)
{
    __shared__ scatter_lds<K, M, kMaxSuper> s;
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads) s.hist[d] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kPartChunk;
    K k[kPartItems];
    uint64_t v[kPartItems];
    uint32_t dr[kPartItems]; // digit << 16 | rank within the chunk's run of that digit
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        const int64_t e = base + j * kPartThreads + threadIdx.x;
        dr[j] = ~0U;
        if (e < n)
        {
            fetch_element<K, M>(values, keys, vals, e, &k[j], &v[j]);
            if (table_ref<K, M>::is_reserved(k[j]))
            {
                // the two reserved key values go to the per-key kernel, from the top of buffer B
                const uint32_t p = atomicAdd(&ctl->reserved, 1U);
                region_elem<K, M>::put(sc.b_keys, sc.b_vals, static_cast<uint64_t>(n) - 1 - p, k[j], v[j]);
            }
            else
            {
                const uint32_t d = region_of(static_cast<uint64_t>(k[j]), t.nbuckets, rb) / kSubBins;
                dr[j] = (d << 16) | atomicAdd(&s.hist[d], 1U);
            }
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nsuper; d += kPartThreads)
    {
        s.pre[d] = s.hist[d];
        if (s.hist[d]) s.gbase[d] = atomicAdd(&ctl->super_cursor[d], s.hist[d]);
    }
    __syncthreads();
    const uint32_t total = lds_exclusive_scan(s.pre, nsuper, s.tmp);
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        if (dr[j] != ~0U)
        {
            const uint32_t p = s.pre[dr[j] >> 16] + (dr[j] & 0xFFFFU);
            s.key[p] = k[j];
            if constexpr (slot_traits<K, M>::is_map) s.val[p] = static_cast<M>(v[j]);
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < total; p += kPartThreads)
    {
        const K kk = s.key[p];
        const uint32_t d = region_of(static_cast<uint64_t>(kk), t.nbuckets, rb) / kSubBins;
        const uint64_t g = static_cast<uint64_t>(s.gbase[d]) + (p - s.pre[d]);
        region_elem<K, M>::put(sc.a_keys, sc.a_vals, g, kk, slot_traits<K, M>::is_map ? static_cast<uint64_t>(s.val[p]) : 0);
    }
}

// ------------------------------------------------------------------ pass 2: by region inside a super-bin
// workgroup -> (super-bin, element range) from the chunk plan; false for workgroups past the plan
__device__ __forceinline__ bool
pass2_range(const region_ctl* ctl, uint32_t nsuper, uint32_t* seg, uint32_t* lo, uint32_t* hi)
{
    const uint32_t c = blockIdx.x;
    if (c >= ctl->chunk_start[nsuper]) return false;
    uint32_t a = 0, b = nsuper - 1; // last super-bin whose first chunk is <= c
    while (a < b)
    {
        const uint32_t mid = (a + b + 1) >> 1;
        if (ctl->chunk_start[mid] <= c)
            a = mid;
        else
            b = mid - 1;
    }
    *seg = a;
    *lo = ctl->super_start[a] + (c - ctl->chunk_start[a]) * static_cast<uint32_t>(kPartChunk);
    const uint32_t end = ctl->super_start[a + 1];
    *hi = (*lo + kPartChunk < end) ? *lo + kPartChunk : end;
    return true;
}

template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part2_count(table_ref<K, M> t, uint32_t rb, uint32_t nsuper, region_ctl* ctl, region_scratch sc)
{
    __shared__ uint32_t hist[kSubBins];
    uint32_t seg = 0, lo = 0, hi = 0;
    if (!pass2_range(ctl, nsuper, &seg, &lo, &hi)) return;
    for (uint32_t d = threadIdx.x; d < kSubBins; d += kPartThreads) hist[d] = 0;
    __syncthreads();
    for (uint32_t e = lo + threadIdx.x; e < hi; e += kPartThreads)
    {
        const K k = static_cast<const K*>(sc.a_keys)[e];
        atomicAdd(&hist[region_of(static_cast<uint64_t>(k), t.nbuckets, rb) % kSubBins], 1U);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kSubBins; d += kPartThreads)
        if (hist[d]) atomicAdd(&sc.region_count[seg * kSubBins + d], hist[d]);
}

// one workgroup of kSubBins threads per super-bin: region offsets
template <int = 0>
__global__ __launch_bounds__(kSubBins) void
k_part2_scan(const region_ctl* ctl, uint32_t nsuper, region_scratch sc)
{
    __shared__ uint32_t tmp[kSubBins / 64 + 1];
    const uint32_t seg = blockIdx.x;
    const uint32_t r = seg * kSubBins + threadIdx.x;
    const uint32_t c = sc.region_count[r];
    const int lane = lane_id();
    const int wid = static_cast<int>(threadIdx.x) >> 6;
    uint32_t inc = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t y = shfl32(inc, lane >= off ? lane - off : lane);
        if (lane >= off) inc += y;
    }
    if (lane == 63) tmp[wid] = inc;
    __syncthreads();
    const uint32_t before = wid == 0 ? 0U : tmp[0];
    const uint32_t start = ctl->super_start[seg] + before + inc - c;
    sc.region_start[r] = start;
    sc.region_cursor[r] = start;
    if (seg == 0 && threadIdx.x == 0) sc.region_start[nsuper * kSubBins] = ctl->super_start[nsuper];
}

template <typename K, typename M>
__global__ __launch_bounds__(kPartThreads) void
k_part2_scatter(table_ref<K, M> t, uint32_t rb, uint32_t nsuper, region_ctl* ctl, region_scratch sc)
{
    __shared__ scatter_lds<K, M, kSubBins> s;
    uint32_t seg = 0, lo = 0, hi = 0;
    if (!pass2_range(ctl, nsuper, &seg, &lo, &hi)) return;
    for (uint32_t d = threadIdx.x; d < kSubBins; d += kPartThreads) s.hist[d] = 0;
    __syncthreads();
    K k[kPartItems];
    uint64_t v[kPartItems];
    uint32_t dr[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        const uint32_t e = lo + j * kPartThreads + threadIdx.x;
        dr[j] = ~0U;
        if (e < hi)
        {
            region_elem<K, M>::get(sc.a_keys, sc.a_vals, e, &k[j], &v[j]);
            const uint32_t d = region_of(static_cast<uint64_t>(k[j]), t.nbuckets, rb) % kSubBins;
            dr[j] = (d << 16) | atomicAdd(&s.hist[d], 1U);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kSubBins; d += kPartThreads)
    {
        s.pre[d] = s.hist[d];
        if (s.hist[d]) s.gbase[d] = atomicAdd(&sc.region_cursor[seg * kSubBins + d], s.hist[d]);
    }
    __syncthreads();
    const uint32_t total = lds_exclusive_scan(s.pre, kSubBins, s.tmp);
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
    {
        if (dr[j] != ~0U)
        {
            const uint32_t p = s.pre[dr[j] >> 16] + (dr[j] & 0xFFFFU);
            s.key[p] = k[j];
            if constexpr (slot_traits<K, M>::is_map) s.val[p] = static_cast<M>(v[j]);
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < total; p += kPartThreads)
    {
        const K kk = s.key[p];
        const uint32_t d = region_of(static_cast<uint64_t>(kk), t.nbuckets, rb) % kSubBins;
        const uint64_t g = static_cast<uint64_t>(s.gbase[d]) + (p - s.pre[d]);
        region_elem<K, M>::put(sc.b_keys, sc.b_vals, g, kk, slot_traits<K, M>::is_map ? static_cast<uint64_t>(s.val[p]) : 0);
    }
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
template <typename K, typename M, int W>
__device__ __forceinline__ region_result
region_insert_one(typename slot_traits<K, M>::value_type* lslots, uint32_t* lovf, uint32_t nb, uint32_t lb, uint32_t o,
                  K key, uint64_t val)
{
    using value_type = typename slot_traits<K, M>::value_type;
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    auto kp = [&](uint32_t b, uint32_t j) -> K* {
        return reinterpret_cast<K*>(lslots + b * W + ((o + j) & (W - 1)));
    };
    auto ovp = [&](uint32_t b) -> uint32_t* {
        if constexpr (kInline)
            return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(lslots + b * W) + slot_traits<K, M>::ovf_offset);
        else
            return lovf + b;
    };
    uint32_t b = lb;
    int32_t fb = -1;
    uint32_t fj = 0;
    while (true)
    {
        bool saw_empty = false;
        for (uint32_t j = 0; j < W; ++j)
        {
            const K kk = lds_load(kp(b, j));
            if (kk == key) return region_result::present;
            if (fb < 0 && (kk == empty_key<K>() || kk == tomb_key<K>()))
            {
                fb = static_cast<int32_t>(b);
                fj = j;
            }
            saw_empty |= kk == empty_key<K>();
        }
        if (saw_empty || lds_load(ovp(b)) == 0) break; // absent
        if (b + 1 >= nb) return region_result::deferred; // the sequence runs on into the next region
        ++b;
    }
    uint32_t cb = fb >= 0 ? static_cast<uint32_t>(fb) : b + 1;
    uint32_t cj = fb >= 0 ? fj : 0U;
    while (cb < nb)
    {
        bool found = false;
        K seen = empty_key<K>();
        for (; cj < W; ++cj)
        {
            const K x = lds_load(kp(cb, cj));
            if (x == key) return region_result::present;
            if (x == empty_key<K>() || x == tomb_key<K>())
            {
                seen = x;
                found = true;
                break;
            }
        }
        if (!found)
        {
            ++cb;
            cj = 0;
            continue;
        }
        const K old = lds_cas(kp(cb, cj), seen, key);
        if (old == seen)
        {
            store_mapped<K, M>(reinterpret_cast<value_type*>(kp(cb, cj)), val);
            for (uint32_t x = lb; x < cb; ++x) atomicAdd(ovp(x), 1U);
            return region_result::inserted;
        }
        if (old == key) return region_result::present;
        ++cj;
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
k_region_insert(table_ref<K, M> t, region_ctl* ctl, region_scratch sc)
{
    using L = region_lds<K, M, G>;
    using value_type = typename slot_traits<K, M>::value_type;
    constexpr uint32_t RB = L::RB;
    constexpr uint32_t W = L::W;
    constexpr bool kInline = slot_traits<K, M>::ovf_in_bucket;
    __shared__ L s;
    const uint32_t r = blockIdx.x;
    const uint32_t lo = sc.region_start[r], hi = sc.region_start[r + 1];
    if (lo == hi) return; // no key of this launch lands here: the region is not touched
    const uint64_t b0 = static_cast<uint64_t>(r) * RB;
    const uint32_t nb = static_cast<uint32_t>((t.nbuckets - b0) < RB ? (t.nbuckets - b0) : RB);
    // region -> LDS (16-byte pieces, coalesced)
    const uint4* g = reinterpret_cast<const uint4*>(t.slots + b0 * W);
    uint4* l = reinterpret_cast<uint4*>(s.slots);
    const uint32_t n16 = nb * W * static_cast<uint32_t>(sizeof(value_type)) / 16;
    for (uint32_t i = threadIdx.x; i < n16; i += kRegionThreads) l[i] = g[i];
    uint32_t* lovf = s.ovf_array();
    if constexpr (!kInline)
        for (uint32_t i = threadIdx.x; i < nb; i += kRegionThreads) lovf[i] = t.ovf[b0 + i];
    __syncthreads();
    const bool ordered = t.nbuckets < kOrderedClaimBuckets;
    int64_t added = 0;
    for (uint32_t e = lo + threadIdx.x; e < hi; e += kRegionThreads)
    {
        K k;
        uint64_t v;
        region_elem<K, M>::get(sc.b_keys, sc.b_vals, e, &k, &v);
        const uint32_t lb = static_cast<uint32_t>(home_bucket(static_cast<uint64_t>(k), t.nbuckets) - b0);
        const uint32_t o = ordered ? claim_offset(static_cast<uint64_t>(k), W) : 0U;
        const region_result res = region_insert_one<K, M, W>(s.slots, lovf, nb, lb, o, k, v);
        if (res == region_result::inserted)
            ++added;
        else if (res == region_result::deferred)
        {
            const uint32_t p = atomicAdd(&ctl->deferred, 1U);
            region_elem<K, M>::put(sc.a_keys, sc.a_vals, p, k, v);
        }
    }
    __syncthreads();
    // LDS -> region
    uint4* gw = reinterpret_cast<uint4*>(t.slots + b0 * W);
    for (uint32_t i = threadIdx.x; i < n16; i += kRegionThreads) gw[i] = l[i];
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

// The per-key kernel over a list whose length lives in device memory: the deferred elements
// (bottom of buffer A) or the reserved-key elements (top of buffer B, from_top).
template <typename K, typename M, int G, int R>
__global__ __launch_bounds__(256) void
k_insert_list(table_ref<K, M> t, const void* __restrict__ keys, const void* __restrict__ vals,
              const uint32_t* __restrict__ count, int64_t n_total, int from_top)
{
    const int64_t cnt = static_cast<int64_t>(*count);
    const int64_t start = from_top ? n_total - cnt : 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < cnt; base += stride)
    {
        const int64_t j = base + threadIdx.x;
        insert_body<K, M, G, R>(t, nullptr, keys, vals, start + j, j < cnt, nullptr, false);
    }
}

} // namespace htab_dev

#endif // STDGPU_AMD_HTAB_REGION_HPP
