// Typed dispatch of the C ABI onto the kernels: one instantiation set per key/value kind, each
// compiled in its own translation unit (kind_*.hip).
#ifndef STDGPU_AMD_HTAB_OPS_HPP
#define STDGPU_AMD_HTAB_OPS_HPP

#include <cmath>
#include <new>
#include <type_traits>

#include "htab_freelist.hpp"
#include "htab_internal.hpp"

namespace htab_host
{

constexpr int
default_rb(int g)
{
    return g < 8 ? g : 8;
}

// Block sizes of the persistent batched kernel: all three for the benchmarked map kind, 256 elsewhere
// (keeps the build small).
template <typename K, typename M>
constexpr bool batch_all_blocks = std::is_same_v<K, uint64_t> && std::is_same_v<M, uint32_t>;

} // namespace htab_host

namespace htab_host
{

using namespace htab_dev;

// ------------------------------------------------------------------ typed dispatch
template <typename K, typename M, int G, int R>
struct ops
{
    using T = table_ref<K, M>;

    // The partitioned path (htab_region.hpp) for a launch without per-element flags on a table whose
    // limit is its slot count: always when the table asks for it (insert_path 2), and by default
    // (insert_path 0) for a launch of half to all of the slots of a table larger than the 256 MB
    // Infinity Cache, whose per-key compare-and-swaps go to HBM.  At 2^26 keys into 2^23 buckets of
    // map<u64,u32> the partitioned launches take 3.5 ms against 4.3 ms; a launch of a fifth of the slots
    // (2^24 keys) takes 1.6 ms against 1.3 ms, since the region copy streams the whole table whatever
    // the launch holds.  A launch of more keys than slots (copies of keys, which the per-key kernel
    // finds in cache) and tables in the cache keep one CAS per key.
    static bool use_regions(const htab* h, int64_t n, const uint8_t* ins)
    {
        if (ins || h->view.nstripes == 1 || n <= 0 || n > 0x7FFFFFFFLL || h->insert_path == 1) return false;
        const uint64_t nreg = (h->view.nbuckets + region_buckets<G>() - 1) / region_buckets<G>();
        if (nreg > kMaxRegions) return false;
        if (h->insert_path == 2) return true;
        const uint64_t slots = T::from_view(h->view).nslots();
        const uint64_t bytes = slots * sizeof(typename T::value_type);
        const uint64_t keys = static_cast<uint64_t>(n);
        return bytes > kRegionDefaultMinBytes && 2 * keys >= slots && keys <= slots;
    }

    static int insert_regions(htab* h, const void* values, const void* keys, const void* vals, int64_t n, hipStream_t s)
    {
        using VT = std::conditional_t<std::is_void_v<M>, char, M>;
        constexpr size_t kb = sizeof(K), vb = slot_traits<K, M>::is_map ? sizeof(VT) : 0;
        const T t = T::from_view(h->view);
        constexpr uint32_t rbk = region_buckets<G>();
        const uint32_t nreg = static_cast<uint32_t>((t.nbuckets + rbk - 1) / rbk);
        const uint32_t nsuper = (nreg + kSubBins - 1) / kSubBins;
        const size_t nrs = static_cast<size_t>(nsuper) * kSubBins;
        // bin capacities: a bin's expected share of uniformly hashed keys, plus 6 % (super-bins) or
        // 12.5 % (regions) and four standard deviations; what a bin cannot take spills
        const double per_region = static_cast<double>(n) / nreg;
        const double e1 = per_region * kSubBins / kSuperCopies;
        const uint32_t cap1 = static_cast<uint32_t>(std::ceil(e1 * 1.0625 + 4.0 * std::sqrt(e1) + 64.0));
        const uint32_t cap2 = static_cast<uint32_t>(std::ceil(per_region * 1.125 + 4.0 * std::sqrt(per_region) + 32.0));
        auto up = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t na = static_cast<size_t>(nsuper) * kSuperCopies * cap1, nbb = nrs * cap2, ns = static_cast<size_t>(n);
        size_t off = up(sizeof(region_ctl));
        const size_t o_fill = off;
        off += up(nrs * 4);
        const size_t o_ak = off;
        off += up(na * kb);
        const size_t o_av = off;
        off += up(na * vb);
        const size_t o_bk = off;
        off += up(nbb * kb);
        const size_t o_bv = off;
        off += up(nbb * vb);
        const size_t o_sk = off;
        off += up(ns * kb);
        const size_t o_sv = off;
        off += up(ns * vb);
        // the deferred keys reuse buffer A once pass 2 has read it: at most every element defers
        if (na < ns) return fail(HTAB_E_LIMIT, "partitioned insert: super-bin capacity below the launch");
        if (h->part_bytes < off)
        {
            if (h->part) (void)hipFree(h->part);
            h->part = nullptr;
            h->part_bytes = 0;
            if (hipMalloc(&h->part, off) != hipSuccess)
            {
                (void)hipGetLastError();
                h->part = nullptr;
                // the default choice falls back to one compare-and-swap per key
                if (h->insert_path == 0) return kRegionsDeclined;
                return fail(HTAB_E_OOM, "partitioned insert scratch");
            }
            h->part_bytes = off;
        }
        char* base = static_cast<char*>(h->part);
        region_ctl* ctl = reinterpret_cast<region_ctl*>(base);
        region_scratch sc{ctl,
                          reinterpret_cast<uint32_t*>(base + o_fill),
                          base + o_ak,
                          vb ? base + o_av : nullptr,
                          base + o_bk,
                          vb ? base + o_bv : nullptr,
                          base + o_sk,
                          vb ? base + o_sv : nullptr,
                          cap1,
                          cap2};
        const uint32_t nchunks = static_cast<uint32_t>((n + kPartChunk - 1) / kPartChunk);
        const uint32_t chunks_per_super = (cap1 + kPartChunk - 1) / kPartChunk;
        uint32_t rgrid = static_cast<uint32_t>((nrs + 255) / 256);
        if (rgrid < 2) rgrid = 2;
        if (rgrid > 1024) rgrid = 1024;
        hipLaunchKernelGGL(k_region_reset<>, dim3(rgrid), dim3(256), 0, s, ctl, sc.region_fill, static_cast<uint32_t>(nrs));
        // the default choice first samples the launch for copies of keys (k_region_sample)
        const bool sampled = h->insert_path == 0;
        if (sampled)
            hipLaunchKernelGGL((k_region_sample<K, M>), dim3(1), dim3(256), 0, s, values, keys, vals, n, ctl);
        hipLaunchKernelGGL((k_part1_scatter<K, M>), dim3(nchunks), dim3(kPartThreads), 0, s, t, values, keys, vals, n,
                           rbk, nsuper, sc);
        hipLaunchKernelGGL((k_part2_scatter<K, M>), dim3(nsuper * kSuperCopies * chunks_per_super), dim3(kPartThreads), 0,
                           s, t, rbk, chunks_per_super, nsuper, sc);
        hipLaunchKernelGGL((k_region_insert<K, M, G>), dim3(nreg), dim3(kRegionThreads), 0, s, t, sc);
        // what the regions could not place, what the bins could not take and the two reserved key
        // values, by the per-key kernel
        hipLaunchKernelGGL((k_insert_list<K, M, G, R>), dim3(1024), dim3(256), 0, s, t, sc.a_keys, sc.a_vals,
                           &ctl->deferred);
        hipLaunchKernelGGL((k_insert_list<K, M, G, R>), dim3(1024), dim3(256), 0, s, t, sc.s_keys, sc.s_vals,
                           &ctl->spilled);
        if (sampled)
        {
            // one-wavefront workgroups, as the per-key kernel runs best, at most 65536 of them: with one
            // workgroup per 64 keys the launch that finds nothing to do cost a uniform 2^26-key insert
            // 0.2 ms
            const int64_t blocks = (n + 63) / 64;
            hipLaunchKernelGGL((k_insert_per_key_if<K, M, G, R>), dim3(static_cast<uint32_t>(blocks < 65536 ? blocks : 65536)),
                               dim3(64), 0, s, t, values, keys, vals, n, ctl);
        }
        return launch_status("partitioned insert launch");
    }

    static int insert(htab* h, const void* values, const void* keys, const void* vals, int64_t n, uint8_t* ins,
                      hipStream_t s)
    {
        if (use_regions(h, n, ins))
        {
            const int rc = insert_regions(h, values, keys, vals, n, s);
            if (rc != kRegionsDeclined) return rc;
        }
        const int64_t grid = (n + h->block - 1) / h->block;
        const T t = T::from_view(h->view);
        const bool one_stripe = h->view.nstripes == 1;
        uint32_t* plan = one_stripe ? plan_word(h) : nullptr;
        if (one_stripe) hipLaunchKernelGGL((k_limit_plan<K, M>), dim3(1), dim3(64), 0, s, t, n, plan);
        hipLaunchKernelGGL((k_insert<K, M, G, R>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s, t, values,
                           keys, vals, n, ins, plan);
        if (one_stripe) hipLaunchKernelGGL((k_limit_fold<K, M>), dim3(1), dim3(64), 0, s, t);
        return launch_status("insert launch");
    }

    static int find(htab* h, const void* keys, int64_t n, uint8_t* hits, void* vals, hipStream_t s)
    {
        const int64_t grid = (n + h->block - 1) / h->block;
        hipLaunchKernelGGL((k_find<K, M, G, R>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s,
                           T::from_view(h->view), static_cast<const K*>(keys), n, hits, vals);
        return launch_status("find launch");
    }

    static int erase(htab* h, const void* keys, int64_t n, uint8_t* del, hipStream_t s)
    {
        const int64_t grid = (n + h->block - 1) / h->block;
        const T t = T::from_view(h->view);
        const bool one_stripe = h->view.nstripes == 1;
        hipLaunchKernelGGL((k_erase<K, M, G, R>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s, t,
                           static_cast<const K*>(keys), n, del, one_stripe ? 1 : 0);
        if (one_stripe) hipLaunchKernelGGL((k_limit_fold<K, M>), dim3(1), dim3(64), 0, s, t);
        return launch_status("erase launch");
    }
};

template <typename K, typename M, int G>
struct batch_ops
{
    template <int R, int B>
    static int launch_rb(htab_batch* b, int32_t nsegs, hipStream_t s)
    {
        hipLaunchKernelGGL((k_batch<K, M, G, R, B>), dim3(b->grid), dim3(B), 0, s,
                           static_cast<const table_ref<K, M>*>(b->tables), b->segs, nsegs, b->chunk_start);
        return launch_status("batch launch");
    }

    template <int B>
    static int launch_b(htab_batch* b, int32_t nsegs, hipStream_t s)
    {
        // every probe batch for the benchmarked kind; the default one elsewhere
        if constexpr (batch_all_blocks<K, M>)
        {
            if constexpr (G >= 8)
            {
                if (b->rb == 8) return launch_rb<8, B>(b, nsegs, s);
            }
            if constexpr (G >= 4)
            {
                if (b->rb == 4) return launch_rb<4, B>(b, nsegs, s);
            }
            if constexpr (G >= 2)
            {
                if (b->rb == 2) return launch_rb<2, B>(b, nsegs, s);
            }
        }
        return launch_rb<default_rb(G), B>(b, nsegs, s);
    }

    static int build(htab_batch* b, htab* const* tables)
    {
        using T = table_ref<K, M>;
        T* host = new (std::nothrow) T[b->ntables];
        if (!host) return fail(HTAB_E_OOM, "htab_batch_create: host allocation");
        for (int32_t t = 0; t < b->ntables; ++t) host[t] = T::from_view(tables[t]->view);
        hipError_t e = hipMemcpy(b->tables, host, sizeof(T) * b->ntables, hipMemcpyHostToDevice);
        delete[] host;
        if (e != hipSuccess) return fail_hip(e, "htab_batch_create: descriptor copy");
        return HTAB_OK;
    }

    static int launch(htab_batch* b, int32_t nsegs, hipStream_t s)
    {
        if constexpr (batch_all_blocks<K, M>)
        {
            if (b->block == 64) return launch_b<64>(b, nsegs, s);
            if (b->block == 512) return launch_b<512>(b, nsegs, s);
            if (b->block == 1024) return launch_b<1024>(b, nsegs, s);
        }
        return launch_b<256>(b, nsegs, s);
    }

    static int run(htab_batch* b, const int8_t* ops, int op, const int64_t* offsets, const void* keys,
                   const void* vals, uint8_t* flags, void* vals_out, hipStream_t s)
    {
        const int vb = std::is_void_v<M> ? 0 : static_cast<int>(sizeof(std::conditional_t<std::is_void_v<M>, char, M>));
        hipLaunchKernelGGL(k_batch_plan_offsets<>, dim3(1), dim3(1024), 0, s, offsets, b->ntables, ops, op,
                           static_cast<const char*>(keys), static_cast<const char*>(vals), flags,
                           static_cast<char*>(vals_out), static_cast<int>(sizeof(K)), vb, b->block, b->segs,
                           b->chunk_start);
        int rc = launch_status("batch plan");
        if (rc) return rc;
        return launch(b, b->ntables, s);
    }

    static int run_segments(htab_batch* b, int32_t nsegs, hipStream_t s)
    {
        hipLaunchKernelGGL(k_batch_plan<>, dim3(1), dim3(1024), 0, s, b->segs, nsegs, b->block, b->chunk_start);
        int rc = launch_status("batch plan");
        if (rc) return rc;
        return launch(b, nsegs, s);
    }
};

template <typename K, typename M>
struct table_ops
{
    using T = table_ref<K, M>;
    using value_type = typename slot_traits<K, M>::value_type;

    static int clear(htab* h, hipStream_t s)
    {
        const T t = T::from_view(h->view);
        const uint64_t area = (t.nslots() + 2) * sizeof(value_type);
        const uint64_t nchunks = (area + 15) / 16;
        uint64_t grid = (nchunks + 255) / 256;
        if (grid > 65536) grid = 65536;
        if (grid == 0) grid = 1;
        hipLaunchKernelGGL((k_clear<K, M>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, t, nchunks);
        int rc = launch_status("clear launch");
        if (rc) return rc;
        hipLaunchKernelGGL((k_clear_tail<K, M>), dim3(1), dim3(1), 0, s, t);
        return launch_status("clear launch");
    }

    // device_range (and contents).  scratch_ext, when given, holds the tile counts instead of the
    // table's own scratch (nothing is allocated); count_dev, when given, receives the entry count on
    // the device and the call returns without waiting for the stream.
    static int range(htab* h, int64_t* idx_out, void* keys_out, void* vals_out, int64_t cap, int64_t* n_out,
                     hipStream_t s, int64_t* scratch_ext = nullptr, int64_t* count_dev = nullptr)
    {
        const T t = T::from_view(h->view);
        const uint64_t total = t.nslots() + 2;
        const uint64_t ntiles = (total + kRangeTile - 1) / kRangeTile;
        int64_t* tiles = scratch_ext;
        if (!tiles)
        {
            const size_t need = (ntiles + 1) * sizeof(int64_t);
            if (h->scratch_bytes < need)
            {
                if (h->scratch) (void)hipFree(h->scratch);
                h->scratch = nullptr;
                h->scratch_bytes = 0;
                HTAB_HIP_TRY(hipMalloc(&h->scratch, need), "range scratch");
                h->scratch_bytes = need;
            }
            tiles = h->scratch;
        }
        hipLaunchKernelGGL((k_range_count<K, M>), dim3(static_cast<uint32_t>(ntiles)), dim3(kRangeThreads), 0, s, t,
                           total, tiles);
        int rc = launch_status("range count");
        if (rc) return rc;
        hipLaunchKernelGGL(k_scan_tiles<>, dim3(1), dim3(1024), 0, s, tiles, static_cast<int64_t>(ntiles));
        rc = launch_status("range scan");
        if (rc) return rc;
        hipLaunchKernelGGL((k_range_scatter<K, M>), dim3(static_cast<uint32_t>(ntiles)), dim3(kRangeThreads), 0, s, t,
                           total, tiles, idx_out, static_cast<K*>(keys_out), vals_out, cap);
        rc = launch_status("range scatter");
        if (rc) return rc;
        if (count_dev)
        {
            hipLaunchKernelGGL(k_copy_count<>, dim3(1), dim3(64), 0, s, tiles + ntiles, count_dev);
            return launch_status("range count store");
        }
        int64_t n = 0;
        HTAB_HIP_TRY(hipMemcpyAsync(&n, tiles + ntiles, sizeof(int64_t), hipMemcpyDeviceToHost, s), "range count copy");
        HTAB_HIP_TRY(hipStreamSynchronize(s), "range sync");
        *n_out = n;
        return HTAB_OK;
    }

    static int valid(htab* h, const uint64_t* hashes, int32_t* out, hipStream_t s)
    {
        const T t = T::from_view(h->view);
        // scratch: two counters, then one expected ovf count per bucket
        const size_t need = 2 * sizeof(unsigned long long) + t.nbuckets * sizeof(uint32_t);
        if (h->scratch_bytes < need)
        {
            if (h->scratch) (void)hipFree(h->scratch);
            h->scratch = nullptr;
            h->scratch_bytes = 0;
            HTAB_HIP_TRY(hipMalloc(&h->scratch, need), "valid scratch");
            h->scratch_bytes = need;
        }
        unsigned long long* bad = reinterpret_cast<unsigned long long*>(h->scratch);
        uint32_t* want = reinterpret_cast<uint32_t*>(bad + 2);
        HTAB_HIP_TRY(hipMemsetAsync(h->scratch, 0, need, s), "valid scratch clear");
        const uint64_t total = t.nslots() + 2;
        uint64_t grid = (total + 255) / 256;
        if (grid > 16384) grid = 16384;
        hipLaunchKernelGGL((k_valid<K, M>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, t, total, bad, want,
                           hashes);
        int rc = launch_status("valid launch");
        if (rc) return rc;
        uint64_t bgrid = (t.nbuckets + 255) / 256;
        if (bgrid > 16384) bgrid = 16384;
        hipLaunchKernelGGL((k_valid_overflow<K, M>), dim3(static_cast<uint32_t>(bgrid)), dim3(256), 0, s, t, want, bad);
        rc = launch_status("valid overflow launch");
        if (rc) return rc;
        if (t.away_counters() != 0)
        {
            hipLaunchKernelGGL((k_valid_away<K, M>), dim3(static_cast<uint32_t>(bgrid)), dim3(256), 0, s, t, bad,
                               hashes);
            rc = launch_status("valid away launch");
            if (rc) return rc;
        }
        unsigned long long res[2] = {0, 0};
        int64_t stripes[kMaxStripes * kStripeStride];
        HTAB_HIP_TRY(hipMemcpyAsync(res, bad, sizeof(res), hipMemcpyDeviceToHost, s), "valid copy");
        HTAB_HIP_TRY(hipMemcpyAsync(stripes, h->view.stripes, sizeof(stripes), hipMemcpyDeviceToHost, s), "valid copy");
        HTAB_HIP_TRY(hipStreamSynchronize(s), "valid sync");
        int64_t size = 0;
        for (uint32_t i = 0; i < kMaxStripes; ++i) size += stripes[i * kStripeStride];
        // a table whose limit is its slot count also holds the two reserved key values in their side
        // slots; a table with an exact limit counts them against it
        const int64_t bound = static_cast<int64_t>(h->view.limit) + (h->view.nstripes == 1 ? 0 : 2);
        *out = (res[0] == 0 && static_cast<int64_t>(res[1]) == size && size >= 0 && size <= bound) ? 1 : 0;
        return HTAB_OK;
    }
};

template <typename K, typename M, int G>
int
dispatch_g(htab_call& c)
{
    switch (c.op)
    {
        case call_op::insert:
        case call_op::find:
        case call_op::erase:
        {
            // probe batch: the table's setting, or per operation by default (find keeps 8 home
            // buckets per lane group in flight; insert and erase 4, which measured faster on gfx950)
            int rb = c.h->rb;
            if (rb == 0) rb = (c.op == call_op::find) ? default_rb(G) : (G < 4 ? G : 4);
            auto go = [&](auto o) -> int {
                using O = decltype(o);
                if (c.op == call_op::insert) return O::insert(c.h, c.values, c.keys, c.vals, c.n, c.flags, c.s);
                if (c.op == call_op::find) return O::find(c.h, c.keys, c.n, c.flags, c.vals_out, c.s);
                return O::erase(c.h, c.keys, c.n, c.flags, c.s);
            };
            if constexpr (G >= 8)
            {
                if (rb == 8) return go(ops<K, M, G, 8>{});
            }
            if constexpr (G >= 4)
            {
                if (rb == 4) return go(ops<K, M, G, 4>{});
                if (rb == 2) return go(ops<K, M, G, 2>{});
            }
            else
            {
                // one lane per bucket: each lane probes its own element
                if (rb == 1) return go(ops<K, M, G, 1>{});
            }
            return fail(HTAB_E_ARG, "bad probe batch");
        }
        case call_op::batch_build: return batch_ops<K, M, G>::build(c.b, c.tables);
        case call_op::batch_run_offsets:
            return batch_ops<K, M, G>::run(c.b, c.ops, c.bop, c.offsets, c.keys, c.vals, c.flags, c.vals_out, c.s);
        case call_op::batch_run_segments: return batch_ops<K, M, G>::run_segments(c.b, c.nsegs, c.s);
        case call_op::serve:
        {
            // per-operation probe batches of the plain calls
            const uint32_t grid = c.serve->chunk_start[c.serve->nsegs];
            hipLaunchKernelGGL((k_serve<K, M, G, default_rb(G), (G < 4 ? G : 4)>), dim3(grid), dim3(kServeBlock), 0,
                               c.s, static_cast<const table_ref<K, M>*>(c.b->tables), *c.serve);
            return launch_status("served launch");
        }
        default: return fail(HTAB_E_ARG, "bad call");
    }
}

template <typename K, typename M>
int
dispatch_kind(htab_call& c)
{
    switch (c.op)
    {
        case call_op::assign:
            if constexpr (std::is_void_v<M>)
            {
                return fail(HTAB_E_ARG, "htab_assign: sets have no mapped values");
            }
            else
            {
                hipLaunchKernelGGL((k_assign<K, M>), dim3(static_cast<uint32_t>((c.n + 255) / 256)), dim3(256), 0,
                                   c.s, table_ref<K, M>::from_view(c.h->view), static_cast<const K*>(c.keys),
                                   static_cast<const M*>(c.vals), c.n, c.flags);
                return launch_status("assign launch");
            }
        case call_op::take:
        case call_op::evict:
            if constexpr (std::is_void_v<M>)
            {
                return fail(HTAB_E_ARG, "free-list calls: sets have no mapped values");
            }
            else
            {
                const table_ref<K, M> t = table_ref<K, M>::from_view(c.h->view);
                int64_t* w = reinterpret_cast<int64_t*>(static_cast<char*>(c.h->mem) + kFreelistOffset);
                const bool take = c.op == call_op::take;
                hipLaunchKernelGGL(k_freelist_begin<>, dim3(1), dim3(64), 0, c.s, *c.ring, w, take ? 0 : 1);
                int rc = launch_status("free-list start");
                if (rc) return rc;
                // at most 2048 workgroups (one counter atomic each) unless a workgroup would own more
                // than kFreelistTiles tiles
                const uint64_t work = take ? static_cast<uint64_t>(c.n) : t.nslots();
                const uint64_t tiles = (work + kFreelistThreads - 1) / kFreelistThreads;
                uint64_t grid = tiles < 2048 ? tiles : 2048;
                if (grid == 0) grid = 1;
                if ((tiles + grid - 1) / grid > static_cast<uint64_t>(kFreelistTiles))
                    grid = (tiles + kFreelistTiles - 1) / kFreelistTiles;
                if (grid > 0x7FFFFFFFULL) return fail(HTAB_E_LIMIT, "free-list call: too many elements");
                if (take)
                    hipLaunchKernelGGL((k_take_values<K, M>), dim3(static_cast<uint32_t>(grid)), dim3(kFreelistThreads), 0,
                                       c.s, t, static_cast<const K*>(c.keys), c.flags, c.n, *c.ring, w);
                else
                    hipLaunchKernelGGL((k_evict_far<K, M>), dim3(static_cast<uint32_t>(grid)), dim3(kFreelistThreads), 0,
                                       c.s, t, *c.region, *c.ring, w);
                rc = launch_status(take ? "take launch" : "evict launch");
                if (rc) return rc;
                if (take)
                    hipLaunchKernelGGL(k_take_end<>, dim3(1), dim3(64), 0, c.s, *c.ring, w, c.count_dev);
                else
                    hipLaunchKernelGGL((k_evict_end<K, M>), dim3(1), dim3(64), 0, c.s, t, *c.region, *c.ring, w,
                                       c.count_dev);
                return launch_status("free-list end");
            }
        case call_op::serve_sizes:
            hipLaunchKernelGGL((k_serve_sizes<K, M>), dim3(1), dim3(256), 0, c.s,
                               static_cast<const table_ref<K, M>*>(c.b->tables), *c.serve_tabs, c.sizes_out);
            return launch_status("served sizes");
        case call_op::clear: return table_ops<K, M>::clear(c.h, c.s);
        case call_op::range: return table_ops<K, M>::range(c.h, c.idx_out, const_cast<void*>(c.keys), c.vals_out, c.cap,
                                                           c.n_out, c.s, c.scratch_ext, c.count_dev);
        case call_op::valid: return table_ops<K, M>::valid(c.h, c.hashes, c.valid_out, c.s);
        default: break;
    }
    const int lanes = c.h ? c.h->lanes : c.b->lanes;
    switch (lanes)
    {
        case 1: return dispatch_g<K, M, 1>(c);
        case 4: return dispatch_g<K, M, 4>(c);
        case 8: return dispatch_g<K, M, 8>(c);
        case 16: return dispatch_g<K, M, 16>(c);
        default: return fail(HTAB_E_ARG, "bad bucket_lanes");
    }
}

} // namespace htab_host

#endif // STDGPU_AMD_HTAB_OPS_HPP
