// C ABI of the MI355X hashed containers: table lifetime, bulk operations, size, range,
// validity check and the free-list deque.  Kernels live in htab_kernels.hpp.
//
// Build: hipcc --offload-arch=gfx950 -O3 -shared -fPIC (see stdgpu_amd/build.py).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include <stdgpu_amd/htab.h>

#include "htab_kernels.hpp"

using namespace htab_dev;

namespace
{

thread_local std::string g_last_error;

int
fail(int code, const char* what)
{
    g_last_error = what;
    return code;
}

int
fail_hip(hipError_t e, const char* what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return HTAB_E_HIP;
}

#define HTAB_HIP_TRY(expr, what)                                                                                       \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return fail_hip(e_, what);                                                               \
    } while (0)

int
launch_status(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(e, what);
    return HTAB_OK;
}

int
value_bytes_of(int kind)
{
    switch (kind)
    {
        case HTAB_SET_U32: return 4;
        case HTAB_SET_U64: return 8;
        case HTAB_MAP_U64_U32: return 16;
        case HTAB_MAP_U32_U32: return 8;
        case HTAB_MAP_U64_U64: return 16;
        default: return 0;
    }
}

int
key_bytes_of(int kind)
{
    return (kind == HTAB_SET_U32 || kind == HTAB_MAP_U32_U32) ? 4 : 8;
}

// Reference sizing (src/stdgpu/impl/unordered_base_detail.cuh:34-48, 1263-1273): bucket_count =
// bit_ceil(ceil(float(capacity) / 1.0f)); excess = ceil(n (1 - (1 - 1/k)^(n-1))) + 1 in long double.
void
reference_sizing(int64_t capacity, int64_t* bucket_count, int64_t* max_size)
{
    const uint64_t want = static_cast<uint64_t>(std::ceil(static_cast<float>(capacity) / 1.0F));
    uint64_t bc = 1;
    while (bc < want) bc <<= 1;
    const long double k = static_cast<long double>(bc);
    const long double n = static_cast<long double>(capacity);
    const int64_t excess = static_cast<int64_t>(std::ceil(n * (1.0L - std::pow(1.0L - (1.0L / k), n - 1.0L)))) + 1;
    *bucket_count = static_cast<int64_t>(bc);
    *max_size = static_cast<int64_t>(bc) + excess;
}

constexpr size_t kCtlBytes = 8192; // stripes (32 x 128 B) + flags + scratch words
constexpr size_t kFlagsOffset = 4096;

} // namespace

struct htab
{
    int kind = 0;
    int lanes = 8;
    int block = 256;
    int64_t capacity = 0;
    htab_view view{};
    htab_info info{};
    void* mem = nullptr;
    size_t bytes = 0;
    int64_t* scratch = nullptr; // range/valid scratch, grown on demand
    size_t scratch_bytes = 0;
    int64_t* range_idx = nullptr; // table-owned device_range() indices (limit + 2 entries)
};

struct htab_batch
{
    int kind = 0;
    int lanes = 8;
    int32_t ntables = 0;
    void* tables = nullptr;            // device array of table_ref<K, M>
    int64_t* chunk_start = nullptr;    // ntables + 1
    unsigned long long* counter = nullptr;
    uint32_t grid = 0;                 // persistent workgroups
};

struct htab_graph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace
{

// ------------------------------------------------------------------ typed dispatch
template <typename K, typename M, int G>
struct ops
{
    using T = table_ref<K, M>;

    static int insert(htab* h, const void* values, const void* keys, const void* vals, int64_t n, uint8_t* ins,
                      hipStream_t s)
    {
        const int64_t grid = (n + h->block - 1) / h->block;
        hipLaunchKernelGGL((k_insert<K, M, G>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s,
                           T::from_view(h->view), values, keys, vals, n, ins);
        return launch_status("insert launch");
    }

    static int find(htab* h, const void* keys, int64_t n, uint8_t* hits, void* vals, hipStream_t s)
    {
        const int64_t grid = (n + h->block - 1) / h->block;
        hipLaunchKernelGGL((k_find<K, M, G>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s,
                           T::from_view(h->view), static_cast<const K*>(keys), n, hits, vals);
        return launch_status("find launch");
    }

    static int erase(htab* h, const void* keys, int64_t n, uint8_t* del, hipStream_t s)
    {
        const int64_t grid = (n + h->block - 1) / h->block;
        hipLaunchKernelGGL((k_erase<K, M, G>), dim3(static_cast<uint32_t>(grid)), dim3(h->block), 0, s,
                           T::from_view(h->view), static_cast<const K*>(keys), n, del);
        return launch_status("erase launch");
    }
};

template <typename K, typename M, int G>
struct batch_ops
{
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

    static int run(htab_batch* b, const int8_t* ops, int op, const int64_t* offsets, const void* keys,
                   const void* vals, uint8_t* flags, void* vals_out, hipStream_t s)
    {
        hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(1024), 0, s, offsets, b->ntables, b->chunk_start, b->counter);
        int rc = launch_status("batch plan");
        if (rc) return rc;
        hipLaunchKernelGGL((k_batch<K, M, G>), dim3(b->grid), dim3(kBatchBlock), 0, s,
                           static_cast<const table_ref<K, M>*>(b->tables), b->ntables, offsets, ops, op,
                           b->chunk_start, b->counter, keys, vals, flags, vals_out);
        return launch_status("batch launch");
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

    static int range(htab* h, int64_t* idx_out, void* keys_out, void* vals_out, int64_t cap, int64_t* n_out,
                     hipStream_t s)
    {
        const T t = T::from_view(h->view);
        const uint64_t total = t.nslots() + 2;
        const uint64_t ntiles = (total + kRangeTile - 1) / kRangeTile;
        const size_t need = (ntiles + 1) * sizeof(int64_t);
        if (h->scratch_bytes < need)
        {
            if (h->scratch) (void)hipFree(h->scratch);
            h->scratch = nullptr;
            h->scratch_bytes = 0;
            HTAB_HIP_TRY(hipMalloc(&h->scratch, need), "range scratch");
            h->scratch_bytes = need;
        }
        hipLaunchKernelGGL((k_range_count<K, M>), dim3(static_cast<uint32_t>(ntiles)), dim3(kRangeThreads), 0, s, t,
                           total, h->scratch);
        int rc = launch_status("range count");
        if (rc) return rc;
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, h->scratch, static_cast<int64_t>(ntiles));
        rc = launch_status("range scan");
        if (rc) return rc;
        hipLaunchKernelGGL((k_range_scatter<K, M>), dim3(static_cast<uint32_t>(ntiles)), dim3(kRangeThreads), 0, s, t,
                           total, h->scratch, idx_out, static_cast<K*>(keys_out), vals_out, cap);
        rc = launch_status("range scatter");
        if (rc) return rc;
        int64_t n = 0;
        HTAB_HIP_TRY(hipMemcpyAsync(&n, h->scratch + ntiles, sizeof(int64_t), hipMemcpyDeviceToHost, s), "range count copy");
        HTAB_HIP_TRY(hipStreamSynchronize(s), "range sync");
        *n_out = n;
        return HTAB_OK;
    }

    static int valid(htab* h, int32_t* out, hipStream_t s)
    {
        const T t = T::from_view(h->view);
        const size_t need = 2 * sizeof(unsigned long long);
        if (h->scratch_bytes < need)
        {
            if (h->scratch) (void)hipFree(h->scratch);
            h->scratch = nullptr;
            h->scratch_bytes = 0;
            HTAB_HIP_TRY(hipMalloc(&h->scratch, need), "valid scratch");
            h->scratch_bytes = need;
        }
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned long long*>(h->scratch), 2);
        int rc = launch_status("valid zero");
        if (rc) return rc;
        const uint64_t total = t.nslots() + 2;
        uint64_t grid = (total + 255) / 256;
        if (grid > 16384) grid = 16384;
        hipLaunchKernelGGL((k_valid<K, M>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0, s, t, total,
                           reinterpret_cast<unsigned long long*>(h->scratch));
        rc = launch_status("valid launch");
        if (rc) return rc;
        unsigned long long res[2] = {0, 0};
        int64_t stripes[kMaxStripes * kStripeStride];
        HTAB_HIP_TRY(hipMemcpyAsync(res, h->scratch, sizeof(res), hipMemcpyDeviceToHost, s), "valid copy");
        HTAB_HIP_TRY(hipMemcpyAsync(stripes, h->view.stripes, sizeof(stripes), hipMemcpyDeviceToHost, s), "valid copy");
        HTAB_HIP_TRY(hipStreamSynchronize(s), "valid sync");
        int64_t size = 0;
        for (uint32_t i = 0; i < kMaxStripes; ++i) size += stripes[i * kStripeStride];
        *out = (res[0] == 0 && static_cast<int64_t>(res[1]) == size && size >= 0 &&
                size <= static_cast<int64_t>(h->view.limit))
                       ? 1
                       : 0;
        return HTAB_OK;
    }
};

template <template <typename, typename, int> class Op, typename F>
int
with_ops(int kind, int lanes, F&& f)
{
#define HTAB_G(K, M)                                                                                                   \
    switch (lanes)                                                                                                     \
    {                                                                                                                  \
        case 4: return f(Op<K, M, 4>{});                                                                               \
        case 8: return f(Op<K, M, 8>{});                                                                               \
        case 16: return f(Op<K, M, 16>{});                                                                             \
        default: return fail(HTAB_E_ARG, "bad bucket_lanes");                                                          \
    }
    switch (kind)
    {
        case HTAB_SET_U32: HTAB_G(uint32_t, void)
        case HTAB_SET_U64: HTAB_G(uint64_t, void)
        case HTAB_MAP_U64_U32: HTAB_G(uint64_t, uint32_t)
        case HTAB_MAP_U32_U32: HTAB_G(uint32_t, uint32_t)
        case HTAB_MAP_U64_U64: HTAB_G(uint64_t, uint64_t)
        default: return fail(HTAB_E_ARG, "bad kind");
    }
#undef HTAB_G
}

template <typename F>
int
with_table_ops(int kind, F&& f)
{
    switch (kind)
    {
        case HTAB_SET_U32: return f(table_ops<uint32_t, void>{});
        case HTAB_SET_U64: return f(table_ops<uint64_t, void>{});
        case HTAB_MAP_U64_U32: return f(table_ops<uint64_t, uint32_t>{});
        case HTAB_MAP_U32_U32: return f(table_ops<uint32_t, uint32_t>{});
        case HTAB_MAP_U64_U64: return f(table_ops<uint64_t, uint64_t>{});
        default: return fail(HTAB_E_ARG, "bad kind");
    }
}

} // namespace

// ====================================================================== C ABI
extern "C" {

const char*
htab_last_error(void)
{
    return g_last_error.c_str();
}

int
htab_create(int kind, int64_t capacity, const htab_tuning* tuning, htab** out)
{
    if (!out) return fail(HTAB_E_ARG, "htab_create: out is null");
    *out = nullptr;
    const int vb = value_bytes_of(kind);
    if (vb == 0) return fail(HTAB_E_ARG, "htab_create: unknown kind");
    if (capacity <= 0) return fail(HTAB_E_ARG, "htab_create: capacity must be positive");
    if (capacity > (int64_t(1) << 34)) return fail(HTAB_E_LIMIT, "htab_create: capacity above 2^34");
    int lanes = 8, block = 256;
    float spc = 1.0F;
    if (tuning)
    {
        if (tuning->bucket_lanes) lanes = tuning->bucket_lanes;
        if (tuning->block_size) block = tuning->block_size;
        if (tuning->slots_per_capacity > 0.0F) spc = tuning->slots_per_capacity;
    }
    if (lanes != 4 && lanes != 8 && lanes != 16) return fail(HTAB_E_ARG, "htab_create: bucket_lanes must be 4, 8 or 16");
    if (block != 256 && block != 512 && block != 1024) return fail(HTAB_E_ARG, "htab_create: block_size must be 256, 512 or 1024");
    if (!(spc >= 0.25F && spc <= 16.0F)) return fail(HTAB_E_ARG, "htab_create: slots_per_capacity out of [0.25, 16]");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(HTAB_E_NODEV, "htab_create: no GPU");

    htab* h = new (std::nothrow) htab();
    if (!h) return fail(HTAB_E_OOM, "htab_create: host allocation");
    h->kind = kind;
    h->lanes = lanes;
    h->block = block;
    h->capacity = capacity;

    const int64_t W = 16LL * lanes / vb;
    const double want = std::ceil(static_cast<double>(capacity) * static_cast<double>(spc));
    int64_t nbuckets = static_cast<int64_t>(std::ceil(want / static_cast<double>(W)));
    if (nbuckets < 1) nbuckets = 1;
    if (nbuckets >= (int64_t(1) << 32)) { delete h; return fail(HTAB_E_LIMIT, "htab_create: too many buckets"); }
    const int64_t nslots = nbuckets * W;
    int64_t ref_bc = 0, ref_max = 0;
    reference_sizing(capacity, &ref_bc, &ref_max);
    const int64_t limit = ref_max < nslots ? ref_max : nslots;

    const size_t area = static_cast<size_t>((nslots + 2) * vb);
    h->bytes = kCtlBytes + ((area + 255) / 256) * 256;
    hipError_t e = hipMalloc(&h->mem, h->bytes);
    if (e != hipSuccess)
    {
        delete h;
        return fail(HTAB_E_OOM, "htab_create: device allocation failed");
    }
    char* base = static_cast<char*>(h->mem);
    h->view.slots = base + kCtlBytes;
    h->view.stripes = reinterpret_cast<int64_t*>(base);
    h->view.flags = reinterpret_cast<uint32_t*>(base + kFlagsOffset);
    h->view.nbuckets = static_cast<uint64_t>(nbuckets);
    h->view.limit = static_cast<uint64_t>(limit);
    h->view.nstripes = (limit < nslots) ? 1U : kMaxStripes;
    h->view.bucket_lanes = static_cast<uint32_t>(lanes);
    h->view.ref_bucket_count = ref_bc;
    h->view.ref_max_size = ref_max;

    h->info.kind = kind;
    h->info.bucket_lanes = lanes;
    h->info.capacity = capacity;
    h->info.bucket_count = ref_bc;
    h->info.max_size = ref_max;
    h->info.physical_buckets = nbuckets;
    h->info.physical_slots = nslots;
    h->info.limit = limit;
    h->info.value_bytes = vb;
    h->info.device_bytes = static_cast<int64_t>(h->bytes);

    e = hipMemset(h->mem, 0, kCtlBytes);
    if (e != hipSuccess)
    {
        (void)hipFree(h->mem);
        delete h;
        return fail_hip(e, "htab_create: control block init");
    }
    int rc = with_table_ops(kind, [&](auto o) { return decltype(o)::clear(h, nullptr); });
    if (rc == HTAB_OK)
    {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = fail_hip(e, "htab_create: init sync");
    }
    if (rc != HTAB_OK)
    {
        (void)hipFree(h->mem);
        delete h;
        return rc;
    }
    *out = h;
    return HTAB_OK;
}

int
htab_destroy(htab* h)
{
    if (!h) return fail(HTAB_E_ARG, "htab_destroy: null handle");
    hipError_t e1 = hipSuccess, e2 = hipSuccess;
    if (h->mem) e1 = hipFree(h->mem);
    if (h->scratch) e2 = hipFree(h->scratch);
    if (h->range_idx && e2 == hipSuccess) e2 = hipFree(h->range_idx);
    delete h;
    if (e1 != hipSuccess) return fail_hip(e1, "htab_destroy");
    if (e2 != hipSuccess) return fail_hip(e2, "htab_destroy scratch");
    return HTAB_OK;
}

static int
insert_common(htab* h, const void* values, const void* keys, const void* vals, int64_t n, uint8_t* ins, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "insert: null handle");
    if (n < 0) return fail(HTAB_E_ARG, "insert: negative count");
    if (n == 0) return HTAB_OK;
    if (!values && !keys) return fail(HTAB_E_ARG, "insert: null input");
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "insert: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_ops<ops>(h->kind, h->lanes, [&](auto o) { return decltype(o)::insert(h, values, keys, vals, n, ins, s); });
}

int
htab_insert(htab* h, const void* values, int64_t n, uint8_t* inserted, void* stream)
{
    if (n > 0 && !values) return fail(HTAB_E_ARG, "htab_insert: null values");
    return insert_common(h, values, nullptr, nullptr, n, inserted, stream);
}

int
htab_insert_kv(htab* h, const void* keys, const void* vals, int64_t n, uint8_t* inserted, void* stream)
{
    if (n > 0 && !keys) return fail(HTAB_E_ARG, "htab_insert_kv: null keys");
    return insert_common(h, nullptr, keys, vals, n, inserted, stream);
}

int
htab_find(htab* h, const void* keys, int64_t n, uint8_t* hits, void* vals_out, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "htab_find: null handle");
    if (n < 0) return fail(HTAB_E_ARG, "htab_find: negative count");
    if (n == 0) return HTAB_OK;
    if (!keys || !hits) return fail(HTAB_E_ARG, "htab_find: null keys or hits");
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "htab_find: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_ops<ops>(h->kind, h->lanes, [&](auto o) { return decltype(o)::find(h, keys, n, hits, vals_out, s); });
}

int
htab_erase(htab* h, const void* keys, int64_t n, uint8_t* erased, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "htab_erase: null handle");
    if (n < 0) return fail(HTAB_E_ARG, "htab_erase: negative count");
    if (n == 0) return HTAB_OK;
    if (!keys) return fail(HTAB_E_ARG, "htab_erase: null keys");
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "htab_erase: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_ops<ops>(h->kind, h->lanes, [&](auto o) { return decltype(o)::erase(h, keys, n, erased, s); });
}

int
htab_size(htab* h, int64_t* out, void* stream)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_size: null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t stripes[kMaxStripes * kStripeStride];
    HTAB_HIP_TRY(hipMemcpyAsync(stripes, h->view.stripes, sizeof(stripes), hipMemcpyDeviceToHost, s), "htab_size copy");
    HTAB_HIP_TRY(hipStreamSynchronize(s), "htab_size sync");
    int64_t n = 0;
    for (uint32_t i = 0; i < kMaxStripes; ++i) n += stripes[i * kStripeStride];
    *out = n;
    return HTAB_OK;
}

int
htab_clear(htab* h, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "htab_clear: null handle");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_table_ops(h->kind, [&](auto o) { return decltype(o)::clear(h, s); });
}

int
htab_range(htab* h, int64_t* idx_out, int64_t cap, int64_t* n_out, void* stream)
{
    if (!h || !n_out) return fail(HTAB_E_ARG, "htab_range: null argument");
    if (cap < 0) return fail(HTAB_E_ARG, "htab_range: negative capacity");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_table_ops(h->kind, [&](auto o) { return decltype(o)::range(h, idx_out, nullptr, nullptr, cap, n_out, s); });
}

int
htab_range_owned(htab* h, const int64_t** idx_out, int64_t* n_out, void* stream)
{
    if (!h || !idx_out || !n_out) return fail(HTAB_E_ARG, "htab_range_owned: null argument");
    const int64_t cap = h->info.limit + 2;
    if (!h->range_idx)
    {
        HTAB_HIP_TRY(hipMalloc(&h->range_idx, static_cast<size_t>(cap) * sizeof(int64_t)), "htab_range_owned alloc");
    }
    *idx_out = h->range_idx;
    return htab_range(h, h->range_idx, cap, n_out, stream);
}

int
htab_contents(htab* h, void* keys_out, void* vals_out, int64_t cap, int64_t* n_out, void* stream)
{
    if (!h || !n_out) return fail(HTAB_E_ARG, "htab_contents: null argument");
    if (cap < 0) return fail(HTAB_E_ARG, "htab_contents: negative capacity");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_table_ops(h->kind,
                          [&](auto o) { return decltype(o)::range(h, nullptr, keys_out, vals_out, cap, n_out, s); });
}

int
htab_valid(htab* h, int32_t* out, void* stream)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_valid: null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_table_ops(h->kind, [&](auto o) { return decltype(o)::valid(h, out, s); });
}

int
htab_bucket(const htab* h, uint64_t key, int64_t* out)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_bucket: null argument");
    // hash<integer> is the identity, widened to size_t (signed keys are not part of the C ABI)
    *out = reference_bucket(key, h->view.ref_bucket_count);
    return HTAB_OK;
}

int
htab_batch_create(htab* const* tables, int32_t ntables, htab_batch** out)
{
    if (!out || !tables || ntables <= 0) return fail(HTAB_E_ARG, "htab_batch_create: bad argument");
    *out = nullptr;
    for (int32_t t = 0; t < ntables; ++t)
    {
        if (!tables[t]) return fail(HTAB_E_ARG, "htab_batch_create: null table");
        if (tables[t]->kind != tables[0]->kind || tables[t]->lanes != tables[0]->lanes)
            return fail(HTAB_E_ARG, "htab_batch_create: tables differ in kind or bucket_lanes");
    }
    htab_batch* b = new (std::nothrow) htab_batch();
    if (!b) return fail(HTAB_E_OOM, "htab_batch_create: host allocation");
    b->kind = tables[0]->kind;
    b->lanes = tables[0]->lanes;
    b->ntables = ntables;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    b->grid = static_cast<uint32_t>(cus) * 8U;
    const size_t desc = 64 * static_cast<size_t>(ntables); // >= sizeof(table_ref<K, M>)
    static_assert(sizeof(table_ref<uint64_t, uint64_t>) <= 64, "descriptor size");
    if (hipMalloc(&b->tables, desc) != hipSuccess ||
        hipMalloc(&b->chunk_start, sizeof(int64_t) * (static_cast<size_t>(ntables) + 1)) != hipSuccess ||
        hipMalloc(&b->counter, 256) != hipSuccess)
    {
        if (b->tables) (void)hipFree(b->tables);
        if (b->chunk_start) (void)hipFree(b->chunk_start);
        delete b;
        return fail(HTAB_E_OOM, "htab_batch_create: device allocation");
    }
    const int rc = with_ops<batch_ops>(b->kind, b->lanes, [&](auto o) { return decltype(o)::build(b, tables); });
    if (rc != HTAB_OK)
    {
        htab_batch_destroy(b);
        return rc;
    }
    *out = b;
    return HTAB_OK;
}

int
htab_batch_destroy(htab_batch* b)
{
    if (!b) return fail(HTAB_E_ARG, "htab_batch_destroy: null handle");
    if (b->tables) (void)hipFree(b->tables);
    if (b->chunk_start) (void)hipFree(b->chunk_start);
    if (b->counter) (void)hipFree(b->counter);
    delete b;
    return HTAB_OK;
}

int
htab_batch_run(htab_batch* b, const int8_t* ops, int op, const int64_t* offsets, const void* keys, const void* vals,
               uint8_t* flags, void* vals_out, void* stream)
{
    if (!b || !offsets || !keys) return fail(HTAB_E_ARG, "htab_batch_run: null argument");
    if (!ops && (op < HTAB_OP_INSERT || op > HTAB_OP_ERASE)) return fail(HTAB_E_ARG, "htab_batch_run: bad op");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return with_ops<batch_ops>(b->kind, b->lanes, [&](auto o) {
        return decltype(o)::run(b, ops, op, offsets, keys, vals, flags, vals_out, s);
    });
}

int
htab_batch_capture(htab_batch* b, const int8_t* ops, int op, const int64_t* offsets, const void* keys,
                   const void* vals, uint8_t* flags, void* vals_out, void* stream, htab_graph** out)
{
    if (!b || !out) return fail(HTAB_E_ARG, "htab_batch_capture: null argument");
    if (!stream) return fail(HTAB_E_ARG, "htab_batch_capture: capture needs a created stream");
    *out = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HTAB_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "htab_batch_capture begin");
    const int rc = htab_batch_run(b, ops, op, offsets, keys, vals, flags, vals_out, stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != HTAB_OK)
    {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail_hip(e, "htab_batch_capture end");
    htab_graph* g = new (std::nothrow) htab_graph();
    if (!g)
    {
        (void)hipGraphDestroy(graph);
        return fail(HTAB_E_OOM, "htab_batch_capture: host allocation");
    }
    g->graph = graph;
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess)
    {
        (void)hipGraphDestroy(graph);
        delete g;
        return fail_hip(ei, "htab_batch_capture instantiate");
    }
    *out = g;
    return HTAB_OK;
}

int
htab_graph_launch(htab_graph* g, void* stream)
{
    if (!g) return fail(HTAB_E_ARG, "htab_graph_launch: null handle");
    HTAB_HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)), "htab_graph_launch");
    return HTAB_OK;
}

int
htab_graph_destroy(htab_graph* g)
{
    if (!g) return fail(HTAB_E_ARG, "htab_graph_destroy: null handle");
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return HTAB_OK;
}

int
htab_get_info(const htab* h, htab_info* out)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_get_info: null argument");
    *out = h->info;
    return HTAB_OK;
}

int
htab_get_view(const htab* h, htab_view* out)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_get_view: null argument");
    *out = h->view;
    return HTAB_OK;
}

} // extern "C"
