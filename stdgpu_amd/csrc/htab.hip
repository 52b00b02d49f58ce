// C ABI of the MI355X hashed containers: table lifetime, bulk operations, size, range,
// validity check and the batched form.  Kernels live in htab_kernels.hpp; their typed dispatch is
// compiled per key/value kind in kind_*.hip.
//
// Build: hipcc --offload-arch=gfx950 -O3 -shared -fPIC (see stdgpu_amd/build.py).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "htab_internal.hpp"

using namespace htab_host;
using namespace htab_dev;

namespace htab_host
{

namespace
{
thread_local std::string g_last_error;
}

int
fail(int code, const char* what)
{
    g_last_error = what;
    return code;
}

const char*
last_error()
{
    return g_last_error.c_str();
}

int
fail_hip(hipError_t e, const char* what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return HTAB_E_HIP;
}


int
launch_status(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_hip(e, what);
    return HTAB_OK;
}

// bytes set to zero (htab_rt_zero): 16-byte stores where the range allows, bytes at its ends
template <int = 0>
__global__ __launch_bounds__(256) void
k_zero_bytes(unsigned char* __restrict__ p, uint64_t n)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t head = (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15;
    const uint64_t h = head < n ? head : n;
    const uint64_t words = (n - h) / 16;
    uint4* w = reinterpret_cast<uint4*>(p + h);
    for (uint64_t i = t; i < words; i += stride) w[i] = make_uint4(0, 0, 0, 0);
    for (uint64_t i = t; i < h; i += stride) p[i] = 0;
    for (uint64_t i = h + words * 16 + t; i < n; i += stride) p[i] = 0;
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

bool
ovf_in_bucket_of(int kind)
{
    switch (kind)
    {
        case HTAB_SET_U32: return htab_dev::slot_traits<uint32_t, void>::ovf_in_bucket;
        case HTAB_SET_U64: return htab_dev::slot_traits<uint64_t, void>::ovf_in_bucket;
        case HTAB_MAP_U64_U32: return htab_dev::slot_traits<uint64_t, uint32_t>::ovf_in_bucket;
        case HTAB_MAP_U32_U32: return htab_dev::slot_traits<uint32_t, uint32_t>::ovf_in_bucket;
        case HTAB_MAP_U64_U64: return htab_dev::slot_traits<uint64_t, uint64_t>::ovf_in_bucket;
        default: return false;
    }
}

int
mapped_bytes_of(int kind)
{
    switch (kind)
    {
        case HTAB_MAP_U64_U32: return 4;
        case HTAB_MAP_U32_U32: return 4;
        case HTAB_MAP_U64_U64: return 8;
        default: return 0;
    }
}

int
key_bytes_of(int kind)
{
    return (kind == HTAB_SET_U32 || kind == HTAB_MAP_U32_U32) ? 4 : 8;
}

// device loads and stores of keys, values and pairs need them naturally aligned (null passes)
static bool
aligned(const void* p, int bytes)
{
    return p == nullptr || bytes <= 1 || reinterpret_cast<uintptr_t>(p) % static_cast<uintptr_t>(bytes) == 0;
}

static int
check_alignment(int kind, const void* values, const void* keys, const void* vals, const char* what)
{
    const int kb = key_bytes_of(kind);
    const int mb = mapped_bytes_of(kind);
    if (!aligned(values, kb > mb ? kb : mb) || !aligned(keys, kb) || !aligned(vals, mb))
        return fail(HTAB_E_ARG, what);
    return HTAB_OK;
}

bool
segment_aligned(int kind, const htab_segment& g)
{
    const int kb = key_bytes_of(kind);
    const int mb = mapped_bytes_of(kind);
    return aligned(g.keys, kb) && aligned(g.vals, mb) && aligned(g.vals_out, mb);
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


} // namespace htab_host


// ====================================================================== C ABI
extern "C" {

const char*
htab_last_error(void)
{
    return htab_host::last_error();
}

// A table of a given geometry: allocation, view, sizes, and every slot EMPTY.  htab_create derives
// nbuckets from the capacity; htab_restore takes it from a snapshot.
static int
build_table(int kind, int64_t capacity, int lanes, int block, int rb, int64_t nbuckets, int payload, htab** out)
{
    const int vb = value_bytes_of(kind);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(HTAB_E_NODEV, "htab_create: no GPU");
    if (nbuckets < 1 || nbuckets >= (int64_t(1) << 32)) return fail(HTAB_E_LIMIT, "htab_create: too many buckets");

    htab* h = new (std::nothrow) htab();
    if (!h) return fail(HTAB_E_OOM, "htab_create: host allocation");
    h->kind = kind;
    h->lanes = lanes;
    h->rb = rb;
    h->block = block;
    h->capacity = capacity;

    const int64_t W = 16LL * lanes / vb;
    const int64_t nslots = nbuckets * W;
    int64_t ref_bc = 0, ref_max = 0;
    reference_sizing(capacity, &ref_bc, &ref_max);
    const int64_t limit = ref_max < nslots ? ref_max : nslots;

    const size_t area = static_cast<size_t>((nslots + 2) * vb);
    const size_t area_padded = ((area + 255) / 256) * 256;
    // the ovf count array (none when the count lives in the first slot's padding)
    const size_t ovf_bytes =
            ovf_in_bucket_of(kind) ? 0 : ((static_cast<size_t>(nbuckets) * sizeof(uint32_t) + 255) / 256) * 256;
    // records of the caller's payload, parallel to the slots (side slots included)
    const size_t payload_area = payload ? ((static_cast<size_t>(nslots + 2) * payload + 255) / 256) * 256 : 0;
    h->bytes = kCtlBytes + area_padded + ovf_bytes + payload_area;
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
    h->view.ovf = ovf_bytes ? reinterpret_cast<uint32_t*>(base + kCtlBytes + area_padded) : nullptr;
    h->view.payload = payload ? base + kCtlBytes + area_padded + ovf_bytes : nullptr;
    h->view.payload_bytes = payload;
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
    htab_call cc;
    cc.op = call_op::clear;
    cc.h = h;
    int rc = dispatch(kind, cc);
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

static int
check_geometry(int kind, int64_t capacity, int lanes, int block, int rb, const char* who)
{
    (void)who;
    if (value_bytes_of(kind) == 0) return fail(HTAB_E_ARG, "htab_create: unknown kind");
    if (capacity <= 0) return fail(HTAB_E_ARG, "htab_create: capacity must be positive");
    if (capacity > (int64_t(1) << 34)) return fail(HTAB_E_LIMIT, "htab_create: capacity above 2^34");
    if (lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16)
        return fail(HTAB_E_ARG, "htab_create: bucket_lanes must be 1, 4, 8 or 16");
    if (rb != 0 && ((rb != 2 && rb != 4 && rb != 8) || rb > lanes))
        return fail(HTAB_E_ARG, "htab_create: probe_batch must be 2, 4 or 8 and at most bucket_lanes");
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
        return fail(HTAB_E_ARG, "htab_create: block_size must be 64, 128, 256, 512 or 1024");
    return HTAB_OK;
}

int
htab_create(int kind, int64_t capacity, const htab_tuning* tuning, htab** out)
{
    if (!out) return fail(HTAB_E_ARG, "htab_create: out is null");
    *out = nullptr;
    int lanes = 8, block = 64, rb = 0, ipath = 0, payload = 0;
    bool hashed = false;
    float spc = 1.0F;
    if (tuning)
    {
        if (tuning->bucket_lanes) lanes = tuning->bucket_lanes;
        if (tuning->block_size) block = tuning->block_size;
        if (tuning->slots_per_capacity > 0.0F) spc = tuning->slots_per_capacity;
        rb = tuning->probe_batch;
        ipath = tuning->insert_path;
        hashed = tuning->caller_hashes != 0;
        payload = tuning->payload_bytes;
    }
    int rc = check_geometry(kind, capacity, lanes, block, rb, "htab_create");
    if (rc != HTAB_OK) return rc;
    if (ipath < 0 || ipath > 2) return fail(HTAB_E_ARG, "htab_create: insert_path must be 0, 1 or 2");
    if (payload < 0 || payload > 256) return fail(HTAB_E_ARG, "htab_create: payload_bytes must be in [0, 256]");
    if (!(spc >= 0.25F && spc <= 16.0F)) return fail(HTAB_E_ARG, "htab_create: slots_per_capacity out of [0.25, 16]");

    const int64_t W = 16LL * lanes / value_bytes_of(kind);
    const double want = std::ceil(static_cast<double>(capacity) * static_cast<double>(spc));
    int64_t nbuckets = static_cast<int64_t>(std::ceil(want / static_cast<double>(W)));
    if (nbuckets < 1) nbuckets = 1;
    // room for two entries at least, as the reference's capacity-1 table (one bucket head plus one
    // excess entry); only one-slot buckets need the second bucket for it
    if (nbuckets * W < 2) nbuckets = (2 + W - 1) / W;
    rc = build_table(kind, capacity, lanes, block, rb, nbuckets, payload, out);
    if (rc == HTAB_OK)
    {
        (*out)->insert_path = ipath;
        (*out)->caller_hashes = hashed;
    }
    return rc;
}

// ---------------------------------------------------------------- snapshot / restore
static int64_t
slot_area_bytes(const htab* h)
{
    return (h->info.physical_slots + 2) * h->info.value_bytes;
}

static int64_t
ovf_array_bytes(const htab* h)
{
    return h->view.ovf ? h->info.physical_buckets * static_cast<int64_t>(sizeof(uint32_t)) : 0;
}

static int64_t
payload_area_bytes(const htab* h)
{
    return (h->info.physical_slots + 2) * h->view.payload_bytes;
}

int
htab_snapshot_bytes(const htab* h, int64_t* bytes)
{
    if (!h || !bytes) return fail(HTAB_E_ARG, "htab_snapshot_bytes: null argument");
    *bytes = static_cast<int64_t>(sizeof(htab_snapshot_header)) + slot_area_bytes(h) + ovf_array_bytes(h) +
             payload_area_bytes(h);
    return HTAB_OK;
}

int
htab_snapshot(htab* h, void* host_buffer, int64_t bytes, void* stream)
{
    if (!h || !host_buffer) return fail(HTAB_E_ARG, "htab_snapshot: null argument");
    int64_t need = 0;
    (void)htab_snapshot_bytes(h, &need);
    if (bytes < need) return fail(HTAB_E_ARG, "htab_snapshot: buffer too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_snapshot_header hd{};
    std::memcpy(hd.magic, "HTABSNP1", 8);
    hd.version = 2;
    hd.kind = h->kind;
    hd.bucket_lanes = h->lanes;
    hd.block_size = h->block;
    hd.probe_batch = h->rb;
    hd.capacity = h->capacity;
    hd.physical_buckets = h->info.physical_buckets;
    hd.slot_bytes = slot_area_bytes(h);
    hd.ovf_bytes = ovf_array_bytes(h);
    hd.payload_bytes = payload_area_bytes(h);
    hd.record_bytes = static_cast<int32_t>(h->view.payload_bytes);
    hd.caller_hashes = h->caller_hashes ? 1 : 0;
    int rc = htab_size(h, &hd.size, stream);
    if (rc != HTAB_OK) return rc;
    char* out = static_cast<char*>(host_buffer);
    HTAB_HIP_TRY(hipMemcpyAsync(&hd.flags, h->view.flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "htab_snapshot flags");
    HTAB_HIP_TRY(hipMemcpyAsync(out + sizeof(hd), h->view.slots, static_cast<size_t>(hd.slot_bytes),
                                hipMemcpyDeviceToHost, s),
                 "htab_snapshot slots");
    if (hd.ovf_bytes)
        HTAB_HIP_TRY(hipMemcpyAsync(out + sizeof(hd) + hd.slot_bytes, h->view.ovf, static_cast<size_t>(hd.ovf_bytes),
                                    hipMemcpyDeviceToHost, s),
                     "htab_snapshot counts");
    if (hd.payload_bytes)
        HTAB_HIP_TRY(hipMemcpyAsync(out + sizeof(hd) + hd.slot_bytes + hd.ovf_bytes, h->view.payload,
                                    static_cast<size_t>(hd.payload_bytes), hipMemcpyDeviceToHost, s),
                     "htab_snapshot payload");
    HTAB_HIP_TRY(hipStreamSynchronize(s), "htab_snapshot sync");
    std::memcpy(out, &hd, sizeof(hd));
    return HTAB_OK;
}

int
htab_restore(const void* host_buffer, int64_t bytes, htab** out)
{
    if (!host_buffer || !out) return fail(HTAB_E_ARG, "htab_restore: null argument");
    *out = nullptr;
    // version 1 headers end before payload_bytes (72 bytes)
    constexpr int64_t kV1Header = static_cast<int64_t>(offsetof(htab_snapshot_header, payload_bytes));
    if (bytes < kV1Header) return fail(HTAB_E_ARG, "htab_restore: truncated");
    htab_snapshot_header hd{};
    std::memcpy(&hd, host_buffer, static_cast<size_t>(kV1Header));
    if (std::memcmp(hd.magic, "HTABSNP1", 8) != 0 || (hd.version != 1 && hd.version != 2))
        return fail(HTAB_E_ARG, "htab_restore: not a table snapshot");
    // version 1 predates the away counters in the slot padding of in-bucket layouts: its tables of
    // such a layout would restore with all counters at zero, which makes stored keys unreachable
    if (hd.version == 1 && ovf_in_bucket_of(hd.kind) && hd.bucket_lanes > 1)
        return fail(HTAB_E_ARG, "htab_restore: version 1 snapshot of a layout with away counters");
    const int64_t header_bytes = hd.version == 1 ? kV1Header : static_cast<int64_t>(sizeof(hd));
    if (hd.version == 2)
    {
        if (bytes < header_bytes) return fail(HTAB_E_ARG, "htab_restore: truncated");
        std::memcpy(&hd, host_buffer, sizeof(hd));
    }
    int rc = check_geometry(hd.kind, hd.capacity, hd.bucket_lanes, hd.block_size, hd.probe_batch, "htab_restore");
    if (rc != HTAB_OK) return rc;
    const int vb = value_bytes_of(hd.kind);
    const int64_t W = 16LL * hd.bucket_lanes / vb;
    if (hd.physical_buckets < 1 || hd.physical_buckets >= (int64_t(1) << 32) ||
        hd.slot_bytes != (hd.physical_buckets * W + 2) * vb ||
        hd.ovf_bytes != (ovf_in_bucket_of(hd.kind) ? 0 : hd.physical_buckets * static_cast<int64_t>(sizeof(uint32_t))) ||
        hd.size < 0 || hd.size > hd.physical_buckets * W + 2 || hd.record_bytes < 0 || hd.record_bytes > 256 ||
        hd.payload_bytes != (hd.physical_buckets * W + 2) * hd.record_bytes ||
        bytes < header_bytes + hd.slot_bytes + hd.ovf_bytes + hd.payload_bytes)
        return fail(HTAB_E_ARG, "htab_restore: inconsistent snapshot");
    htab* h = nullptr;
    rc = build_table(hd.kind, hd.capacity, hd.bucket_lanes, hd.block_size, hd.probe_batch, hd.physical_buckets,
                     hd.record_bytes, &h);
    if (rc != HTAB_OK) return rc;
    h->caller_hashes = hd.caller_hashes != 0;
    const char* in = static_cast<const char*>(host_buffer) + header_bytes;
    hipError_t e = hipMemcpy(h->view.slots, in, static_cast<size_t>(hd.slot_bytes), hipMemcpyHostToDevice);
    if (e == hipSuccess && hd.ovf_bytes)
        e = hipMemcpy(h->view.ovf, in + hd.slot_bytes, static_cast<size_t>(hd.ovf_bytes), hipMemcpyHostToDevice);
    if (e == hipSuccess && hd.payload_bytes)
        e = hipMemcpy(h->view.payload, in + hd.slot_bytes + hd.ovf_bytes, static_cast<size_t>(hd.payload_bytes),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->view.flags, &hd.flags, sizeof(uint32_t), hipMemcpyHostToDevice);
    // the size goes into the first stripe (the others are zero after creation)
    if (e == hipSuccess) e = hipMemcpy(h->view.stripes, &hd.size, sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        (void)htab_destroy(h);
        return fail_hip(e, "htab_restore: copy");
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
    if (h->part && e2 == hipSuccess) e2 = hipFree(h->part);
    if (h->hash_buf && e2 == hipSuccess) e2 = hipFree(h->hash_buf);
    delete h;
    if (e1 != hipSuccess) return fail_hip(e1, "htab_destroy");
    if (e2 != hipSuccess) return fail_hip(e2, "htab_destroy scratch");
    return HTAB_OK;
}

static int
insert_common(htab* h, const void* values, const void* keys, const void* vals, int64_t n, uint8_t* ins, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "insert: null handle");
    if (refuse_caller_hashed(h, "insert: the table places keys by caller hashes (native calls only)")) return HTAB_E_ARG;
    if (n < 0) return fail(HTAB_E_ARG, "insert: negative count");
    if (n == 0) return HTAB_OK;
    if (!values && !keys) return fail(HTAB_E_ARG, "insert: null input");
    if (check_alignment(h->kind, values, keys, vals, "insert: keys, values or pairs not aligned to their size"))
        return HTAB_E_ARG;
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "insert: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::insert;
    c.h = h;
    c.values = values;
    c.keys = keys;
    c.vals = vals;
    c.n = n;
    c.flags = ins;
    c.s = s;
    return dispatch(h->kind, c);
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
    if (refuse_caller_hashed(h, "htab_find: the table places keys by caller hashes (native calls only)")) return HTAB_E_ARG;
    if (n < 0) return fail(HTAB_E_ARG, "htab_find: negative count");
    if (n == 0) return HTAB_OK;
    if (!keys || !hits) return fail(HTAB_E_ARG, "htab_find: null keys or hits");
    if (check_alignment(h->kind, nullptr, keys, vals_out, "htab_find: keys or values not aligned to their size"))
        return HTAB_E_ARG;
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "htab_find: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::find;
    c.h = h;
    c.keys = keys;
    c.n = n;
    c.flags = hits;
    c.vals_out = vals_out;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_erase(htab* h, const void* keys, int64_t n, uint8_t* erased, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "htab_erase: null handle");
    if (refuse_caller_hashed(h, "htab_erase: the table places keys by caller hashes (native calls only)")) return HTAB_E_ARG;
    if (n < 0) return fail(HTAB_E_ARG, "htab_erase: negative count");
    if (n == 0) return HTAB_OK;
    if (!keys) return fail(HTAB_E_ARG, "htab_erase: null keys");
    if (check_alignment(h->kind, nullptr, keys, nullptr, "htab_erase: keys not aligned to their size"))
        return HTAB_E_ARG;
    if ((n + h->block - 1) / h->block > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "htab_erase: too many elements");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::erase;
    c.h = h;
    c.keys = keys;
    c.n = n;
    c.flags = erased;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_assign(htab* h, const void* keys, const void* vals, int64_t n, uint8_t* found, void* stream)
{
    if (!h) return fail(HTAB_E_ARG, "htab_assign: null handle");
    if (refuse_caller_hashed(h, "htab_assign: the table places keys by caller hashes (native calls only)")) return HTAB_E_ARG;
    if (n < 0) return fail(HTAB_E_ARG, "htab_assign: negative count");
    if (n == 0) return HTAB_OK;
    if (!keys || !vals) return fail(HTAB_E_ARG, "htab_assign: null keys or values");
    if (check_alignment(h->kind, nullptr, keys, vals, "htab_assign: keys or values not aligned to their size"))
        return HTAB_E_ARG;
    if ((n + 255) / 256 > 0x7FFFFFFF) return fail(HTAB_E_LIMIT, "htab_assign: too many elements");
    htab_call c;
    c.op = call_op::assign;
    c.h = h;
    c.keys = keys;
    c.vals = vals;
    c.n = n;
    c.flags = found;
    c.s = static_cast<hipStream_t>(stream);
    return dispatch(h->kind, c);
}

// the free-list calls' common checks: a map whose mapped size is the deque's element size
static int
freelist_args(htab* h, htab_deque* q, htab_ring_view* view, const char* who)
{
    if (!h || !q) return fail(HTAB_E_ARG, who);
    if (refuse_caller_hashed(h, who)) return HTAB_E_ARG;
    if (htab_deque_get_view(q, view) != HTAB_OK) return fail(HTAB_E_ARG, who);
    const int mb = mapped_bytes_of(h->kind);
    if (mb <= 0 || view->elem_bytes != mb) return fail(HTAB_E_ARG, who);
    return HTAB_OK;
}

int
htab_assign_from_deque(htab* h, const void* keys, uint8_t* flags, int64_t n, htab_deque* q, int64_t* count,
                       void* stream)
{
    htab_ring_view view{};
    if (freelist_args(h, q, &view, "htab_assign_from_deque: needs a map and a deque of its mapped size"))
        return HTAB_E_ARG;
    if (n < 0) return fail(HTAB_E_ARG, "htab_assign_from_deque: negative count");
    if (n > 0 && (!keys || !flags)) return fail(HTAB_E_ARG, "htab_assign_from_deque: null keys or flags");
    if (check_alignment(h->kind, nullptr, keys, nullptr, "htab_assign_from_deque: keys not aligned to their size"))
        return HTAB_E_ARG;
    htab_call c;
    c.op = call_op::take;
    c.h = h;
    c.keys = keys;
    c.flags = flags;
    c.n = n;
    c.ring = &view;
    c.count_dev = count;
    c.s = static_cast<hipStream_t>(stream);
    return dispatch(h->kind, c);
}

int
htab_erase_far_blocks(htab* h, const htab_block_region* region, htab_deque* q, int64_t* count, void* stream)
{
    htab_ring_view view{};
    if (freelist_args(h, q, &view, "htab_erase_far_blocks: needs a map and a deque of its mapped size"))
        return HTAB_E_ARG;
    if (!region || region->bits < 1 || region->bits > 21 || region->reserved != 0 || !(region->radius >= 0.0))
        return fail(HTAB_E_ARG, "htab_erase_far_blocks: bad region (bits 1..21, reserved 0, radius >= 0)");
    if (key_bytes_of(h->kind) * 8 < 3 * region->bits)
        return fail(HTAB_E_ARG, "htab_erase_far_blocks: the key holds fewer than 3 * bits bits");
    htab_call c;
    c.op = call_op::evict;
    c.h = h;
    c.ring = &view;
    c.region = region;
    c.count_dev = count;
    c.s = static_cast<hipStream_t>(stream);
    return dispatch(h->kind, c);
}

// ------------------------------------------------------------------ runtime helpers (C++ facade)
int
htab_rt_alloc(int64_t bytes, void** out)
{
    if (!out || bytes < 0) return fail(HTAB_E_ARG, "htab_rt_alloc: bad argument");
    *out = nullptr;
    if (hipMalloc(out, static_cast<size_t>(bytes > 0 ? bytes : 1)) != hipSuccess)
    {
        *out = nullptr;
        return fail(HTAB_E_OOM, "htab_rt_alloc: device allocation failed");
    }
    return HTAB_OK;
}

int
htab_rt_free(void* p)
{
    if (p && hipFree(p) != hipSuccess) return fail(HTAB_E_HIP, "htab_rt_free");
    return HTAB_OK;
}

int
htab_rt_copy(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream)
{
    if (bytes < 0 || kind < 0 || kind > 2 || (bytes > 0 && (!dst || !src))) return fail(HTAB_E_ARG, "htab_rt_copy: bad argument");
    if (bytes == 0) return HTAB_OK;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyDeviceToHost : (kind == 1 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice);
    hipStream_t s = static_cast<hipStream_t>(stream);
    HTAB_HIP_TRY(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), k, s), "htab_rt_copy");
    HTAB_HIP_TRY(hipStreamSynchronize(s), "htab_rt_copy: sync");
    return HTAB_OK;
}

int
htab_rt_zero(void* p, int64_t bytes, void* stream)
{
    if (bytes < 0 || (bytes > 0 && !p)) return fail(HTAB_E_ARG, "htab_rt_zero: bad argument");
    if (bytes == 0) return HTAB_OK;
    const uint64_t n = static_cast<uint64_t>(bytes);
    uint64_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_zero_bytes<>, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<unsigned char*>(p), n);
    return launch_status("htab_rt_zero");
}

int
htab_rt_sync(void* stream)
{
    HTAB_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "htab_rt_sync");
    return HTAB_OK;
}

int
htab_rt_launch_status(void)
{
    return launch_status("kernel launch");
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
    htab_call c;
    c.op = call_op::clear;
    c.h = h;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_range(htab* h, int64_t* idx_out, int64_t cap, int64_t* n_out, void* stream)
{
    if (!h || !n_out) return fail(HTAB_E_ARG, "htab_range: null argument");
    if (cap < 0) return fail(HTAB_E_ARG, "htab_range: negative capacity");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::range;
    c.h = h;
    c.idx_out = idx_out;
    c.cap = cap;
    c.n_out = n_out;
    c.s = s;
    return dispatch(h->kind, c);
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
    htab_call c;
    c.op = call_op::range;
    c.h = h;
    c.keys = keys_out;
    c.vals_out = vals_out;
    c.cap = cap;
    c.n_out = n_out;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_valid(htab* h, int32_t* out, void* stream)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_valid: null argument");
    if (refuse_caller_hashed(h, "htab_valid: the table places keys by caller hashes (htab_valid_hashed)"))
        return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::valid;
    c.h = h;
    c.valid_out = out;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_valid_hashed(htab* h, const uint64_t* hashes, int32_t* out, void* stream)
{
    if (!h || !hashes || !out) return fail(HTAB_E_ARG, "htab_valid_hashed: null argument");
    if (!aligned(hashes, 8)) return fail(HTAB_E_ARG, "htab_valid_hashed: hashes not aligned to 8 bytes");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::valid;
    c.h = h;
    c.valid_out = out;
    c.hashes = hashes;
    c.s = s;
    return dispatch(h->kind, c);
}

int
htab_hash_buffer(htab* h, uint64_t** out)
{
    if (!h || !out) return fail(HTAB_E_ARG, "htab_hash_buffer: null argument");
    if (!h->hash_buf)
    {
        const size_t bytes = static_cast<size_t>(h->info.physical_slots + 2) * sizeof(uint64_t);
        HTAB_HIP_TRY(hipMalloc(&h->hash_buf, bytes), "htab_hash_buffer");
    }
    *out = h->hash_buf;
    return HTAB_OK;
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
        if (refuse_caller_hashed(tables[t], "htab_batch_create: a table places keys by caller hashes"))
            return HTAB_E_ARG;
        if (tables[t]->kind != tables[0]->kind || tables[t]->lanes != tables[0]->lanes ||
            tables[t]->block != tables[0]->block || tables[t]->rb != tables[0]->rb)
            return fail(HTAB_E_ARG, "htab_batch_create: tables differ in kind, bucket_lanes, block_size or probe_batch");
    }
    htab_batch* b = new (std::nothrow) htab_batch();
    if (!b) return fail(HTAB_E_OOM, "htab_batch_create: host allocation");
    b->kind = tables[0]->kind;
    b->lanes = tables[0]->lanes;
    // the persistent kernel's workgroup size is also its chunk size, so it is the size the launch
    // uses: the tables' own where that kernel exists (64, 512 and 1024 for unordered_map<u64, u32>,
    // batch_all_blocks in htab_ops.hpp), else 256
    const int tb = tables[0]->block;
    b->block = (b->kind == HTAB_MAP_U64_U32 && (tb == 64 || tb == 512 || tb == 1024)) ? tb : 256;
    b->rb = tables[0]->rb;
    b->ntables = ntables;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    b->grid = static_cast<uint32_t>(cus) * static_cast<uint32_t>(2048 / b->block);
    b->seg_cap = ntables < 4096 ? 4096 : ntables;
    b->handles = new (std::nothrow) htab*[ntables];
    if (!b->handles)
    {
        delete b;
        return fail(HTAB_E_OOM, "htab_batch_create: host allocation");
    }
    for (int32_t t = 0; t < ntables; ++t) b->handles[t] = tables[t];
    const size_t desc = 64 * static_cast<size_t>(ntables); // >= sizeof(table_ref<K, M>)
    static_assert(sizeof(table_ref<uint64_t, uint64_t>) <= 64, "descriptor size");
    static_assert(sizeof(htab_dev::batch_segment) == sizeof(htab_segment), "segment layout");
    if (hipMalloc(&b->tables, desc) != hipSuccess ||
        hipMalloc(&b->segs, sizeof(htab_dev::batch_segment) * static_cast<size_t>(b->seg_cap)) != hipSuccess ||
        hipMalloc(&b->chunk_start, sizeof(int64_t) * (static_cast<size_t>(b->seg_cap) + 1)) != hipSuccess)
    {
        htab_batch_destroy(b);
        return fail(HTAB_E_OOM, "htab_batch_create: device allocation");
    }
    htab_call c;
    c.op = call_op::batch_build;
    c.b = b;
    c.tables = tables;
    const int rc = dispatch(b->kind, c);
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
    if (b->segs) (void)hipFree(b->segs);
    if (b->chunk_start) (void)hipFree(b->chunk_start);
    delete[] b->handles;
    delete b;
    return HTAB_OK;
}

int
htab_batch_run_segments(htab_batch* b, const htab_segment* segs, int32_t nsegs, void* stream)
{
    if (!b || (nsegs > 0 && !segs) || nsegs < 0) return fail(HTAB_E_ARG, "htab_batch_run_segments: bad argument");
    if (nsegs > b->seg_cap) return fail(HTAB_E_LIMIT, "htab_batch_run_segments: more segments than capacity");
    // the kernels take each table's work in a launch to be one operation (an insert relies on no
    // slot being freed meanwhile, an erase may free slots outright), so segments of one table that
    // name different operations are refused before any device work
    std::vector<int8_t> table_op(static_cast<size_t>(b->ntables), static_cast<int8_t>(-1));
    for (int32_t g = 0; g < nsegs; ++g)
    {
        if (segs[g].table < 0 || segs[g].table >= b->ntables || segs[g].n < 0 || segs[g].op < HTAB_OP_INSERT ||
            segs[g].op > HTAB_OP_ERASE || (segs[g].n > 0 && !segs[g].keys) || !segment_aligned(b->kind, segs[g]))
            return fail(HTAB_E_ARG, "htab_batch_run_segments: bad segment");
        int8_t& o = table_op[static_cast<size_t>(segs[g].table)];
        if (o >= 0 && o != static_cast<int8_t>(segs[g].op))
            return fail(HTAB_E_ARG, "htab_batch_run_segments: segments of one table name different operations");
        o = static_cast<int8_t>(segs[g].op);
    }
    if (nsegs == 0) return HTAB_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HTAB_HIP_TRY(hipMemcpyAsync(b->segs, segs, sizeof(htab_segment) * static_cast<size_t>(nsegs), hipMemcpyHostToDevice, s),
                 "htab_batch_run_segments copy");
    htab_call c;
    c.op = call_op::batch_run_segments;
    c.b = b;
    c.nsegs = nsegs;
    c.s = s;
    return dispatch(b->kind, c);
}

int
htab_batch_run_per_table(htab_batch* b, int op, const int64_t* off, const void* keys, const void* vals, uint8_t* flags,
                         void* vals_out, void* stream)
{
    if (!b || !off || !keys) return fail(HTAB_E_ARG, "htab_batch_run_per_table: null argument");
    if (op == HTAB_OP_FIND && !flags) return fail(HTAB_E_ARG, "htab_batch_run_per_table: find needs flags (hits)");
    const int kb = key_bytes_of(b->kind);
    const int vb = mapped_bytes_of(b->kind);
    for (int32_t t = 0; t < b->ntables; ++t)
    {
        const int64_t a = off[t], n = off[t + 1] - off[t];
        if (n <= 0) continue;
        const char* k = static_cast<const char*>(keys) + a * kb;
        int rc = HTAB_OK;
        if (op == HTAB_OP_INSERT)
            rc = htab_insert_kv(b->handles[t], k, vals ? static_cast<const char*>(vals) + a * vb : nullptr, n,
                                flags ? flags + a : nullptr, stream);
        else if (op == HTAB_OP_FIND)
            rc = htab_find(b->handles[t], k, n, flags + a, vals_out ? static_cast<char*>(vals_out) + a * vb : nullptr,
                           stream);
        else if (op == HTAB_OP_ERASE)
            rc = htab_erase(b->handles[t], k, n, flags ? flags + a : nullptr, stream);
        else
            rc = fail(HTAB_E_ARG, "htab_batch_run_per_table: bad op");
        if (rc != HTAB_OK) return rc;
    }
    return HTAB_OK;
}

int
htab_batch_run(htab_batch* b, const int8_t* ops, int op, const int64_t* offsets, const void* keys, const void* vals,
               uint8_t* flags, void* vals_out, void* stream)
{
    if (!b || !offsets || !keys) return fail(HTAB_E_ARG, "htab_batch_run: null argument");
    if (!ops && (op < HTAB_OP_INSERT || op > HTAB_OP_ERASE)) return fail(HTAB_E_ARG, "htab_batch_run: bad op");
    hipStream_t s = static_cast<hipStream_t>(stream);
    htab_call c;
    c.op = call_op::batch_run_offsets;
    c.b = b;
    c.ops = ops;
    c.bop = op;
    c.offsets = offsets;
    c.keys = keys;
    c.vals = vals;
    c.flags = flags;
    c.vals_out = vals_out;
    c.s = s;
    return dispatch(b->kind, c);
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
