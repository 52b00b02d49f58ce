// Internal declarations shared by the C ABI (htab.hip) and the per-kind translation units
// (kind_*.hip): the opaque handle structs, error helpers and the call record a kind dispatcher
// executes.  Each key/value kind is compiled in its own translation unit so the library builds
// in parallel.
#ifndef STDGPU_AMD_HTAB_INTERNAL_HPP
#define STDGPU_AMD_HTAB_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include <stdgpu_amd/htab.h>

#include "htab_kernels.hpp"
#include "htab_region.hpp"

namespace htab_host
{

int fail(int code, const char* what);
const char* last_error();
int fail_hip(hipError_t e, const char* what);
int launch_status(const char* what);
int value_bytes_of(int kind);
int mapped_bytes_of(int kind);
bool ovf_in_bucket_of(int kind);
// a segment's keys, values and value outputs are aligned to their element sizes for the table kind
bool segment_aligned(int kind, const htab_segment& g);
int key_bytes_of(int kind);
void reference_sizing(int64_t capacity, int64_t* bucket_count, int64_t* max_size);

constexpr size_t kCtlBytes = 8192; // stripes (32 x 128 B) + flags + scratch words
constexpr size_t kFlagsOffset = 4096;
constexpr size_t kPlanOffset = 4224; // one-stripe tables: whether a bulk insert can reach the limit
constexpr size_t kFreelistOffset = 4352; // three int64 words of a free-list call (htab_freelist.hpp)

#define HTAB_HIP_TRY(expr, what)                                                                                       \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess) return ::htab_host::fail_hip(e_, what);                                                  \
    } while (0)

} // namespace htab_host

struct htab
{
    int kind = 0;
    int lanes = 8;
    int rb = 0;                   // elements a lane group probes together; 0 = per-operation default
    int block = 256;
    int64_t capacity = 0;
    htab_view view{};
    htab_info info{};
    void* mem = nullptr;
    size_t bytes = 0;
    int64_t* scratch = nullptr; // range/valid scratch, grown on demand
    size_t scratch_bytes = 0;
    int64_t* range_idx = nullptr; // table-owned device_range() indices (limit + 2 entries)
    int insert_path = 0;          // htab_tuning::insert_path
    void* part = nullptr;         // partitioned-insert scratch, grown on demand
    size_t part_bytes = 0;
    bool caller_hashes = false;   // htab_tuning::caller_hashes: only native calls and htab_valid_hashed
    uint64_t* hash_buf = nullptr; // htab_hash_buffer's array, allocated on demand
};

// the bulk entry points that place keys by their bits refuse a table placed by caller hashes
inline int
refuse_caller_hashed(const htab* h, const char* who)
{
    if (h && h->caller_hashes)
        return htab_host::fail(HTAB_E_ARG, who);
    return HTAB_OK;
}

// the device word in the table's control block where k_limit_plan records its decision
inline uint32_t*
plan_word(htab* h)
{
    return reinterpret_cast<uint32_t*>(static_cast<char*>(h->mem) + htab_host::kPlanOffset);
}

struct htab_batch
{
    int kind = 0;
    int lanes = 8;
    int block = 256;                   // workgroup size = chunk size of the persistent kernel
    int rb = 0;                        // probe batch of the tables (0 = default)
    int32_t ntables = 0;
    htab** handles = nullptr;          // host copies of the table handles (per-table baseline)
    void* tables = nullptr;            // device array of table_ref<K, M>
    htab_dev::batch_segment* segs = nullptr; // device segments, seg_cap entries
    int32_t seg_cap = 0;
    int64_t* chunk_start = nullptr;    // seg_cap + 1
    uint32_t grid = 0;                 // persistent workgroups
};

struct htab_graph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};


namespace htab_host
{

enum class call_op
{
    insert,
    find,
    erase,
    assign,
    clear,
    range,
    valid,
    batch_build,
    batch_run_offsets,
    batch_run_segments,
    serve,
    serve_sizes,
    take,
    evict
};

// One request to a kind dispatcher; fields unused by an operation stay zero.
struct htab_call
{
    call_op op{};
    htab* h = nullptr;
    htab_batch* b = nullptr;
    const void* values = nullptr;
    const void* keys = nullptr;
    const void* vals = nullptr;
    int64_t n = 0;
    uint8_t* flags = nullptr;
    void* vals_out = nullptr;
    int64_t* idx_out = nullptr;
    int64_t cap = 0;
    int64_t* n_out = nullptr;
    int32_t* valid_out = nullptr;
    htab* const* tables = nullptr;
    const int8_t* ops = nullptr;
    int bop = 0;
    const int64_t* offsets = nullptr;
    int32_t nsegs = 0;
    const htab_dev::serve_args* serve = nullptr;       // served launch: segments and chunk plan
    const htab_dev::serve_tables* serve_tabs = nullptr; // served launch: its tables (size readback)
    int64_t* sizes_out = nullptr;                      // served launch: host-mapped table sizes
    int64_t* scratch_ext = nullptr;                    // range: tile scratch given by the caller (no allocation)
    int64_t* count_dev = nullptr;                      // range: entry count written to device memory (no sync)
    const uint64_t* hashes = nullptr;                  // valid: the caller's hash of each slot's key (or null)
    const htab_ring_view* ring = nullptr;              // take / evict: the deque's device view
    const htab_block_region* region = nullptr;         // evict: the blocks to erase
    hipStream_t s = nullptr;
};

int dispatch_set_u32(htab_call& c);
int dispatch_set_u64(htab_call& c);
int dispatch_map_u64_u32(htab_call& c);
int dispatch_map_u32_u32(htab_call& c);
int dispatch_map_u64_u64(htab_call& c);

inline int
dispatch(int kind, htab_call& c)
{
    switch (kind)
    {
        case HTAB_SET_U32: return dispatch_set_u32(c);
        case HTAB_SET_U64: return dispatch_set_u64(c);
        case HTAB_MAP_U64_U32: return dispatch_map_u64_u32(c);
        case HTAB_MAP_U32_U32: return dispatch_map_u32_u32(c);
        case HTAB_MAP_U64_U64: return dispatch_map_u64_u64(c);
        default: return fail(HTAB_E_ARG, "bad kind");
    }
}

} // namespace htab_host

#endif // STDGPU_AMD_HTAB_INTERNAL_HPP
