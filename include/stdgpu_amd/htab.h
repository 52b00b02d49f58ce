/*
 * C ABI of the MI355X hashed-container library (libstdgpu_amd.so).
 *
 * Plain C: pointers, sizes and status codes only.  HIP stays inside the shared
 * library; a stream is passed as an opaque `void*` (a hipStream_t, or NULL for the
 * legacy default stream).  Every entry point returns 0 on success or a negative
 * HTAB_E* code; htab_last_error() describes the last failure of the calling thread.
 * Nothing prints, aborts or throws across this boundary.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * stdgpu source tree).  Bulk calls are asynchronous on the given stream; a caller that
 * needs the reference's blocking semantics synchronises the stream (the C++ facade in
 * include/stdgpu/unordered_{map,set}.cuh does).
 */
#ifndef STDGPU_AMD_HTAB_H
#define STDGPU_AMD_HTAB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HTAB_OK 0
#define HTAB_E_ARG (-1)     /* invalid argument (null handle, bad kind, size out of range, a key / value
                               array not aligned to its element size) */
#define HTAB_E_OOM (-2)     /* device allocation failed */
#define HTAB_E_HIP (-3)     /* a HIP runtime call failed */
#define HTAB_E_NODEV (-4)   /* no usable GPU */
#define HTAB_E_LIMIT (-5)   /* request exceeds a documented maximum */

/* container kinds: key type -> mapped type (value_type layout follows stdgpu::pair) */
#define HTAB_SET_U32 0     /* unordered_set<uint32_t>             value_type  4 B */
#define HTAB_SET_U64 1     /* unordered_set<uint64_t>             value_type  8 B */
#define HTAB_MAP_U64_U32 2 /* unordered_map<uint64_t, uint32_t>   value_type 16 B (key, value, 4 B pad) */
#define HTAB_MAP_U32_U32 3 /* unordered_map<uint32_t, uint32_t>   value_type  8 B */
#define HTAB_MAP_U64_U64 4 /* unordered_map<uint64_t, uint64_t>   value_type 16 B */

typedef struct htab htab; /* opaque table handle; owns all device memory of one table */

/* Layout/launch knobs (all optional; zero fields take the defaults).  No reference
 * counterpart: the reference sizes itself from the capacity alone. */
typedef struct htab_tuning
{
    int32_t bucket_lanes;      /* lanes that probe one bucket together: 1, 4, 8 or 16 (bucket = 16*lanes bytes); default 8 */
    int32_t block_size;        /* threads per workgroup of the bulk kernels: 64, 128, 256, 512 or 1024;
                                  default 64 (the batched form's persistent kernel takes 64, 512 and 1024
                                  for unordered_map<u64, u32>, and 256 otherwise) */
    float slots_per_capacity;  /* physical slots = ceil(capacity * this); default 1.0 (load 0.8 at 0.8*capacity keys) */
    int32_t probe_batch;       /* elements a lane group probes together: 2, 4 or 8 (at most bucket_lanes);
                                * default: 8 for find, 4 for insert and erase (1 with one-lane buckets) */
    int32_t insert_path;       /* bulk inserts without per-element flags on tables whose limit is their
                                * slot count: 1 = one compare-and-swap per key; 2 = keys partitioned
                                * by table region and each region inserted in LDS (tables of at most
                                * 65536 regions of 32 KB; the launch allocates its scratch the first
                                * time, about 40 bytes per key for unordered_map<u64, u32>); 0 (default) = 2 for a launch of half
                                * to all of the slots of a table larger than 256 MB, unless a 1024-key
                                * sample of the launch shows copies of keys (then 1, decided on the
                                * device) or its scratch cannot be allocated (then 1), else 1 */
    int32_t caller_hashes;     /* nonzero: keys are placed by a 64-bit hash their caller computes (the C++
                                * facade's hasher, for transparent lookups of struct keys), so only the
                                * native per-thread calls (which take that hash) and htab_valid_hashed
                                * work on the table; the bulk insert / find / erase / assign calls, htab_valid
                                * and the batched and served forms refuse it */
    int32_t payload_bytes;     /* 0, or 1..256: bytes of a caller-owned record per slot, kept in a second
                                * array of (physical_slots + 2) records parallel to the slots (the facade
                                * stores a value_type there when it cannot alias the slot layout) */
} htab_tuning;

/* Sizes a caller can query.  bucket_count / max_size follow the reference formulas so
 * the logical API is unchanged; physical_slots is this library's own layout. */
typedef struct htab_info
{
    int32_t kind;
    int32_t bucket_lanes;
    int64_t capacity;          /* as passed to htab_create */
    int64_t bucket_count;      /* reference: bit_ceil(capacity)                   unordered_base_detail.cuh:1263-1266 */
    int64_t max_size;          /* reference: bucket_count + excess_count          unordered_base_detail.cuh:1271-1273 */
    int64_t physical_buckets;  /* buckets of 16*bucket_lanes bytes */
    int64_t physical_slots;    /* physical_buckets * slots per bucket */
    int64_t limit;             /* most keys the table accepts: min(max_size, physical_slots) */
    int64_t value_bytes;       /* sizeof(value_type) */
    int64_t device_bytes;      /* device memory owned by the table */
} htab_info;

/* Device view handed to user kernels by the C++ facade (native find/insert/erase). */
typedef struct htab_view
{
    void* slots;               /* (physical_slots + 2) value_type entries, AoS */
    int64_t* stripes;          /* size counter stripes, 16 int64 apart */
    uint32_t* flags;           /* presence bits of the two reserved key values */
    uint64_t nbuckets;         /* physical buckets */
    uint64_t limit;            /* logical size limit */
    uint32_t nstripes;         /* 1 when limit < physical_slots (exact limit), else 32 */
    uint32_t bucket_lanes;
    int64_t ref_bucket_count;  /* reference bucket_count(), for bucket() */
    int64_t ref_max_size;      /* reference max_size() */
    uint32_t* ovf;             /* per bucket: stored keys whose probe sequence runs through it; NULL
                                  when the count lives in the padding of each bucket's first slot */
    void* payload;             /* (physical_slots + 2) records of payload_bytes parallel to the slots, or NULL */
    int64_t payload_bytes;     /* htab_tuning::payload_bytes */
} htab_view;

/* unordered_map::createDeviceObject / unordered_set::createDeviceObject
 * (unordered_map.cuh:106-120, unordered_set.cuh:95-109; engine unordered_base_detail.cuh:1254-1308) */
int htab_create(int kind, int64_t capacity, const htab_tuning* tuning, htab** out);

/* destroyDeviceObject (unordered_map.cuh:126-138; unordered_base_detail.cuh:1310-1348) */
int htab_destroy(htab* table);

/* Bulk insert of value_type entries (AoS, sizeof(value_type) each), never overwriting an
 * existing key: insert(begin, end) (unordered_map.cuh:359-375; unordered_base_detail.cuh:996-1017).
 * inserted[i] (optional) = 1 iff element i was inserted, the bool of device insert(). */
int htab_insert(htab* table, const void* values, int64_t n, uint8_t* inserted, void* stream);

/* Same as htab_insert with keys and mapped values in two arrays (vals is NULL for sets). */
int htab_insert_kv(htab* table, const void* keys, const void* vals, int64_t n, uint8_t* inserted, void* stream);

/* Bulk find: hits[i] = contains(keys[i]); vals_out[i] = find(keys[i])->second on a hit
 * (untouched on a miss).  The reference has no bulk find; this is the device
 * find/contains (unordered_map.cuh:288-335; unordered_base_detail.cuh:653-706) over a range. */
int htab_find(htab* table, const void* keys, int64_t n, uint8_t* hits, void* vals_out, void* stream);

/* Bulk erase: erase(begin, end) (unordered_map.cuh:390-406; unordered_base_detail.cuh:1045-1066).
 * erased[i] (optional) = 1 iff keys[i] was removed by this call (device erase() == 1). */
int htab_erase(htab* table, const void* keys, int64_t n, uint8_t* erased, void* stream);

/* Overwrites the mapped value of every present key: what user code does in the reference with
 * find(key)->second = value (unordered_map.cuh:288-317) over a range.  Maps only; each key at most
 * once per call.  found[i] (optional) = 1 iff keys[i] was present. */
int htab_assign(htab* table, const void* keys, const void* vals, int64_t n, uint8_t* found, void* stream);

/* valid() of a table made with caller_hashes (unordered_base_detail.cuh:1192-1218): hashes[s] is the
 * caller's hash of the key stored in slot s, for every slot s < physical_slots that holds a key (other
 * entries are not read), typically written by the caller into htab_hash_buffer's array.  Also accepted
 * for a table without caller hashes, whose hash of a key is then its bits. */
int htab_valid_hashed(htab* table, const uint64_t* hashes, int32_t* out, void* stream);

/* A device array of physical_slots + 2 uint64 owned by the table (allocated on the first call, freed
 * with the table), for the per-slot hashes htab_valid_hashed reads.  No reference counterpart. */
int htab_hash_buffer(htab* table, uint64_t** out);

/* size() (unordered_base_detail.cuh:1110-1132): synchronises the stream, then reads the count. */
int htab_size(htab* table, int64_t* out, void* stream);

/* clear() (unordered_base_detail.cuh:1220-1252), as device work on the stream. */
int htab_clear(htab* table, void* stream);

/* device_range() (unordered_base_detail.cuh:160-208): writes the slot indices of all stored
 * entries, in ascending slot order, to idx_out (capacity `cap` entries, device memory) and
 * their number to *n_out.  Synchronises the stream.  Entry i is view.slots[idx_out[i]]. */
int htab_range(htab* table, int64_t* idx_out, int64_t cap, int64_t* n_out, void* stream);

/* device_range() with the index buffer owned by the table, as the reference keeps _range_indices
 * inside the table (unordered_base.cuh:526-554): *idx_out receives a device pointer to the
 * ascending slot indices, valid until the next htab_range_owned call on this table or its
 * destruction.  Synchronises the stream. */
int htab_range_owned(htab* table, const int64_t** idx_out, int64_t* n_out, void* stream);

/* Largest table htab_range_noalloc serves: HTAB_NOALLOC_RANGE_TILES * 4096 slots (134,213,632). */
#define HTAB_NOALLOC_RANGE_TILES 32767

/* device_range() that allocates nothing and does not synchronise.  No reference counterpart: for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Writes the ascending slot indices of all stored entries to idx_out (cap
 * entries, device memory) and their number to *n_out_dev (device memory) once the stream reaches
 * the call; the caller synchronises when it needs them.  Scratch: a static array of
 * HTAB_NOALLOC_RANGE_TILES + 1 int64 (256 KiB) per GPU, held until the process ends; a table with
 * more than HTAB_NOALLOC_RANGE_TILES * 4096 slots is refused with HTAB_E_LIMIT before any device
 * work.  Calls that overlap on one GPU share that array and race. */
int htab_range_noalloc(htab* table, int64_t* idx_out, int64_t cap, int64_t* n_out_dev, void* stream);

/* Contents as two dense arrays (keys, mapped values; vals_out NULL for sets) in slot order,
 * the data the reference's device_range() exposes.  Synchronises the stream. */
int htab_contents(htab* table, void* keys_out, void* vals_out, int64_t cap, int64_t* n_out, void* stream);

/* valid() (unordered_base_detail.cuh:1192-1218) for this layout: every stored key is found at
 * its own slot, no key is stored twice, and the size counter equals the number of stored
 * entries.  *out = 1 if valid.  Synchronises the stream. */
int htab_valid(htab* table, int32_t* out, void* stream);

/* bucket(key) (unordered_base_detail.cuh:547-564): the reference's Fibonacci bucket over
 * bucket_count(), so callers that assert on bucket numbers keep working. */
int htab_bucket(const htab* table, uint64_t key, int64_t* out);

int htab_get_info(const htab* table, htab_info* out);
int htab_get_view(const htab* table, htab_view* out);

/* ------------------------------------------------------------------------------------------
 * Batched form: many tables of one kind and bucket_lanes served by ONE persistent launch.  No
 * reference counterpart (the reference serves one table per bulk call); every table gets the same
 * results as a plain call on it alone.  Table t owns elements [offsets[t], offsets[t+1]) of the
 * flattened device arrays keys / vals / flags / vals_out and runs ops[t] (device int8 array, one
 * of HTAB_OP_*) or, when ops is NULL, `op` for every table.  flags receives inserted / hit /
 * erased per element; vals_out the mapped values of find hits.  offsets is a device int64 array
 * of ntables + 1 entries.  Elements of one table in one launch must share an operation.
 * ------------------------------------------------------------------------------------------ */
#define HTAB_OP_INSERT 0
#define HTAB_OP_FIND 1
#define HTAB_OP_ERASE 2

typedef struct htab_batch htab_batch;
typedef struct htab_graph htab_graph;

/* One unit of work of a batched launch: `op` on `n` elements of table `table` (index into the
 * batch's tables), with its own device arrays. */
typedef struct htab_segment
{
    const void* keys;
    const void* vals;     /* mapped values to insert (maps), else NULL */
    uint8_t* flags;       /* inserted / hit / erased per element, or NULL */
    void* vals_out;       /* mapped values of find hits, or NULL */
    int64_t n;
    int32_t table;
    int32_t op;
} htab_segment;

int htab_batch_create(htab* const* tables, int32_t ntables, htab_batch** out);
int htab_batch_destroy(htab_batch* batch);
/* One persistent launch over nsegs segments given in host memory (copied to the device on `stream`). */
int htab_batch_run_segments(htab_batch* batch, const htab_segment* segs, int32_t nsegs, void* stream);
/* Plain per-table calls for the same work (one htab_insert_kv / htab_find / htab_erase per table),
 * the baseline the batched launch is compared with.  offsets: host int64[ntables + 1]. */
int htab_batch_run_per_table(htab_batch* batch, int op, const int64_t* host_offsets, const void* keys,
                             const void* vals, uint8_t* flags, void* vals_out, void* stream);
int htab_batch_run(htab_batch* batch, const int8_t* ops, int op, const int64_t* offsets, const void* keys,
                   const void* vals, uint8_t* flags, void* vals_out, void* stream);
/* Records htab_batch_run with these arguments into a HIP graph (captured on `stream`, which must not
 * be the legacy default stream); htab_graph_launch replays it on any stream. */
int htab_batch_capture(htab_batch* batch, const int8_t* ops, int op, const int64_t* offsets, const void* keys,
                       const void* vals, uint8_t* flags, void* vals_out, void* stream, htab_graph** out);
int htab_graph_launch(htab_graph* graph, void* stream);
int htab_graph_destroy(htab_graph* graph);

/* ------------------------------------------------------------------------------------------
 * Served form: a request loop over resident tables with continuous batching (no reference
 * counterpart).  Requests are htab_segment records (table index, op, n <= 2^32-1, device buffers).
 * While requests arrive, a scheduler keeps up to two launches per stream in flight; each launch
 * takes the longest same-operation run at the head of every table's queue.  A launch waits on the
 * device for the previous launch of each of its tables, so per table the requests take effect in
 * arrival order while launches run back to back.  An insert that may reach the table's limit waits
 * until the table has no launch in flight and then runs with no other request of that table: it
 * inserts what a plain call would (keys past the limit are not inserted); no request is refused.
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_server htab_server;

typedef struct htab_server_stats
{
    int64_t launches;
    int64_t segments;
    int64_t max_concurrent_launches;
    int64_t wall_ns;
} htab_server_stats;

/* nstreams streams (1..16) with two launches in flight each; max_inflight_elems bounds the elements
 * in flight over all of them (a launch takes at most max_inflight_elems / (2 * nstreams) elements,
 * and at least one request).  ntables <= 32767. */
int htab_server_create(htab* const* tables, int32_t ntables, int32_t nstreams, int64_t max_inflight_elems,
                       htab_server** out);
int htab_server_destroy(htab_server* server);
/* Open-loop replay of a request log: request i is submitted at arrival_ns[i] after the start.
 * Returns per request its submit and completion times (ns since the start) and status (0 once
 * the request has taken effect).  Blocks until every request has completed. */
int htab_server_run(htab_server* server, const htab_segment* requests, const int64_t* arrival_ns, int64_t nreq,
                    int64_t* submit_ns, int64_t* done_ns, int32_t* status, htab_server_stats* stats);
/* The same requests as plain calls, one per request in log order on `stream` (skipping requests whose
 * status is non-zero when status is given); blocking = 1 synchronises after every call, as the
 * reference's bulk calls do. */
int htab_server_replay_plain(htab_server* server, const htab_segment* requests, const int32_t* status, int64_t nreq,
                             int blocking, void* stream);

/* ------------------------------------------------------------------------------------------
 * Snapshot and restore of a table (no reference counterpart: restart / resume of resident tables).
 * Byte layout, little-endian: htab_snapshot_header, then the slot area
 * ((physical_buckets * slots per bucket + 2) * value_bytes bytes: buckets in order, the two side slots
 * of the reserved key values last), then the overflow-count array (physical_buckets uint32, present
 * when ovf_bytes > 0), then the payload records (present when payload_bytes > 0).  A restored table
 * has the snapshot's geometry, options, contents and size.
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_snapshot_header
{
    char magic[8];            /* "HTABSNP1" */
    int32_t version;          /* 2 (1: before the away counters; restored only for layouts without them) */
    int32_t kind;
    int32_t bucket_lanes;
    int32_t block_size;
    int32_t probe_batch;
    uint32_t flags;           /* presence bits of the two reserved key values */
    int64_t capacity;
    int64_t physical_buckets;
    int64_t size;
    int64_t slot_bytes;
    int64_t ovf_bytes;
    int64_t payload_bytes;    /* bytes of the payload area after the counts (version 2; 0 without one) */
    int32_t record_bytes;     /* htab_tuning::payload_bytes (version 2) */
    int32_t caller_hashes;    /* htab_tuning::caller_hashes (version 2) */
} htab_snapshot_header;

/* bytes a snapshot of the table takes (header included) */
int htab_snapshot_bytes(const htab* table, int64_t* bytes);
/* writes the snapshot into host memory of at least htab_snapshot_bytes bytes; waits for the copies */
int htab_snapshot(htab* table, void* host_buffer, int64_t bytes, void* stream);
/* a new table from a snapshot in host memory (HTAB_E_ARG for a buffer that is not a whole snapshot).
   The header and sizes are checked; the slot bytes are taken as they are (a snapshot altered after
   htab_snapshot wrote it gives a table that htab_valid may reject, at a cost that grows with how far
   its keys lie from their home buckets) */
int htab_restore(const void* host_buffer, int64_t bytes, htab** out);

/* ------------------------------------------------------------------------------------------
 * vector<T> and deque<T> for elements of 1..256 bytes (src/stdgpu/vector.cuh, deque.cuh), on one ring
 * of slots: the elements are the positions [begin, end), position p in slot p mod capacity (a vector
 * keeps begin at 0, so its positions are the indices of its data).  Bulk calls push or pop n elements
 * at one end and reserve their positions at once, so their result is that of a loop of the reference's
 * single calls: push_back puts element i at end + i, pop_back takes end - 1 - i, push_front puts it at
 * begin - 1 - i, pop_front takes begin + i; size follows the reference's clamped size().  User kernels
 * push and pop per thread, at both ends and mixed in one kernel, through the view
 * (include/stdgpu_amd/ring_device.hpp; the C++ stdgpu::vector and stdgpu::deque).
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_ring_view
{
    void* data;          /* capacity elements of elem_bytes */
    uint32_t* state;     /* capacity slot states: the low two bits are free, writing, full, reading */
    uint64_t* begin;     /* front position */
    uint64_t* end;       /* back position; size = end - begin */
    int64_t* ready;      /* elements published and not yet reserved by a pop */
    int64_t* free_slots; /* slots released and not yet reserved by a push */
    int64_t capacity;
    int64_t elem_bytes;
} htab_ring_view;

/* vector<T> (vector_detail.cuh): push_back / pop_back :181-302, clear_iota :361-368, size :454-512 */
typedef struct htab_vector htab_vector;
typedef htab_ring_view htab_vector_view;

/* createDeviceObject(capacity) of a vector of elem_bytes-byte elements (vector_detail.cuh:40-80) */
int htab_vector_create_sized(int64_t capacity, int64_t elem_bytes, htab_vector** out);
/* the same with 4-byte elements (the excess free list's vector<index_t>) */
int htab_vector_create(int64_t capacity, htab_vector** out);
int htab_vector_destroy(htab_vector* vec);
/* push_back of n elements; pushed[i] (optional) = 1 iff element i found room (vector_detail.cuh:181-240) */
int htab_vector_push_back(htab_vector* vec, const void* values, int64_t n, uint8_t* pushed, void* stream);
/* pop_back of up to n elements from the top: out[i] = the element, or all-ones bytes (0xFFFFFFFF for
 * 4-byte elements) when the vector ran empty (out may be null: the elements are dropped, as
 * erase(end - n, end) does); popped[i] (optional) = 1 iff element i got one (vector_detail.cuh:242-302) */
int htab_vector_pop_back(htab_vector* vec, void* out, int64_t n, uint8_t* popped, void* stream);
/* size = 0 (vector_detail.cuh:579-614 clear); kernels only, no memset (safe inside a captured graph) */
int htab_vector_clear(htab_vector* vec, void* stream);
/* device addresses for user kernels (include/stdgpu_amd/ring_device.hpp) */
int htab_vector_get_view(htab_vector* vec, htab_vector_view* out);
/* vector_clear_iota of 4-byte elements: element i = first + i, size = capacity (vector_detail.cuh:361-368) */
int htab_vector_fill_iota(htab_vector* vec, uint32_t first, void* stream);
/* size() clamped to [0, capacity] (vector_detail.cuh:454-512) */
int htab_vector_size(htab_vector* vec, int64_t* out, void* stream);
/* the elements from bottom to top into out (device memory, capacity elements); *n_out = size */
int htab_vector_contents(htab_vector* vec, void* out, int64_t* n_out, void* stream);
/* valid() (vector_detail.cuh:514-577): the positions [0, size) full, every other slot free, the counts
 * consistent; for a vector no kernel is changing */
int htab_vector_valid(htab_vector* vec, int32_t* out, void* stream);

/* ------------------------------------------------------------------------------------------
 * bitset over 32-bit blocks (src/stdgpu/bitset.cuh).  Bulk forms of the per-bit device calls:
 * set / reset / flip of bit positions[i] return its previous value in old_out[i]
 * (bitset_detail.cuh:47-112, 265-329; positions outside [0, size) are skipped, in_range[i] = 0);
 * set() / reset() / flip() of every bit (:247-319); count() of the set bits below size (:131-160).
 * all() = count == size, any() = count > 0, none() = count == 0.
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_bitset htab_bitset;

#define HTAB_BIT_SET 0
#define HTAB_BIT_RESET 1
#define HTAB_BIT_FLIP 2

int htab_bitset_create(int64_t nbits, htab_bitset** out);
int htab_bitset_destroy(htab_bitset* bits);
int htab_bitset_update(htab_bitset* bits, int op, const int64_t* positions, int64_t n, uint8_t* old_out,
                       uint8_t* in_range, void* stream);
int htab_bitset_test(htab_bitset* bits, const int64_t* positions, int64_t n, uint8_t* out, void* stream);
int htab_bitset_all_bits(htab_bitset* bits, int op, void* stream);
int htab_bitset_count(htab_bitset* bits, int64_t* out, void* stream);

typedef struct htab_bitset_view
{
    uint32_t* blocks; /* ceil(nbits / 32) blocks; bit n is bit n % 32 of block n / 32 */
    int64_t nbits;
} htab_bitset_view;

/* device pointer and size for user kernels that set / reset / flip / test single bits (the C++
 * stdgpu::bitset); no reference counterpart: the reference's bitset object is itself the device handle */
int htab_bitset_get_view(const htab_bitset* bits, htab_bitset_view* out);

/* ------------------------------------------------------------------------------------------
 * atomic<T> / atomic_ref<T> (src/stdgpu/atomic.cuh) for T = int, unsigned int,
 * unsigned long long int, float (atomic.cuh:137-139).  A bulk call applies n operations, each as the
 * device call of the same name would (atomic_detail.cuh:428-760; semantics of the OpenMP backend,
 * openmp/impl/atomic_detail.h:40-217): operation i applies operands[i] to element targets[i]
 * (element 0 for every i when targets is NULL) and writes the value it replaced into old_out[i]
 * (may be NULL; nothing is written for STORE).  done_out[i] (may be NULL) is 1 when target i is in
 * [0, count) and, for COMPARE_EXCHANGE, when the value equalled expected[i] and was replaced; out of
 * range targets are skipped.  The n operations take effect in some order, as n concurrent device calls
 * do.  FETCH_INC_MOD / FETCH_DEC_MOD take the modulus arg of fetch_inc_mod(arg) / fetch_dec_mod(arg)
 * (unsigned int only, atomic_detail.cuh:679-702); AND / OR / XOR exist for the integer types only
 * (HTAB_E_ARG otherwise).
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_atomic htab_atomic;

#define HTAB_ATOMIC_I32 0
#define HTAB_ATOMIC_U32 1
#define HTAB_ATOMIC_U64 2
#define HTAB_ATOMIC_F32 3

#define HTAB_ATOMIC_LOAD 0
#define HTAB_ATOMIC_STORE 1
#define HTAB_ATOMIC_EXCHANGE 2
#define HTAB_ATOMIC_COMPARE_EXCHANGE 3
#define HTAB_ATOMIC_FETCH_ADD 4
#define HTAB_ATOMIC_FETCH_SUB 5
#define HTAB_ATOMIC_FETCH_MIN 6
#define HTAB_ATOMIC_FETCH_MAX 7
#define HTAB_ATOMIC_FETCH_AND 8
#define HTAB_ATOMIC_FETCH_OR 9
#define HTAB_ATOMIC_FETCH_XOR 10
#define HTAB_ATOMIC_FETCH_INC_MOD 11
#define HTAB_ATOMIC_FETCH_DEC_MOD 12

/* count device values of the given type, all 0 (createDeviceObject, atomic_detail.cuh:143-151) */
int htab_atomic_create(int type, int64_t count, htab_atomic** out);
int htab_atomic_destroy(htab_atomic* a);
/* the device address of element 0, for device code that takes the values by atomic_ref */
int htab_atomic_data(const htab_atomic* a, void** out);
int htab_atomic_apply(htab_atomic* a, int op, const int64_t* targets, const void* operands, const void* expected,
                      int64_t n, void* old_out, uint8_t* done_out, void* stream);
/* the same over count values of the given type at data (device memory, aligned to the value size;
 * HTAB_E_ARG otherwise): atomic_ref<T> */
int htab_atomic_ref_apply(int type, void* data, int64_t count, int op, const int64_t* targets, const void* operands,
                          const void* expected, int64_t n, void* old_out, uint8_t* done_out, void* stream);
/* host-side load() / store() of element index (atomic_detail.cuh:461-526); value in host or device memory */
int htab_atomic_load(const htab_atomic* a, int64_t index, void* value_out, void* stream);
int htab_atomic_store(htab_atomic* a, int64_t index, const void* value, void* stream);

/* Message for the last non-zero status returned to the calling thread. */
const char* htab_last_error(void);

/* ------------------------------------------------------------------------------------------
 * Runtime helpers for the header-only C++ facade (include/stdgpu), so that its host code reaches
 * HIP only through this library (the facade still launches its own kernels, which are instantiated
 * for the caller's key and value types).  No reference counterpart: the reference's facade calls its
 * backend directly (memory_detail.cuh, the backends' memory.h / device.h).
 * ------------------------------------------------------------------------------------------ */
/* device allocation of bytes (at least one byte); *out = null on failure (HTAB_E_OOM) */
int htab_rt_alloc(int64_t bytes, void** out);
int htab_rt_free(void* p);
/* copy of bytes between host and device memory, then a wait for the stream: kind 0 device to host,
 * 1 host to device, 2 device to device */
int htab_rt_copy(void* dst, const void* src, int64_t bytes, int32_t kind, void* stream);
/* bytes of device memory set to zero by a kernel (no memset node: safe inside a captured graph) */
int htab_rt_zero(void* p, int64_t bytes, void* stream);
/* waits for the stream's work; HTAB_E_HIP with the runtime's message if any of it failed */
int htab_rt_sync(void* stream);
/* the status of the calling thread's last kernel launch (HTAB_OK or HTAB_E_HIP) */
int htab_rt_launch_status(void);

/* ------------------------------------------------------------------------------------------
 * deque<T> (deque.cuh; deque_detail.cuh:202-448 push_back / pop_back / push_front / pop_front,
 * :482-512 size), the free list of voxel-block pipelines among others; see htab_ring_view above.
 * ------------------------------------------------------------------------------------------ */
typedef struct htab_deque htab_deque;
typedef htab_ring_view htab_deque_view;

/* device pointers of the deque for user kernels (no reference counterpart: the reference's deque
 * object is itself the device handle) */
int htab_deque_get_view(const htab_deque* q, htab_deque_view* out);
/* clear() (deque_detail.cuh:575-635): no elements, every slot free */
int htab_deque_clear(htab_deque* q, void* stream);
/* createDeviceObject(capacity) of a deque of elem_bytes-byte elements (deque_detail.cuh:40-80) */
int htab_deque_create_sized(int64_t capacity, int64_t elem_bytes, htab_deque** out);
/* the same with 4-byte elements (a free list of slot numbers) */
int htab_deque_create(int64_t capacity, htab_deque** out);
int htab_deque_destroy(htab_deque* q);
/* push_back / push_front of n elements; pushed[i] (optional) = 1 iff element i found room */
int htab_deque_push_back(htab_deque* q, const void* values, int64_t n, uint8_t* pushed, void* stream);
int htab_deque_push_front(htab_deque* q, const void* values, int64_t n, uint8_t* pushed, void* stream);
/* pop_front / pop_back of up to n elements: out[i] = the element, or all-ones bytes when the deque ran
 * empty (out may be null); popped[i] (optional) = 1 iff element i got one */
int htab_deque_pop_front(htab_deque* q, void* out, int64_t n, uint8_t* popped, void* stream);
int htab_deque_pop_back(htab_deque* q, void* out, int64_t n, uint8_t* popped, void* stream);
/* 4-byte elements: the deque holds 0, 1, ..., capacity - 1 front to back (the voxel free list's start) */
int htab_deque_fill_iota(htab_deque* q, void* stream);
int htab_deque_size(htab_deque* q, int64_t* out, void* stream);
/* the elements front to back into out (device memory, capacity elements); *n_out = size */
int htab_deque_contents(htab_deque* q, void* out, int64_t* n_out, void* stream);
/* valid() (deque_detail.cuh:514-573): the positions [begin, end) full, every other slot free, the counts
 * consistent; for a deque no kernel is changing */
int htab_deque_valid(htab_deque* q, int32_t* out, void* stream);

/* ------------------------------------------------------------------------------------------
 * Free-list calls between a map and a deque of its mapped values: the voxel-hashing caller of the
 * reference's containers (per point: deque::pop_front a slot, insert (block, slot), push the slot back
 * when the block existed; later, erase the blocks out of range and deque::push_back their slots) as
 * two bulk calls with no host synchronisation.  No reference counterpart as library calls: each
 * replaces a caller loop over deque_detail.cuh:390-448 (pop_front), unordered_base_detail.cuh:653-794
 * (insert / erase) and deque_detail.cuh:204-260 (push_back).  Maps only, the deque's element size
 * equal to the mapped size; like the deque's bulk calls, each runs alone on its table and deque.
 * ------------------------------------------------------------------------------------------ */

/* Keys that pack three block coordinates of `bits` bits each (x highest, z lowest), each stored as
 * coordinate + offset; a block's centre is (coordinate + 0.5) * block_size on each axis. */
typedef struct htab_block_region
{
    int32_t bits;      /* 1..21 */
    int32_t reserved;  /* 0 */
    int64_t offset;
    double block_size;
    double center[3];
    double radius;     /* blocks whose centre is farther than this from center are selected */
} htab_block_region;

/* After a bulk insert wrote its per-element flags (htab_insert's `inserted`): every keys[i] with
 * flags[i] == 1 gets the front element of q as its mapped value, popped in rank order; flags[i]
 * becomes 2 when q ran empty first (the key keeps the value its insert wrote) and 3 when keys[i] is
 * not in the table (nothing popped for it).  count (device int64, optional) receives the number of
 * elements popped. */
int htab_assign_from_deque(htab* table, const void* keys, uint8_t* flags, int64_t n, htab_deque* q, int64_t* count,
                           void* stream);
/* Erases every entry whose key's block centre lies farther than region->radius from region->center
 * and pushes its mapped value onto the back of q; once q is full, the remaining such entries stay.
 * count (device int64, optional) receives the number erased (= pushed). */
int htab_erase_far_blocks(htab* table, const htab_block_region* region, htab_deque* q, int64_t* count,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* STDGPU_AMD_HTAB_H */
