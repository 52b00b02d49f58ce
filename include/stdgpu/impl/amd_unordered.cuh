// Shared engine of the source-compatible unordered_map / unordered_set facade.
//
// The object is a trivially copyable handle (a table pointer for host calls plus the device
// view for native calls), passed by value into kernels like the reference's containers
// (src/stdgpu/impl/unordered_base.cuh:56-583).  Host bulk calls go through the C ABI of
// libstdgpu_amd.so and complete before returning; native device calls run the per-thread
// probing of include/stdgpu_amd/htab_device.hpp on the same layout.
//
// Supported keys: integers of up to 8 bytes, and trivially copyable types of up to 8 bytes whose value
// is their bits (unique object representations: structs of integers without padding, whose operator==
// compares every member, such as the reference tests' 6-byte vec3int16), compared with equal_to<Key>
// or the transparent equal_to<>; any hasher callable with a Key.  A key is stored in a 4- or 8-byte
// slot key, zero-extended.  Keys of more than 8 bytes whose copy is a byte copy (a trivial copy
// assignment; the reference tests' 12-byte vec3int32, whose destructor is user-provided) are kept whole in the slot's payload record
// and compared with key_equal; the slot key is then an id unique to the entry, the table is placed by
// the hasher's value, and presence is decided under the home bucket's lock (htab_device.hpp
// insert_if / find_if / erase_if).  Payload records are bytes: no destructor runs on them.  Mapped
// types: any trivially copyable type, empty ones included.
//
// Two storage forms:
//   * aliased: value_type is laid out exactly as the table's slot (4- or 8-byte keys with 4- or 8-byte
//     mapped types of the library's kinds, sets of 4- or 8-byte keys, and maps with an empty mapped
//     type whose pair is 4 or 8 bytes, such as pair<const vec3int16, dummy>): iterators point into
//     the slots;
//   * payload: otherwise (sets of 6-byte keys, uint32 -> uint64 maps, 12-byte pairs, ...), the table
//     keeps the keys in a set kind and a value_type record per slot in a parallel payload array
//     (htab_tuning::payload_bytes); iterators point into the payload.
// Placement: by the key's bits (htab_device.hpp), so the hasher only feeds bucket(), which keeps the
// reference's value for it; a found entry is not at begin() + bucket(key).  With a transparent hasher
// and key_equal, find / contains / count take a KeyLike (unordered_base.cuh:241-301): an integer
// KeyLike of an integer key by conversion; for any other key the table is placed by the hasher's value
// instead (htab_tuning::caller_hashes), since hash(k) == hash(KeyLike) is what ties a KeyLike to its
// key.  Such tables, and payload tables, run their bulk insert / erase / find as per-element native
// calls in kernels of this header, as the reference's bulk calls do (unordered_base_detail.cuh:996-1066).
#ifndef STDGPU_AMD_FACADE_UNORDERED_CUH
#define STDGPU_AMD_FACADE_UNORDERED_CUH

#include <stdgpu/impl/amd_core.h>

namespace stdgpu::detail
{

template <std::size_t N>
using storage_uint_t = std::conditional_t<N == 4, std::uint32_t, std::uint64_t>;

// the slot key of a Key: 4 bytes for keys of up to 4 bytes, else 8
template <typename Key>
using slot_key_t = std::conditional_t<(sizeof(Key) <= 4), std::uint32_t, std::uint64_t>;

// mapped types the engine stores in the slot itself (the library's map kinds): 4 or 8 trivially
// copyable bytes, not uint32 -> uint64
template <typename Key, typename Mapped>
constexpr bool
slot_mapped()
{
    if constexpr (std::is_void_v<Mapped>)
        return false;
    else if constexpr (std::is_empty_v<Mapped>)
        return false;
    else
        return std::is_trivially_copyable_v<Mapped> && (sizeof(Mapped) == 4 || sizeof(Mapped) == 8) &&
               sizeof(Key) == sizeof(slot_key_t<Key>) && !(sizeof(Key) == 4 && sizeof(Mapped) == 8);
}

template <typename Key, typename Mapped, bool kSlotMapped = slot_mapped<Key, Mapped>()>
struct storage_mapped
{
    using type = void; // keys only in the slot (sets, empty or payload-held mapped values)
};
template <typename Key, typename Mapped>
struct storage_mapped<Key, Mapped, true>
{
    using type = storage_uint_t<sizeof(Mapped)>;
};

template <typename Key, typename Mapped>
constexpr int
htab_kind()
{
    if constexpr (!slot_mapped<Key, Mapped>())
        return sizeof(slot_key_t<Key>) == 4 ? HTAB_SET_U32 : HTAB_SET_U64;
    else if constexpr (sizeof(Key) == 8 && sizeof(Mapped) == 4)
        return HTAB_MAP_U64_U32;
    else if constexpr (sizeof(Key) == 4 && sizeof(Mapped) == 4)
        return HTAB_MAP_U32_U32;
    else
        return HTAB_MAP_U64_U64;
}

template <typename Table, typename Value>
__global__ void
k_facade_insert(Table t, const Value* values, std::int64_t n)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) (void)t.insert(values[i]);
}

template <typename Table, typename Key>
__global__ void
k_facade_erase(Table t, const Key* keys, std::int64_t n)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) (void)t.erase(keys[i]);
}

template <typename Table, typename Key, typename Mapped>
__global__ void
k_facade_find(Table t, const Key* keys, std::int64_t n, std::uint8_t* hits, Mapped* values)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto it = t.find(keys[i]);
    const bool hit = it != t.end();
    hits[i] = hit ? 1 : 0;
    if constexpr (!std::is_void_v<Mapped>)
    {
        if (hit && values) values[i] = it->second;
    }
}

// the caller's hash of every slot's key (0 for free slots), for valid() of a table placed by hashes
template <typename Table>
__global__ void
k_facade_slot_hashes(Table t, std::uint64_t* out)
{
    const std::uint64_t n = t.eng().nslots();
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t s = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < n; s += stride)
        out[s] = t.slot_hash(s);
}

template <typename Key, typename Value, typename Mapped, typename Hash, typename KeyEqual, typename Allocator>
class amd_unordered_base
{
public:
    using key_type = Key;
    using value_type = Value;
    using index_type = index_t;
    using difference_type = std::ptrdiff_t;
    using key_equal = KeyEqual;
    using hasher = Hash;
    using allocator_type = Allocator;
    using reference = value_type&;
    using const_reference = const value_type&;
    using pointer = value_type*;
    using const_pointer = const value_type*;
    using iterator = pointer;
    using const_iterator = const_pointer;

    // keys wider than a slot key: whole in the payload record, an entry id in the slot
    static constexpr bool kWide = sizeof(Key) > 8;
    static_assert(((std::is_integral_v<Key> ||
                    (std::is_trivially_copyable_v<Key> && std::has_unique_object_representations_v<Key>)) &&
                   sizeof(Key) <= 8) ||
                          (kWide && std::is_copy_constructible_v<Key> && std::is_trivially_copy_assignable_v<Key>),
                  "stdgpu (MI355X library): keys are integers, trivially copyable types without padding bits of "
                  "at most 8 bytes, or wider types whose copy is a byte copy");
    static constexpr bool mapped_ok()
    {
        if constexpr (std::is_void_v<Mapped>)
            return true;
        else
            return std::is_trivially_copyable_v<Mapped>;
    }
    static_assert(mapped_ok(), "stdgpu (MI355X library): mapped types are trivially copyable");
    static_assert(std::is_trivially_copyable_v<Value> || !std::is_void_v<Mapped> ||
                          (kWide && std::is_trivially_copy_assignable_v<Value>),
                  "stdgpu (MI355X library): value types are trivially copyable");
    static_assert(std::is_same_v<KeyEqual, equal_to<Key>> || std::is_same_v<KeyEqual, equal_to<void>>,
                  "stdgpu (MI355X library): keys compare with stdgpu::equal_to<Key> or equal_to<>");
    static_assert(std::is_invocable_r_v<std::size_t, const Hash&, const Key&>,
                  "stdgpu (MI355X library): the hasher maps a Key to std::size_t");

    template <typename H, typename = void>
    struct is_transparent_t : std::false_type
    {
    };
    template <typename H>
    struct is_transparent_t<H, std::void_t<typename H::is_transparent>> : std::true_type
    {
    };
    static constexpr bool kTransparent = is_transparent_t<Hash>::value && is_transparent_t<KeyEqual>::value;
    // the KeyLike overloads exist for a transparent hasher and key_equal (unordered_base.cuh:241-301):
    // an integer KeyLike of an integer key, or any KeyLike the hasher and key_equal take for a table
    // placed by the hasher's value
    template <typename KeyLike>
    static constexpr bool keylike_ok = kTransparent && std::is_integral_v<Key> && std::is_integral_v<KeyLike> &&
                                       !std::is_same_v<KeyLike, Key>;
    // a table placed by the hasher's value: transparent lookups of a non-integer key
    static constexpr bool kHashed = (kTransparent && !std::is_integral_v<Key>) || kWide;
    template <typename KeyLike>
    static constexpr bool keylike_hashed_ok =
            kHashed && !std::is_same_v<std::remove_cv_t<KeyLike>, Key> &&
            std::is_invocable_r_v<std::size_t, const Hash&, const KeyLike&> &&
            std::is_invocable_r_v<bool, const KeyEqual&, const Key&, const KeyLike&>;

    using skey = slot_key_t<Key>;
    using smapped = typename storage_mapped<Key, Mapped>::type;
    using engine = htab_dev::table_ref<skey, smapped>;
    using slot_type = typename engine::value_type;
    // aliased: value_type is the slot (the key at offset 0, a slot-held mapped value at the slot's);
    // otherwise each slot has a value_type record in the payload array
    static constexpr bool aliased()
    {
        if constexpr (std::is_void_v<Mapped>)
            return sizeof(Key) == sizeof(skey);
        else if constexpr (std::is_void_v<smapped>)
            return std::is_empty_v<Mapped> && sizeof(value_type) == sizeof(skey);
        else
            return sizeof(slot_type) == sizeof(value_type);
    }
    static constexpr bool kAliased = aliased();
    static_assert(kAliased || sizeof(value_type) <= 256, "stdgpu (MI355X library): value_type of at most 256 bytes");
    // bulk calls run as native calls in this header's kernels (tables placed by hashes, payload tables)
    static constexpr bool kNativeBulk = kHashed || !kAliased;

    amd_unordered_base() = default;

    // ---------------------------------------------------------------- lifetime
    static void create(amd_unordered_base& out, const index_t capacity, const Allocator& allocator)
    {
        if (capacity <= 0) throw std::invalid_argument("stdgpu: capacity must be positive");
        htab* h = nullptr;
        htab_tuning tun{};
        tun.caller_hashes = kHashed ? 1 : 0;
        tun.payload_bytes = kAliased ? 0 : static_cast<std::int32_t>(sizeof(value_type));
        check(htab_create(htab_kind<Key, Mapped>(), capacity, &tun, &h), "createDeviceObject");
        out._h = h;
        check(htab_get_view(h, &out._view), "createDeviceObject");
        out._allocator = allocator;
        if constexpr (kWide)
        {
            // the entry ids: a 64-bit counter, first id 1 (the busy claim value 0 is never an id)
            void* ids = nullptr;
            check(htab_rt_alloc(sizeof(std::uint64_t), &ids), "createDeviceObject");
            out._ids = static_cast<std::uint64_t*>(ids);
            check(htab_rt_zero(ids, sizeof(std::uint64_t), nullptr), "createDeviceObject");
            check(htab_rt_sync(nullptr), "createDeviceObject");
        }
    }

    // takes over a table made through the C ABI (for the tunings createDeviceObject has no parameter
    // for); the kind must match Key and Mapped, and destroyDeviceObject frees it
    static void adopt(amd_unordered_base& out, htab* h, const Allocator& allocator)
    {
        htab_info info{};
        check(htab_get_info(h, &info), "adopt");
        if (info.kind != htab_kind<Key, Mapped>()) throw std::invalid_argument("stdgpu: adopt: table of another kind");
        if (kNativeBulk) throw std::invalid_argument("stdgpu: adopt: this key / value type needs its own table layout");
        out._h = h;
        check(htab_get_view(h, &out._view), "adopt");
        out._allocator = allocator;
    }

    static void destroy(amd_unordered_base& obj)
    {
        if (obj._h) check(htab_destroy(obj._h), "destroyDeviceObject");
        if (obj._ids) check(htab_rt_free(obj._ids), "destroyDeviceObject");
        obj._h = nullptr;
        obj._ids = nullptr;
        obj._view = htab_view{};
    }

    // ---------------------------------------------------------------- native (device) API
    __device__ iterator begin() noexcept { return entries(); }
    __device__ const_iterator begin() const noexcept { return entries(); }
    __device__ const_iterator cbegin() const noexcept { return begin(); }
    __device__ iterator end() noexcept { return entry(eng().end_ptr()); }
    __device__ const_iterator end() const noexcept { return entry(eng().end_ptr()); }
    __device__ const_iterator cend() const noexcept { return end(); }

    __device__ iterator find(const key_type& key) { return entry(find_slot(key)); }
    __device__ const_iterator find(const key_type& key) const { return entry(find_slot(key)); }

    // transparent lookups of a table placed by the hasher's value: the KeyLike's hash gives the home
    // bucket, and key_equal compares it with each stored key on the probe sequence
    template <typename KeyLike, std::enable_if_t<keylike_hashed_ok<KeyLike>, int> = 0>
    __device__ iterator find(const KeyLike& k)
    {
        return entry(find_like(k));
    }
    template <typename KeyLike, std::enable_if_t<keylike_hashed_ok<KeyLike>, int> = 0>
    __device__ const_iterator find(const KeyLike& k) const
    {
        return entry(find_like(k));
    }
    template <typename KeyLike, std::enable_if_t<keylike_hashed_ok<KeyLike>, int> = 0>
    __device__ bool contains(const KeyLike& k) const
    {
        return find(k) != end();
    }
    template <typename KeyLike, std::enable_if_t<keylike_hashed_ok<KeyLike>, int> = 0>
    __device__ index_type count(const KeyLike& k) const
    {
        return contains(k) ? 1 : 0;
    }
    __device__ bool contains(const key_type& key) const { return find(key) != end(); }
    __device__ index_type count(const key_type& key) const { return contains(key) ? 1 : 0; }

    // KeyLike lookups.  Integral conversion to the common type of Key and KeyLike is one-to-one on Key,
    // so at most one key compares equal to k, and only Key(k) can: it is looked up when it does.
    template <typename KeyLike, std::enable_if_t<keylike_ok<KeyLike>, int> = 0>
    __device__ iterator find(const KeyLike& k)
    {
        const key_type c = static_cast<key_type>(k);
        return KeyEqual{}(c, k) ? find(c) : end();
    }
    template <typename KeyLike, std::enable_if_t<keylike_ok<KeyLike>, int> = 0>
    __device__ const_iterator find(const KeyLike& k) const
    {
        const key_type c = static_cast<key_type>(k);
        return KeyEqual{}(c, k) ? find(c) : end();
    }
    template <typename KeyLike, std::enable_if_t<keylike_ok<KeyLike>, int> = 0>
    __device__ bool contains(const KeyLike& k) const
    {
        return find(k) != end();
    }
    template <typename KeyLike, std::enable_if_t<keylike_ok<KeyLike>, int> = 0>
    __device__ index_type count(const KeyLike& k) const
    {
        return contains(k) ? 1 : 0;
    }

    __device__ pair<iterator, bool> insert(const value_type& value)
    {
        bool ins = false;
        const engine e = eng();
        const auto write = [&](slot_type* slot) {
            if constexpr (!kAliased)
                __builtin_memcpy(static_cast<void*>(entry(slot)), &value, sizeof(value_type));
            else if constexpr (!std::is_void_v<smapped>)
            {
                smapped bits;
                __builtin_memcpy(&bits, &value.second, sizeof(bits));
                slot->second = bits;
            }
        };
        slot_type* p = nullptr;
        if constexpr (kWide)
        {
            const skey id = static_cast<skey>(
                    __hip_atomic_fetch_add(_ids, std::uint64_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1);
            p = e.insert_if(id, kBusy, home_of(key_of(value)), holds(key_of(value)), write, &ins);
        }
        else
        {
            const skey k = to_skey(key_of(value));
            p = kHashed ? e.insert_at(k, home_of(key_of(value)), write, &ins) : e.insert(k, write, &ins);
        }
        return pair<iterator, bool>(entry(p), ins);
    }

    template <class... Args>
    __device__ pair<iterator, bool> emplace(Args&&... args)
    {
        return insert(value_type(std::forward<Args>(args)...));
    }

    __device__ index_type erase(const key_type& key)
    {
        const engine e = eng();
        bool done = false;
        if constexpr (kWide)
            done = e.erase_if(home_of(key), holds(key));
        else
            done = kHashed ? e.erase_at(to_skey(key), home_of(key)) : e.erase(to_skey(key));
        return done ? 1 : 0;
    }

    // bucket(): the reference's Fibonacci bucket over bucket_count() (unordered_base_detail.cuh:547-564)
    __host__ __device__ index_type bucket(const key_type& key) const
    {
        return static_cast<index_type>(htab_dev::reference_bucket(hasher{}(key), _view.ref_bucket_count));
    }

    // bucket_size(n) (unordered_base_detail.cuh:566-593): the number of stored keys k with
    // bucket(k) == n, which is what the reference counts along bucket n's chain.  This layout places
    // keys by their bits, not by bucket(), so the call reads every slot: O(max_size()) work per call.
    // All lanes of a wavefront read the same slots in the same order, so a kernel that asks for many
    // buckets at once (as the reference's bucket_size_sum case does) shares each load.
    __device__ index_type bucket_size(index_type n) const
    {
        const engine e = eng();
        const std::uint64_t ns = e.nslots();
        const skey* keys = reinterpret_cast<const skey*>(e.slots);
        constexpr std::uint64_t step = sizeof(slot_type) / sizeof(skey);
        index_type c = 0;
        for (std::uint64_t s = 0; s < ns; ++s)
        {
            const skey k = keys[s * step];
            if (k == htab_dev::empty_key<skey>() || k == htab_dev::tomb_key<skey>()) continue;
            if constexpr (kWide)
                c += (k != kBusy && bucket(key_of(entries()[s])) == n) ? 1 : 0;
            else
                c += bucket(from_skey(k)) == n ? 1 : 0;
        }
        if constexpr (!kWide)
        {
            // the two reserved key values live in side slots with a presence bit each (entry ids of
            // wide keys never take them)
            const std::uint32_t f = htab_dev::load_agent(e.flags);
            if ((f & 1U) && bucket(from_skey(htab_dev::empty_key<skey>())) == n) ++c;
            if ((f & 2U) && bucket(from_skey(htab_dev::tomb_key<skey>())) == n) ++c;
        }
        return c;
    }

    // ---------------------------------------------------------------- host (agnostic) API
    template <typename ValueIterator>
    void insert(ValueIterator first, ValueIterator last)
    {
        insert(execution::device, first, last);
    }

    template <typename ExecutionPolicy, typename ValueIterator,
              std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void insert(ExecutionPolicy&& policy, ValueIterator first, ValueIterator last)
    {
        const std::int64_t n = static_cast<std::int64_t>(last - first);
        if constexpr (kNativeBulk)
        {
            if (n > 0)
                hipLaunchKernelGGL((k_facade_insert<amd_unordered_base, value_type>), dim3(grid_of(n)), dim3(256), 0,
                                   policy.stream, *this, static_cast<const value_type*>(raw(first)), n);
            check(htab_rt_launch_status(), "insert launch");
        }
        else
            check(htab_insert(_h, raw(first), n, nullptr, policy.stream), "insert");
        check(htab_rt_sync(policy.stream), "insert sync");
    }

    template <typename KeyIterator>
    void erase(KeyIterator first, KeyIterator last)
    {
        erase(execution::device, first, last);
    }

    template <typename ExecutionPolicy, typename KeyIterator,
              std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void erase(ExecutionPolicy&& policy, KeyIterator first, KeyIterator last)
    {
        const std::int64_t n = static_cast<std::int64_t>(last - first);
        if constexpr (kNativeBulk)
        {
            if (n > 0)
                hipLaunchKernelGGL((k_facade_erase<amd_unordered_base, key_type>), dim3(grid_of(n)), dim3(256), 0,
                                   policy.stream, *this, static_cast<const key_type*>(raw(first)), n);
            check(htab_rt_launch_status(), "erase launch");
        }
        else
            check(htab_erase(_h, raw(first), n, nullptr, policy.stream), "erase");
        check(htab_rt_sync(policy.stream), "erase sync");
    }

    // Bulk lookups over a key range (the reference has none: its callers run find per thread in a
    // functor, unordered_datastructure.inc:2180-2226).  hits[i] = contains(first[i]); for maps,
    // values[i] = find(first[i])->second where hits[i] (untouched elsewhere; values may be null).
    // hits and values are device arrays of last - first entries.
    template <typename KeyIterator>
    void contains(KeyIterator first, KeyIterator last, std::uint8_t* hits) const
    {
        contains(execution::device, first, last, hits);
    }

    template <typename ExecutionPolicy, typename KeyIterator,
              std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void contains(ExecutionPolicy&& policy, KeyIterator first, KeyIterator last, std::uint8_t* hits) const
    {
        bulk_find(policy.stream, first, last, hits, nullptr, "contains");
    }

    template <typename KeyIterator, typename U = Mapped, std::enable_if_t<!std::is_void_v<U>, int> = 0>
    void find(KeyIterator first, KeyIterator last, std::uint8_t* hits, U* values) const
    {
        find(execution::device, first, last, hits, values);
    }

    template <typename ExecutionPolicy, typename KeyIterator, typename U = Mapped,
              std::enable_if_t<is_execution_policy_v<ExecutionPolicy> && !std::is_void_v<U>, int> = 0>
    void find(ExecutionPolicy&& policy, KeyIterator first, KeyIterator last, std::uint8_t* hits, U* values) const
    {
        bulk_find(policy.stream, first, last, hits, values, "find");
    }

    device_indexed_range<value_type> device_range() { return device_range(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    device_indexed_range<value_type> device_range(ExecutionPolicy&& policy)
    {
        const std::int64_t* idx = nullptr;
        std::int64_t n = 0;
        if (_h) check(htab_range_owned(_h, &idx, &n, policy.stream), "device_range");
        return device_indexed_range<value_type>(idx, static_cast<index_t>(n), entries_host());
    }
    device_indexed_range<const value_type> device_range() const { return device_range(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    device_indexed_range<const value_type> device_range(ExecutionPolicy&& policy) const
    {
        const std::int64_t* idx = nullptr;
        std::int64_t n = 0;
        if (_h) check(htab_range_owned(_h, &idx, &n, policy.stream), "device_range");
        return device_indexed_range<const value_type>(idx, static_cast<index_t>(n), entries_host());
    }

    void clear() { clear(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void clear(ExecutionPolicy&& policy)
    {
        if (!_h) return;
        check(htab_clear(_h, policy.stream), "clear");
        check(htab_rt_sync(policy.stream), "clear sync");
    }

    bool valid() const { return valid(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool valid(ExecutionPolicy&& policy) const
    {
        if (!_h) return true; // an empty (never created) container is valid
        std::int32_t ok = 0;
        if constexpr (kHashed)
        {
            // the hasher's value of every stored key, then the engine's checks against those
            std::uint64_t* hashes = nullptr;
            check(htab_hash_buffer(_h, &hashes), "valid");
            hipLaunchKernelGGL((k_facade_slot_hashes<amd_unordered_base>), dim3(1024), dim3(256), 0, policy.stream, *this,
                               hashes);
            check(htab_rt_launch_status(), "valid launch");
            check(htab_valid_hashed(_h, hashes, &ok, policy.stream), "valid");
        }
        else
            check(htab_valid(_h, &ok, policy.stream), "valid");
        return ok != 0;
    }

    // size() on the host is a device-to-host read; on the device an atomic sum of the counters
    __host__ __device__ index_t size() const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return static_cast<index_t>(eng().size());
#else
        return size(execution::device);
#endif
    }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    index_t size(ExecutionPolicy&& policy) const
    {
        if (!_h) return 0;
        std::int64_t n = 0;
        check(htab_size(_h, &n, policy.stream), "size");
        return static_cast<index_t>(n);
    }
    __host__ __device__ bool empty() const { return size() == 0; }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool empty(ExecutionPolicy&& policy) const
    {
        return size(policy) == 0;
    }
    __host__ __device__ bool full() const { return size() == max_size(); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool full(ExecutionPolicy&& policy) const
    {
        return size(policy) == max_size();
    }
    __host__ __device__ index_t max_size() const noexcept { return static_cast<index_t>(_view.ref_max_size); }
    __host__ __device__ index_t bucket_count() const { return static_cast<index_t>(_view.ref_bucket_count); }
    __host__ __device__ float load_factor() const
    {
        return bucket_count() ? static_cast<float>(size()) / static_cast<float>(bucket_count()) : 0.0F;
    }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    float load_factor(ExecutionPolicy&& policy) const
    {
        return bucket_count() ? static_cast<float>(size(policy)) / static_cast<float>(bucket_count()) : 0.0F;
    }
    __host__ __device__ float max_load_factor() const { return 1.0F; }
    __host__ __device__ hasher hash_function() const { return hasher{}; }
    __host__ __device__ key_equal key_eq() const { return key_equal{}; }
    __host__ __device__ allocator_type get_allocator() const noexcept { return _allocator; }

    // access for callers that hand the table to the C ABI directly
    htab* native_handle() const { return _h; }
    __host__ __device__ const htab_view& native_view() const { return _view; }

    __host__ __device__ engine eng() const { return engine::from_view(_view); }

    // the caller's hash of the key in slot s (0 for a free slot): what a table placed by hashes was
    // placed by (valid())
    __device__ std::uint64_t slot_hash(std::uint64_t s) const
    {
        const skey k = *eng().key_ptr(s);
        if (k == htab_dev::empty_key<skey>() || k == htab_dev::tomb_key<skey>()) return 0;
        if constexpr (kWide)
            return static_cast<std::uint64_t>(hasher{}(key_of(entries()[s])));
        else
            return static_cast<std::uint64_t>(hasher{}(from_skey(k)));
    }

protected:
    static unsigned grid_of(std::int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

    // home bucket of a key in a table placed by the hasher's value
    __device__ std::uint64_t home_of(const key_type& key) const
    {
        return eng().home_of_hash(static_cast<std::uint64_t>(hasher{}(key)));
    }

    __device__ slot_type* find_slot(const key_type& key) const
    {
        const engine e = eng();
        if constexpr (kWide)
            return e.find_if(home_of(key), holds(key));
        else
            return kHashed ? e.find_at(to_skey(key), home_of(key)) : e.find(to_skey(key));
    }

    template <typename KeyLike>
    __device__ slot_type* find_like(const KeyLike& k) const
    {
        const engine e = eng();
        const std::uint64_t h = e.home_of_hash(static_cast<std::uint64_t>(hasher{}(k)));
        if constexpr (kWide)
            return e.find_if(h, holds(k));
        else
            return e.find_match(h, [&k](skey bits) { return key_equal{}(from_skey(bits), k); });
    }

    // wide keys: the slot key an entry holds while its record is written (never an entry id), and
    // whether occupied slot s (slot key k) holds a key comparing equal to `key` (a Key or a KeyLike).
    // The record's key is read with agent-scope loads: a plain load may be served by this CU's L1
    // from before another call wrote the record, while the entry id that call stored after writing it
    // (a release store) is only ever read from L2, and so is the record then.
    static constexpr skey kBusy = 0;
    template <typename KeyLike>
    __device__ auto holds(const KeyLike& key) const
    {
        const value_type* rec = entries();
        return [rec, &key](std::uint64_t s, skey k) {
            if (k == kBusy) return false;
            alignas(Key) unsigned char buf[sizeof(Key)];
            load_coherent(buf, &key_of(rec[s]));
            return key_equal{}(*reinterpret_cast<const Key*>(buf), key);
        };
    }

    // the bytes of *p read with agent-scope loads, four at a time where the key's alignment allows
    __device__ static void load_coherent(unsigned char* out, const Key* p)
    {
        if constexpr (sizeof(Key) % 4 == 0 && alignof(Key) % 4 == 0)
        {
            const std::uint32_t* w = reinterpret_cast<const std::uint32_t*>(p);
#pragma unroll
            for (std::size_t i = 0; i < sizeof(Key) / 4; ++i)
            {
                const std::uint32_t v = __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_memcpy(out + 4 * i, &v, 4);
            }
        }
        else
        {
            const unsigned char* b = reinterpret_cast<const unsigned char*>(p);
            for (std::size_t i = 0; i < sizeof(Key); ++i)
                out[i] = __hip_atomic_load(b + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // the entry of a slot: the slot itself (aliased) or its payload record
    __host__ __device__ value_type* entries() const
    {
        if constexpr (kAliased)
            return reinterpret_cast<value_type*>(_view.slots);
        else
            return static_cast<value_type*>(_view.payload);
    }
    __host__ value_type* entries_host() const { return entries(); }
    __host__ __device__ value_type* entry(const slot_type* p) const
    {
        return entries() + (p - static_cast<const slot_type*>(_view.slots));
    }

    template <typename KeyIterator>
    void bulk_find(hipStream_t s, KeyIterator first, KeyIterator last, std::uint8_t* hits, void* values,
                   const char* what) const
    {
        const std::int64_t n = static_cast<std::int64_t>(last - first);
        if (n > 0 && !hits) throw std::invalid_argument(std::string("stdgpu::") + what + ": hits is null");
        if constexpr (kNativeBulk)
        {
            using M = std::conditional_t<std::is_void_v<Mapped>, void, std::remove_cv_t<Mapped>>;
            if (n > 0)
                hipLaunchKernelGGL((k_facade_find<amd_unordered_base, key_type, M>), dim3(grid_of(n)), dim3(256), 0, s,
                                   *this, static_cast<const key_type*>(raw(first)), n, hits, static_cast<M*>(values));
            check(htab_rt_launch_status(), what);
        }
        else
            check(htab_find(_h, raw(first), n, hits, values, s), what);
        check(htab_rt_sync(s), what);
    }

    __host__ __device__ static const Key& key_of(const value_type& v)
    {
        if constexpr (std::is_void_v<Mapped>)
            return v;
        else
            return v.first;
    }

    // the stored bits of a key: integers by value (sign bits included), other keys as they lie in
    // memory, zero-extended to the slot key
    __host__ __device__ static skey to_skey(const Key& k)
    {
        if constexpr (std::is_integral_v<Key>)
            return static_cast<skey>(static_cast<std::make_unsigned_t<Key>>(k));
        else if constexpr (sizeof(Key) == sizeof(skey))
            return __builtin_bit_cast(skey, k);
        else
        {
            skey bits = 0;
            __builtin_memcpy(&bits, &k, sizeof(Key));
            return bits;
        }
    }

    // a key from its stored bits (the inverse of to_skey)
    __host__ __device__ static Key from_skey(const skey k)
    {
        if constexpr (std::is_integral_v<Key>)
            return static_cast<Key>(static_cast<std::make_unsigned_t<Key>>(k));
        else if constexpr (sizeof(Key) == sizeof(skey))
            return __builtin_bit_cast(Key, k);
        else
        {
            alignas(Key) unsigned char raw_key[sizeof(Key)];
            __builtin_memcpy(raw_key, &k, sizeof(Key));
            return *reinterpret_cast<const Key*>(raw_key);
        }
    }

    htab* _h = nullptr;
    htab_view _view{};
    Allocator _allocator{};
    std::uint64_t* _ids = nullptr; // wide keys: the entry-id counter
};

} // namespace stdgpu::detail

#endif // STDGPU_AMD_FACADE_UNORDERED_CUH
