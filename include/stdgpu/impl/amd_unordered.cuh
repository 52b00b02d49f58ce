// Shared engine of the source-compatible unordered_map / unordered_set facade.
//
// The object is a trivially copyable handle (a table pointer for host calls plus the device
// view for native calls), passed by value into kernels like the reference's containers
// (src/stdgpu/impl/unordered_base.cuh:56-583).  Host bulk calls go through the C ABI of
// libstdgpu_amd.so and complete before returning; native device calls run the per-thread
// probing of include/stdgpu_amd/htab_device.hpp on the same layout.
//
// Supported keys: 4- or 8-byte integers, and trivially copyable 4- or 8-byte types whose value is their
// bits (unique object representations: structs of integers without padding, whose operator== compares
// every member), compared with equal_to<Key> or the transparent equal_to<>; any hasher callable with a
// Key; (maps) trivially copyable 4- or 8-byte mapped types, with pair<const Key, T> laid out as the
// slot (an 8-byte struct key then needs 8-byte alignment).  The layout places keys by their bits
// (htab_device.hpp), so the hasher only feeds bucket(), which keeps the reference's value for it; a
// found entry is not at begin() + bucket(key).  With a transparent hasher and key_equal, find /
// contains / count of an integer key also take an integer KeyLike (unordered_base.cuh:241-301).
#ifndef STDGPU_AMD_FACADE_UNORDERED_CUH
#define STDGPU_AMD_FACADE_UNORDERED_CUH

#include <stdgpu/impl/amd_core.h>

namespace stdgpu::detail
{

template <std::size_t N>
using storage_uint_t = std::conditional_t<N == 4, std::uint32_t, std::uint64_t>;

template <typename Key, typename Mapped>
constexpr int
htab_kind()
{
    if constexpr (std::is_void_v<Mapped>)
        return sizeof(Key) == 4 ? HTAB_SET_U32 : HTAB_SET_U64;
    else if constexpr (sizeof(Key) == 8 && sizeof(Mapped) == 4)
        return HTAB_MAP_U64_U32;
    else if constexpr (sizeof(Key) == 4 && sizeof(Mapped) == 4)
        return HTAB_MAP_U32_U32;
    else
        return HTAB_MAP_U64_U64;
}

template <typename Mapped>
struct storage_mapped
{
    using type = storage_uint_t<sizeof(Mapped)>;
};
template <>
struct storage_mapped<void>
{
    using type = void;
};

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

    static_assert((std::is_integral_v<Key> ||
                   (std::is_trivially_copyable_v<Key> && std::has_unique_object_representations_v<Key>)) &&
                          (sizeof(Key) == 4 || sizeof(Key) == 8),
                  "stdgpu (MI355X library): keys are 4- or 8-byte integers, or 4- or 8-byte trivially copyable "
                  "types without padding bits");
    static constexpr bool mapped_ok()
    {
        if constexpr (std::is_void_v<Mapped>)
            return true;
        else
            return std::is_trivially_copyable_v<Mapped> && (sizeof(Mapped) == 4 || sizeof(Mapped) == 8) &&
                   !(sizeof(Key) == 4 && sizeof(Mapped) == 8);
    }
    static_assert(mapped_ok(), "stdgpu (MI355X library): mapped types are trivially copyable, 4 or 8 bytes "
                               "(uint32->uint64 maps are not provided)");
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
    // the KeyLike overloads exist for a transparent hasher and key_equal (unordered_base.cuh:241-301)
    template <typename KeyLike>
    static constexpr bool keylike_ok = is_transparent_t<Hash>::value && is_transparent_t<KeyEqual>::value &&
                                       std::is_integral_v<Key> && std::is_integral_v<KeyLike> &&
                                       !std::is_same_v<KeyLike, Key>;

    using skey = storage_uint_t<sizeof(Key)>;
    using smapped = typename storage_mapped<Mapped>::type;
    using engine = htab_dev::table_ref<skey, smapped>;
    using slot_type = typename engine::value_type;
    static_assert(sizeof(slot_type) == sizeof(value_type), "value_type layout differs from the slot layout");

    amd_unordered_base() = default;

    // ---------------------------------------------------------------- lifetime
    static void create(amd_unordered_base& out, const index_t capacity, const Allocator& allocator)
    {
        if (capacity <= 0) throw std::invalid_argument("stdgpu: capacity must be positive");
        htab* h = nullptr;
        check(htab_create(htab_kind<Key, Mapped>(), capacity, nullptr, &h), "createDeviceObject");
        out._h = h;
        check(htab_get_view(h, &out._view), "createDeviceObject");
        out._allocator = allocator;
    }

    // takes over a table made through the C ABI (for the tunings createDeviceObject has no parameter
    // for); the kind must match Key and Mapped, and destroyDeviceObject frees it
    static void adopt(amd_unordered_base& out, htab* h, const Allocator& allocator)
    {
        htab_info info{};
        check(htab_get_info(h, &info), "adopt");
        if (info.kind != htab_kind<Key, Mapped>()) throw std::invalid_argument("stdgpu: adopt: table of another kind");
        out._h = h;
        check(htab_get_view(h, &out._view), "adopt");
        out._allocator = allocator;
    }

    static void destroy(amd_unordered_base& obj)
    {
        if (obj._h) check(htab_destroy(obj._h), "destroyDeviceObject");
        obj._h = nullptr;
        obj._view = htab_view{};
    }

    // ---------------------------------------------------------------- native (device) API
    __device__ iterator begin() noexcept { return reinterpret_cast<iterator>(_view.slots); }
    __device__ const_iterator begin() const noexcept { return reinterpret_cast<const_iterator>(_view.slots); }
    __device__ const_iterator cbegin() const noexcept { return begin(); }
    __device__ iterator end() noexcept { return reinterpret_cast<iterator>(eng().end_ptr()); }
    __device__ const_iterator end() const noexcept { return reinterpret_cast<const_iterator>(eng().end_ptr()); }
    __device__ const_iterator cend() const noexcept { return end(); }

    __device__ iterator find(const key_type& key) { return reinterpret_cast<iterator>(eng().find(to_skey(key))); }
    __device__ const_iterator find(const key_type& key) const
    {
        return reinterpret_cast<const_iterator>(eng().find(to_skey(key)));
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
        slot_type* p = eng().insert(
                to_skey(key_of(value)),
                [&value](slot_type* slot) {
                    if constexpr (!std::is_void_v<Mapped>)
                    {
                        smapped bits;
                        __builtin_memcpy(&bits, &value.second, sizeof(bits));
                        slot->second = bits;
                    }
                },
                &ins);
        return pair<iterator, bool>(reinterpret_cast<iterator>(p), ins);
    }

    template <class... Args>
    __device__ pair<iterator, bool> emplace(Args&&... args)
    {
        return insert(value_type(std::forward<Args>(args)...));
    }

    __device__ index_type erase(const key_type& key) { return eng().erase(to_skey(key)) ? 1 : 0; }

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
            c += bucket(from_skey(k)) == n ? 1 : 0;
        }
        // the two reserved key values live in side slots with a presence bit each
        const std::uint32_t f = htab_dev::load_agent(e.flags);
        if ((f & 1U) && bucket(from_skey(htab_dev::empty_key<skey>())) == n) ++c;
        if ((f & 2U) && bucket(from_skey(htab_dev::tomb_key<skey>())) == n) ++c;
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
        check(htab_insert(_h, raw(first), n, nullptr, policy.stream), "insert");
        check(static_cast<int>(hipStreamSynchronize(policy.stream)) == 0 ? HTAB_OK : HTAB_E_HIP, "insert sync");
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
        check(htab_erase(_h, raw(first), n, nullptr, policy.stream), "erase");
        check(static_cast<int>(hipStreamSynchronize(policy.stream)) == 0 ? HTAB_OK : HTAB_E_HIP, "erase sync");
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
        return device_indexed_range<value_type>(idx, static_cast<index_t>(n), reinterpret_cast<value_type*>(_view.slots));
    }
    device_indexed_range<const value_type> device_range() const { return device_range(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    device_indexed_range<const value_type> device_range(ExecutionPolicy&& policy) const
    {
        const std::int64_t* idx = nullptr;
        std::int64_t n = 0;
        if (_h) check(htab_range_owned(_h, &idx, &n, policy.stream), "device_range");
        return device_indexed_range<const value_type>(idx, static_cast<index_t>(n),
                                                      reinterpret_cast<const value_type*>(_view.slots));
    }

    void clear() { clear(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void clear(ExecutionPolicy&& policy)
    {
        if (!_h) return;
        check(htab_clear(_h, policy.stream), "clear");
        check(static_cast<int>(hipStreamSynchronize(policy.stream)) == 0 ? HTAB_OK : HTAB_E_HIP, "clear sync");
    }

    bool valid() const { return valid(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool valid(ExecutionPolicy&& policy) const
    {
        if (!_h) return true; // an empty (never created) container is valid
        std::int32_t ok = 0;
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

protected:
    __host__ __device__ engine eng() const { return engine::from_view(_view); }

    template <typename KeyIterator>
    void bulk_find(hipStream_t s, KeyIterator first, KeyIterator last, std::uint8_t* hits, void* values,
                   const char* what) const
    {
        const std::int64_t n = static_cast<std::int64_t>(last - first);
        if (n > 0 && !hits) throw std::invalid_argument(std::string("stdgpu::") + what + ": hits is null");
        check(htab_find(_h, raw(first), n, hits, values, s), what);
        check(static_cast<int>(hipStreamSynchronize(s)) == 0 ? HTAB_OK : HTAB_E_HIP, what);
    }

    __host__ __device__ static const Key& key_of(const value_type& v)
    {
        if constexpr (std::is_void_v<Mapped>)
            return v;
        else
            return v.first;
    }

    // the stored bits of a key: integers by value (sign bits included), other keys as they lie in memory
    __host__ __device__ static skey to_skey(const Key& k)
    {
        if constexpr (std::is_integral_v<Key>)
            return static_cast<skey>(static_cast<std::make_unsigned_t<Key>>(k));
        else
            return __builtin_bit_cast(skey, k);
    }

    // a key from its stored bits (the inverse of to_skey)
    __host__ __device__ static Key from_skey(const skey k)
    {
        if constexpr (std::is_integral_v<Key>)
            return static_cast<Key>(static_cast<std::make_unsigned_t<Key>>(k));
        else
            return __builtin_bit_cast(Key, k);
    }

    htab* _h = nullptr;
    htab_view _view{};
    Allocator _allocator{};
};

} // namespace stdgpu::detail

#endif // STDGPU_AMD_FACADE_UNORDERED_CUH
