// Source-compatible stdgpu::vector<T> (src/stdgpu/vector.cuh) for trivially copyable T of up to 256
// bytes, on gfx950.
//
// Host: createDeviceObject / destroyDeviceObject, insert at the end and erase up to the end from
// device ranges (the only positions the reference implements, vector.cuh:73-75), clear, size / empty
// / full / capacity / max_size, shrink_to_fit, device_begin / device_end / device_range, valid.  Device
// (user kernels, per thread): push_back / emplace_back / pop_back, size / empty / full, operator[] /
// at / front / back (include/stdgpu_amd/ring_device.hpp: the back end of the ring, which never
// wraps).  The calling lanes of a wavefront take their positions with one reservation and one atomic;
// each push or pop then claims its slot by compare-and-swap on the slot's state, as the reference
// claims its per-slot lock, so pushes and pops may be mixed in one kernel.  Bulk pushes and pops of
// many elements per call: htab_vector_* in <stdgpu_amd/htab.h>.
#ifndef STDGPU_VECTOR_CUH
#define STDGPU_VECTOR_CUH

#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include <stdgpu/impl/amd_core.h>
#include <stdgpu_amd/ring_device.hpp>

namespace stdgpu
{

template <typename T, typename Allocator = safe_device_allocator<T>>
class vector
{
public:
    static_assert(std::is_trivially_copyable_v<T> && sizeof(T) <= 256,
                  "stdgpu::vector (MI355X library): trivially copyable value types of at most 256 bytes");

    using value_type = T;
    using allocator_type = Allocator;
    using index_type = index_t;
    using difference_type = std::ptrdiff_t;
    using reference = value_type&;
    using const_reference = const value_type&;
    using pointer = value_type*;
    using const_pointer = const value_type*;

    [[nodiscard]] static vector createDeviceObject(const index_t& capacity, const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, capacity, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static vector createDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, const index_t& capacity,
                                                   const Allocator& allocator = Allocator())
    {
        if (capacity <= 0) throw std::invalid_argument("stdgpu::vector: capacity must be positive");
        vector v;
        v._allocator = allocator;
        detail::check(htab_vector_create_sized(capacity, static_cast<std::int64_t>(sizeof(T)), &v._h),
                      "vector::createDeviceObject");
        detail::check(htab_vector_get_view(v._h, &v._view), "vector::createDeviceObject");
        return v;
    }

    static void destroyDeviceObject(vector& device_object) { destroyDeviceObject(execution::device, device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, vector& device_object)
    {
        if (device_object._h) detail::check(htab_vector_destroy(device_object._h), "vector::destroyDeviceObject");
        device_object._h = nullptr;
        device_object._view = htab_vector_view{};
    }

    vector() = default;

    [[nodiscard]] __host__ __device__ allocator_type get_allocator() const noexcept { return _allocator; }

    // ---------------------------------------------------------------- device element access
    __device__ reference at(const index_type n) { return data()[n]; }
    __device__ const_reference at(const index_type n) const { return data()[n]; }
    __device__ reference operator[](const index_type n) { return at(n); }
    __device__ const_reference operator[](const index_type n) const { return at(n); }
    __device__ reference front() { return at(0); }
    __device__ const_reference front() const { return at(0); }
    __device__ reference back() { return at(size() - 1); }
    __device__ const_reference back() const { return at(size() - 1); }

    // ---------------------------------------------------------------- device modifiers
    __device__ bool push_back(const T& element) { return ref().push_back(element); }

    template <class... Args>
    __device__ bool emplace_back(Args&&... args)
    {
        return push_back(T(std::forward<Args>(args)...));
    }

    // the popped element and true, or T() and false (vector.cuh:234-239)
    __device__ pair<T, bool> pop_back()
    {
        T v{};
        const bool ok = ref().pop_back(&v);
        return pair<T, bool>(ok ? v : T(), ok);
    }

    // ---------------------------------------------------------------- host modifiers
    // insert(device_end(), first, last): appends [first, last) (vector_detail.cuh:374-414).  As in the
    // reference, a range that would pass the capacity, or a position other than device_end() (the only
    // one the reference implements), leaves the vector unchanged.
    template <typename ValueIterator>
    void insert(device_ptr<const T> position, ValueIterator first, ValueIterator last)
    {
        insert(execution::device, position, first, last);
    }

    template <typename ExecutionPolicy, typename ValueIterator,
              std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void insert(ExecutionPolicy&& policy, device_ptr<const T> position, ValueIterator first, ValueIterator last)
    {
        const index_t n0 = size(policy);
        if (position.get() != data() + n0) return;
        const std::int64_t n = static_cast<std::int64_t>(last - first);
        if (n <= 0 || n0 + n > capacity()) return;
        detail::check(htab_vector_push_back(_h, static_cast<const void*>(detail::raw(first)), n, nullptr,
                                            policy.stream),
                      "vector::insert");
        sync(policy.stream, "vector::insert");
    }

    // erase(first, device_end()): drops the elements from first on (vector_detail.cuh:417-451).  As in
    // the reference, a range that does not end at device_end() or starts before device_begin() leaves
    // the vector unchanged.
    void erase(device_ptr<const T> first, device_ptr<const T> last) { erase(execution::device, first, last); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void erase(ExecutionPolicy&& policy, device_ptr<const T> first, device_ptr<const T> last)
    {
        const index_t n = size(policy);
        if (last.get() != data() + n || first.get() < data() || first.get() > last.get()) return;
        const std::int64_t k = static_cast<std::int64_t>(last.get() - first.get());
        if (k == 0) return;
        detail::check(htab_vector_pop_back(_h, nullptr, k, nullptr, policy.stream), "vector::erase");
        sync(policy.stream, "vector::erase");
    }

    void clear() { clear(execution::device); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void clear(ExecutionPolicy&& policy)
    {
        if (!_h) return;
        detail::check(htab_vector_clear(_h, policy.stream), "vector::clear");
        sync(policy.stream, "vector::clear");
    }

    // the capacity stays (the reference may or may not shrink, vector.inc shrink_to_fit)
    void shrink_to_fit() {}

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void shrink_to_fit([[maybe_unused]] ExecutionPolicy&& policy)
    {
    }

    // ---------------------------------------------------------------- capacity
    [[nodiscard]] __host__ __device__ bool empty() const { return size() == 0; }
    [[nodiscard]] __host__ __device__ bool full() const { return size() == max_size(); }

    __host__ __device__ index_t size() const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return static_cast<index_t>(ref().size());
#else
        return size(execution::device);
#endif
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    index_t size(ExecutionPolicy&& policy) const
    {
        std::int64_t n = 0;
        if (_h) detail::check(htab_vector_size(_h, &n, policy.stream), "vector::size");
        return static_cast<index_t>(n);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool empty(ExecutionPolicy&& policy) const
    {
        return size(std::forward<ExecutionPolicy>(policy)) == 0;
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool full(ExecutionPolicy&& policy) const
    {
        return size(std::forward<ExecutionPolicy>(policy)) == max_size();
    }

    __host__ __device__ index_t max_size() const noexcept { return capacity(); }
    __host__ __device__ index_t capacity() const noexcept { return static_cast<index_t>(_view.capacity); }

    __host__ __device__ const T* data() const noexcept { return reinterpret_cast<const T*>(_view.data); }
    __host__ __device__ T* data() noexcept { return reinterpret_cast<T*>(_view.data); }

    // device ranges over the elements (host calls; device_end reads the size)
    device_ptr<T> device_begin() { return device_ptr<T>(data()); }
    device_ptr<T> device_end() { return device_ptr<T>(data() + size()); }
    device_ptr<const T> device_begin() const { return device_ptr<const T>(data()); }
    device_ptr<const T> device_end() const { return device_ptr<const T>(data() + size()); }
    device_ptr<const T> device_cbegin() const { return device_begin(); }
    device_ptr<const T> device_cend() const { return device_end(); }

    // the elements as a range (vector.cuh device_range; a host call, it reads the size)
    stdgpu::device_range<T> device_range() { return stdgpu::device_range<T>(data(), size()); }
    stdgpu::device_range<const T> device_range() const { return stdgpu::device_range<const T>(data(), size()); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    stdgpu::device_range<T> device_range(ExecutionPolicy&& policy)
    {
        return stdgpu::device_range<T>(data(), size(std::forward<ExecutionPolicy>(policy)));
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    stdgpu::device_range<const T> device_range(ExecutionPolicy&& policy) const
    {
        return stdgpu::device_range<const T>(data(), size(std::forward<ExecutionPolicy>(policy)));
    }

    // the elements' slots full, every other slot free, the counts consistent (vector_detail.cuh:514-577)
    [[nodiscard]] bool valid() const { return valid(execution::device); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] bool valid(ExecutionPolicy&& policy) const
    {
        if (!_h) return true;
        std::int32_t ok = 0;
        detail::check(htab_vector_valid(_h, &ok, policy.stream), "vector::valid");
        return ok != 0;
    }

    htab_vector* native_handle() const { return _h; }

private:
    __host__ __device__ htab_dev::ring_ref ref() const { return htab_dev::ring_ref::from_view(_view); }

    static void sync(hipStream_t s, const char* what)
    {
        detail::check(htab_rt_sync(s), what);
    }

    htab_vector* _h = nullptr;
    htab_vector_view _view{};
    Allocator _allocator{};
};

} // namespace stdgpu

#endif // STDGPU_VECTOR_CUH
