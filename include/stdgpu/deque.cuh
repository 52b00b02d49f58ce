// Source-compatible stdgpu::deque<T> (src/stdgpu/deque.cuh) for trivially copyable T of up to 256
// bytes, on gfx950.
//
// Host: createDeviceObject / destroyDeviceObject, size / empty / full / capacity / max_size, clear,
// valid, shrink_to_fit, data, device_range (front to back over the ring).  Device (user kernels, per
// thread): push_back / emplace_back / push_front / emplace_front / pop_back / pop_front, size / empty /
// full, operator[] / at / front / back (include/stdgpu_amd/ring_device.hpp).  The calling lanes of a
// wavefront reserve from the published-element or released-slot count with one atomic and take
// consecutive positions with one more; a push or pop then claims its slot by compare-and-swap on the
// slot's state, waiting while another call on the slot finishes, as the reference waits on its
// per-slot lock.  Pushes and pops of both ends may be mixed in one kernel.  Bulk pushes and pops of
// many elements per call: htab_deque_* in <stdgpu_amd/htab.h>.
#ifndef STDGPU_DEQUE_CUH
#define STDGPU_DEQUE_CUH

#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include <stdgpu/impl/amd_core.h>
#include <stdgpu_amd/ring_device.hpp>

namespace stdgpu
{

// Random-access iterator over a ring of `capacity` slots: element i is data[(start + i) mod capacity]
template <typename T>
class ring_iterator
{
public:
    using value_type = std::remove_cv_t<T>;
    using difference_type = std::ptrdiff_t;
    using reference = T&;
    using pointer = T*;
    using iterator_category = std::random_access_iterator_tag;

    constexpr __host__ __device__ ring_iterator(T* data, std::uint64_t capacity, std::uint64_t pos)
      : _data(data), _cap(capacity), _pos(pos)
    {
    }
    __device__ T& operator*() const { return _data[_pos % _cap]; }
    __device__ T& operator[](difference_type i) const { return _data[(_pos + static_cast<std::uint64_t>(i)) % _cap]; }
    constexpr __host__ __device__ ring_iterator operator+(difference_type d) const
    {
        return ring_iterator(_data, _cap, _pos + static_cast<std::uint64_t>(d));
    }
    constexpr __host__ __device__ difference_type operator-(const ring_iterator& o) const
    {
        return static_cast<difference_type>(_pos - o._pos);
    }
    constexpr __host__ __device__ ring_iterator& operator++()
    {
        ++_pos;
        return *this;
    }
    constexpr __host__ __device__ bool operator==(const ring_iterator& o) const { return _pos == o._pos; }
    constexpr __host__ __device__ bool operator!=(const ring_iterator& o) const { return _pos != o._pos; }

private:
    T* _data;
    std::uint64_t _cap;
    std::uint64_t _pos;
};

// the elements of a deque front to back, as a range over its ring (deque.cuh device_range)
template <typename T>
class ring_range
{
public:
    constexpr __host__ __device__ ring_range(T* data, std::uint64_t capacity, std::uint64_t first, index_t n)
      : _data(data), _cap(capacity), _first(first), _n(n)
    {
    }
    constexpr __host__ __device__ ring_iterator<T> begin() const { return ring_iterator<T>(_data, _cap, _first); }
    constexpr __host__ __device__ ring_iterator<T> end() const
    {
        return ring_iterator<T>(_data, _cap, _first + static_cast<std::uint64_t>(_n));
    }
    constexpr __host__ __device__ index_t size() const { return _n; }
    constexpr __host__ __device__ bool empty() const { return _n == 0; }

private:
    T* _data;
    std::uint64_t _cap;
    std::uint64_t _first;
    index_t _n;
};

template <typename T, typename Allocator = safe_device_allocator<T>>
class deque
{
public:
    static_assert(std::is_trivially_copyable_v<T> && sizeof(T) <= 256,
                  "stdgpu::deque (MI355X library): trivially copyable value types of at most 256 bytes");

    using value_type = T;
    using allocator_type = Allocator;
    using index_type = index_t;
    using difference_type = std::ptrdiff_t;
    using reference = value_type&;
    using const_reference = const value_type&;
    using pointer = value_type*;
    using const_pointer = const value_type*;

    [[nodiscard]] static deque createDeviceObject(const index_t& capacity, const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, capacity, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static deque createDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, const index_t& capacity,
                                                  const Allocator& allocator = Allocator())
    {
        if (capacity <= 0) throw std::invalid_argument("stdgpu::deque: capacity must be positive");
        deque d;
        d._allocator = allocator;
        detail::check(htab_deque_create_sized(capacity, static_cast<std::int64_t>(sizeof(T)), &d._h),
                      "deque::createDeviceObject");
        detail::check(htab_deque_get_view(d._h, &d._view), "deque::createDeviceObject");
        return d;
    }

    static void destroyDeviceObject(deque& device_object) { destroyDeviceObject(execution::device, device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, deque& device_object)
    {
        if (device_object._h) detail::check(htab_deque_destroy(device_object._h), "deque::destroyDeviceObject");
        device_object._h = nullptr;
        device_object._view = htab_deque_view{};
    }

    deque() = default;

    [[nodiscard]] __host__ __device__ allocator_type get_allocator() const noexcept { return _allocator; }

    // ---------------------------------------------------------------- device element access
    __device__ reference at(const index_type n) { return *ref().template at<T>(n); }
    __device__ const_reference at(const index_type n) const { return *ref().template at<T>(n); }
    __device__ reference operator[](const index_type n) { return at(n); }
    __device__ const_reference operator[](const index_type n) const { return at(n); }
    __device__ reference front() { return at(0); }
    __device__ const_reference front() const { return at(0); }
    __device__ reference back() { return at(static_cast<index_type>(ref().size()) - 1); }
    __device__ const_reference back() const { return at(static_cast<index_type>(ref().size()) - 1); }

    // ---------------------------------------------------------------- device modifiers
    __device__ bool push_back(const T& element) { return ref().push_back(element); }
    __device__ bool push_front(const T& element) { return ref().push_front(element); }

    template <class... Args>
    __device__ bool emplace_back(Args&&... args)
    {
        return push_back(T(std::forward<Args>(args)...));
    }

    template <class... Args>
    __device__ bool emplace_front(Args&&... args)
    {
        return push_front(T(std::forward<Args>(args)...));
    }

    // the popped element and true, or T() and false (deque.cuh:213-239)
    __device__ pair<T, bool> pop_back()
    {
        T v{};
        const bool ok = ref().pop_back(&v);
        return pair<T, bool>(ok ? v : T(), ok);
    }
    __device__ pair<T, bool> pop_front()
    {
        T v{};
        const bool ok = ref().pop_front(&v);
        return pair<T, bool>(ok ? v : T(), ok);
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
        if (_h) detail::check(htab_deque_size(_h, &n, policy.stream), "deque::size");
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

    // ---------------------------------------------------------------- host modifiers and checks
    void clear() { clear(execution::device); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void clear(ExecutionPolicy&& policy)
    {
        if (!_h) return;
        detail::check(htab_deque_clear(_h, policy.stream), "deque::clear");
        detail::check(htab_rt_sync(policy.stream), "deque::clear");
    }

    // the capacity stays (the reference may or may not shrink, deque.inc shrink_to_fit)
    void shrink_to_fit() {}

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void shrink_to_fit([[maybe_unused]] ExecutionPolicy&& policy)
    {
    }

    // the elements' slots full, every other slot free, the counts consistent (deque_detail.cuh:514-573)
    [[nodiscard]] bool valid() const { return valid(execution::device); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] bool valid(ExecutionPolicy&& policy) const
    {
        if (!_h) return true;
        std::int32_t ok = 0;
        detail::check(htab_deque_valid(_h, &ok, policy.stream), "deque::valid");
        return ok != 0;
    }

    // the elements front to back (a host call: it reads the front counter and the size)
    ring_range<T> device_range() { return device_range(execution::device); }
    ring_range<const T> device_range() const { return device_range(execution::device); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    ring_range<T> device_range(ExecutionPolicy&& policy)
    {
        const std::uint64_t first = front_position(policy.stream);
        return ring_range<T>(data(), static_cast<std::uint64_t>(_view.capacity), first, size(policy));
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    ring_range<const T> device_range(ExecutionPolicy&& policy) const
    {
        const std::uint64_t first = front_position(policy.stream);
        return ring_range<const T>(data(), static_cast<std::uint64_t>(_view.capacity), first, size(policy));
    }

    htab_deque* native_handle() const { return _h; }

private:
    __host__ __device__ htab_dev::ring_ref ref() const { return htab_dev::ring_ref::from_view(_view); }

    std::uint64_t front_position(hipStream_t s) const
    {
        std::uint64_t b = 0;
        if (_h)
        {
            detail::check(htab_rt_copy(&b, _view.begin, sizeof(b), 0, s), "deque::device_range");
        }
        return b;
    }

    htab_deque* _h = nullptr;
    htab_deque_view _view{};
    Allocator _allocator{};
};

} // namespace stdgpu

#endif // STDGPU_DEQUE_CUH
