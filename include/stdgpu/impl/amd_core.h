// Shared vocabulary of the source-compatible facade: index_t, pair, hash, equal_to,
// execution policies, allocator tag, device_ptr, device_range and device_indexed_range.
//
// Names and meanings follow the reference headers they stand in for (paths relative to the
// stdgpu source tree): cstddef.h:55-59 (index_t), utility.h:43 (pair), functional.h and
// impl/functional_detail.h:60-125 (hash, equal_to), execution.h:43-92 (execution::device),
// memory.h:340 (safe_device_allocator), iterator.h:43-184 (device_ptr, make_device) and
// ranges.h (device_range, device_indexed_range :285).  Compiled with hipcc (gfx950).
#ifndef STDGPU_AMD_FACADE_CORE_H
#define STDGPU_AMD_FACADE_CORE_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include <stdgpu_amd/htab.h>
#include <stdgpu_amd/htab_device.hpp>

namespace stdgpu
{

// 32-bit index type, the reference default (STDGPU_USE_32_BIT_INDEX)
using index_t = std::int32_t;

template <typename T1, typename T2>
struct pair
{
    using first_type = T1;
    using second_type = T2;

    T1 first;
    T2 second;

    constexpr __host__ __device__ pair() : first(), second() {}
    constexpr __host__ __device__ pair(const T1& x, const T2& y) : first(x), second(y) {}
    template <class U1, class U2>
    constexpr __host__ __device__ pair(U1&& x, U2&& y) : first(std::forward<U1>(x)), second(std::forward<U2>(y))
    {
    }
    template <class U1, class U2>
    constexpr __host__ __device__ pair(const pair<U1, U2>& p) : first(p.first), second(p.second)
    {
    }
    pair(const pair&) = default;
    pair(pair&&) = default;
};

// hash<T>: the identity for integers, widened to size_t (functional_detail.h:60-80)
template <typename T>
struct hash
{
    static_assert(std::is_integral_v<T>, "stdgpu::hash: the MI355X library hashes integral keys");
    constexpr __host__ __device__ std::size_t operator()(const T& key) const { return static_cast<std::size_t>(key); }
};

template <typename T = void>
struct equal_to
{
    constexpr __host__ __device__ bool operator()(const T& a, const T& b) const { return a == b; }
};

template <>
struct equal_to<void>
{
    using is_transparent = void;
    template <class A, class B>
    constexpr __host__ __device__ bool operator()(const A& a, const B& b) const
    {
        return a == b;
    }
};

namespace execution
{
// The device policy carries the stream bulk calls run on; calls complete before returning,
// as the reference's thrust-dispatched bulk calls do.
struct device_policy
{
    hipStream_t stream = nullptr;
};
inline constexpr device_policy device{};
} // namespace execution

template <typename T>
inline constexpr bool is_execution_policy_v = std::is_base_of_v<execution::device_policy, std::decay_t<T>>;

// The tables allocate through the C ABI (hipMalloc inside libstdgpu_amd.so); the allocator is
// carried for API compatibility, returned by get_allocator(), and allocates device arrays itself
// (memory.h:340-380: allocate(n) / deallocate(p, n)).
template <typename T>
struct safe_device_allocator
{
    using value_type = T;
    safe_device_allocator() = default;
    template <typename U>
    constexpr __host__ __device__ safe_device_allocator(const safe_device_allocator<U>&) noexcept
    {
    }

    [[nodiscard]] T* allocate(std::int64_t n)
    {
        if (n < 0) throw std::invalid_argument("stdgpu: negative allocation size");
        void* p = nullptr;
        if (htab_rt_alloc(n * static_cast<std::int64_t>(sizeof(T)), &p) != HTAB_OK)
            throw std::runtime_error("stdgpu: device allocation failed");
        return static_cast<T*>(p);
    }

    void deallocate(T* p, std::int64_t /*n*/)
    {
        if (p) (void)htab_rt_free(p);
    }
};

template <typename T>
class device_ptr
{
public:
    using value_type = T;
    using difference_type = std::ptrdiff_t;
    using pointer = T*;
    using reference = T&;
    using iterator_category = std::random_access_iterator_tag;

    constexpr __host__ __device__ device_ptr() = default;
    constexpr __host__ __device__ explicit device_ptr(T* p) : _p(p) {}
    // device_ptr<T> converts to device_ptr<const T> (iterator.h: the const overloads take either)
    template <typename U, std::enable_if_t<std::is_convertible_v<U*, T*> && !std::is_same_v<U, T>, int> = 0>
    constexpr __host__ __device__ device_ptr(const device_ptr<U>& other) : _p(other.get())
    {
    }
    constexpr __host__ __device__ T* get() const { return _p; }
    __device__ T& operator*() const { return *_p; }
    __device__ T& operator[](difference_type i) const { return _p[i]; }
    constexpr __host__ __device__ device_ptr operator+(difference_type d) const { return device_ptr(_p + d); }
    constexpr __host__ __device__ device_ptr operator-(difference_type d) const { return device_ptr(_p - d); }
    constexpr __host__ __device__ difference_type operator-(const device_ptr& o) const { return _p - o._p; }
    constexpr __host__ __device__ device_ptr& operator++()
    {
        ++_p;
        return *this;
    }
    constexpr __host__ __device__ bool operator==(const device_ptr& o) const { return _p == o._p; }
    constexpr __host__ __device__ bool operator!=(const device_ptr& o) const { return _p != o._p; }

private:
    T* _p = nullptr;
};

template <typename T>
constexpr __host__ __device__ device_ptr<T>
make_device(T* p)
{
    return device_ptr<T>(p);
}

namespace detail
{
template <typename It>
constexpr auto
raw(It it)
{
    if constexpr (std::is_pointer_v<It>)
        return it;
    else
        return it.get();
}

inline void
check(int rc, const char* what)
{
    if (rc != HTAB_OK)
    {
        throw std::runtime_error(std::string(what) + ": " + htab_last_error());
    }
}
} // namespace detail

// Random-access iterator over the stored entries: element i is values[idx[i]].
template <typename T>
class indexed_iterator
{
public:
    using value_type = std::remove_cv_t<T>;
    using difference_type = std::ptrdiff_t;
    using reference = T&;
    using pointer = T*;
    using iterator_category = std::random_access_iterator_tag;

    constexpr __host__ __device__ indexed_iterator(const std::int64_t* idx, T* values) : _idx(idx), _values(values) {}
    __device__ T& operator*() const { return _values[*_idx]; }
    __device__ T& operator[](difference_type i) const { return _values[_idx[i]]; }
    constexpr __host__ __device__ indexed_iterator operator+(difference_type d) const
    {
        return indexed_iterator(_idx + d, _values);
    }
    constexpr __host__ __device__ difference_type operator-(const indexed_iterator& o) const { return _idx - o._idx; }
    constexpr __host__ __device__ indexed_iterator& operator++()
    {
        ++_idx;
        return *this;
    }
    constexpr __host__ __device__ bool operator==(const indexed_iterator& o) const { return _idx == o._idx; }
    constexpr __host__ __device__ bool operator!=(const indexed_iterator& o) const { return _idx != o._idx; }

private:
    const std::int64_t* _idx;
    T* _values;
};

// The stored entries of a table in ascending slot order (reference: transform_range over the
// collected indices, ranges.h:285).  Indices live in table-owned device memory.
template <typename T>
class device_indexed_range
{
public:
    constexpr __host__ __device__ device_indexed_range(const std::int64_t* idx, index_t n, T* values)
      : _idx(idx), _n(n), _values(values)
    {
    }
    constexpr __host__ __device__ indexed_iterator<T> begin() const { return indexed_iterator<T>(_idx, _values); }
    constexpr __host__ __device__ indexed_iterator<T> end() const { return indexed_iterator<T>(_idx + _n, _values); }
    constexpr __host__ __device__ index_t size() const { return _n; }
    constexpr __host__ __device__ const std::int64_t* indices() const { return _idx; }

private:
    const std::int64_t* _idx;
    index_t _n;
    T* _values;
};

// a contiguous device range [begin, begin + size) (ranges.h device_range), e.g. a vector's elements
template <typename T>
class device_range
{
public:
    constexpr __host__ __device__ device_range() = default;
    constexpr __host__ __device__ device_range(T* first, index_t n) : _first(first), _n(n) {}
    constexpr __host__ __device__ device_ptr<T> begin() const { return device_ptr<T>(_first); }
    constexpr __host__ __device__ device_ptr<T> end() const { return device_ptr<T>(_first + _n); }
    constexpr __host__ __device__ index_t size() const { return _n; }
    constexpr __host__ __device__ bool empty() const { return _n == 0; }

private:
    T* _first = nullptr;
    index_t _n = 0;
};

} // namespace stdgpu

#endif // STDGPU_AMD_FACADE_CORE_H
