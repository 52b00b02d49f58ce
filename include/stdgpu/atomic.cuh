// Source-compatible stdgpu::atomic<T> / stdgpu::atomic_ref<T> (src/stdgpu/atomic.cuh) for
// T = int, unsigned int, unsigned long long int, float, on gfx950.
//
// Device member functions are single device-scope atomics (include/stdgpu_amd/atomic_device.hpp), with
// the reference's semantics for each (src/stdgpu/impl/atomic_detail.cuh:428-760); host load() /
// store() copy the value (atomic_detail.cuh:461-526).  atomic<T> owns one value allocated by its
// allocator and filled with 0 (atomic_detail.cuh:143-151); it is copied by value into kernels, as in the
// reference.  The memory orders map to the compiler's; an operation at seq_cst is a seq_cst atomic rather
// than the reference's relaxed atomic between two fences, which orders at least as strongly.
#ifndef STDGPU_ATOMIC_CUH
#define STDGPU_ATOMIC_CUH

#include <cstdint>
#include <type_traits>
#include <utility>

#include <stdgpu/impl/amd_core.h>
#include <stdgpu_amd/atomic_device.hpp>

namespace stdgpu
{

enum class memory_order : std::int8_t
{
    relaxed,
    consume,
    acquire,
    release,
    acq_rel,
    seq_cst
};
inline constexpr memory_order memory_order_relaxed = memory_order::relaxed;
inline constexpr memory_order memory_order_consume = memory_order::consume;
inline constexpr memory_order memory_order_acquire = memory_order::acquire;
inline constexpr memory_order memory_order_release = memory_order::release;
inline constexpr memory_order memory_order_acq_rel = memory_order::acq_rel;
inline constexpr memory_order memory_order_seq_cst = memory_order::seq_cst;

namespace detail
{
constexpr __host__ __device__ htab_dev::mem_order
to_dev(memory_order o)
{
    return static_cast<htab_dev::mem_order>(static_cast<int>(o));
}

inline void
check_rt(int rc, const char* what)
{
    if (rc != HTAB_OK) throw std::runtime_error(std::string("stdgpu::atomic: ") + what + ": " + htab_last_error());
}
} // namespace detail

// a fence at device scope with the given order (relaxed: none)
__device__ inline void
atomic_thread_fence(const memory_order order) noexcept
{
    switch (order)
    {
    case memory_order::relaxed:
        break;
    case memory_order::consume:
    case memory_order::acquire:
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        break;
    case memory_order::release:
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        break;
    case memory_order::acq_rel:
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        break;
    default:
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        break;
    }
}

// a fence between a thread and code interrupting it on the same thread: orders the compiler only
__device__ inline void
atomic_signal_fence(const memory_order order) noexcept
{
    if (order != memory_order::relaxed) __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <typename T, typename Allocator>
class atomic;

template <typename T>
class atomic_ref
{
public:
    static_assert(htab_dev::is_atomic_value_v<T>, "stdgpu::atomic_ref: No support for type T");

    using value_type = T;
    using difference_type = value_type;

    atomic_ref() = delete;

    __host__ __device__ explicit atomic_ref(T& obj) noexcept : _value(&obj) {}

    [[nodiscard]] __host__ __device__ bool is_lock_free() const noexcept { return true; }

    [[nodiscard]] __host__ __device__ T load(const memory_order order = memory_order_seq_cst) const
    {
        if (_value == nullptr) return T(0);
#if defined(__HIP_DEVICE_COMPILE__)
        return htab_dev::atomic_load_op(_value, detail::to_dev(order));
#else
        return load(execution::device, order);
#endif
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] T load(ExecutionPolicy&& policy, [[maybe_unused]] const memory_order order = memory_order_seq_cst) const
    {
        if (_value == nullptr) return T(0);
        T v{};
        detail::check_rt(htab_rt_copy(&v, _value, sizeof(T), 0, policy.stream), "load");
        return v;
    }

    __host__ __device__ operator T() const { return load(); } // NOLINT(hicpp-explicit-conversions)

    __host__ __device__ void store(const T desired, const memory_order order = memory_order_seq_cst)
    {
        if (_value == nullptr) return;
#if defined(__HIP_DEVICE_COMPILE__)
        htab_dev::atomic_store_op(_value, desired, detail::to_dev(order));
#else
        store(execution::device, desired, order);
#endif
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void store(ExecutionPolicy&& policy, const T desired, [[maybe_unused]] const memory_order order = memory_order_seq_cst)
    {
        if (_value == nullptr) return;
        T v = desired;
        detail::check_rt(htab_rt_copy(_value, &v, sizeof(T), 1, policy.stream), "store");
    }

    __host__ __device__ T operator=(const T desired) // NOLINT(misc-unconventional-assign-operator)
    {
        store(desired);
        return desired;
    }

    __device__ T exchange(const T desired, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_exchange_op(_value, desired, detail::to_dev(order));
    }

    // expected takes the value seen when it differs (atomic_detail.cuh:583-610)
    __device__ bool compare_exchange_weak(T& expected, const T desired,
                                          const memory_order order = memory_order_seq_cst) noexcept
    {
        return compare_exchange_strong(expected, desired, order);
    }

    __device__ bool compare_exchange_strong(T& expected, const T desired,
                                            const memory_order order = memory_order_seq_cst) noexcept
    {
        const T old = htab_dev::atomic_cas_op(_value, expected, desired, detail::to_dev(order));
        const bool changed = old == expected;
        if (!changed) expected = old;
        return changed;
    }

    __device__ T fetch_add(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_add_op(_value, arg, detail::to_dev(order));
    }

    __device__ T fetch_sub(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_sub_op(_value, arg, detail::to_dev(order));
    }

    __device__ T fetch_min(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_min_op(_value, arg, detail::to_dev(order));
    }

    __device__ T fetch_max(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_max_op(_value, arg, detail::to_dev(order));
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_and(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_and_op(_value, arg, detail::to_dev(order));
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_or(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_or_op(_value, arg, detail::to_dev(order));
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_xor(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_fetch_xor_op(_value, arg, detail::to_dev(order));
    }

    // the value counts modulo arg: 0, 1, ..., arg - 1, 0, ... (atomic_detail.cuh:679-688)
    template <typename U = T, std::enable_if_t<std::is_same_v<U, unsigned int>, int> = 0>
    __device__ T fetch_inc_mod(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_inc_wrap_op(_value, arg - 1U, detail::to_dev(order));
    }

    template <typename U = T, std::enable_if_t<std::is_same_v<U, unsigned int>, int> = 0>
    __device__ T fetch_dec_mod(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return htab_dev::atomic_dec_wrap_op(_value, arg - 1U, detail::to_dev(order));
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator++() noexcept
    {
        return fetch_add(1) + 1;
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator++(int) noexcept
    {
        return fetch_add(1);
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator--() noexcept
    {
        return fetch_sub(1) - 1;
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator--(int) noexcept
    {
        return fetch_sub(1);
    }

    __device__ T operator+=(const T arg) noexcept { return fetch_add(arg) + arg; }
    __device__ T operator-=(const T arg) noexcept { return fetch_sub(arg) - arg; }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator&=(const T arg) noexcept
    {
        return fetch_and(arg) & arg;
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator|=(const T arg) noexcept
    {
        return fetch_or(arg) | arg;
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator^=(const T arg) noexcept
    {
        return fetch_xor(arg) ^ arg;
    }

private:
    template <typename, typename>
    friend class atomic;

    __host__ __device__ explicit atomic_ref(T* value) noexcept : _value(value) {}

    T* _value = nullptr;
};

template <typename T, typename Allocator = safe_device_allocator<T>>
class atomic
{
public:
    static_assert(htab_dev::is_atomic_value_v<T>, "stdgpu::atomic: No support for type T");

    using value_type = T;
    using difference_type = value_type;
    using allocator_type = Allocator;

    [[nodiscard]] static atomic createDeviceObject(const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static atomic createDeviceObject(ExecutionPolicy&& policy, const Allocator& allocator = Allocator())
    {
        atomic result(allocator);
        T* p = result._allocator.allocate(1);
        detail::check_rt(htab_rt_zero(p, sizeof(T), policy.stream), "createDeviceObject");
        detail::check_rt(htab_rt_sync(policy.stream), "createDeviceObject");
        result._value = p;
        return result;
    }

    static void destroyDeviceObject(atomic& device_object) { destroyDeviceObject(execution::device, device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, atomic& device_object)
    {
        device_object._allocator.deallocate(device_object._value, 1);
        device_object._value = nullptr;
    }

    atomic() noexcept = default;

    [[nodiscard]] __host__ __device__ allocator_type get_allocator() const noexcept { return _allocator; }
    [[nodiscard]] __host__ __device__ bool is_lock_free() const noexcept { return ref().is_lock_free(); }

    [[nodiscard]] __host__ __device__ T load(const memory_order order = memory_order_seq_cst) const
    {
        return ref().load(order);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] T load(ExecutionPolicy&& policy, const memory_order order = memory_order_seq_cst) const
    {
        return ref().load(std::forward<ExecutionPolicy>(policy), order);
    }

    __host__ __device__ operator T() const { return load(); } // NOLINT(hicpp-explicit-conversions)

    __host__ __device__ void store(const T desired, const memory_order order = memory_order_seq_cst)
    {
        ref().store(desired, order);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void store(ExecutionPolicy&& policy, const T desired, const memory_order order = memory_order_seq_cst)
    {
        ref().store(std::forward<ExecutionPolicy>(policy), desired, order);
    }

    __host__ __device__ T operator=(const T desired) // NOLINT(misc-unconventional-assign-operator)
    {
        store(desired);
        return desired;
    }

    __device__ T exchange(const T desired, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().exchange(desired, order);
    }
    __device__ bool compare_exchange_weak(T& expected, const T desired,
                                          const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().compare_exchange_weak(expected, desired, order);
    }
    __device__ bool compare_exchange_strong(T& expected, const T desired,
                                            const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().compare_exchange_strong(expected, desired, order);
    }
    __device__ T fetch_add(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_add(arg, order);
    }
    __device__ T fetch_sub(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_sub(arg, order);
    }
    __device__ T fetch_min(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_min(arg, order);
    }
    __device__ T fetch_max(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_max(arg, order);
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_and(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_and(arg, order);
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_or(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_or(arg, order);
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T fetch_xor(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_xor(arg, order);
    }
    template <typename U = T, std::enable_if_t<std::is_same_v<U, unsigned int>, int> = 0>
    __device__ T fetch_inc_mod(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_inc_mod(arg, order);
    }
    template <typename U = T, std::enable_if_t<std::is_same_v<U, unsigned int>, int> = 0>
    __device__ T fetch_dec_mod(const T arg, const memory_order order = memory_order_seq_cst) noexcept
    {
        return ref().fetch_dec_mod(arg, order);
    }

    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator++() noexcept
    {
        return ++ref();
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator++(int) noexcept
    {
        return ref()++;
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator--() noexcept
    {
        return --ref();
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator--(int) noexcept
    {
        return ref()--;
    }
    __device__ T operator+=(const T arg) noexcept { return ref() += arg; }
    __device__ T operator-=(const T arg) noexcept { return ref() -= arg; }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator&=(const T arg) noexcept
    {
        return ref() &= arg;
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator|=(const T arg) noexcept
    {
        return ref() |= arg;
    }
    template <typename U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
    __device__ T operator^=(const T arg) noexcept
    {
        return ref() ^= arg;
    }

private:
    explicit atomic(const Allocator& allocator) : _allocator(allocator) {}

    __host__ __device__ atomic_ref<T> ref() const { return atomic_ref<T>(_value); }

    T* _value = nullptr;
    Allocator _allocator = {};
};

// ------------------------------------------------------------------ non-member functions
// (atomic.cuh:786-1066)

template <typename T, typename Allocator>
__host__ __device__ bool
atomic_is_lock_free(const atomic<T, Allocator>* obj) noexcept
{
    return obj->is_lock_free();
}

template <typename T, typename Allocator>
__host__ __device__ T
atomic_load(const atomic<T, Allocator>* obj) noexcept
{
    return obj->load();
}

template <typename T, typename Allocator>
__host__ __device__ T
atomic_load_explicit(const atomic<T, Allocator>* obj, const memory_order order) noexcept
{
    return obj->load(order);
}

template <typename T, typename Allocator>
__host__ __device__ void
atomic_store(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::value_type desired) noexcept
{
    obj->store(desired);
}

template <typename T, typename Allocator>
__host__ __device__ void
atomic_store_explicit(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::value_type desired,
                      const memory_order order) noexcept
{
    obj->store(desired, order);
}

template <typename T, typename Allocator>
__device__ T
atomic_exchange(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::value_type desired) noexcept
{
    return obj->exchange(desired);
}

template <typename T, typename Allocator>
__device__ T
atomic_exchange_explicit(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::value_type desired,
                         const memory_order order) noexcept
{
    return obj->exchange(desired, order);
}

template <typename T, typename Allocator>
__device__ bool
atomic_compare_exchange_weak(atomic<T, Allocator>* obj, typename atomic<T, Allocator>::value_type* expected,
                             const typename atomic<T, Allocator>::value_type desired) noexcept
{
    return obj->compare_exchange_weak(*expected, desired);
}

template <typename T, typename Allocator>
__device__ bool
atomic_compare_exchange_strong(atomic<T, Allocator>* obj, typename atomic<T, Allocator>::value_type* expected,
                               const typename atomic<T, Allocator>::value_type desired) noexcept
{
    return obj->compare_exchange_strong(*expected, desired);
}

#define STDGPU_AMD_ATOMIC_NONMEMBER(NAME, ARG)                                                                       \
    template <typename T, typename Allocator>                                                                        \
    __device__ T atomic_##NAME(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::ARG arg) noexcept     \
    {                                                                                                                \
        return obj->NAME(arg);                                                                                       \
    }                                                                                                                \
    template <typename T, typename Allocator>                                                                        \
    __device__ T atomic_##NAME##_explicit(atomic<T, Allocator>* obj, const typename atomic<T, Allocator>::ARG arg,   \
                                          const memory_order order) noexcept                                         \
    {                                                                                                                \
        return obj->NAME(arg, order);                                                                                \
    }

STDGPU_AMD_ATOMIC_NONMEMBER(fetch_add, difference_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_sub, difference_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_min, value_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_max, value_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_and, value_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_or, value_type)
STDGPU_AMD_ATOMIC_NONMEMBER(fetch_xor, value_type)

#undef STDGPU_AMD_ATOMIC_NONMEMBER

} // namespace stdgpu

#endif // STDGPU_ATOMIC_CUH
