// Device-side atomic operations with the reference's semantics, shared by the bulk atomic kernels
// (stdgpu_amd/csrc/atomic.hip) and the C++ facade's stdgpu::atomic / atomic_ref
// (include/stdgpu/atomic.cuh).  gfx950 only.
//
// Value types: int, unsigned int, unsigned long long int and float, as src/stdgpu/atomic.cuh:137-139
// allows.  Every operation is one device-scope (agent) atomic on global memory; the semantics of each
// follow the reference's OpenMP backend, which states them as plain critical sections
// (src/stdgpu/openmp/impl/atomic_detail.h:40-217):
//   load / store                   the value (:40-55)
//   exchange                       old = *p, *p = arg (:58-68)
//   compare_exchange               old = *p, *p = (old == expected) ? desired : old (:71-81)
//   fetch_add / sub / and / or / xor   old = *p, *p = old op arg (:84-166)
//   fetch_min / fetch_max          old = *p, *p = std::min(old, arg) / std::max(old, arg) (:169-191)
//   fetch_inc_mod(limit)           old = *p, *p = (old >= limit) ? 0 : old + 1 (:195-204)
//   fetch_dec_mod(limit)           old = *p, *p = (old == 0 || old > limit) ? limit : old - 1 (:208-217)
// and the public fetch_inc_mod(arg) / fetch_dec_mod(arg) pass limit = arg - 1
// (src/stdgpu/impl/atomic_detail.cuh:679-702).  Integer min / max / and / or / xor / inc / dec map to
// the hardware's atomic instructions; float min / max are compare-and-swap loops with std::min /
// std::max's comparison, so NaN arguments behave as in the reference, not as IEEE fmin / fmax.
#ifndef STDGPU_AMD_ATOMIC_DEVICE_HPP
#define STDGPU_AMD_ATOMIC_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

namespace htab_dev
{

// the memory orders of std::memory_order, mapped to the compiler's; consume is taken as acquire
enum class mem_order : int
{
    relaxed = 0,
    consume = 1,
    acquire = 2,
    release = 3,
    acq_rel = 4,
    seq_cst = 5
};

template <typename T>
inline constexpr bool is_atomic_value_v = std::is_same_v<T, int> || std::is_same_v<T, unsigned int> ||
                                          std::is_same_v<T, unsigned long long int> || std::is_same_v<T, float>;

#define HTAB_ATOMIC_ORDERED(expr)                                                                                    \
    switch (o)                                                                                                       \
    {                                                                                                                \
    case mem_order::relaxed: {                                                                                       \
        constexpr int ord = __ATOMIC_RELAXED;                                                                        \
        return expr;                                                                                                 \
    }                                                                                                                \
    case mem_order::consume:                                                                                         \
    case mem_order::acquire: {                                                                                       \
        constexpr int ord = __ATOMIC_ACQUIRE;                                                                        \
        return expr;                                                                                                 \
    }                                                                                                                \
    case mem_order::release: {                                                                                       \
        constexpr int ord = __ATOMIC_RELEASE;                                                                        \
        return expr;                                                                                                 \
    }                                                                                                                \
    case mem_order::acq_rel: {                                                                                       \
        constexpr int ord = __ATOMIC_ACQ_REL;                                                                        \
        return expr;                                                                                                 \
    }                                                                                                                \
    default: {                                                                                                       \
        constexpr int ord = __ATOMIC_SEQ_CST;                                                                        \
        return expr;                                                                                                 \
    }                                                                                                                \
    }

template <typename T>
__device__ inline T
atomic_load_op(const T* p, mem_order o = mem_order::seq_cst)
{
    // a load cannot have release semantics: release / acq_rel loads are taken as acquire
    if (o == mem_order::release || o == mem_order::acq_rel) o = mem_order::acquire;
    HTAB_ATOMIC_ORDERED(__hip_atomic_load(p, ord, __HIP_MEMORY_SCOPE_AGENT))
}

template <typename T>
__device__ inline void
atomic_store_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    // a store cannot have acquire semantics: acquire / consume / acq_rel stores are taken as release
    if (o == mem_order::acquire || o == mem_order::consume || o == mem_order::acq_rel) o = mem_order::release;
    HTAB_ATOMIC_ORDERED(__hip_atomic_store(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

template <typename T>
__device__ inline T
atomic_exchange_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__hip_atomic_exchange(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

// the value *p held; *p = desired when that value == expected.  For float the comparison is the
// reference's value comparison (+0 equals -0, NaN equals nothing), done as a loop of bitwise
// compare-and-swaps on the value seen.
template <typename T>
__device__ inline T
atomic_cas_op(T* p, T expected, T desired, mem_order o = mem_order::seq_cst)
{
    if constexpr (std::is_same_v<T, float>)
    {
        uint32_t* q = reinterpret_cast<uint32_t*>(p);
        uint32_t seen = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;)
        {
            if (!(__builtin_bit_cast(float, seen) == expected))
            {
                (void)atomic_load_op(q, o);
                return __builtin_bit_cast(float, seen);
            }
            const uint32_t got = atomic_cas_op(q, seen, __builtin_bit_cast(uint32_t, desired), o);
            if (got == seen) return __builtin_bit_cast(float, seen);
            seen = got;
        }
    }
    else
    {
        T e = expected;
        // the failure order of a compare-exchange may not be release / acq_rel
        switch (o)
        {
        case mem_order::relaxed:
            __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
        case mem_order::consume:
        case mem_order::acquire:
            __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_ACQUIRE, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
        case mem_order::release:
            __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
        case mem_order::acq_rel:
            __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
        default:
            __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        return e;
    }
}

template <typename T>
__device__ inline T
atomic_fetch_add_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_add(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

template <typename T>
__device__ inline T
atomic_fetch_sub_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    if constexpr (std::is_same_v<T, float>)
    {
        HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_add(p, -v, ord, __HIP_MEMORY_SCOPE_AGENT))
    }
    else
    {
        // two's complement wrap for int as for the unsigned types: subtract in the unsigned type
        using U = std::make_unsigned_t<T>;
        HTAB_ATOMIC_ORDERED(static_cast<T>(__hip_atomic_fetch_sub(reinterpret_cast<U*>(p), static_cast<U>(v), ord,
                                                                  __HIP_MEMORY_SCOPE_AGENT)))
    }
}

// float min / max: compare-and-swap loop with std::min(old, v) = (v < old ? v : old) and
// std::max(old, v) = (old < v ? v : old)
template <bool Max>
__device__ inline float
atomic_fetch_minmax_float(float* p, float v, mem_order o)
{
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    uint32_t seen = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;)
    {
        const float old = __builtin_bit_cast(float, seen);
        const float nv = Max ? (old < v ? v : old) : (v < old ? v : old);
        if (__builtin_bit_cast(uint32_t, nv) == seen)
        {
            // no change: the operation reads only; its order is that of a load
            (void)atomic_load_op(q, o);
            return old;
        }
        const uint32_t got = atomic_cas_op(q, seen, __builtin_bit_cast(uint32_t, nv), o);
        if (got == seen) return old;
        seen = got;
    }
}

template <typename T>
__device__ inline T
atomic_fetch_min_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    if constexpr (std::is_same_v<T, float>)
        return atomic_fetch_minmax_float<false>(p, v, o);
    else
    {
        HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_min(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
    }
}

template <typename T>
__device__ inline T
atomic_fetch_max_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    if constexpr (std::is_same_v<T, float>)
        return atomic_fetch_minmax_float<true>(p, v, o);
    else
    {
        HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_max(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
    }
}

template <typename T>
__device__ inline T
atomic_fetch_and_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_and(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

template <typename T>
__device__ inline T
atomic_fetch_or_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_or(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

template <typename T>
__device__ inline T
atomic_fetch_xor_op(T* p, T v, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__hip_atomic_fetch_xor(p, v, ord, __HIP_MEMORY_SCOPE_AGENT))
}

// *p = (old >= limit) ? 0 : old + 1 (the hardware's wrapping increment)
__device__ inline unsigned int
atomic_inc_wrap_op(unsigned int* p, unsigned int limit, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__builtin_amdgcn_atomic_inc32(p, limit, ord, "agent"))
}

// *p = (old == 0 || old > limit) ? limit : old - 1 (the hardware's wrapping decrement)
__device__ inline unsigned int
atomic_dec_wrap_op(unsigned int* p, unsigned int limit, mem_order o = mem_order::seq_cst)
{
    HTAB_ATOMIC_ORDERED(__builtin_amdgcn_atomic_dec32(p, limit, ord, "agent"))
}

#undef HTAB_ATOMIC_ORDERED

} // namespace htab_dev

#endif // STDGPU_AMD_ATOMIC_DEVICE_HPP
