// Source-compatibility tests of stdgpu::atomic / atomic_ref, written as the reference's own test cases
// are (tests/stdgpu/atomic.inc): a functor holding the atomic by value runs once per index in a
// device launch and calls a member, non-member or non-member explicit function, or an operator; the
// host checks the final value against the closed form.  Prints one line per case; exits non-zero if
// any case fails.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <limits>
#include <vector>

#include <stdgpu/atomic.cuh>

static int g_failures = 0;
static const char* g_case = "";

#define CHECK(cond)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            std::printf("FAIL %s: %s (line %d)\n", g_case, #cond, __LINE__);                                          \
            ++g_failures;                                                                                              \
            return;                                                                                                    \
        }                                                                                                              \
    } while (0)

template <typename F>
__global__ void
k_each(stdgpu::index_t n, F f)
{
    const stdgpu::index_t i = static_cast<stdgpu::index_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) f(i);
}

// for_each_index(execution::device, n, f) of the reference (algorithm.h), for these tests
template <typename F>
void
for_each_index(stdgpu::index_t n, F f)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_each<F>, dim3((n + 255) / 256), dim3(256), 0, nullptr, n, f);
    (void)hipDeviceSynchronize();
}

// the sequence 1..N in device memory (iota from 1, atomic.inc:850-866)
template <typename T>
T*
iota_from_one(stdgpu::index_t n)
{
    std::vector<T> h(static_cast<std::size_t>(n));
    for (stdgpu::index_t i = 0; i < n; ++i) h[static_cast<std::size_t>(i)] = static_cast<T>(i + 1);
    T* d = nullptr;
    (void)hipMalloc(&d, h.size() * sizeof(T));
    (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return d;
}

enum class form
{
    member,
    nonmember,
    nonmember_explicit,
    op
};

template <typename T, form F>
struct add_seq
{
    stdgpu::atomic<T> value;
    T* seq;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        if constexpr (F == form::member) v.fetch_add(seq[i]);
        else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_add(&v, seq[i]);
        else if constexpr (F == form::nonmember_explicit)
            stdgpu::atomic_fetch_add_explicit(&v, seq[i], stdgpu::memory_order_relaxed);
        else v += seq[i];
    }
};

template <typename T, form F>
struct sub_seq
{
    stdgpu::atomic<T> value;
    T* seq;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        if constexpr (F == form::member) v.fetch_sub(seq[i]);
        else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_sub(&v, seq[i]);
        else if constexpr (F == form::nonmember_explicit)
            stdgpu::atomic_fetch_sub_explicit(&v, seq[i], stdgpu::memory_order_relaxed);
        else v -= seq[i];
    }
};

template <typename T, bool Max, form F>
struct minmax_seq
{
    stdgpu::atomic<T> value;
    T* seq;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        if constexpr (Max)
        {
            if constexpr (F == form::member) v.fetch_max(seq[i]);
            else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_max(&v, seq[i]);
            else stdgpu::atomic_fetch_max_explicit(&v, seq[i], stdgpu::memory_order_relaxed);
        }
        else
        {
            if constexpr (F == form::member) v.fetch_min(seq[i]);
            else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_min(&v, seq[i]);
            else stdgpu::atomic_fetch_min_explicit(&v, seq[i], stdgpu::memory_order_relaxed);
        }
    }
};

// one bit per index: or / xor set it, and clears it (atomic.inc:1479-1495, 1669-1697, 1871-1887)
template <typename T, int Which, form F>
struct bits_seq
{
    stdgpu::atomic<T> value;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        const T pattern = static_cast<T>(std::make_unsigned_t<T>(1) << i);
        if constexpr (Which == 0)
        {
            if constexpr (F == form::member) v.fetch_or(pattern);
            else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_or(&v, pattern);
            else if constexpr (F == form::nonmember_explicit)
                stdgpu::atomic_fetch_or_explicit(&v, pattern, stdgpu::memory_order_relaxed);
            else v |= pattern;
        }
        else if constexpr (Which == 1)
        {
            if constexpr (F == form::member) v.fetch_xor(pattern);
            else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_xor(&v, pattern);
            else if constexpr (F == form::nonmember_explicit)
                stdgpu::atomic_fetch_xor_explicit(&v, pattern, stdgpu::memory_order_relaxed);
            else v ^= pattern;
        }
        else
        {
            const T clear = static_cast<T>(~pattern);
            if constexpr (F == form::member) v.fetch_and(clear);
            else if constexpr (F == form::nonmember) stdgpu::atomic_fetch_and(&v, clear);
            else if constexpr (F == form::nonmember_explicit)
                stdgpu::atomic_fetch_and_explicit(&v, clear, stdgpu::memory_order_relaxed);
            else v &= clear;
        }
    }
};

// the compare-exchange loop of atomic.inc:584-670
template <typename T, bool Weak, bool Member>
struct cas_add_seq
{
    stdgpu::atomic<T> value;
    T* seq;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        T old = v.load();
        if constexpr (Member)
        {
            if constexpr (Weak)
                while (!v.compare_exchange_weak(old, old + seq[i])) {}
            else
                while (!v.compare_exchange_strong(old, old + seq[i])) {}
        }
        else
        {
            if constexpr (Weak)
                while (!stdgpu::atomic_compare_exchange_weak(&v, &old, old + seq[i])) {}
            else
                while (!stdgpu::atomic_compare_exchange_strong(&v, &old, old + seq[i])) {}
        }
    }
};

template <typename T, int Which>
struct incdec_op
{
    stdgpu::atomic<T> value;
    __device__ void operator()(stdgpu::index_t) const
    {
        stdgpu::atomic<T> v = value;
        if constexpr (Which == 0) ++v;
        else if constexpr (Which == 1) v++;
        else if constexpr (Which == 2) --v;
        else v--;
    }
};

template <typename T, bool Inc>
struct mod_seq
{
    stdgpu::atomic<T> value;
    T modulus;
    __device__ void operator()(stdgpu::index_t) const
    {
        stdgpu::atomic<T> v = value;
        if constexpr (Inc) v.fetch_inc_mod(modulus);
        else v.fetch_dec_mod(modulus);
    }
};

template <typename T>
struct exchange_seq
{
    stdgpu::atomic<T> value;
    unsigned int* seen; // a count per value 0..N: every value written is replaced at most once
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic<T> v = value;
        const T old = (i % 2) ? v.exchange(static_cast<T>(i + 1)) : stdgpu::atomic_exchange(&v, static_cast<T>(i + 1));
        atomicAdd(&seen[static_cast<std::size_t>(old)], 1U);
    }
};

template <typename T>
struct ref_add
{
    T* data;
    __device__ void operator()(stdgpu::index_t i) const
    {
        stdgpu::atomic_ref<T> r(data[i % 7]);
        r.fetch_add(static_cast<T>(1));
    }
};

template <typename T>
void
sequences()
{
    // float sums of integers are exact below 2^24: a shorter sequence for float
    const stdgpu::index_t n = std::is_same_v<T, float> ? 4000 : 40000;
    T* seq = iota_from_one<T>(n);
    const T total = static_cast<T>(static_cast<long long>(n) * (n + 1) / 2);
    stdgpu::atomic<T> value = stdgpu::atomic<T>::createDeviceObject();
    g_case = "atomic sequences";
    CHECK(value.load() == T(0));
    CHECK(value.is_lock_free() == stdgpu::atomic_is_lock_free(&value));

    value.store(T(0));
    for_each_index(n, add_seq<T, form::member>{ value, seq });
    CHECK(value.load() == total);
    value.store(T(0));
    for_each_index(n, add_seq<T, form::nonmember>{ value, seq });
    CHECK(value.load() == total);
    value.store(T(0));
    for_each_index(n, add_seq<T, form::nonmember_explicit>{ value, seq });
    CHECK(value.load() == total);
    value.store(T(0));
    for_each_index(n, add_seq<T, form::op>{ value, seq });
    CHECK(value.load() == total);

    for_each_index(n, sub_seq<T, form::member>{ value, seq });
    CHECK(value.load() == T(0));
    value = total;
    for_each_index(n, sub_seq<T, form::nonmember>{ value, seq });
    CHECK(static_cast<T>(value) == T(0));
    value.store(total);
    for_each_index(n, sub_seq<T, form::nonmember_explicit>{ value, seq });
    CHECK(value.load() == T(0));
    value.store(total);
    for_each_index(n, sub_seq<T, form::op>{ value, seq });
    CHECK(value.load() == T(0));

    value.store(std::numeric_limits<T>::max());
    for_each_index(n, minmax_seq<T, false, form::member>{ value, seq });
    CHECK(value.load() == T(1));
    value.store(std::numeric_limits<T>::max());
    for_each_index(n, minmax_seq<T, false, form::nonmember>{ value, seq });
    CHECK(value.load() == T(1));
    value.store(std::numeric_limits<T>::max());
    for_each_index(n, minmax_seq<T, false, form::nonmember_explicit>{ value, seq });
    CHECK(value.load() == T(1));
    value.store(std::numeric_limits<T>::lowest());
    for_each_index(n, minmax_seq<T, true, form::member>{ value, seq });
    CHECK(value.load() == static_cast<T>(n));
    value.store(std::numeric_limits<T>::lowest());
    for_each_index(n, minmax_seq<T, true, form::nonmember>{ value, seq });
    CHECK(value.load() == static_cast<T>(n));
    value.store(std::numeric_limits<T>::lowest());
    for_each_index(n, minmax_seq<T, true, form::nonmember_explicit>{ value, seq });
    CHECK(value.load() == static_cast<T>(n));

    // compare-exchange loops (a shorter sequence: every round of a contended loop retries)
    const stdgpu::index_t m = 4000;
    const T mtotal = static_cast<T>(static_cast<long long>(m) * (m + 1) / 2);
    value.store(T(0));
    for_each_index(m, cas_add_seq<T, true, true>{ value, seq });
    CHECK(value.load() == mtotal);
    value.store(T(0));
    for_each_index(m, cas_add_seq<T, false, true>{ value, seq });
    CHECK(value.load() == mtotal);
    value.store(T(0));
    for_each_index(m, cas_add_seq<T, true, false>{ value, seq });
    CHECK(value.load() == mtotal);
    value.store(T(0));
    for_each_index(m, cas_add_seq<T, false, false>{ value, seq });
    CHECK(value.load() == mtotal);

    // exchange: the start value and every written value but the last are each replaced exactly once
    unsigned int* seen = nullptr;
    (void)hipMalloc(&seen, (n + 1) * sizeof(unsigned int));
    (void)hipMemset(seen, 0, (n + 1) * sizeof(unsigned int));
    value.store(T(0));
    for_each_index(n, exchange_seq<T>{ value, seen });
    std::vector<unsigned int> hs(static_cast<std::size_t>(n) + 1);
    (void)hipMemcpy(hs.data(), seen, hs.size() * sizeof(unsigned int), hipMemcpyDeviceToHost);
    (void)hipFree(seen);
    const auto last = static_cast<std::size_t>(value.load());
    CHECK(last >= 1 && last <= static_cast<std::size_t>(n));
    for (std::size_t k = 0; k < hs.size(); ++k) CHECK(hs[k] == (k == last ? 0U : 1U));

    if constexpr (std::is_integral_v<T>)
    {
        const stdgpu::index_t nb = std::numeric_limits<T>::digits + std::numeric_limits<T>::is_signed;
        const T ones = static_cast<T>(~static_cast<std::make_unsigned_t<T>>(0));
        value.store(T(0));
        for_each_index(nb, bits_seq<T, 0, form::member>{ value });
        CHECK(value.load() == ones);
        value.store(T(0));
        for_each_index(nb, bits_seq<T, 0, form::nonmember>{ value });
        CHECK(value.load() == ones);
        value.store(T(0));
        for_each_index(nb, bits_seq<T, 0, form::nonmember_explicit>{ value });
        CHECK(value.load() == ones);
        value.store(T(0));
        for_each_index(nb, bits_seq<T, 0, form::op>{ value });
        CHECK(value.load() == ones);
        for_each_index(nb, bits_seq<T, 2, form::member>{ value });
        CHECK(value.load() == T(0));
        value.store(ones);
        for_each_index(nb, bits_seq<T, 2, form::nonmember>{ value });
        CHECK(value.load() == T(0));
        value.store(ones);
        for_each_index(nb, bits_seq<T, 2, form::nonmember_explicit>{ value });
        CHECK(value.load() == T(0));
        value.store(ones);
        for_each_index(nb, bits_seq<T, 2, form::op>{ value });
        CHECK(value.load() == T(0));
        for_each_index(nb, bits_seq<T, 1, form::member>{ value });
        CHECK(value.load() == ones);
        for_each_index(nb, bits_seq<T, 1, form::nonmember>{ value });
        CHECK(value.load() == T(0));
        for_each_index(nb, bits_seq<T, 1, form::nonmember_explicit>{ value });
        CHECK(value.load() == ones);
        for_each_index(nb, bits_seq<T, 1, form::op>{ value });
        CHECK(value.load() == T(0));

        // pre / post increment and decrement (atomic.inc post_inc_operator ...): n of each
        value.store(T(0));
        for_each_index(n, incdec_op<T, 0>{ value });
        for_each_index(n, incdec_op<T, 1>{ value });
        CHECK(value.load() == static_cast<T>(2 * n));
        for_each_index(n, incdec_op<T, 2>{ value });
        for_each_index(n, incdec_op<T, 3>{ value });
        CHECK(value.load() == T(0));
    }

    // atomic_ref over a plain device array
    T* arr = nullptr;
    (void)hipMalloc(&arr, 7 * sizeof(T));
    (void)hipMemset(arr, 0, 7 * sizeof(T));
    for_each_index(7000, ref_add<T>{ arr });
    std::vector<T> ha(7);
    (void)hipMemcpy(ha.data(), arr, 7 * sizeof(T), hipMemcpyDeviceToHost);
    (void)hipFree(arr);
    for (T x : ha) CHECK(x == T(1000));

    stdgpu::atomic<T>::destroyDeviceObject(value);
    (void)hipFree(seq);
}

void
wrapping_increment_decrement()
{
    g_case = "fetch_inc_mod / fetch_dec_mod";
    const stdgpu::index_t n = 50000;
    const unsigned int modulus = static_cast<unsigned int>(n / 10);
    auto value = stdgpu::atomic<unsigned int>::createDeviceObject();
    value.store(42U);
    for_each_index(n, mod_seq<unsigned int, true>{ value, modulus });
    CHECK(value.load() == 42U);
    for_each_index(n, mod_seq<unsigned int, false>{ value, modulus });
    CHECK(value.load() == 42U);
    stdgpu::atomic<unsigned int>::destroyDeviceObject(value);
}

void
empty_and_allocator()
{
    g_case = "empty atomic, allocator, policy";
    stdgpu::atomic<int> empty;
    CHECK(empty.load() == 0); // a default-constructed atomic reads 0 (atomic_detail.cuh:436-441)
    empty.store(5);           // and ignores stores
    CHECK(empty.load() == 0);
    stdgpu::execution::device_policy policy{};
    auto v = stdgpu::atomic<unsigned long long int>::createDeviceObject(policy);
    v.store(policy, (1ULL << 63) + 9);
    CHECK(v.load(policy) == (1ULL << 63) + 9);
    auto alloc = v.get_allocator();
    (void)alloc;
    stdgpu::atomic<unsigned long long int>::destroyDeviceObject(policy, v);
}

int
main()
{
    std::setvbuf(stdout, nullptr, _IOLBF, 0); // a case that never ends still names itself
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    {
        std::printf("FAIL no GPU\n");
        return 2;
    }
    void (*cases[])() = { sequences<int>, sequences<unsigned int>, sequences<unsigned long long int>,
                          sequences<float>, wrapping_increment_decrement, empty_and_allocator };
    const char* names[] = { "int", "unsigned int", "unsigned long long int", "float", "mod", "empty" };
    int k = 0;
    for (auto c : cases)
    {
        const int before = g_failures;
        try
        {
            c();
        }
        catch (const std::exception& e)
        {
            std::printf("FAIL %s: exception %s\n", g_case, e.what());
            ++g_failures;
        }
        std::printf("%s %s [%s]\n", g_failures == before ? "ok  " : "FAIL", g_case, names[k++]);
    }
    std::printf("OK: %d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
