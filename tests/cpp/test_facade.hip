// Source-compatibility tests of the stdgpu::unordered_map / unordered_set facade, written as the
// reference's own test cases are (tests/stdgpu/unordered_datastructure.inc): user kernels call the
// native device API (insert / emplace / find / contains / count / erase / bucket_size) on a table
// passed by value; host code uses createDeviceObject, bulk insert / erase, device_range, size,
// clear and valid.  Prints one line per case; exits non-zero if any case fails.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include <stdgpu/atomic.cuh>
#include <stdgpu/unordered_map.cuh>
#include <stdgpu/unordered_set.cuh>

using map_t = stdgpu::unordered_map<std::uint64_t, std::uint32_t>;
using set_t = stdgpu::unordered_set<std::uint32_t>;
using imap_t = stdgpu::unordered_map<int, int>;
using iset_t = stdgpu::unordered_set<int>;

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

#define HIP_OK(x) CHECK((x) == hipSuccess)

template <typename T>
T*
dev_alloc(std::size_t n)
{
    T* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T) + 1) != hipSuccess) return nullptr;
    return p;
}

template <typename T>
std::vector<T>
to_host(const T* d, std::size_t n)
{
    std::vector<T> h(n);
    if (n) (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
    return h;
}

template <typename T>
T*
to_dev(const std::vector<T>& h)
{
    T* d = dev_alloc<T>(h.size());
    if (!h.empty()) (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return d;
}

static unsigned
blocks(std::size_t n)
{
    return static_cast<unsigned>((n + 255) / 256);
}

// ------------------------------------------------------------------ kernels (user code)
template <typename Table, typename Key>
__global__ void
k_insert_keys(Table t, const Key* keys, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (std::is_same_v<typename Table::value_type, Key>)
        ok[i] = t.insert(keys[i]).second ? 1 : 0;
    else
        ok[i] = t.insert(typename Table::value_type(keys[i], static_cast<typename Table::mapped_type>(keys[i] * 3 + 1)))
                                .second
                        ? 1
                        : 0;
}

template <typename Table, typename Key>
__global__ void
k_emplace_keys(Table t, const Key* keys, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ok[i] = t.emplace(keys[i], static_cast<typename Table::mapped_type>(keys[i] * 3 + 1)).second ? 1 : 0;
}

template <typename Table, typename Key>
__global__ void
k_erase_keys(Table t, const Key* keys, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ok[i] = static_cast<int>(t.erase(keys[i]));
}

template <typename Table, typename Key>
__global__ void
k_find_keys(Table t, const Key* keys, std::size_t n, int* found, std::uint64_t* val)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto it = t.find(keys[i]);
    found[i] = (it != t.end()) ? 1 : 0;
    if (t.contains(keys[i]) != (it != t.end())) found[i] = 99;
    if (t.count(keys[i]) != found[i]) found[i] = 98;
    if constexpr (!std::is_same_v<typename Table::value_type, Key>)
    {
        if (it != t.end())
        {
            val[i] = static_cast<std::uint64_t>(it->second);
            if (it->first != keys[i]) found[i] = 97;
        }
    }
}

template <typename Table>
__global__ void
k_bucket_sizes(Table t, int* sizes)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= t.bucket_count()) return;
    sizes[n] = t.bucket_size(n);
}

template <typename Table, typename Key>
__global__ void
k_insert_then_erase(Table t, const Key* keys, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool a = t.insert(typename Table::value_type(keys[i], 1)).second;
    const bool b = t.erase(keys[i]) == 1;
    ok[i] = (a && b) ? 1 : 0;
}

template <typename Range, typename T>
__global__ void
k_range_copy(Range r, T* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.size()) return;
    new (&out[i]) T(*(r.begin() + i));
}

template <typename Table>
__global__ void
k_device_size(Table t, int* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = t.size();
}

// the reference's transparent hasher (tests/stdgpu/unordered_map.inc:21-30): stdgpu::hash of the
// type it is given
struct int_hash
{
    using is_transparent = void;
    template <typename T>
    __host__ __device__ std::size_t operator()(const T& key) const
    {
        return stdgpu::hash<T>{}(key);
    }
};

// transparent lookups against the regular ones (unordered_datastructure.inc:392-456, 509-640,
// 2204-2300): out[i] = 1 found, 0 not found, 9 when the transparent and regular answers differ
template <typename Table, typename KeyLike>
__global__ void
k_transparent(Table t, const KeyLike* q, std::size_t n, int* out)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Table& ct = t;
    const KeyLike k = q[i];
    using key_type = typename Table::key_type;
    const key_type kk = static_cast<key_type>(k);
    const bool equal = stdgpu::equal_to<>{}(kk, k);
    const auto regular = equal ? t.find(kk) : t.end();
    const auto it = t.find(k);
    const auto cit = ct.find(k);
    const bool c = ct.contains(k);
    const auto cnt = ct.count(k);
    const bool agree = it == regular && cit == regular && c == (regular != t.end()) && cnt == (c ? 1 : 0);
    out[i] = agree ? (c ? 1 : 0) : 9;
}

template <typename Table, typename KeyLike>
std::vector<int>
run_transparent(Table& t, const std::vector<KeyLike>& q)
{
    KeyLike* dq = to_dev(q);
    int* out = dev_alloc<int>(q.size());
    hipLaunchKernelGGL((k_transparent<Table, KeyLike>), dim3(blocks(q.size())), dim3(256), 0, 0, t, dq, q.size(), out);
    std::vector<int> h = to_host(out, q.size());
    (void)hipFree(dq);
    (void)hipFree(out);
    return h;
}

// ------------------------------------------------------------------ helpers
template <typename Table, typename Key>
std::vector<int>
run_insert(Table& t, const std::vector<Key>& keys, bool emplace = false)
{
    Key* dk = to_dev(keys);
    int* ok = dev_alloc<int>(keys.size());
    if constexpr (!std::is_same_v<typename Table::value_type, Key>)
    {
        if (emplace)
        {
            hipLaunchKernelGGL((k_emplace_keys<Table, Key>), dim3(blocks(keys.size())), dim3(256), 0, 0, t, dk,
                               keys.size(), ok);
        }
        else
        {
            hipLaunchKernelGGL((k_insert_keys<Table, Key>), dim3(blocks(keys.size())), dim3(256), 0, 0, t, dk,
                               keys.size(), ok);
        }
    }
    else
    {
        hipLaunchKernelGGL((k_insert_keys<Table, Key>), dim3(blocks(keys.size())), dim3(256), 0, 0, t, dk, keys.size(),
                           ok);
    }
    std::vector<int> h = to_host(ok, keys.size());
    (void)hipFree(dk);
    (void)hipFree(ok);
    return h;
}

template <typename Table, typename Key>
std::vector<int>
run_erase(Table& t, const std::vector<Key>& keys)
{
    Key* dk = to_dev(keys);
    int* ok = dev_alloc<int>(keys.size());
    hipLaunchKernelGGL((k_erase_keys<Table, Key>), dim3(blocks(keys.size())), dim3(256), 0, 0, t, dk, keys.size(), ok);
    std::vector<int> h = to_host(ok, keys.size());
    (void)hipFree(dk);
    (void)hipFree(ok);
    return h;
}

template <typename Table, typename Key>
std::vector<int>
run_find(Table& t, const std::vector<Key>& keys, std::vector<std::uint64_t>* vals = nullptr)
{
    Key* dk = to_dev(keys);
    int* f = dev_alloc<int>(keys.size());
    std::uint64_t* v = dev_alloc<std::uint64_t>(keys.size());
    hipLaunchKernelGGL((k_find_keys<Table, Key>), dim3(blocks(keys.size())), dim3(256), 0, 0, t, dk, keys.size(), f, v);
    std::vector<int> h = to_host(f, keys.size());
    if (vals) *vals = to_host(v, keys.size());
    (void)hipFree(dk);
    (void)hipFree(f);
    (void)hipFree(v);
    return h;
}

static std::vector<std::uint64_t>
unique_keys(std::size_t n, std::uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::set<std::uint64_t> s;
    std::vector<std::uint64_t> out;
    while (out.size() < n)
    {
        const std::uint64_t k = rng() >> 8;
        if (s.insert(k).second) out.push_back(k);
    }
    return out;
}

static int
sum(const std::vector<int>& v)
{
    return std::accumulate(v.begin(), v.end(), 0);
}

// ------------------------------------------------------------------ cases
static void
empty_container()
{
    g_case = "empty_container";
    map_t m;
    CHECK(m.empty());
    CHECK(m.full());
    CHECK(m.size() == 0);
    CHECK(m.valid());
}

static void
size_limits_and_bucket()
{
    g_case = "size_limits_and_bucket";
    map_t m = map_t::createDeviceObject(100000);
    CHECK(m.bucket_count() == 131072);
    CHECK(m.max_size() == 184444);
    CHECK(m.bucket(1) == 81006 && m.bucket(2) == 30941 && m.bucket(3) == 111948);
    CHECK(m.bucket(0xFFFFFFFEULL) == 34231 && m.bucket(0x0123456789ABCDEFULL) == 28855);
    CHECK(m.max_load_factor() == 1.0F && m.load_factor() == 0.0F);
    map_t::destroyDeviceObject(m);
    map_t t = map_t::createDeviceObject(1);
    CHECK(t.bucket_count() == 1 && t.max_size() == 2);
    map_t::destroyDeviceObject(t);
}

static void
insert_find_erase_single()
{
    g_case = "insert_find_erase_single";
    map_t m = map_t::createDeviceObject(100000);
    const std::vector<std::uint64_t> k{ 7, 42 };
    CHECK(sum(run_insert(m, k)) == 2);
    CHECK(m.size() == 2 && m.valid());
    std::vector<std::uint64_t> vals;
    CHECK(run_find(m, std::vector<std::uint64_t>{ 7, 42, 9 }, &vals) == (std::vector<int>{ 1, 1, 0 }));
    CHECK(vals[0] == 22 && vals[1] == 127);
    // insert_double: a second insert of the same key reports false and keeps the value
    CHECK(sum(run_insert(m, std::vector<std::uint64_t>{ 7 })) == 0);
    CHECK(m.size() == 2);
    // erase_double
    CHECK(sum(run_erase(m, std::vector<std::uint64_t>{ 7 })) == 1);
    CHECK(sum(run_erase(m, std::vector<std::uint64_t>{ 7 })) == 0);
    CHECK(m.size() == 1 && m.valid());
    CHECK(run_find(m, std::vector<std::uint64_t>{ 7, 42 }) == (std::vector<int>{ 0, 1 }));
    map_t::destroyDeviceObject(m);
}

static void
insert_erase_multiple_same_key()
{
    g_case = "insert_erase_multiple_same_key";
    map_t m = map_t::createDeviceObject(100000);
    const std::size_t N = 100000;
    CHECK(sum(run_insert(m, std::vector<std::uint64_t>(N, 1234))) == 1);
    CHECK(m.size() == 1 && m.valid());
    CHECK(sum(run_erase(m, std::vector<std::uint64_t>(N, 1234))) == 1);
    CHECK(m.size() == 0 && m.valid());
    map_t::destroyDeviceObject(m);
}

static void
tiny_table_cases()
{
    g_case = "tiny_table_cases";
    map_t m = map_t::createDeviceObject(1);
    CHECK(sum(run_insert(m, std::vector<std::uint64_t>{ 1 })) == 1);
    CHECK(!m.full() && m.size() == 1);
    // insert_parallel_while_one_free: 100000 random keys, exactly one wins
    std::vector<std::uint64_t> r = unique_keys(100000, 77);
    CHECK(sum(run_insert(m, r)) == 1);
    CHECK(m.full() && m.size() == 2 && m.valid());
    // insert_while_full / insert_multiple_while_full
    CHECK(sum(run_insert(m, std::vector<std::uint64_t>(100000, 999999))) == 0);
    CHECK(m.size() == 2);
    map_t::destroyDeviceObject(m);
    // emplace_parallel_while_one_free
    map_t e = map_t::createDeviceObject(1);
    CHECK(sum(run_insert(e, std::vector<std::uint64_t>{ 1 })) == 1);
    CHECK(sum(run_insert(e, r, true)) == 1);
    CHECK(e.size() == 2 && e.full());
    map_t::destroyDeviceObject(e);
}

static void
unique_parallel_native()
{
    g_case = "unique_parallel_native";
    const std::size_t N = 100000;
    map_t m = map_t::createDeviceObject(static_cast<stdgpu::index_t>(N));
    const std::vector<std::uint64_t> k = unique_keys(N, 5);
    CHECK(sum(run_insert(m, k, true)) == static_cast<int>(N));
    CHECK(m.size() == static_cast<stdgpu::index_t>(N) && m.valid());
    std::vector<std::uint64_t> vals;
    const std::vector<int> f = run_find(m, k, &vals);
    CHECK(sum(f) == static_cast<int>(N));
    for (std::size_t i = 0; i < N; ++i) CHECK(vals[i] == static_cast<std::uint32_t>(k[i] * 3 + 1));
    // count_sum / bucket_size_sum (unordered_datastructure.inc:2229-2300)
    int* sizes = dev_alloc<int>(m.bucket_count());
    hipLaunchKernelGGL((k_bucket_sizes<map_t>), dim3(blocks(m.bucket_count())), dim3(256), 0, 0, m, sizes);
    const std::vector<int> bs = to_host(sizes, m.bucket_count());
    (void)hipFree(sizes);
    CHECK(sum(bs) == static_cast<int>(N));
    int* ds = dev_alloc<int>(1);
    hipLaunchKernelGGL((k_device_size<map_t>), dim3(1), dim3(64), 0, 0, m, ds);
    CHECK(to_host(ds, 1)[0] == static_cast<int>(N));
    (void)hipFree(ds);
    CHECK(sum(run_erase(m, k)) == static_cast<int>(N));
    CHECK(m.size() == 0 && m.valid());
    map_t::destroyDeviceObject(m);
}

static void
range_insert_erase_and_device_range()
{
    g_case = "range_insert_erase_and_device_range";
    const std::size_t N = 100000;
    map_t m = map_t::createDeviceObject(static_cast<stdgpu::index_t>(N));
    const std::vector<std::uint64_t> k = unique_keys(N, 9);
    std::vector<map_t::value_type> host_values;
    for (auto x : k) host_values.emplace_back(x, static_cast<std::uint32_t>(x ^ 0x5555));
    map_t::value_type* dv = dev_alloc<map_t::value_type>(N);
    HIP_OK(hipMemcpy(dv, host_values.data(), N * sizeof(map_t::value_type), hipMemcpyHostToDevice));
    m.insert(stdgpu::make_device(dv), stdgpu::make_device(dv + N));
    CHECK(m.size() == static_cast<stdgpu::index_t>(N) && m.valid());
    auto range = m.device_range();
    CHECK(range.size() == static_cast<stdgpu::index_t>(N));
    map_t::value_type* out = dev_alloc<map_t::value_type>(N);
    hipLaunchKernelGGL((k_range_copy<decltype(range), map_t::value_type>), dim3(blocks(N)), dim3(256), 0, 0, range, out);
    std::vector<std::pair<std::uint64_t, std::uint32_t>> got, want;
    std::vector<map_t::value_type> hv(N);
    HIP_OK(hipMemcpy(hv.data(), out, N * sizeof(map_t::value_type), hipMemcpyDeviceToHost));
    for (auto& p : hv) got.emplace_back(p.first, p.second);
    for (auto& p : host_values) want.emplace_back(p.first, p.second);
    std::sort(got.begin(), got.end());
    std::sort(want.begin(), want.end());
    CHECK(got == want);
    // erase(begin, end) of the first half through a custom policy with its own stream
    hipStream_t s;
    HIP_OK(hipStreamCreate(&s));
    stdgpu::execution::device_policy pol{ s };
    std::uint64_t* dk = to_dev(k);
    m.erase(pol, stdgpu::make_device(dk), stdgpu::make_device(dk + N / 2));
    CHECK(m.size(pol) == static_cast<stdgpu::index_t>(N - N / 2));
    const std::vector<int> f = run_find(m, k);
    for (std::size_t i = 0; i < N; ++i) CHECK(f[i] == (i < N / 2 ? 0 : 1));
    CHECK(m.valid(pol));
    m.clear(pol);
    CHECK(m.size() == 0 && m.empty() && m.valid());
    CHECK(m.device_range().size() == 0);
    HIP_OK(hipStreamDestroy(s));
    (void)hipFree(dk);
    (void)hipFree(dv);
    (void)hipFree(out);
    map_t::destroyDeviceObject(m);
}

static void
insert_and_erase_unique_parallel()
{
    g_case = "insert_and_erase_unique_parallel";
    const std::size_t N = 100000;
    map_t m = map_t::createDeviceObject(static_cast<stdgpu::index_t>(N));
    const std::vector<std::uint64_t> k = unique_keys(N, 13);
    std::uint64_t* dk = to_dev(k);
    int* ok = dev_alloc<int>(N);
    hipLaunchKernelGGL((k_insert_then_erase<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, dk, N, ok);
    CHECK(sum(to_host(ok, N)) == static_cast<int>(N));
    CHECK(m.size() == 0 && m.valid());
    (void)hipFree(dk);
    (void)hipFree(ok);
    map_t::destroyDeviceObject(m);
}

static void
set_u32_native_and_bulk()
{
    g_case = "set_u32_native_and_bulk";
    const std::size_t N = 50000;
    set_t s = set_t::createDeviceObject(static_cast<stdgpu::index_t>(2 * N));
    std::vector<std::uint32_t> k;
    for (auto x : unique_keys(N, 21)) k.push_back(static_cast<std::uint32_t>(x));
    std::sort(k.begin(), k.end());
    k.erase(std::unique(k.begin(), k.end()), k.end());
    CHECK(sum(run_insert(s, k)) == static_cast<int>(k.size()));
    CHECK(sum(run_insert(s, k)) == 0);
    CHECK(s.size() == static_cast<stdgpu::index_t>(k.size()) && s.valid());
    std::uint32_t* dk = to_dev(k);
    s.erase(dk, dk + k.size());
    CHECK(s.size() == 0);
    s.insert(dk, dk + k.size());
    CHECK(s.size() == static_cast<stdgpu::index_t>(k.size()));
    CHECK(sum(run_find(s, k)) == static_cast<int>(k.size()));
    (void)hipFree(dk);
    set_t::destroyDeviceObject(s);
}

// insert_multiple_collision / erase_multiple_collision_{head,tail}_first
// (unordered_datastructure.inc:1295-1373) on this layout: keys sharing the last home bucket, so their
// probe sequence runs over the table's end; native inserts and erases, valid() after every step
// (valid() recomputes each bucket's overflow count)
static std::vector<std::uint64_t>
keys_homed_at(std::uint64_t bucket, std::uint64_t nbuckets, std::size_t count, std::uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::vector<std::uint64_t> out;
    std::set<std::uint64_t> seen;
    while (out.size() < count)
    {
        const std::uint64_t k = rng() >> 4;
        if (htab_dev::home_bucket(k, nbuckets) == bucket && seen.insert(k).second) out.push_back(k);
    }
    return out;
}

static void
collision_chain_native()
{
    g_case = "collision_chain_native";
    const stdgpu::index_t cap = 4096; // 512 buckets of 8 pairs
    const std::uint64_t nb = 512;
    for (int order = 0; order < 2; ++order)
    {
        map_t m = map_t::createDeviceObject(cap);
        const std::vector<std::uint64_t> chain = keys_homed_at(nb - 1, nb, 40, 31 + order);
        const std::vector<std::uint64_t> other = unique_keys(1000, 41);
        CHECK(sum(run_insert(m, other)) == 1000);
        CHECK(sum(run_insert(m, chain, true)) == 40);
        CHECK(m.size() == 1040 && m.valid());
        std::vector<std::uint64_t> vals;
        CHECK(sum(run_find(m, chain, &vals)) == 40);
        for (std::size_t i = 0; i < chain.size(); ++i) CHECK(vals[i] == static_cast<std::uint32_t>(chain[i] * 3 + 1));
        std::vector<std::uint64_t> seq = chain;
        if (order == 1) std::reverse(seq.begin(), seq.end());
        for (std::size_t part = 0; part < 4; ++part)
        {
            const std::vector<std::uint64_t> go(seq.begin() + part * 10, seq.begin() + (part + 1) * 10);
            CHECK(sum(run_erase(m, go)) == 10);
            CHECK(m.valid());
            const std::vector<int> f = run_find(m, seq);
            for (std::size_t i = 0; i < seq.size(); ++i) CHECK(f[i] == (i < (part + 1) * 10 ? 0 : 1));
        }
        CHECK(m.size() == 1000);
        CHECK(sum(run_insert(m, chain)) == 40);
        CHECK(m.size() == 1040 && m.valid());
        map_t::destroyDeviceObject(m);
    }
}

// clear_empty, get_allocator, hash_objects, custom_execution_policy
// (unordered_datastructure.inc:2450-2560, 267-290)
static void
accessors_and_policies()
{
    g_case = "accessors_and_policies";
    hipStream_t st;
    HIP_OK(hipStreamCreate(&st));
    stdgpu::execution::device_policy pol{ st };
    map_t m = map_t::createDeviceObject(pol, 10000);
    CHECK(m.size(pol) == 0 && m.valid(pol));
    m.clear(pol); // clear_empty
    CHECK(m.size(pol) == 0 && m.empty() && m.valid(pol));
    map_t::allocator_type a = m.get_allocator();
    map_t::value_type* arr = a.allocate(10000);
    CHECK(arr != nullptr);
    HIP_OK(hipMemset(arr, 0, 10000 * sizeof(map_t::value_type)));
    a.deallocate(arr, 10000);
    const map_t::hasher h = m.hash_function();
    CHECK(h(std::uint64_t{ 42 }) == 42); // the reference's identity hash on integers
    const map_t::key_equal eq = m.key_eq();
    CHECK(eq(std::uint64_t{ 7 }, std::uint64_t{ 7 }) && !eq(std::uint64_t{ 7 }, std::uint64_t{ 8 }));
    map_t::destroyDeviceObject(pol, m);
    set_t s = set_t::createDeviceObject(pol, 100);
    set_t::allocator_type sa = s.get_allocator();
    std::uint32_t* sarr = sa.allocate(100);
    sa.deallocate(sarr, 100);
    set_t::destroyDeviceObject(pol, s);
    HIP_OK(hipStreamDestroy(st));
}

// examples/openmp/unordered_set.cpp and unordered_map.cpp: duplicate-free neighbour sets checked
// against closed-form sums
template <typename Range, typename T>
__global__ void
k_sum_range(Range r, long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.size()) return;
    const auto& v = *(r.begin() + i);
    if constexpr (std::is_same_v<std::remove_cv_t<std::remove_reference_t<decltype(v)>>, int>)
        atomicAdd(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(v));
    else
    {
        atomicAdd(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(v.first));
        atomicAdd(reinterpret_cast<unsigned long long*>(out + 1), static_cast<unsigned long long>(v.second));
    }
}

template <typename Table>
__global__ void
k_insert_neighbors(Table t, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int num = 2 * i + 1;
    for (int x : { num - 1, num, num + 1 })
    {
        if constexpr (std::is_same_v<typename Table::value_type, int>)
            t.insert(x);
        else
            t.emplace(x, x * x);
    }
}

static void
examples_closed_form()
{
    g_case = "examples_closed_form";
    const int n = 100;
    iset_t s = iset_t::createDeviceObject(n);
    hipLaunchKernelGGL((k_insert_neighbors<iset_t>), dim3(1), dim3(64), 0, 0, s, n / 2);
    CHECK(s.size() == n + 1);
    long long* acc = dev_alloc<long long>(2);
    HIP_OK(hipMemset(acc, 0, 2 * sizeof(long long)));
    auto rs = s.device_range();
    hipLaunchKernelGGL((k_sum_range<decltype(rs), int>), dim3(1), dim3(128), 0, 0, rs, acc);
    CHECK(to_host(acc, 1)[0] == n * (n + 1) / 2);
    iset_t::destroyDeviceObject(s);

    imap_t m = imap_t::createDeviceObject(n);
    hipLaunchKernelGGL((k_insert_neighbors<imap_t>), dim3(1), dim3(64), 0, 0, m, n / 2);
    CHECK(m.size() == n + 1);
    HIP_OK(hipMemset(acc, 0, 2 * sizeof(long long)));
    auto rm = m.device_range();
    hipLaunchKernelGGL((k_sum_range<decltype(rm), int>), dim3(1), dim3(128), 0, 0, rm, acc);
    const std::vector<long long> h = to_host(acc, 2);
    CHECK(h[0] == n * (n + 1) / 2);
    CHECK(h[1] == static_cast<long long>(n) * (n + 1) * (2 * n + 1) / 6);
    (void)hipFree(acc);
    imap_t::destroyDeviceObject(m);
}

// An 8-byte voxel-block key: the reference tests' vec3int16 (tests/stdgpu/unordered_map.inc:53-71)
// with an explicit fourth member, so that the key has no padding bits and its value is its bits
struct alignas(8) block_key
{
    std::int16_t x = 0;
    std::int16_t y = 0;
    std::int16_t z = 0;
    std::int16_t w = 0;
    block_key() = default;
    __host__ __device__ block_key(std::int16_t a, std::int16_t b, std::int16_t c) : x(a), y(b), z(c) {}
};

__host__ __device__ inline bool
operator==(const block_key& a, const block_key& b)
{
    return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}

static bool
operator<(const block_key& a, const block_key& b)
{
    return std::make_tuple(a.x, a.y, a.z, a.w) < std::make_tuple(b.x, b.y, b.z, b.w);
}

// the reference tests' vec_hash (tests/stdgpu/unordered_map.inc:181-206)
struct block_hash
{
    __host__ __device__ std::size_t operator()(const block_key& k) const
    {
        return (static_cast<std::size_t>(k.x) * static_cast<std::size_t>(73856093U)) ^
               (static_cast<std::size_t>(k.y) * static_cast<std::size_t>(19349669U)) ^
               (static_cast<std::size_t>(k.z) * static_cast<std::size_t>(83492791U));
    }
};

using block_set_t = stdgpu::unordered_set<block_key, block_hash, stdgpu::equal_to<>>;
using block_map_t = stdgpu::unordered_map<block_key, std::uint32_t, block_hash, stdgpu::equal_to<>>;

template <typename Table>
__global__ void
k_block_bucket(Table t, const block_key* keys, std::size_t n, int* out)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = t.bucket(keys[i]);
}

__global__ void
k_block_map_emplace(block_map_t m, const block_key* keys, std::size_t n, int* ins)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ins[i] = m.emplace(keys[i], static_cast<std::uint32_t>(i)).second ? 1 : 0;
}

// a later launch than the emplaces: a find racing an insert of its key in the same kernel may see the
// key before its value
__global__ void
k_block_map_find(block_map_t m, const block_key* keys, std::size_t n, std::uint32_t* vals)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto it = m.find(keys[i]);
    vals[i] = (it != m.end() && it->first == keys[i]) ? it->second : 0xFFFFFFFFU;
}

static void
struct_keys()
{
    g_case = "8-byte struct keys (voxel blocks) with the reference tests' vec_hash";
    block_set_t s = block_set_t::createDeviceObject(100000); // the reference fixture's capacity
    // the pinned pairs of insert_collision / insert_no_collision (unordered_datastructure.inc:899-934)
    const std::vector<block_key> pinned = { { -7, -3, 15 }, { 7, 3, 15 }, { -5, -15, 13 }, { 5, 15, 13 } };
    {
        block_key* dp = to_dev(pinned);
        int* db = dev_alloc<int>(pinned.size());
        hipLaunchKernelGGL((k_block_bucket<block_set_t>), dim3(1), dim3(64), 0, 0, s, dp, pinned.size(), db);
        const std::vector<int> b = to_host(db, pinned.size());
        CHECK(b[0] == b[1] && b[2] == b[3] && b[0] != b[2]);
        CHECK(b[0] == s.bucket(pinned[0])); // host and device agree
        (void)hipFree(dp);
        (void)hipFree(db);
    }
    // random blocks with repeats: native inserts, finds, erases against std::set
    std::mt19937 rng(5);
    std::uniform_int_distribution<int> d(-40, 40), dz(-5, 5); // 72,171 blocks: draws repeat often
    std::vector<block_key> keys;
    for (int i = 0; i < 60000; ++i)
        keys.emplace_back(static_cast<std::int16_t>(d(rng)), static_cast<std::int16_t>(d(rng)),
                          static_cast<std::int16_t>(dz(rng)));
    std::set<block_key> want(keys.begin(), keys.end());
    CHECK(sum(run_insert(s, keys)) == static_cast<int>(want.size()));
    CHECK(s.size() == static_cast<stdgpu::index_t>(want.size()) && s.valid());
    CHECK(sum(run_find(s, keys)) == static_cast<int>(keys.size()));
    std::vector<block_key> absent;
    for (int i = 0; i < 2000; ++i) absent.emplace_back(static_cast<std::int16_t>(1000 + i), 0, 0);
    CHECK(sum(run_find(s, absent)) == 0);
    const std::vector<block_key> uniq(want.begin(), want.end());
    const std::vector<block_key> half(uniq.begin(), uniq.begin() + uniq.size() / 2);
    CHECK(sum(run_erase(s, half)) == static_cast<int>(half.size()));
    for (const auto& k : half) want.erase(k);
    // device_range holds exactly the remaining blocks
    auto range = s.device_range();
    CHECK(range.size() == static_cast<stdgpu::index_t>(want.size()));
    block_key* out = dev_alloc<block_key>(want.size());
    hipLaunchKernelGGL((k_range_copy<decltype(range), block_key>), dim3(blocks(want.size())), dim3(256), 0, 0, range,
                       out);
    std::vector<block_key> got = to_host(out, want.size());
    std::sort(got.begin(), got.end());
    CHECK(got == std::vector<block_key>(want.begin(), want.end()));
    (void)hipFree(out);
    // bulk erase(begin, end) of the rest through the agnostic API
    const std::vector<block_key> rest(want.begin(), want.end());
    block_key* dr = to_dev(rest);
    s.erase(stdgpu::make_device(dr), stdgpu::make_device(dr + rest.size()));
    CHECK(s.size() == 0 && s.valid());
    (void)hipFree(dr);
    block_set_t::destroyDeviceObject(s);

    // a map from blocks to slot numbers: emplace, find through the returned iterator
    block_map_t m = block_map_t::createDeviceObject(100000);
    block_key* dk = to_dev(keys);
    int* dins = dev_alloc<int>(keys.size());
    std::uint32_t* dval = dev_alloc<std::uint32_t>(keys.size());
    hipLaunchKernelGGL(k_block_map_emplace, dim3(blocks(keys.size())), dim3(256), 0, 0, m, dk, keys.size(), dins);
    hipLaunchKernelGGL(k_block_map_find, dim3(blocks(keys.size())), dim3(256), 0, 0, m, dk, keys.size(), dval);
    const std::vector<int> ins = to_host(dins, keys.size());
    const std::vector<std::uint32_t> val = to_host(dval, keys.size());
    const std::set<block_key> all(keys.begin(), keys.end());
    CHECK(sum(ins) == static_cast<int>(all.size()) && m.size() == static_cast<stdgpu::index_t>(all.size()));
    // every occurrence of a block reads the index of the one emplace that inserted it
    std::map<block_key, std::uint32_t> winner;
    for (std::size_t i = 0; i < keys.size(); ++i)
        if (ins[i]) winner[keys[i]] = static_cast<std::uint32_t>(i);
    for (std::size_t i = 0; i < keys.size(); ++i) CHECK(val[i] == winner[keys[i]]);
    CHECK(m.valid());
    (void)hipFree(dk);
    (void)hipFree(dins);
    (void)hipFree(dval);
    block_map_t::destroyDeviceObject(m);
}

static void
bulk_find_and_contains()
{
    g_case = "bulk find / contains over key ranges";
    const std::size_t N = 50000;
    map_t m = map_t::createDeviceObject(static_cast<stdgpu::index_t>(N));
    const std::vector<std::uint64_t> k = unique_keys(2 * N, 21);
    const std::vector<std::uint64_t> present(k.begin(), k.begin() + N);
    CHECK(sum(run_insert(m, present)) == static_cast<int>(N));
    // queries: the N present keys, N absent ones, then both reserved key values (absent)
    std::vector<std::uint64_t> q = k;
    q.push_back(~0ULL);
    q.push_back(~0ULL - 1);
    const std::size_t Q = q.size();
    std::uint64_t* dq = to_dev(q);
    std::uint8_t* dh = dev_alloc<std::uint8_t>(Q);
    std::uint32_t* dv = dev_alloc<std::uint32_t>(Q);
    HIP_OK(hipMemset(dv, 0xFF, Q * sizeof(std::uint32_t)));
    m.find(stdgpu::make_device(dq), stdgpu::make_device(dq + Q), dh, dv);
    const std::vector<std::uint8_t> hits = to_host(dh, Q);
    const std::vector<std::uint32_t> vals = to_host(dv, Q);
    for (std::size_t i = 0; i < Q; ++i)
    {
        const bool in = i < N;
        CHECK(hits[i] == (in ? 1 : 0));
        CHECK(vals[i] == (in ? static_cast<std::uint32_t>(q[i] * 3 + 1) : 0xFFFFFFFFU)); // misses untouched
    }
    // the native per-thread find agrees
    const std::vector<int> native = run_find(m, q);
    for (std::size_t i = 0; i < Q; ++i) CHECK(native[i] == hits[i]);
    // contains through a policy with its own stream, after erasing the first half of the present keys
    hipStream_t st;
    HIP_OK(hipStreamCreate(&st));
    stdgpu::execution::device_policy pol{ st };
    m.erase(pol, stdgpu::make_device(dq), stdgpu::make_device(dq + N / 2));
    m.contains(pol, stdgpu::make_device(dq), stdgpu::make_device(dq + Q), dh);
    const std::vector<std::uint8_t> h2 = to_host(dh, Q);
    for (std::size_t i = 0; i < Q; ++i) CHECK(h2[i] == ((i >= N / 2 && i < N) ? 1 : 0));
    HIP_OK(hipStreamDestroy(st));
    // a set answers contains the same way
    set_t s = set_t::createDeviceObject(1000);
    const std::vector<std::uint32_t> sk = { 1U, 5U, 9U, 0xFFFFFFFFU };
    CHECK(sum(run_insert(s, sk)) == 4);
    const std::vector<std::uint32_t> sq = { 0U, 1U, 5U, 6U, 9U, 0xFFFFFFFFU, 0xFFFFFFFEU };
    std::uint32_t* dsq = to_dev(sq);
    s.contains(stdgpu::make_device(dsq), stdgpu::make_device(dsq + sq.size()), dh);
    CHECK(to_host(dh, sq.size()) == (std::vector<std::uint8_t>{ 0, 1, 1, 0, 1, 1, 0 }));
    set_t::destroyDeviceObject(s);
    (void)hipFree(dsq);
    (void)hipFree(dq);
    (void)hipFree(dh);
    (void)hipFree(dv);
    map_t::destroyDeviceObject(m);
}

// a key matches a KeyLike exactly when key == keylike under C++'s usual arithmetic conversions
template <typename Key, typename KeyLike>
static std::vector<int>
expected_matches(const std::vector<Key>& keys, const std::vector<KeyLike>& q)
{
    std::vector<int> out;
    for (const KeyLike& x : q)
    {
        bool hit = false;
        for (const Key& k : keys) hit = hit || stdgpu::equal_to<>{}(k, x);
        out.push_back(hit ? 1 : 0);
    }
    return out;
}

static void
transparent_keylike_lookups()
{
    g_case = "transparent KeyLike find / contains / count";
    using tmap = stdgpu::unordered_map<int, float, int_hash, stdgpu::equal_to<>>;
    tmap m = tmap::createDeviceObject(2000);
    std::vector<int> keys;
    for (int k = -500; k < 500; k += 3) keys.push_back(k);
    CHECK(sum(run_insert(m, keys)) == static_cast<int>(keys.size()));
    std::vector<short> qs;
    for (int k = -600; k < 600; ++k) qs.push_back(static_cast<short>(k));
    const std::vector<int> got_s = run_transparent(m, qs);
    CHECK(got_s == expected_matches(keys, qs));
    CHECK(sum(got_s) == static_cast<int>(keys.size())); // the count_sum case: every key is found once
    // a wider KeyLike whose value does not fit Key matches nothing, even where its low bits would
    std::vector<long long> ql = { 1, -2, (1LL << 32) + 1, -(1LL << 32) - 2, (1LL << 40) - 500, 4, -497 };
    const std::vector<int> got_l = run_transparent(m, ql);
    CHECK(got_l == expected_matches(keys, ql));
    CHECK(got_l[2] == 0 && got_l[3] == 0);
    tmap::destroyDeviceObject(m);

    // unsigned keys: int -1 equals 4294967295u (converted to unsigned), long long -1 does not
    using tset = stdgpu::unordered_set<std::uint32_t, int_hash, stdgpu::equal_to<>>;
    tset s = tset::createDeviceObject(100);
    const std::vector<std::uint32_t> sk = { 0U, 7U, 4294967295U, 4294967294U };
    CHECK(sum(run_insert(s, sk)) == 4);
    const std::vector<int> qi = { -1, -2, 0, 7, 8, -7 };
    const std::vector<int> got_i = run_transparent(s, qi);
    CHECK(got_i == expected_matches(sk, qi));
    CHECK(got_i[0] == 1 && got_i[1] == 1);
    const std::vector<long long> qll = { -1, 4294967295LL, 7, (1LL << 32) + 7 };
    const std::vector<int> got_ll = run_transparent(s, qll);
    CHECK(got_ll == expected_matches(sk, qll));
    CHECK(got_ll[0] == 0 && got_ll[1] == 1 && got_ll[3] == 0);
    tset::destroyDeviceObject(s);
}

// The reference fixture's remaining cases with its own key (three int16 coordinates) and hash
// (vec_hash): hash_inside_range, bucket_number_hits, bucket_number_collisions, empty_size_limits
// (unordered_datastructure.inc:60-256), insert_while_excess_empty (:1443-1465), the const-range
// variants insert_const_range / erase_const_range / const_device_range (:1959-2041, 2376-2410) and
// custom_allocator (:2518-2546).
__global__ void
k_count_bucket_hits(block_set_t s, const block_key* keys, std::size_t n, int* hits)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) stdgpu::atomic_ref<int>(hits[s.bucket(keys[i])]).fetch_add(1);
}

// a stateful allocator that counts what the table asks of it
template <typename T>
struct counting_allocator
{
    using value_type = T;
    static inline int allocations = 0;
    static inline int deallocations = 0;
    int id = 0;
    counting_allocator() = default;
    explicit counting_allocator(int i) : id(i) {}
    template <typename U>
    counting_allocator(const counting_allocator<U>& o) : id(o.id)
    {
    }
    T* allocate(std::int64_t n)
    {
        ++allocations;
        T* p = nullptr;
        return hipMalloc(&p, static_cast<std::size_t>(n) * sizeof(T)) == hipSuccess ? p : nullptr;
    }
    void deallocate(T* p, std::int64_t /*n*/)
    {
        ++deallocations;
        (void)hipFree(p);
    }
};

static std::vector<int>
bucket_hits(block_set_t& s, std::size_t n, std::mt19937& rng)
{
    std::uniform_int_distribution<int> d16(-32768, 32767);
    std::vector<block_key> keys;
    for (std::size_t i = 0; i < n; ++i)
        keys.emplace_back(static_cast<std::int16_t>(d16(rng)), static_cast<std::int16_t>(d16(rng)),
                          static_cast<std::int16_t>(d16(rng)));
    block_key* dk = to_dev(keys);
    std::vector<int> zero(static_cast<std::size_t>(s.bucket_count()), 0);
    int* dh = to_dev(zero);
    hipLaunchKernelGGL(k_count_bucket_hits, dim3(blocks(n)), dim3(256), 0, 0, s, dk, n, dh);
    std::vector<int> h = to_host(dh, zero.size());
    (void)hipFree(dk);
    (void)hipFree(dh);
    return h;
}

static void
reference_fixture_cases()
{
    g_case = "hash distribution, size limits, const ranges, struct-key tiny table, custom allocator";
    block_set_t s = block_set_t::createDeviceObject(100000);
    const stdgpu::index_t bc = s.bucket_count();
    // empty_size_limits
    CHECK(s.empty() && !s.full() && s.size() <= s.max_size() && s.load_factor() <= s.max_load_factor());
    // hash_inside_range: 2^21 random keys on the host
    std::mt19937 rng(17);
    std::uniform_int_distribution<int> d16(-32768, 32767);
    for (int i = 0; i < (1 << 21); ++i)
    {
        const stdgpu::index_t b = s.bucket(block_key(static_cast<std::int16_t>(d16(rng)), static_cast<std::int16_t>(d16(rng)),
                                                     static_cast<std::int16_t>(d16(rng))));
        CHECK(b >= 0 && b < bc);
    }
    // bucket_number_hits: 2 * bucket_count keys hit more than 80% of the buckets
    {
        const std::vector<int> h = bucket_hits(s, 2 * static_cast<std::size_t>(bc), rng);
        CHECK(sum(h) == 2 * bc);
        const auto hit = std::count_if(h.begin(), h.end(), [](int c) { return c > 0; });
        CHECK(hit > static_cast<long>(static_cast<float>(bc) * 80.0F / 100.0F));
    }
    // bucket_number_collisions: bucket_count keys leave fewer than 40% of the buckets with two or more
    {
        const std::vector<int> h = bucket_hits(s, static_cast<std::size_t>(bc), rng);
        CHECK(sum(h) == bc);
        const auto coll = std::count_if(h.begin(), h.end(), [](int c) { return c > 1; });
        CHECK(coll < static_cast<long>(static_cast<float>(bc) * 40.0F / 100.0F));
    }
    block_set_t::destroyDeviceObject(s);

    // insert_while_excess_empty: the one-capacity table takes two positions, not a third
    block_set_t tiny = block_set_t::createDeviceObject(1);
    const std::vector<block_key> p12 = { { 1, 2, 3 }, { -1, 2, 3 } };
    CHECK(sum(run_insert(tiny, p12)) == 2);
    CHECK(tiny.valid() && tiny.size() == 2);
    CHECK(sum(run_insert(tiny, std::vector<block_key>{ { 1, -2, 3 } })) == 0);
    CHECK(tiny.valid() && tiny.size() == 2);
    block_set_t::destroyDeviceObject(tiny);

    // insert_const_range / const_device_range / erase_const_range
    const std::size_t N = 100000;
    map_t m = map_t::createDeviceObject(static_cast<stdgpu::index_t>(N));
    const std::vector<std::uint64_t> k = unique_keys(N, 23);
    std::vector<map_t::value_type> hv;
    for (auto x : k) hv.emplace_back(x, static_cast<std::uint32_t>(x * 5));
    map_t::value_type* dv = dev_alloc<map_t::value_type>(N);
    HIP_OK(hipMemcpy(dv, hv.data(), N * sizeof(map_t::value_type), hipMemcpyHostToDevice));
    const map_t::value_type* cdv = dv;
    m.insert(stdgpu::make_device(cdv), stdgpu::make_device(cdv + N));
    CHECK(!m.empty() && m.size() == static_cast<stdgpu::index_t>(N) && m.valid());
    auto crange = static_cast<const map_t&>(m).device_range();
    CHECK(crange.size() == static_cast<stdgpu::index_t>(N));
    map_t::value_type* out = dev_alloc<map_t::value_type>(N);
    hipLaunchKernelGGL((k_range_copy<decltype(crange), map_t::value_type>), dim3(blocks(N)), dim3(256), 0, 0, crange, out);
    std::vector<map_t::value_type> got(N);
    HIP_OK(hipMemcpy(got.data(), out, N * sizeof(map_t::value_type), hipMemcpyDeviceToHost));
    std::vector<std::uint64_t> gk, wk(k);
    for (auto& p : got)
    {
        gk.push_back(p.first);
        CHECK(p.second == static_cast<std::uint32_t>(p.first * 5));
    }
    std::sort(gk.begin(), gk.end());
    std::sort(wk.begin(), wk.end());
    CHECK(gk == wk);
    std::uint64_t* dk = to_dev(k);
    const std::uint64_t* cdk = dk;
    m.erase(stdgpu::make_device(cdk), stdgpu::make_device(cdk + N));
    CHECK(m.empty() && m.size() == 0 && m.valid());
    (void)hipFree(dk);
    (void)hipFree(dv);
    (void)hipFree(out);
    map_t::destroyDeviceObject(m);

    // custom_allocator: the table keeps the allocator it was given, and get_allocator() returns it
    using calloc_t = counting_allocator<map_t::value_type>;
    using cmap_t = stdgpu::unordered_map<std::uint64_t, std::uint32_t, stdgpu::hash<std::uint64_t>,
                                         stdgpu::equal_to<std::uint64_t>, calloc_t>;
    calloc_t::allocations = 0;
    calloc_t::deallocations = 0;
    cmap_t c = cmap_t::createDeviceObject(10000, calloc_t(42));
    calloc_t a = c.get_allocator();
    CHECK(a.id == 42);
    map_t::value_type* arr = a.allocate(10000);
    CHECK(arr != nullptr);
    a.deallocate(arr, 10000);
    CHECK(calloc_t::allocations == 1 && calloc_t::deallocations == 1);
    CHECK(sum(run_insert(c, std::vector<std::uint64_t>{ 5, 6, 5 })) == 2 && c.size() == 2 && c.valid());
    cmap_t::destroyDeviceObject(c);
}

// Rates of the native per-thread API from a user kernel (one call per thread, no cooperation):
// emplace, find of present and of absent keys, erase; 2^22 keys at load 0.8.  Also find of keys that
// the bulk insert placed (its claim order starts at a key-dependent slot of the bucket).
template <typename Table, typename Key>
__global__ void
k_find_flag(Table t, const Key* keys, std::size_t n, int* found)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) found[i] = t.find(keys[i]) != t.end() ? 1 : 0;
}

static void
native_rates()
{
    g_case = "native per-thread emplace / find / erase rates, 2^22 keys at load 0.8";
    const std::size_t N = 1 << 22;
    const stdgpu::index_t cap = static_cast<stdgpu::index_t>(std::ceil(static_cast<double>(N) / 0.8));
    // distinct keys: a bijective 64-bit mix (splitmix64's finaliser) of 0 .. 2N-1
    std::vector<std::uint64_t> all(2 * N);
    for (std::size_t i = 0; i < 2 * N; ++i)
    {
        std::uint64_t z = i + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        all[i] = z ^ (z >> 31);
    }
    const std::vector<std::uint64_t> k(all.begin(), all.begin() + N), absent(all.begin() + N, all.end());
    std::uint64_t* dk = to_dev(k);
    std::uint64_t* da = to_dev(absent);
    int* ok = dev_alloc<int>(N);
    hipEvent_t e[5];
    for (auto& x : e) HIP_OK(hipEventCreate(&x));
    float ms[4] = { 0, 0, 0, 0 };
    map_t m = map_t::createDeviceObject(cap);
    for (int rep = 0; rep < 3; ++rep) // the last repetition is reported
    {
        m.clear();
        HIP_OK(hipEventRecord(e[0], nullptr));
        hipLaunchKernelGGL((k_emplace_keys<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, dk, N, ok);
        HIP_OK(hipEventRecord(e[1], nullptr));
        hipLaunchKernelGGL((k_find_flag<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, dk, N, ok);
        HIP_OK(hipEventRecord(e[2], nullptr));
        hipLaunchKernelGGL((k_find_flag<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, da, N, ok);
        HIP_OK(hipEventRecord(e[3], nullptr));
        hipLaunchKernelGGL((k_erase_keys<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, dk, N, ok);
        HIP_OK(hipEventRecord(e[4], nullptr));
        HIP_OK(hipEventSynchronize(e[4]));
        for (int j = 0; j < 4; ++j) HIP_OK(hipEventElapsedTime(&ms[j], e[j], e[j + 1]));
    }
    CHECK(sum(to_host(ok, N)) == static_cast<int>(N) && m.size() == 0 && m.valid());
    // keys placed by the bulk insert, then found per thread
    std::vector<map_t::value_type> hv;
    for (auto x : k) hv.emplace_back(x, static_cast<std::uint32_t>(x));
    map_t::value_type* dv = dev_alloc<map_t::value_type>(N);
    HIP_OK(hipMemcpy(dv, hv.data(), N * sizeof(map_t::value_type), hipMemcpyHostToDevice));
    m.clear();
    m.insert(stdgpu::make_device(dv), stdgpu::make_device(dv + N));
    float bulk_ms = 0;
    for (int rep = 0; rep < 3; ++rep)
    {
        HIP_OK(hipEventRecord(e[0], nullptr));
        hipLaunchKernelGGL((k_find_flag<map_t, std::uint64_t>), dim3(blocks(N)), dim3(256), 0, 0, m, dk, N, ok);
        HIP_OK(hipEventRecord(e[1], nullptr));
        HIP_OK(hipEventSynchronize(e[1]));
        HIP_OK(hipEventElapsedTime(&bulk_ms, e[0], e[1]));
    }
    CHECK(sum(to_host(ok, N)) == static_cast<int>(N));
    const double g = static_cast<double>(N) / 1e6;
    std::printf("rate: native emplace %.2f, find (present) %.2f, find (absent) %.2f, erase %.2f, find of bulk-inserted "
                "keys %.2f G/s (%zu keys, capacity %d)\n",
                g / ms[0], g / ms[1], g / ms[2], g / ms[3], g / bulk_ms, N, cap);
    for (auto& x : e) (void)hipEventDestroy(x);
    (void)hipFree(dk);
    (void)hipFree(da);
    (void)hipFree(ok);
    (void)hipFree(dv);
    map_t::destroyDeviceObject(m);
}

// hash of the integer types (functional.cpp:120-211): the identity widened to size_t, signed types
// sign-extended (functional_detail.h:60-80), so every value of a small type hashes apart and a million
// random values of a wide type keep more than 90% distinct hashes, as the reference tests require;
// equal_to and its transparent form (functional.cpp:457-506)
template <typename T>
static bool
hash_small_type_ok()
{
    stdgpu::hash<T> h;
    std::set<std::size_t> seen;
    for (long long v = static_cast<long long>(std::numeric_limits<T>::lowest());
         v <= static_cast<long long>(std::numeric_limits<T>::max()); ++v)
    {
        const T x = static_cast<T>(v);
        if (h(x) != static_cast<std::size_t>(x)) return false;
        seen.insert(h(x));
    }
    return seen.size() == static_cast<std::size_t>(static_cast<long long>(std::numeric_limits<T>::max()) -
                                                   static_cast<long long>(std::numeric_limits<T>::lowest()) + 1);
}

template <typename T>
static bool
hash_wide_type_ok()
{
    stdgpu::hash<T> h;
    std::mt19937_64 rng(sizeof(T) * 7 + std::is_signed_v<T>);
    std::uniform_int_distribution<T> dist(std::numeric_limits<T>::lowest(), std::numeric_limits<T>::max());
    std::set<std::size_t> seen;
    const int N = 1000000;
    for (int i = 0; i < N; ++i)
    {
        const T x = dist(rng);
        if (h(x) != static_cast<std::size_t>(x)) return false;
        seen.insert(h(x));
    }
    if (h(static_cast<T>(-1)) != static_cast<std::size_t>(static_cast<T>(-1))) return false; // sign extension
    return static_cast<int>(seen.size()) > N * 90 / 100;
}

static void
hash_and_equal_to_of_integers()
{
    g_case = "hash of every integer type and equal_to (functional.cpp:120-211, 457-506)";
    CHECK(hash_small_type_ok<bool>() && hash_small_type_ok<char>() && hash_small_type_ok<signed char>());
    CHECK(hash_small_type_ok<unsigned char>() && hash_small_type_ok<short>() && hash_small_type_ok<unsigned short>());
    CHECK(hash_wide_type_ok<int>() && hash_wide_type_ok<unsigned int>() && hash_wide_type_ok<long>());
    CHECK(hash_wide_type_ok<unsigned long>() && hash_wide_type_ok<long long>() &&
          hash_wide_type_ok<unsigned long long>());
    CHECK(stdgpu::hash<int>{}(-5) == ~std::size_t{ 4 }); // -5 sign-extended
    stdgpu::equal_to<int> eq;
    stdgpu::equal_to<> teq;
    CHECK(eq(3, 3) && !eq(3, 4) && teq(3, 3L) && !teq(3U, 4ULL));
}

// ------------------------------------------------------------------ native churn, bulk / native mixing
template <typename Table, typename Key>
__global__ void
k_mixed_emplace_erase(Table t, const Key* ins, const Key* del, std::size_t n, int* ok)
{
    // lane i of an even-numbered pair emplaces ins[i / 2], of an odd one erases del[i / 2], in one kernel
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n) return;
    if (i % 2 == 0)
        ok[i] = t.emplace(ins[i / 2], static_cast<typename Table::mapped_type>(ins[i / 2] * 3 + 1)).second ? 1 : 0;
    else
        ok[i] = static_cast<int>(t.erase(del[i / 2]));
}

static std::vector<std::uint64_t>
mixed_keys(std::size_t n, std::uint64_t salt)
{
    // distinct keys: a bijective 64-bit mix of salt * 2^32 + i
    std::vector<std::uint64_t> out(n);
    for (std::size_t i = 0; i < n; ++i)
    {
        std::uint64_t z = (salt << 32) + i + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        out[i] = (z ^ (z >> 31)) >> 2; // clear of the reserved key values
    }
    return out;
}

// the table equals the model: size, valid(), every model key found with its value, erased keys absent
static bool
same_as_model(map_t& m, const std::map<std::uint64_t, std::uint32_t>& model, const std::vector<std::uint64_t>& gone)
{
    if (m.size() != static_cast<stdgpu::index_t>(model.size()) || !m.valid()) return false;
    std::vector<std::uint64_t> keys;
    for (const auto& kv : model) keys.push_back(kv.first);
    std::vector<std::uint64_t> vals;
    const std::vector<int> f = run_find(m, keys, &vals);
    std::size_t i = 0;
    for (const auto& kv : model)
    {
        if (f[i] != 1 || vals[i] != kv.second) return false;
        ++i;
    }
    if (!gone.empty() && sum(run_find(m, gone)) != 0) return false;
    return true;
}

static void
native_churn_reuses_freed_slots()
{
    g_case = "native emplace / erase churn of 4x the physical slots without clear()";
    const stdgpu::index_t cap = 1 << 16;
    map_t m = map_t::createDeviceObject(cap);
    htab_info info{};
    CHECK(htab_get_info(m.native_handle(), &info) == HTAB_OK);
    const std::size_t slots = static_cast<std::size_t>(info.physical_slots);
    const std::size_t live = slots / 2, batch = slots / 8;
    const std::vector<std::uint64_t> pool = mixed_keys(live + 4 * slots + batch, 11);
    std::map<std::uint64_t, std::uint32_t> model;
    std::size_t next = 0, oldest = 0;
    auto value_of = [](std::uint64_t k) { return static_cast<std::uint32_t>(k * 3 + 1); };
    // fill to half the slots
    {
        std::vector<std::uint64_t> first(pool.begin(), pool.begin() + live);
        CHECK(sum(run_insert(m, first, true)) == static_cast<int>(live));
        for (auto k : first) model[k] = value_of(k);
        next = live;
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    std::vector<float> round_ms;
    std::size_t inserted_total = 0;
    int round = 0;
    while (inserted_total < 4 * slots)
    {
        // one kernel: emplace `batch` fresh keys in even lanes, erase the `batch` oldest in odd lanes
        std::vector<std::uint64_t> ins(pool.begin() + next, pool.begin() + next + batch);
        std::vector<std::uint64_t> del(pool.begin() + oldest, pool.begin() + oldest + batch);
        std::uint64_t* di = to_dev(ins);
        std::uint64_t* dd = to_dev(del);
        int* ok = dev_alloc<int>(2 * batch);
        HIP_OK(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL((k_mixed_emplace_erase<map_t, std::uint64_t>), dim3(blocks(2 * batch)), dim3(256), 0, 0, m, di,
                           dd, batch, ok);
        HIP_OK(hipEventRecord(e1, nullptr));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        round_ms.push_back(ms);
        const std::vector<int> f = to_host(ok, 2 * batch);
        (void)hipFree(di);
        (void)hipFree(dd);
        (void)hipFree(ok);
        CHECK(sum(f) == static_cast<int>(2 * batch)); // every fresh key inserted, every old key erased
        for (auto k : ins) model[k] = value_of(k);
        for (auto k : del) model.erase(k);
        next += batch;
        oldest += batch;
        inserted_total += batch;
        // present keys are refused, absent ones are not erased
        std::vector<std::uint64_t> again(pool.begin() + oldest, pool.begin() + oldest + batch / 8);
        CHECK(sum(run_insert(m, again, true)) == 0);
        std::vector<std::uint64_t> gone(pool.begin() + oldest - batch, pool.begin() + oldest - batch + batch / 8);
        CHECK(sum(run_erase(m, gone)) == 0);
        if (round % 4 == 0 || inserted_total >= 4 * slots) CHECK(same_as_model(m, model, gone));
        ++round;
    }
    CHECK(same_as_model(m, model, std::vector<std::uint64_t>(pool.begin(), pool.begin() + oldest)));
    // flat: the last quarter of the rounds is not slower than twice the first quarter
    const std::size_t q = round_ms.size() / 4;
    double a = 0, b = 0;
    for (std::size_t i = 1; i <= q; ++i) a += round_ms[i];
    for (std::size_t i = round_ms.size() - q; i < round_ms.size(); ++i) b += round_ms[i];
    std::printf("rate: churn rounds %zu (%zu inserts into %zu slots), first quarter %.3f ms, last quarter %.3f ms per round\n",
                round_ms.size(), inserted_total, slots, a / q, b / q);
    CHECK(b <= 2.0 * a + 0.05 * q);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    map_t::destroyDeviceObject(m);
}

static void
native_insert_after_bulk_erase()
{
    g_case = "native insert after a bulk erase (duplicate-free, every kind of freed slot)";
    const std::size_t N = 1 << 16;
    const stdgpu::index_t cap = static_cast<stdgpu::index_t>(std::ceil(static_cast<double>(N) / 0.8));
    for (int lanes : { 1, 4, 8 })
    {
        htab_tuning tun{};
        tun.bucket_lanes = lanes;
        tun.block_size = 64;
        tun.slots_per_capacity = 1.0F;
        htab* h = nullptr;
        CHECK(htab_create(HTAB_MAP_U64_U32, cap, &tun, &h) == HTAB_OK);
        const std::vector<std::uint64_t> k = mixed_keys(N, 23);
        std::vector<map_t::value_type> hv;
        for (auto x : k) hv.emplace_back(x, static_cast<std::uint32_t>(x * 3 + 1));
        map_t::value_type* dv = dev_alloc<map_t::value_type>(N);
        HIP_OK(hipMemcpy(dv, hv.data(), N * sizeof(map_t::value_type), hipMemcpyHostToDevice));
        CHECK(htab_insert(h, dv, static_cast<std::int64_t>(N), nullptr, nullptr) == HTAB_OK);
        std::vector<std::uint64_t> half;
        for (std::size_t i = 0; i < N; i += 2) half.push_back(k[i]);
        std::uint64_t* dh = to_dev(half);
        CHECK(htab_erase(h, dh, static_cast<std::int64_t>(half.size()), nullptr, nullptr) == HTAB_OK);
        HIP_OK(hipDeviceSynchronize());
        (void)hipFree(dh);
        (void)hipFree(dv);
        map_t m = map_t::adopt(h);
        // every key once per thread, plus each key by three more threads of other wavefronts
        std::vector<std::uint64_t> all = k;
        for (int r = 0; r < 3; ++r) all.insert(all.end(), k.begin(), k.end());
        const std::vector<int> ok = run_insert(m, all, true);
        std::map<std::uint64_t, int> wins;
        for (std::size_t i = 0; i < all.size(); ++i) wins[all[i]] += ok[i];
        for (std::size_t i = 0; i < N; ++i) CHECK(wins[k[i]] == (i % 2 == 0 ? 1 : 0));
        std::map<std::uint64_t, std::uint32_t> model;
        for (auto x : k) model[x] = static_cast<std::uint32_t>(x * 3 + 1);
        CHECK(same_as_model(m, model, {}));
        map_t::destroyDeviceObject(m);
    }
}

static void
bucket_size_is_the_reference_count()
{
    g_case = "bucket_size(n) equals the number of stored keys with bucket(k) == n";
    // the reference fixture's table (capacity 100000) holding 10^4 keys, the two reserved values included
    map_t m = map_t::createDeviceObject(100000);
    std::vector<std::uint64_t> k = unique_keys(10000 - 2, 41);
    k.push_back(~std::uint64_t(0));
    k.push_back(~std::uint64_t(0) - 1);
    CHECK(sum(run_insert(m, k, true)) == 10000);
    std::vector<int> want(m.bucket_count(), 0);
    for (auto x : k) ++want[m.bucket(x)];
    int* sizes = dev_alloc<int>(m.bucket_count());
    hipLaunchKernelGGL((k_bucket_sizes<map_t>), dim3(blocks(m.bucket_count())), dim3(256), 0, 0, m, sizes);
    const std::vector<int> got = to_host(sizes, m.bucket_count());
    (void)hipFree(sizes);
    CHECK(got == want);
    // after erasing a third of the keys
    std::vector<std::uint64_t> gone;
    for (std::size_t i = 0; i < k.size(); i += 3) gone.push_back(k[i]);
    CHECK(sum(run_erase(m, gone)) == static_cast<int>(gone.size()));
    for (auto x : gone) --want[m.bucket(x)];
    sizes = dev_alloc<int>(m.bucket_count());
    hipLaunchKernelGGL((k_bucket_sizes<map_t>), dim3(blocks(m.bucket_count())), dim3(256), 0, 0, m, sizes);
    CHECK(to_host(sizes, m.bucket_count()) == want);
    (void)hipFree(sizes);
    map_t::destroyDeviceObject(m);
    // signed keys (sign-extended identity hash) in a set
    iset_t s = iset_t::createDeviceObject(20000);
    std::vector<int> ik;
    for (int i = -5000; i < 5000; ++i) ik.push_back(i * 7919);
    CHECK(sum(run_insert(s, ik)) == 10000);
    std::vector<int> iw(s.bucket_count(), 0);
    for (auto x : ik) ++iw[s.bucket(x)];
    sizes = dev_alloc<int>(s.bucket_count());
    hipLaunchKernelGGL((k_bucket_sizes<iset_t>), dim3(blocks(s.bucket_count())), dim3(256), 0, 0, s, sizes);
    CHECK(to_host(sizes, s.bucket_count()) == iw);
    (void)hipFree(sizes);
    iset_t::destroyDeviceObject(s);
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
    void (*cases[])() = { empty_container,
                          size_limits_and_bucket,
                          insert_find_erase_single,
                          insert_erase_multiple_same_key,
                          tiny_table_cases,
                          unique_parallel_native,
                          range_insert_erase_and_device_range,
                          insert_and_erase_unique_parallel,
                          set_u32_native_and_bulk,
                          collision_chain_native,
                          accessors_and_policies,
                          examples_closed_form,
                          transparent_keylike_lookups,
                          struct_keys,
                          bulk_find_and_contains,
                          reference_fixture_cases,
                          native_rates,
                          hash_and_equal_to_of_integers,
                          native_churn_reuses_freed_slots,
                          native_insert_after_bulk_erase,
                          bucket_size_is_the_reference_count };
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
        if (g_failures == before) std::printf("PASS %s\n", g_case);
        std::fflush(stdout);
    }
    std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
    return g_failures ? 1 : 0;
}
