// Source-compatibility tests of stdgpu::vector used from user kernels and host code, written as the
// reference's own cases are (tests/stdgpu/vector.inc): per-thread push_back / emplace_back / pop_back,
// also mixed in one kernel, insert at and erase up to device_end(), clear, element access, the
// capacity limits, element types of other sizes.  Every case checks the elements against what the
// calls reported.  Prints one line per case; exits non-zero if any case fails.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <vector>

#include <stdgpu/vector.cuh>

using ivec = stdgpu::vector<int>;

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

template <typename T>
T*
dev_alloc(std::size_t n)
{
    T* p = nullptr;
    (void)hipMalloc(&p, n * sizeof(T) + 4);
    (void)hipMemset(p, 0, n * sizeof(T) + 4);
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

static int
sum(const std::vector<int>& v)
{
    return std::accumulate(v.begin(), v.end(), 0);
}

__global__ void
k_push(ivec v, const int* values, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ok[i] = v.push_back(values[i]) ? 1 : 0;
}

__global__ void
k_emplace(ivec v, const int* values, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ok[i] = v.emplace_back(values[i]) ? 1 : 0;
}

__global__ void
k_pop(ivec v, std::size_t n, int* out, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const stdgpu::pair<int, bool> p = v.pop_back();
    out[i] = p.first;
    ok[i] = p.second ? 1 : 0;
}

// pushes from the odd lanes only (a diverged branch); the even lanes do nothing
__global__ void
k_push_odd(ivec v, const int* values, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (i & 1)) ok[i] = v.push_back(values[i]) ? 1 : 0;
}

template <typename Range>
__global__ void
k_copy_range(Range r, int* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r.size()) out[i] = r.begin()[i];
}

__global__ void
k_access(ivec v, int* out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const ivec& c = v;
    out[0] = v.front();
    out[1] = v.back();
    out[2] = v.at(3);
    out[3] = v[4];
    out[4] = c.front() + c.back() + c.at(3) + c[4];
    out[5] = static_cast<int>(v.size());
    out[6] = v.empty() ? 1 : 0;
    out[7] = v.full() ? 1 : 0;
}

// the reference's fill_vector (vector.inc:204-226): 1 .. n pushed per thread, then put in order
static void
fill(ivec& v, stdgpu::index_t n)
{
    std::vector<int> h(static_cast<std::size_t>(n));
    std::iota(h.begin(), h.end(), 1);
    int* d = to_dev(h);
    int* ok = dev_alloc<int>(h.size());
    hipLaunchKernelGGL(k_push, dim3(blocks(h.size())), dim3(256), 0, 0, v, d, h.size(), ok);
    std::vector<int> got = to_host(v.data(), h.size());
    std::sort(got.begin(), got.end());
    (void)hipMemcpy(v.data(), got.data(), got.size() * sizeof(int), hipMemcpyHostToDevice);
    (void)hipFree(d);
    (void)hipFree(ok);
}

static std::vector<int>
pop(ivec& v, std::size_t n, std::vector<int>* ok_out = nullptr)
{
    int* out = dev_alloc<int>(n);
    int* ok = dev_alloc<int>(n);
    hipLaunchKernelGGL(k_pop, dim3(blocks(n)), dim3(256), 0, 0, v, n, out, ok);
    std::vector<int> h = to_host(out, n);
    if (ok_out) *ok_out = to_host(ok, n);
    (void)hipFree(out);
    (void)hipFree(ok);
    return h;
}

static std::vector<int>
push(ivec& v, const std::vector<int>& values, bool emplace = false)
{
    int* d = to_dev(values);
    int* ok = dev_alloc<int>(values.size());
    if (emplace)
        hipLaunchKernelGGL(k_emplace, dim3(blocks(values.size())), dim3(256), 0, 0, v, d, values.size(), ok);
    else
        hipLaunchKernelGGL(k_push, dim3(blocks(values.size())), dim3(256), 0, 0, v, d, values.size(), ok);
    std::vector<int> h = to_host(ok, values.size());
    (void)hipFree(d);
    (void)hipFree(ok);
    return h;
}

static std::vector<int>
iota_from(int first, std::size_t n)
{
    std::vector<int> h(n);
    std::iota(h.begin(), h.end(), first);
    return h;
}

// ------------------------------------------------------------------ cases
static void
empty_container()
{
    g_case = "empty_container (vector.inc:63-70)";
    ivec e;
    CHECK(e.empty() && e.full() && e.size() == 0 && e.valid());
}

static void
pop_back_cases()
{
    g_case = "pop_back_some / _all / _too_many (vector.inc:282-345)";
    const stdgpu::index_t N = 10000;
    ivec v = ivec::createDeviceObject(N);
    fill(v, N);
    CHECK(v.size() == N && v.full() && v.valid());
    std::vector<int> ok;
    std::vector<int> got = pop(v, 1000, &ok);
    CHECK(sum(ok) == 1000 && v.size() == N - 1000 && !v.empty() && !v.full() && v.valid());
    std::vector<int> rest = to_host(v.data(), static_cast<std::size_t>(N - 1000));
    for (stdgpu::index_t i = 0; i < N - 1000; ++i) CHECK(rest[static_cast<std::size_t>(i)] == i + 1);
    std::sort(got.begin(), got.end()); // the top 1000 came off
    CHECK(got == iota_from(N - 999, 1000));
    got = pop(v, static_cast<std::size_t>(N - 1000) + 1, &ok); // one more than remain
    CHECK(sum(ok) == N - 1000 && v.size() == 0 && v.empty() && !v.full() && v.valid());
    std::vector<int> popped;
    for (std::size_t i = 0; i < got.size(); ++i)
        if (ok[i]) popped.push_back(got[i]);
    std::sort(popped.begin(), popped.end());
    CHECK(popped == iota_from(1, static_cast<std::size_t>(N - 1000)));
    got = pop(v, 64, &ok); // empty: every call fails
    CHECK(sum(ok) == 0 && v.size() == 0 && v.valid());
    ivec::destroyDeviceObject(v);
}

static void
push_back_cases()
{
    g_case = "push_back / emplace_back _some / _all / _too_many (vector.inc:404-512, 556-664)";
    const stdgpu::index_t N = 10000;
    for (int emplace = 0; emplace < 2; ++emplace)
    {
        ivec v = ivec::createDeviceObject(N);
        fill(v, N);
        (void)pop(v, 1000);
        std::vector<int> ok = push(v, iota_from(N - 999, 1000), emplace != 0);
        CHECK(sum(ok) == 1000 && v.size() == N && v.full() && v.valid());
        std::vector<int> all = to_host(v.data(), static_cast<std::size_t>(N));
        std::sort(all.begin(), all.end());
        CHECK(all == iota_from(1, static_cast<std::size_t>(N)));
        // too many: N + 1 pushes into an emptied vector; exactly N succeed and the stored values are
        // exactly the pushed ones that reported success
        (void)pop(v, static_cast<std::size_t>(N));
        CHECK(v.size() == 0);
        const std::vector<int> vals = iota_from(1, static_cast<std::size_t>(N) + 1);
        ok = push(v, vals, emplace != 0);
        CHECK(sum(ok) == N && v.size() == N && v.full() && v.valid());
        std::vector<int> want;
        for (std::size_t i = 0; i < vals.size(); ++i)
            if (ok[i]) want.push_back(vals[i]);
        all = to_host(v.data(), static_cast<std::size_t>(N));
        std::sort(all.begin(), all.end());
        CHECK(all == want);
        ok = push(v, iota_from(1, 64), emplace != 0); // full: every call fails
        CHECK(sum(ok) == 0 && v.size() == N);
        ivec::destroyDeviceObject(v);
    }
}

static void
diverged_lanes()
{
    g_case = "push_back from the odd lanes only, pop_back of all";
    const std::size_t n = 100000;
    ivec v = ivec::createDeviceObject(static_cast<stdgpu::index_t>(n));
    const std::vector<int> vals = iota_from(1, n);
    int* d = to_dev(vals);
    int* ok = dev_alloc<int>(n);
    hipLaunchKernelGGL(k_push_odd, dim3(blocks(n)), dim3(256), 0, 0, v, d, n, ok);
    CHECK(sum(to_host(ok, n)) == static_cast<int>(n / 2) && v.size() == static_cast<stdgpu::index_t>(n / 2));
    std::vector<int> got = to_host(v.data(), n / 2);
    std::sort(got.begin(), got.end());
    std::vector<int> want;
    for (std::size_t i = 1; i < n; i += 2) want.push_back(vals[i]);
    CHECK(got == want);
    std::vector<int> pok;
    got = pop(v, n, &pok);
    CHECK(sum(pok) == static_cast<int>(n / 2) && v.empty() && v.valid());
    (void)hipFree(d);
    (void)hipFree(ok);
    ivec::destroyDeviceObject(v);
}

static void
insert_erase_clear()
{
    g_case = "insert / erase at device_end, their non-end and too-many forms, clear (vector.inc:708-958)";
    const stdgpu::index_t N = 10000;
    ivec v = ivec::createDeviceObject(N);
    fill(v, N / 2);
    const std::vector<int> more = iota_from(N / 2 + 1, static_cast<std::size_t>(N / 4));
    int* dm = to_dev(more);
    v.insert(v.device_end(), stdgpu::make_device(static_cast<const int*>(dm)),
             stdgpu::make_device(static_cast<const int*>(dm) + more.size()));
    CHECK(v.size() == N / 2 + N / 4 && !v.empty() && !v.full() && v.valid());
    std::vector<int> all = to_host(v.data(), static_cast<std::size_t>(v.size()));
    CHECK(all == iota_from(1, all.size()));
    // insert_non_end: no effect
    v.insert(v.device_end() - 1, stdgpu::make_device(static_cast<const int*>(dm)),
             stdgpu::make_device(static_cast<const int*>(dm) + more.size()));
    CHECK(v.size() == N / 2 + N / 4);
    // insert_too_many: a range past the capacity, no effect
    std::vector<int> big = iota_from(1, static_cast<std::size_t>(N));
    int* db = to_dev(big);
    v.insert(v.device_end(), stdgpu::make_device(static_cast<const int*>(db)),
             stdgpu::make_device(static_cast<const int*>(db) + big.size()));
    CHECK(v.size() == N / 2 + N / 4 && v.valid());
    // erase the last N / 4
    v.erase(v.device_end() - N / 4, v.device_end());
    CHECK(v.size() == N / 2 && v.valid());
    all = to_host(v.data(), static_cast<std::size_t>(v.size()));
    CHECK(all == iota_from(1, all.size()));
    // erase_non_end and erase_too_many: no effect
    v.erase(v.device_begin(), v.device_end() - 1);
    CHECK(v.size() == N / 2);
    v.erase(v.device_end() - (N + 1), v.device_end());
    CHECK(v.size() == N / 2 && v.valid());
    // the same with a custom execution policy on its own stream
    hipStream_t st;
    CHECK(hipStreamCreate(&st) == hipSuccess);
    stdgpu::execution::device_policy pol{ st };
    v.insert(pol, v.device_end(), stdgpu::make_device(static_cast<const int*>(dm)),
             stdgpu::make_device(static_cast<const int*>(dm) + more.size()));
    CHECK(v.size(pol) == N / 2 + N / 4);
    v.erase(pol, v.device_end() - N / 4, v.device_end());
    CHECK(v.size(pol) == N / 2 && v.valid(pol));
    v.clear(pol);
    CHECK(v.size(pol) == 0 && v.empty(pol) && !v.full(pol) && v.valid(pol));
    fill(v, N);
    v.clear();
    CHECK(v.size() == 0 && v.empty() && !v.full() && v.valid());
    (void)hipStreamDestroy(st);
    (void)hipFree(dm);
    (void)hipFree(db);
    ivec::destroyDeviceObject(v);
}

static void
element_access()
{
    g_case = "front / back / at / operator[] / data / device_begin / device_end / device_range / shrink_to_fit (vector.inc:1079-1200, 1558-1755)";
    const stdgpu::index_t N = 10000;
    ivec v = ivec::createDeviceObject(N);
    fill(v, N);
    int* out = dev_alloc<int>(8);
    hipLaunchKernelGGL(k_access, dim3(1), dim3(64), 0, 0, v, out);
    const std::vector<int> h = to_host(out, 8);
    CHECK(h[0] == 1 && h[1] == N && h[2] == 4 && h[3] == 5 && h[4] == 1 + N + 4 + 5);
    CHECK(h[5] == N && h[6] == 0 && h[7] == 1);
    const ivec& c = v;
    CHECK(v.data() == c.data());
    CHECK(v.device_begin().get() == v.data() && v.device_end().get() == v.data() + N);
    CHECK(c.device_end() == c.device_cend() && c.device_begin() == c.device_cbegin());
    // non_const_device_range / const_device_range (vector.inc:1696-1755)
    int* copy = dev_alloc<int>(static_cast<std::size_t>(N));
    auto r = v.device_range();
    CHECK(r.size() == N && !r.empty() && r.begin().get() == v.data() && r.end().get() == v.data() + N);
    hipLaunchKernelGGL(k_copy_range<decltype(r)>, dim3(blocks(static_cast<std::size_t>(N))), dim3(256), 0, 0, r, copy);
    CHECK(to_host(copy, static_cast<std::size_t>(N)) == iota_from(1, static_cast<std::size_t>(N)));
    auto cr = c.device_range();
    (void)hipMemset(copy, 0, static_cast<std::size_t>(N) * sizeof(int));
    hipLaunchKernelGGL(k_copy_range<decltype(cr)>, dim3(blocks(static_cast<std::size_t>(N))), dim3(256), 0, 0, cr, copy);
    CHECK(to_host(copy, static_cast<std::size_t>(N)) == iota_from(1, static_cast<std::size_t>(N)));
    (void)hipFree(copy);
    (void)pop(v, 100);
    v.shrink_to_fit();
    CHECK(v.size() == N - 100 && (v.capacity() == N - 100 || v.capacity() == N) && v.valid());
    (void)hipFree(out);
    ivec::destroyDeviceObject(v);
}

static void
allocator_and_policies()
{
    g_case = "get_allocator, createDeviceObject / destroyDeviceObject with a policy, other value types (vector.inc:1756-1830)";
    hipStream_t st;
    CHECK(hipStreamCreate(&st) == hipSuccess);
    stdgpu::execution::device_policy pol{ st };
    ivec v = ivec::createDeviceObject(pol, 1000);
    CHECK(v.capacity() == 1000 && v.empty(pol) && v.valid(pol));
    ivec::allocator_type a = v.get_allocator();
    int* arr = a.allocate(1000);
    CHECK(arr != nullptr);
    a.deallocate(arr, 1000);
    ivec::destroyDeviceObject(pol, v);
    auto vf = stdgpu::vector<float>::createDeviceObject(10);
    auto vu = stdgpu::vector<std::uint32_t>::createDeviceObject(10);
    CHECK(vf.empty() && vu.empty() && vf.max_size() == 10);
    stdgpu::vector<float>::destroyDeviceObject(vf);
    stdgpu::vector<std::uint32_t>::destroyDeviceObject(vu);
    (void)hipStreamDestroy(st);
}

static void
per_thread_rates()
{
    g_case = "per-thread push_back / pop_back rates, 2^22 elements";
    const std::size_t C = 1 << 22;
    ivec v = ivec::createDeviceObject(static_cast<stdgpu::index_t>(C));
    const std::vector<int> vals = iota_from(0, C);
    int* d = to_dev(vals);
    int* out = dev_alloc<int>(C);
    int* ok = dev_alloc<int>(C);
    hipEvent_t e0, e1, e2;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    float push_ms = 0, pop_ms = 0;
    for (int rep = 0; rep < 3; ++rep) // the last repetition is reported
    {
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_push, dim3(blocks(C)), dim3(256), 0, 0, v, d, C, ok);
        (void)hipEventRecord(e1, nullptr);
        hipLaunchKernelGGL(k_pop, dim3(blocks(C)), dim3(256), 0, 0, v, C, out, ok);
        (void)hipEventRecord(e2, nullptr);
        (void)hipEventSynchronize(e2);
        (void)hipEventElapsedTime(&push_ms, e0, e1);
        (void)hipEventElapsedTime(&pop_ms, e1, e2);
    }
    CHECK(sum(to_host(ok, C)) == static_cast<int>(C) && v.empty() && v.valid());
    std::vector<int> got = to_host(out, C);
    std::sort(got.begin(), got.end());
    CHECK(got == vals);
    std::printf("rate: push_back %.1f G/s, pop_back %.1f G/s (%zu elements, one atomic per wavefront)\n",
                static_cast<double>(C) / push_ms / 1e6, static_cast<double>(C) / pop_ms / 1e6, C);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipFree(d);
    (void)hipFree(out);
    (void)hipFree(ok);
    ivec::destroyDeviceObject(v);
}


// simultaneous_push_back_and_pop_back (vector.inc:985-1077): each thread pushes its value, pops one,
// and pushes what it popped onto a second vector
template <typename T>
__global__ void
k_simultaneous(stdgpu::vector<T> pool, stdgpu::vector<T> validation, std::size_t n)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pool.push_back(static_cast<T>(i + 1));
    const stdgpu::pair<T, bool> r = pool.pop_back();
    if (r.second) validation.push_back(r.first);
}

template <typename T>
static void
simultaneous_push_back_and_pop_back()
{
    g_case = "simultaneous_push_back_and_pop_back (vector.inc:985-1077)";
    const std::size_t N = 100000;
    auto pool = stdgpu::vector<T>::createDeviceObject(static_cast<stdgpu::index_t>(N));
    auto validation = stdgpu::vector<T>::createDeviceObject(static_cast<stdgpu::index_t>(N));
    hipLaunchKernelGGL(k_simultaneous<T>, dim3(static_cast<unsigned>((N + 255) / 256)), dim3(256), 0, 0, pool,
                       validation, N);
    CHECK(hipDeviceSynchronize() == hipSuccess);
    CHECK(pool.size() == 0 && pool.empty() && !pool.full() && pool.valid());
    CHECK(validation.size() == static_cast<stdgpu::index_t>(N) && !validation.empty() && validation.full() &&
          validation.valid());
    std::vector<T> h(N);
    CHECK(hipMemcpy(h.data(), validation.data(), N * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess);
    std::sort(h.begin(), h.end());
    for (std::size_t i = 0; i < N; ++i) CHECK(h[i] == static_cast<T>(i + 1));
    stdgpu::vector<T>::destroyDeviceObject(pool);
    stdgpu::vector<T>::destroyDeviceObject(validation);
}

// a 16-byte element pushed and popped per thread, and inserted from the host
struct quad
{
    std::int32_t v[4];
};

__global__ void
k_push_quads(stdgpu::vector<quad> q, std::size_t n, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const std::int32_t x = static_cast<std::int32_t>(i);
    ok[i] = q.emplace_back(quad{ { x, x + 1, x * 2, -x } }) ? 1 : 0;
}

__global__ void
k_pop_quads(stdgpu::vector<quad> q, std::size_t n, quad* out, int* ok)
{
    const std::size_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto r = q.pop_back();
    out[i] = r.first;
    ok[i] = r.second ? 1 : 0;
}

static void
sixteen_byte_elements()
{
    g_case = "vector<16-byte struct>: push_back_some / _too_many, pop_back_all / _too_many, insert, erase";
    const std::size_t C = 50000;
    auto q = stdgpu::vector<quad>::createDeviceObject(static_cast<stdgpu::index_t>(C));
    int* ok = dev_alloc<int>(2 * C);
    quad* out = dev_alloc<quad>(2 * C);
    hipLaunchKernelGGL(k_push_quads, dim3(static_cast<unsigned>((2 * C + 255) / 256)), dim3(256), 0, 0, q, 2 * C, ok);
    std::vector<int> hok = to_host(ok, 2 * C);
    CHECK(std::accumulate(hok.begin(), hok.end(), 0) == static_cast<int>(C) && q.full() && q.valid());
    std::vector<quad> h(C);
    CHECK(hipMemcpy(h.data(), q.data(), C * sizeof(quad), hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<int> seen;
    for (const quad& x : h)
    {
        CHECK(x.v[1] == x.v[0] + 1 && x.v[2] == 2 * x.v[0] && x.v[3] == -x.v[0]);
        CHECK(hok[static_cast<std::size_t>(x.v[0])] == 1);
        seen.push_back(x.v[0]);
    }
    std::sort(seen.begin(), seen.end());
    CHECK(std::adjacent_find(seen.begin(), seen.end()) == seen.end());
    // erase the top half from the host, then pop everything and more
    q.erase(q.device_begin() + static_cast<std::ptrdiff_t>(C / 2), q.device_end());
    CHECK(q.size() == static_cast<stdgpu::index_t>(C / 2) && q.valid());
    hipLaunchKernelGGL(k_pop_quads, dim3(static_cast<unsigned>((C + 255) / 256)), dim3(256), 0, 0, q, C, out, ok);
    hok = to_host(ok, C);
    CHECK(std::accumulate(hok.begin(), hok.end(), 0) == static_cast<int>(C / 2) && q.empty() && q.valid());
    // insert a host range back at the end
    quad* src = dev_alloc<quad>(C);
    CHECK(hipMemcpy(src, h.data(), C * sizeof(quad), hipMemcpyHostToDevice) == hipSuccess);
    q.insert(q.device_end(), stdgpu::make_device(src), stdgpu::make_device(src + C));
    CHECK(q.size() == static_cast<stdgpu::index_t>(C) && q.valid());
    std::vector<quad> back(C);
    CHECK(hipMemcpy(back.data(), q.data(), C * sizeof(quad), hipMemcpyDeviceToHost) == hipSuccess);
    for (std::size_t i = 0; i < C; ++i) CHECK(back[i].v[0] == h[i].v[0] && back[i].v[3] == h[i].v[3]);
    (void)hipFree(ok);
    (void)hipFree(out);
    (void)hipFree(src);
    stdgpu::vector<quad>::destroyDeviceObject(q);
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
    void (*cases[])() = { empty_container, pop_back_cases,    push_back_cases,        diverged_lanes,
                          insert_erase_clear, element_access, allocator_and_policies, per_thread_rates,
                          sixteen_byte_elements, simultaneous_push_back_and_pop_back<int>,
                          simultaneous_push_back_and_pop_back<std::int64_t> };
    int index = 0;
    for (auto c : cases)
    {
        std::printf("case %d starts\n", index++);
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
        std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", g_case);
    }
    std::printf("OK: %d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
