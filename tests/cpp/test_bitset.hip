// Source-compatibility tests of stdgpu::bitset<> from user kernels, written as the reference's own
// cases are (tests/stdgpu/bitset.inc): 2^20 bits; per-thread set / reset / flip of every bit or of a
// shuffled third of them, read back through operator[] and test(); host set() / reset() / flip() of
// every bit; count / all / any / none.  Prints one line per case; exits non-zero if any case fails.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

#include <stdgpu/bitset.cuh>

using bits_t = stdgpu::bitset<>;

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

static unsigned
blocks(std::size_t n)
{
    return static_cast<unsigned>((n + 255) / 256);
}

// op 0 set, 1 reset, 2 set then reset if set, 3 flip; of bit pos[i] (or i when pos is null).  old[i]
// receives the call's return value, now[pos] the bit read back through operator[] and test()
__global__ void
k_bits(bits_t b, const stdgpu::index_t* pos, stdgpu::index_t n, int op, std::uint8_t* old, std::uint8_t* now)
{
    const stdgpu::index_t i = static_cast<stdgpu::index_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const stdgpu::index_t p = pos ? pos[i] : i;
    bool r = false;
    if (op == 0)
        r = b.set(p);
    else if (op == 1)
        r = b.reset(p);
    else if (op == 2)
    {
        r = b.set(p);
        if (b[p]) b.reset(p);
    }
    else
        r = b.flip(p);
    old[i] = r ? 1 : 0;
    const bits_t& cb = b;
    now[p] = static_cast<std::uint8_t>(b[p] && cb[p] && b.test(p) ? 1 : (b[p] || b.test(p) ? 2 : 0));
}

// the proxy: assignment returns the previous value, flip() flips, ~ reads the inverse
__global__ void
k_reference(bits_t b, stdgpu::index_t n, std::uint8_t* out)
{
    const stdgpu::index_t i = static_cast<stdgpu::index_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    bits_t::reference r = b[i];
    const bool a = (r = true);   // was false
    const bool c = (r = false);  // was true
    const bool d = r.flip();     // was false, now true
    const bool e = ~r;           // now true: ~ is false
    out[i] = static_cast<std::uint8_t>((!a) && c && (!d) && (!e) && static_cast<bool>(r));
}

static std::vector<stdgpu::index_t>
shuffled(stdgpu::index_t n, unsigned seed)
{
    std::vector<stdgpu::index_t> v(static_cast<std::size_t>(n));
    std::iota(v.begin(), v.end(), 0);
    std::mt19937 rng(seed);
    std::shuffle(v.begin(), v.end(), rng);
    return v;
}

static void
componentwise_and_random(stdgpu::index_t size)
{
    static char name[128];
    std::snprintf(name, sizeof(name), "per-thread set / reset / flip, %d bits", size);
    g_case = name;
    bits_t b = bits_t::createDeviceObject(size);
    CHECK(b.size() == size && !b.empty() && b.count() == 0 && !b.all() && !b.any() && b.none());
    const std::size_t S = static_cast<std::size_t>(size);
    std::uint8_t* old = dev_alloc<std::uint8_t>(S);
    std::uint8_t* now = dev_alloc<std::uint8_t>(S);
    auto run = [&](const stdgpu::index_t* pos, stdgpu::index_t n, int op) {
        hipLaunchKernelGGL(k_bits, dim3(blocks(static_cast<std::size_t>(n))), dim3(256), 0, 0, b, pos, n, op, old, now);
        (void)hipDeviceSynchronize();
    };
    auto all_equal = [&](std::uint8_t* d, std::size_t n, std::uint8_t v) {
        const std::vector<std::uint8_t> h = to_host(d, n);
        return std::all_of(h.begin(), h.end(), [v](std::uint8_t x) { return x == v; });
    };
    // set_all_bits_componentwise: every bit set, each call saw it reset
    run(nullptr, size, 0);
    CHECK(all_equal(now, S, 1) && all_equal(old, S, 0) && b.count() == size && b.all());
    // flip_all_bits_previously_set_componentwise
    run(nullptr, size, 3);
    CHECK(all_equal(now, S, 0) && all_equal(old, S, 1) && b.count() == 0 && b.none());
    // flip_all_bits_previously_reset_componentwise
    run(nullptr, size, 3);
    CHECK(all_equal(now, S, 1) && all_equal(old, S, 0) && b.all());
    // reset_all_bits_componentwise
    run(nullptr, size, 1);
    CHECK(all_equal(now, S, 0) && all_equal(old, S, 1) && b.none());
    // set_and_reset_all_bits_componentwise
    run(nullptr, size, 2);
    CHECK(all_equal(now, S, 0) && b.none());
    // random thirds (set_random_bits, reset_random_bits, flip_random_bits_previously_*)
    const std::vector<stdgpu::index_t> seq = shuffled(size, static_cast<unsigned>(size));
    stdgpu::index_t* dpos = dev_alloc<stdgpu::index_t>(S);
    (void)hipMemcpy(dpos, seq.data(), S * sizeof(stdgpu::index_t), hipMemcpyHostToDevice);
    const stdgpu::index_t N = size / 3;
    std::vector<std::uint8_t> in_third(S, 0);
    for (stdgpu::index_t i = 0; i < N; ++i) in_third[static_cast<std::size_t>(seq[static_cast<std::size_t>(i)])] = 1;
    (void)hipMemset(now, 0, S);
    run(dpos, N, 0);
    CHECK(b.count() == N && b.any() && !b.all());
    CHECK(to_host(now, S) == in_third);
    run(dpos, N, 3); // flip the set third back
    CHECK(b.count() == 0);
    run(dpos, N, 3); // flip it again
    CHECK(b.count() == N && to_host(now, S) == in_third);
    run(dpos, N, 1);
    CHECK(b.count() == 0 && b.none());
    (void)hipFree(dpos);
    (void)hipFree(old);
    (void)hipFree(now);
    bits_t::destroyDeviceObject(b);
}

static void
whole_set_operations_and_reference()
{
    g_case = "set() / reset() / flip() of every bit, with and without a policy; the bit proxy";
    const stdgpu::index_t size = 1048576;
    bits_t b = bits_t::createDeviceObject(size);
    b.set();
    CHECK(b.count() == size && b.all());
    b.flip();
    CHECK(b.count() == 0 && b.none());
    b.flip();
    CHECK(b.all());
    b.reset();
    CHECK(b.none());
    hipStream_t s;
    CHECK(hipStreamCreate(&s) == hipSuccess);
    stdgpu::execution::device_policy pol{ s };
    b.set(pol);
    CHECK(b.count(pol) == size && b.all(pol));
    b.flip(pol);
    CHECK(b.none(pol));
    b.flip(pol);
    b.reset(pol);
    CHECK(b.none(pol) && !b.any(pol));
    (void)hipStreamDestroy(s);
    std::uint8_t* out = dev_alloc<std::uint8_t>(static_cast<std::size_t>(size));
    hipLaunchKernelGGL(k_reference, dim3(blocks(static_cast<std::size_t>(size))), dim3(256), 0, 0, b, size, out);
    const std::vector<std::uint8_t> h = to_host(out, static_cast<std::size_t>(size));
    CHECK(std::all_of(h.begin(), h.end(), [](std::uint8_t x) { return x == 1; }) && b.all());
    (void)hipFree(out);
    auto alloc = b.get_allocator();
    (void)alloc;
    bits_t::destroyDeviceObject(b);
    bits_t e;
    CHECK(e.size() == 0 && e.empty() && e.count() == 0);
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
    const int sizes[] = { 1048576, 1000003, 33 }; // the reference fixture's size; a tail block; one block + 1
    for (int sz : sizes)
    {
        const int before = g_failures;
        componentwise_and_random(sz);
        std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", g_case);
    }
    const int before = g_failures;
    whole_set_operations_and_reference();
    std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", g_case);
    std::printf("OK: %d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
