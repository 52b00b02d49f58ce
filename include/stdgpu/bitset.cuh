// Source-compatible stdgpu::bitset<> (src/stdgpu/bitset.cuh) with 32-bit blocks, on gfx950.
//
// Host: createDeviceObject / destroyDeviceObject, set() / reset() / flip() of every bit, count / all /
// any / none (bulk kernels of libstdgpu_amd.so).  Device (user kernels, per thread): set(n, value) /
// reset(n) / flip(n) returning the previous value (one device-scope atomic or / and / xor on the bit's
// block, bitset_detail.cuh:265-329), test(n), operator[] with the reference's proxy (operator=
// returns the previous value, flip(), ~, conversion to bool; bitset.cuh:60-150), size / empty.
#ifndef STDGPU_BITSET_CUH
#define STDGPU_BITSET_CUH

#include <cstdint>
#include <stdexcept>
#include <type_traits>

#include <stdgpu/impl/amd_core.h>

namespace stdgpu
{

template <typename Block = unsigned int, typename Allocator = safe_device_allocator<Block>>
class bitset
{
public:
    static_assert(std::is_same_v<Block, unsigned int>,
                  "stdgpu::bitset (MI355X library): 32-bit blocks (unsigned int)");

    using block_type = Block;
    using allocator_type = Allocator;

    // a bit of the set: assignment and flip are atomic and return the previous value
    class reference
    {
    public:
        reference(const reference&) = default;
        ~reference() = default;

        __device__ bool operator=(bool x) noexcept // NOLINT(misc-unconventional-assign-operator)
        {
            return x ? bitset::fetch_or(_bit_block, _bit_n) : bitset::fetch_and_not(_bit_block, _bit_n);
        }

        __device__ bool operator=(const reference& x) noexcept // NOLINT(misc-unconventional-assign-operator)
        {
            return operator=(static_cast<bool>(x));
        }

        reference& operator=(reference&&) = delete;

        __device__ operator bool() const noexcept // NOLINT(hicpp-explicit-conversions)
        {
            return bitset::load_bit(_bit_block, _bit_n);
        }

        __device__ bool operator~() const noexcept { return !static_cast<bool>(*this); }

        __device__ bool flip() noexcept { return bitset::fetch_xor(_bit_block, _bit_n); }

    private:
        friend bitset;

        __host__ __device__ reference(block_type* bit_block, const index_t bit_n) : _bit_block(bit_block), _bit_n(bit_n)
        {
        }

        block_type* _bit_block = nullptr;
        index_t _bit_n = -1;
    };

    [[nodiscard]] static bitset createDeviceObject(const index_t& size, const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, size, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static bitset createDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, const index_t& size,
                                                   const Allocator& allocator = Allocator())
    {
        if (size <= 0) throw std::invalid_argument("stdgpu::bitset: size must be positive");
        bitset b;
        b._allocator = allocator;
        detail::check(htab_bitset_create(size, &b._h), "bitset::createDeviceObject");
        detail::check(htab_bitset_get_view(b._h, &b._view), "bitset::createDeviceObject");
        return b;
    }

    static void destroyDeviceObject(bitset& device_object) { destroyDeviceObject(execution::device, device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject([[maybe_unused]] ExecutionPolicy&& policy, bitset& device_object)
    {
        if (device_object._h) detail::check(htab_bitset_destroy(device_object._h), "bitset::destroyDeviceObject");
        device_object._h = nullptr;
        device_object._view = htab_bitset_view{};
    }

    bitset() noexcept = default;

    [[nodiscard]] __host__ __device__ allocator_type get_allocator() const noexcept { return _allocator; }

    // ---------------------------------------------------------------- every bit (host)
    void set() { set(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void set(ExecutionPolicy&& policy)
    {
        all_bits(HTAB_BIT_SET, policy.stream, "bitset::set");
    }

    void reset() { reset(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void reset(ExecutionPolicy&& policy)
    {
        all_bits(HTAB_BIT_RESET, policy.stream, "bitset::reset");
    }

    void flip() { flip(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    void flip(ExecutionPolicy&& policy)
    {
        all_bits(HTAB_BIT_FLIP, policy.stream, "bitset::flip");
    }

    // ---------------------------------------------------------------- one bit (device)
    __device__ bool set(const index_t n, const bool value = true)
    {
        return value ? fetch_or(block_of(n), n) : fetch_and_not(block_of(n), n);
    }
    __device__ bool reset(const index_t n) { return fetch_and_not(block_of(n), n); }
    __device__ bool flip(const index_t n) { return fetch_xor(block_of(n), n); }
    __device__ bool operator[](const index_t n) const { return test(n); }
    __device__ reference operator[](const index_t n) { return reference(block_of(n), n); }
    __device__ bool test(const index_t n) const { return load_bit(block_of(n), n); }

    // ---------------------------------------------------------------- size and counts
    [[nodiscard]] __host__ __device__ bool empty() const noexcept { return size() == 0; }
    __host__ __device__ index_t size() const noexcept { return static_cast<index_t>(_view.nbits); }

    index_t count() const { return count(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    index_t count(ExecutionPolicy&& policy) const
    {
        if (!_h) return 0;
        std::int64_t c = 0;
        detail::check(htab_bitset_count(_h, &c, policy.stream), "bitset::count");
        return static_cast<index_t>(c);
    }

    bool all() const { return all(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool all(ExecutionPolicy&& policy) const
    {
        return size() > 0 && count(std::forward<ExecutionPolicy>(policy)) == size();
    }

    bool any() const { return any(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool any(ExecutionPolicy&& policy) const
    {
        return count(std::forward<ExecutionPolicy>(policy)) > 0;
    }

    bool none() const { return none(execution::device); }
    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    bool none(ExecutionPolicy&& policy) const
    {
        return count(std::forward<ExecutionPolicy>(policy)) == 0;
    }

    htab_bitset* native_handle() const { return _h; }

private:
    static constexpr index_t kBitsPerBlock = 32;

    __host__ __device__ block_type* block_of(const index_t n) const { return _view.blocks + n / kBitsPerBlock; }

    __device__ static block_type mask(const index_t n) { return block_type(1) << (n % kBitsPerBlock); }

    __device__ static bool fetch_or(block_type* b, const index_t n)
    {
        return (__hip_atomic_fetch_or(b, mask(n), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask(n)) != 0;
    }

    __device__ static bool fetch_and_not(block_type* b, const index_t n)
    {
        return (__hip_atomic_fetch_and(b, ~mask(n), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask(n)) != 0;
    }

    __device__ static bool fetch_xor(block_type* b, const index_t n)
    {
        return (__hip_atomic_fetch_xor(b, mask(n), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask(n)) != 0;
    }

    __device__ static bool load_bit(const block_type* b, const index_t n)
    {
        return (__hip_atomic_load(b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask(n)) != 0;
    }

    void all_bits(int op, hipStream_t s, const char* what)
    {
        if (!_h) return;
        detail::check(htab_bitset_all_bits(_h, op, s), what);
        detail::check(htab_rt_sync(s), what);
    }

    htab_bitset* _h = nullptr;
    htab_bitset_view _view{};
    Allocator _allocator{};
};

} // namespace stdgpu

#endif // STDGPU_BITSET_CUH
