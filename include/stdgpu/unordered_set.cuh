// stdgpu::unordered_set for MI355X: source-compatible with the reference's
// src/stdgpu/unordered_set.cuh:61-525, backed by libstdgpu_amd.so (see unordered_map.cuh).
#ifndef STDGPU_UNORDERED_SET_H
#define STDGPU_UNORDERED_SET_H

#include <stdgpu/impl/amd_unordered.cuh>

namespace stdgpu
{

template <typename Key,
          typename Hash = hash<Key>,
          typename KeyEqual = equal_to<Key>,
          typename Allocator = safe_device_allocator<Key>>
class unordered_set : public detail::amd_unordered_base<Key, Key, void, Hash, KeyEqual, Allocator>
{
    using base = detail::amd_unordered_base<Key, Key, void, Hash, KeyEqual, Allocator>;

public:
    using typename base::allocator_type;
    using typename base::value_type;

    unordered_set() noexcept = default;

    [[nodiscard]] static unordered_set createDeviceObject(const index_t& capacity, const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, capacity, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static unordered_set createDeviceObject(ExecutionPolicy&&, const index_t& capacity,
                                                          const Allocator& allocator = Allocator())
    {
        unordered_set s;
        base::create(s, capacity, allocator);
        return s;
    }

    // not in the reference: a facade object over a table made with htab_create (own tuning);
    // destroyDeviceObject frees it
    [[nodiscard]] static unordered_set adopt(htab* table, const Allocator& allocator = Allocator())
    {
        unordered_set m;
        base::adopt(m, table, allocator);
        return m;
    }

    static void destroyDeviceObject(unordered_set& device_object) { base::destroy(device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject(ExecutionPolicy&&, unordered_set& device_object)
    {
        base::destroy(device_object);
    }
};

} // namespace stdgpu

#endif // STDGPU_UNORDERED_SET_H
