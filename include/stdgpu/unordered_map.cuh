// stdgpu::unordered_map for MI355X: source-compatible with the reference's
// src/stdgpu/unordered_map.cuh:60-538 (createDeviceObject / destroyDeviceObject, bulk
// insert / erase, device_range, size queries; native find / contains / count / insert /
// emplace / erase / bucket / bucket_size in user kernels), backed by libstdgpu_amd.so.
#ifndef STDGPU_UNORDERED_MAP_H
#define STDGPU_UNORDERED_MAP_H

#include <stdgpu/impl/amd_unordered.cuh>

namespace stdgpu
{

template <typename Key,
          typename T,
          typename Hash = hash<Key>,
          typename KeyEqual = equal_to<Key>,
          typename Allocator = safe_device_allocator<pair<const Key, T>>>
class unordered_map : public detail::amd_unordered_base<Key, pair<const Key, T>, T, Hash, KeyEqual, Allocator>
{
    using base = detail::amd_unordered_base<Key, pair<const Key, T>, T, Hash, KeyEqual, Allocator>;

public:
    using mapped_type = T;
    using typename base::allocator_type;
    using typename base::value_type;

    unordered_map() noexcept = default;

    [[nodiscard]] static unordered_map createDeviceObject(const index_t& capacity, const Allocator& allocator = Allocator())
    {
        return createDeviceObject(execution::device, capacity, allocator);
    }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    [[nodiscard]] static unordered_map createDeviceObject(ExecutionPolicy&&, const index_t& capacity,
                                                          const Allocator& allocator = Allocator())
    {
        unordered_map m;
        base::create(m, capacity, allocator);
        return m;
    }

    // not in the reference: a facade object over a table made with htab_create (own tuning);
    // destroyDeviceObject frees it
    [[nodiscard]] static unordered_map adopt(htab* table, const Allocator& allocator = Allocator())
    {
        unordered_map m;
        base::adopt(m, table, allocator);
        return m;
    }

    static void destroyDeviceObject(unordered_map& device_object) { base::destroy(device_object); }

    template <typename ExecutionPolicy, std::enable_if_t<is_execution_policy_v<ExecutionPolicy>, int> = 0>
    static void destroyDeviceObject(ExecutionPolicy&&, unordered_map& device_object)
    {
        base::destroy(device_object);
    }
};

} // namespace stdgpu

#endif // STDGPU_UNORDERED_MAP_H
