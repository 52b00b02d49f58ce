// Kernel instantiations and dispatch for map u64 u32 tables (one translation unit per kind).
#include "htab_ops.hpp"

namespace htab_host
{

int
dispatch_map_u64_u32(htab_call& c)
{
    return dispatch_kind<uint64_t, uint32_t>(c);
}

} // namespace htab_host
