// Kernel instantiations and dispatch for map u32 u32 tables (one translation unit per kind).
#include "htab_ops.hpp"

namespace htab_host
{

int
dispatch_map_u32_u32(htab_call& c)
{
    return dispatch_kind<uint32_t, uint32_t>(c);
}

} // namespace htab_host
