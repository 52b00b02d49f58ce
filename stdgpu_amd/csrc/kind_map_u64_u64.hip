// Kernel instantiations and dispatch for map u64 u64 tables (one translation unit per kind).
#include "htab_ops.hpp"

namespace htab_host
{

int
dispatch_map_u64_u64(htab_call& c)
{
    return dispatch_kind<uint64_t, uint64_t>(c);
}

} // namespace htab_host
