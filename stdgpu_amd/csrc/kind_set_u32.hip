// Kernel instantiations and dispatch for set u32 tables (one translation unit per kind).
#include "htab_ops.hpp"

namespace htab_host
{

int
dispatch_set_u32(htab_call& c)
{
    return dispatch_kind<uint32_t, void>(c);
}

} // namespace htab_host
