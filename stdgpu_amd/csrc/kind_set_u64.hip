// Kernel instantiations and dispatch for set u64 tables (one translation unit per kind).
#include "htab_ops.hpp"

namespace htab_host
{

int
dispatch_set_u64(htab_call& c)
{
    return dispatch_kind<uint64_t, void>(c);
}

} // namespace htab_host
