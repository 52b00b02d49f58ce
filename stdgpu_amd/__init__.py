"""MI355X-native hashed containers with stdgpu's unordered_map / unordered_set API.

The compute path is the HIP library libstdgpu_amd.so (C ABI: include/stdgpu_amd/htab.h);
this package is its host-side mirror for Python callers, tests and the benchmark.
"""
from .containers import (  # noqa: F401
    Tuning,
    atomic,
    atomic_ref_apply,
    bitset,
    deque,
    deque_u32,
    unordered_map,
    unordered_map_u32_u32,
    unordered_map_u64_u32,
    unordered_map_u64_u64,
    unordered_set,
    unordered_set_u32,
    unordered_set_u64,
    vector,
    vector_u32,
)
from ._lib import HtabError, load  # noqa: F401
