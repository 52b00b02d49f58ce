"""ctypes binding of libstdgpu_amd.so (the C ABI declared in include/stdgpu_amd/htab.h).

The library is built in-tree by `python -m stdgpu_amd.build`.  Loading fails loudly when it
is missing: there is no CPU fallback for any operation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstdgpu_amd.so")

HTAB_OK = 0
HTAB_E_ARG = -1
HTAB_E_OOM = -2
HTAB_E_HIP = -3
HTAB_E_NODEV = -4
HTAB_E_LIMIT = -5

HTAB_NOALLOC_RANGE_TILES = 32767  # htab.h: largest table htab_range_noalloc serves, in 4096-slot tiles

HTAB_SET_U32 = 0
HTAB_SET_U64 = 1
HTAB_MAP_U64_U32 = 2
HTAB_MAP_U32_U32 = 3
HTAB_MAP_U64_U64 = 4

HTAB_BIT_SET = 0
HTAB_BIT_RESET = 1
HTAB_BIT_FLIP = 2

HTAB_ATOMIC_I32 = 0
HTAB_ATOMIC_U32 = 1
HTAB_ATOMIC_U64 = 2
HTAB_ATOMIC_F32 = 3
HTAB_ATOMIC_OPS = {
    "load": 0, "store": 1, "exchange": 2, "compare_exchange": 3, "fetch_add": 4, "fetch_sub": 5, "fetch_min": 6,
    "fetch_max": 7, "fetch_and": 8, "fetch_or": 9, "fetch_xor": 10, "fetch_inc_mod": 11, "fetch_dec_mod": 12,
}

HTAB_OP_INSERT = 0
HTAB_OP_FIND = 1
HTAB_OP_ERASE = 2


class htab_tuning(ctypes.Structure):
    _fields_ = [
        ("bucket_lanes", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("slots_per_capacity", ctypes.c_float),
        ("probe_batch", ctypes.c_int32),
        ("insert_path", ctypes.c_int32),
        ("caller_hashes", ctypes.c_int32),
        ("payload_bytes", ctypes.c_int32),
    ]


class htab_info(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("bucket_lanes", ctypes.c_int32),
        ("capacity", ctypes.c_int64),
        ("bucket_count", ctypes.c_int64),
        ("max_size", ctypes.c_int64),
        ("physical_buckets", ctypes.c_int64),
        ("physical_slots", ctypes.c_int64),
        ("limit", ctypes.c_int64),
        ("value_bytes", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
    ]


class htab_segment(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
        ("flags", ctypes.c_void_p),
        ("vals_out", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("table", ctypes.c_int32),
        ("op", ctypes.c_int32),
    ]


class htab_server_stats(ctypes.Structure):
    _fields_ = [
        ("launches", ctypes.c_int64),
        ("segments", ctypes.c_int64),
        ("max_concurrent_launches", ctypes.c_int64),
        ("wall_ns", ctypes.c_int64),
    ]


class htab_snapshot_header(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 8),
        ("version", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("bucket_lanes", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("probe_batch", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("capacity", ctypes.c_int64),
        ("physical_buckets", ctypes.c_int64),
        ("size", ctypes.c_int64),
        ("slot_bytes", ctypes.c_int64),
        ("ovf_bytes", ctypes.c_int64),
        ("payload_bytes", ctypes.c_int64),
        ("record_bytes", ctypes.c_int32),
        ("caller_hashes", ctypes.c_int32),
    ]


class htab_bitset_view(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_void_p), ("nbits", ctypes.c_int64)]


class htab_ring_view(ctypes.Structure):
    """the device view of a vector or deque (htab_ring_view in htab.h)"""

    _fields_ = [
        ("data", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
        ("begin", ctypes.c_void_p),
        ("end", ctypes.c_void_p),
        ("ready", ctypes.c_void_p),
        ("free_slots", ctypes.c_void_p),
        ("capacity", ctypes.c_int64),
        ("elem_bytes", ctypes.c_int64),
    ]


htab_deque_view = htab_ring_view
htab_vector_view = htab_ring_view


class htab_block_region(ctypes.Structure):
    """packed 3-D block keys and a sphere around a point (htab_block_region in htab.h)"""

    _fields_ = [
        ("bits", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("offset", ctypes.c_int64),
        ("block_size", ctypes.c_double),
        ("center", ctypes.c_double * 3),
        ("radius", ctypes.c_double),
    ]


class htab_view(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_void_p),
        ("stripes", ctypes.c_void_p),
        ("flags", ctypes.c_void_p),
        ("nbuckets", ctypes.c_uint64),
        ("limit", ctypes.c_uint64),
        ("nstripes", ctypes.c_uint32),
        ("bucket_lanes", ctypes.c_uint32),
        ("ref_bucket_count", ctypes.c_int64),
        ("ref_max_size", ctypes.c_int64),
        ("ovf", ctypes.c_void_p),
        ("payload", ctypes.c_void_p),
        ("payload_bytes", ctypes.c_int64),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int

# name -> (restype, argtypes); every function declared in htab.h
SIGNATURES = {
    "htab_create": (_I32, [_I32, _I64, ctypes.POINTER(htab_tuning), ctypes.POINTER(_P)]),
    "htab_destroy": (_I32, [_P]),
    "htab_insert": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_insert_kv": (_I32, [_P, _P, _P, _I64, _P, _P]),
    "htab_find": (_I32, [_P, _P, _I64, _P, _P, _P]),
    "htab_erase": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_assign": (_I32, [_P, _P, _P, _I64, _P, _P]),
    "htab_size": (_I32, [_P, ctypes.POINTER(_I64), _P]),
    "htab_clear": (_I32, [_P, _P]),
    "htab_range": (_I32, [_P, _P, _I64, ctypes.POINTER(_I64), _P]),
    "htab_range_owned": (_I32, [_P, ctypes.POINTER(_P), ctypes.POINTER(_I64), _P]),
    "htab_range_noalloc": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_contents": (_I32, [_P, _P, _P, _I64, ctypes.POINTER(_I64), _P]),
    "htab_valid": (_I32, [_P, ctypes.POINTER(ctypes.c_int32), _P]),
    "htab_valid_hashed": (_I32, [_P, _P, ctypes.POINTER(ctypes.c_int32), _P]),
    "htab_hash_buffer": (_I32, [_P, ctypes.POINTER(_P)]),
    "htab_bucket": (_I32, [_P, ctypes.c_uint64, ctypes.POINTER(_I64)]),
    "htab_get_info": (_I32, [_P, ctypes.POINTER(htab_info)]),
    "htab_get_view": (_I32, [_P, ctypes.POINTER(htab_view)]),
    "htab_last_error": (ctypes.c_char_p, []),
    "htab_batch_create": (_I32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(_P)]),
    "htab_batch_destroy": (_I32, [_P]),
    "htab_batch_run": (_I32, [_P, _P, _I32, _P, _P, _P, _P, _P, _P]),
    "htab_batch_run_segments": (_I32, [_P, _P, ctypes.c_int32, _P]),
    "htab_batch_run_per_table": (_I32, [_P, _I32, ctypes.POINTER(_I64), _P, _P, _P, _P, _P]),
    "htab_batch_capture": (_I32, [_P, _P, _I32, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "htab_graph_launch": (_I32, [_P, _P]),
    "htab_graph_destroy": (_I32, [_P]),
    "htab_server_create": (_I32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, _I64, ctypes.POINTER(_P)]),
    "htab_server_destroy": (_I32, [_P]),
    "htab_server_run": (_I32, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "htab_server_replay_plain": (_I32, [_P, _P, _P, _I64, _I32, _P]),
    "htab_deque_create": (_I32, [_I64, ctypes.POINTER(_P)]),
    "htab_deque_create_sized": (_I32, [_I64, _I64, ctypes.POINTER(_P)]),
    "htab_deque_destroy": (_I32, [_P]),
    "htab_deque_push_back": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_deque_push_front": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_deque_pop_front": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_deque_pop_back": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_deque_valid": (_I32, [_P, ctypes.POINTER(ctypes.c_int32), _P]),
    "htab_deque_fill_iota": (_I32, [_P, _P]),
    "htab_deque_size": (_I32, [_P, ctypes.POINTER(_I64), _P]),
    "htab_deque_contents": (_I32, [_P, _P, ctypes.POINTER(_I64), _P]),
    "htab_deque_get_view": (_I32, [_P, ctypes.POINTER(htab_deque_view)]),
    "htab_deque_clear": (_I32, [_P, _P]),
    "htab_snapshot_bytes": (_I32, [_P, ctypes.POINTER(_I64)]),
    "htab_snapshot": (_I32, [_P, _P, _I64, _P]),
    "htab_restore": (_I32, [_P, _I64, ctypes.POINTER(_P)]),
    "htab_bitset_create": (_I32, [_I64, ctypes.POINTER(_P)]),
    "htab_bitset_destroy": (_I32, [_P]),
    "htab_bitset_update": (_I32, [_P, _I32, _P, _I64, _P, _P, _P]),
    "htab_bitset_test": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_bitset_all_bits": (_I32, [_P, _I32, _P]),
    "htab_bitset_count": (_I32, [_P, ctypes.POINTER(_I64), _P]),
    "htab_bitset_get_view": (_I32, [_P, ctypes.POINTER(htab_bitset_view)]),
    "htab_atomic_create": (_I32, [_I32, _I64, ctypes.POINTER(_P)]),
    "htab_atomic_destroy": (_I32, [_P]),
    "htab_atomic_data": (_I32, [_P, ctypes.POINTER(_P)]),
    "htab_atomic_apply": (_I32, [_P, _I32, _P, _P, _P, _I64, _P, _P, _P]),
    "htab_atomic_ref_apply": (_I32, [_I32, _P, _I64, _I32, _P, _P, _P, _I64, _P, _P, _P]),
    "htab_atomic_load": (_I32, [_P, _I64, _P, _P]),
    "htab_atomic_store": (_I32, [_P, _I64, _P, _P]),
    "htab_vector_create": (_I32, [_I64, ctypes.POINTER(_P)]),
    "htab_vector_create_sized": (_I32, [_I64, _I64, ctypes.POINTER(_P)]),
    "htab_vector_valid": (_I32, [_P, ctypes.POINTER(ctypes.c_int32), _P]),
    "htab_vector_destroy": (_I32, [_P]),
    "htab_vector_push_back": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_vector_pop_back": (_I32, [_P, _P, _I64, _P, _P]),
    "htab_vector_fill_iota": (_I32, [_P, ctypes.c_uint32, _P]),
    "htab_vector_size": (_I32, [_P, ctypes.POINTER(_I64), _P]),
    "htab_vector_contents": (_I32, [_P, _P, ctypes.POINTER(_I64), _P]),
    "htab_vector_clear": (_I32, [_P, _P]),
    "htab_vector_get_view": (_I32, [_P, ctypes.POINTER(htab_vector_view)]),
    "htab_assign_from_deque": (_I32, [_P, _P, _P, _I64, _P, _P, _P]),
    "htab_erase_far_blocks": (_I32, [_P, ctypes.POINTER(htab_block_region), _P, _P, _P]),
    "htab_rt_alloc": (_I32, [_I64, ctypes.POINTER(_P)]),
    "htab_rt_free": (_I32, [_P]),
    "htab_rt_copy": (_I32, [_P, _P, _I64, ctypes.c_int32, _P]),
    "htab_rt_zero": (_I32, [_P, _I64, _P]),
    "htab_rt_sync": (_I32, [_P]),
    "htab_rt_launch_status": (_I32, []),
}

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m stdgpu_amd.build` (hipcc, gfx950). "
            "There is no CPU fallback."
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HtabError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != HTAB_OK:
        msg = load().htab_last_error()
        raise HtabError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
