"""CPU checks of the C ABI boundary: the HIP library loads and exports exactly what
include/stdgpu_amd/htab.h declares (no compute calls: this container has no GPU)."""
import ctypes
import os
import re

import pytest

from stdgpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stdgpu_amd", "htab.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(htab_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ["htab_create", "htab_destroy", "htab_insert", "htab_find", "htab_erase", "htab_size",
                 "htab_range", "htab_clear", "htab_valid", "htab_bucket", "htab_last_error",
                 "htab_deque_push_back", "htab_deque_pop_front"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_python_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_struct_layouts_match_header():
    # field order/size of the C structs as seen by ctypes (plain 64-bit ABI)
    assert ctypes.sizeof(_lib.htab_tuning) == 28
    assert ctypes.sizeof(_lib.htab_info) == 8 + 8 * 8
    assert ctypes.sizeof(_lib.htab_view) == 3 * 8 + 2 * 8 + 2 * 4 + 2 * 8 + 8 + 2 * 8
    assert ctypes.sizeof(_lib.htab_ring_view) == 8 * 8
    assert _lib.htab_deque_view is _lib.htab_ring_view and _lib.htab_vector_view is _lib.htab_ring_view
    assert ctypes.sizeof(_lib.htab_bitset_view) == 2 * 8
    assert ctypes.sizeof(_lib.htab_block_region) == 4 + 4 + 8 + 8 + 3 * 8 + 8


def test_argument_errors_without_gpu():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.htab_create(99, 10, None, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert lib.htab_create(_lib.HTAB_MAP_U64_U32, 0, None, ctypes.byref(h)) == _lib.HTAB_E_ARG
    # workgroup sizes: 64, 128, 256, 512 or 1024 (checked before any device work)
    for bad in (32, 96, 100, 2048):
        tun = _lib.htab_tuning(8, bad, 1.0, 0)
        assert lib.htab_create(_lib.HTAB_MAP_U64_U32, 1000, ctypes.byref(tun), ctypes.byref(h)) == _lib.HTAB_E_ARG
        assert b"block_size" in lib.htab_last_error()
    assert lib.htab_destroy(None) == _lib.HTAB_E_ARG
    assert lib.htab_find(None, None, 0, None, None, None) == _lib.HTAB_E_ARG
    assert b"null" in lib.htab_last_error()
    # runtime helpers refuse bad arguments before any device work
    out = ctypes.c_void_p()
    assert lib.htab_rt_alloc(-1, ctypes.byref(out)) == _lib.HTAB_E_ARG
    assert lib.htab_rt_copy(None, None, 8, 0, None) == _lib.HTAB_E_ARG
    assert lib.htab_rt_copy(ctypes.c_void_p(0x1000), ctypes.c_void_p(0x1000), 8, 3, None) == _lib.HTAB_E_ARG
    assert lib.htab_rt_zero(None, 8, None) == _lib.HTAB_E_ARG
    # the free-list calls refuse a missing table or deque before any device work
    r = _lib.htab_block_region(21, 0, 1 << 20, 0.064)
    assert lib.htab_assign_from_deque(None, None, None, 0, None, None, None) == _lib.HTAB_E_ARG
    assert lib.htab_erase_far_blocks(None, ctypes.byref(r), None, None, None) == _lib.HTAB_E_ARG


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(ImportError):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


def test_device_workload_generator_matches_numpy():
    from stdgpu_amd import workloads_device as wd

    assert wd.check_against_numpy()
    assert wd.check_zipf_against_numpy(1 << 12)
    assert wd.check_against_numpy(ntables=40, log2_lo=5, log2_hi=11)


def test_restore_rejects_what_is_not_a_snapshot():
    # checked before any device work, so this runs without a GPU
    import numpy as np

    lib = _lib.load()
    h = ctypes.c_void_p()
    hs = ctypes.sizeof(_lib.htab_snapshot_header)
    assert hs == 8 + 6 * 4 + 6 * 8 + 2 * 4
    junk = np.zeros(4096, dtype=np.uint8)
    assert lib.htab_restore(junk.ctypes.data_as(ctypes.c_void_p), junk.size, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert lib.htab_restore(junk.ctypes.data_as(ctypes.c_void_p), 10, ctypes.byref(h)) == _lib.HTAB_E_ARG
    hd = _lib.htab_snapshot_header(b"HTABSNP1", 2, _lib.HTAB_MAP_U64_U32, 8, 256, 0, 0, 1000, 125, 0,
                                   (125 * 8 + 2) * 16, 0, 0, 0, 0)
    raw = np.frombuffer(bytes(hd), dtype=np.uint8)
    # version 1 predates the away counters that this layout keeps in its slot padding
    hd.version = 1
    old = np.frombuffer(bytes(hd), dtype=np.uint8)
    assert lib.htab_restore(old.ctypes.data_as(ctypes.c_void_p), old.size, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert b"version 1" in lib.htab_last_error()
    hd.version = 2
    # a header that promises more bytes than the buffer holds
    assert lib.htab_restore(raw.ctypes.data_as(ctypes.c_void_p), raw.size, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert b"inconsistent" in lib.htab_last_error()
    # inconsistent slot bytes for the geometry
    hd.slot_bytes += 16
    buf = np.zeros(hs + hd.slot_bytes, dtype=np.uint8)
    buf[:hs] = np.frombuffer(bytes(hd), dtype=np.uint8)
    assert lib.htab_restore(buf.ctypes.data_as(ctypes.c_void_p), buf.size, ctypes.byref(h)) == _lib.HTAB_E_ARG
    # wrong magic
    hd.slot_bytes -= 16
    hd.magic = b"HTABSNPX"
    buf[:hs] = np.frombuffer(bytes(hd), dtype=np.uint8)
    assert lib.htab_restore(buf.ctypes.data_as(ctypes.c_void_p), buf.size, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert b"not a table snapshot" in lib.htab_last_error()


def test_container_argument_errors_without_gpu():
    """vector / deque / bitset / atomic entry points refuse bad arguments before any device work"""
    lib = _lib.load()
    h = ctypes.c_void_p()
    E = _lib.HTAB_E_ARG
    assert lib.htab_vector_create(0, ctypes.byref(h)) == E
    assert lib.htab_vector_create((1 << 32) + 1, ctypes.byref(h)) == E
    assert lib.htab_vector_push_back(None, None, 1, None, None) == E
    assert lib.htab_vector_clear(None, None) == E
    assert lib.htab_vector_get_view(None, None) == E
    assert lib.htab_deque_create(0, ctypes.byref(h)) == E
    assert lib.htab_bitset_create(0, ctypes.byref(h)) == E
    assert lib.htab_bitset_create(10, None) == E
    assert lib.htab_bitset_update(None, _lib.HTAB_BIT_SET, None, 1, None, None, None) == E
    assert lib.htab_bitset_count(None, None, None) == E
    assert lib.htab_atomic_create(7, 1, ctypes.byref(h)) == E  # no such value type
    assert lib.htab_atomic_create(_lib.HTAB_ATOMIC_U32, 0, ctypes.byref(h)) == E
    assert lib.htab_atomic_apply(None, 4, None, None, None, 1, None, None, None) == E
    buf = ctypes.c_void_p(0x1000)  # never dereferenced: every call below is refused first
    ops = _lib.HTAB_ATOMIC_OPS
    ref = lib.htab_atomic_ref_apply
    assert ref(9, buf, 4, ops["fetch_add"], None, buf, None, 1, None, None, None) == E  # type
    assert ref(_lib.HTAB_ATOMIC_U32, None, 4, ops["fetch_add"], None, buf, None, 1, None, None, None) == E  # data
    assert ref(_lib.HTAB_ATOMIC_U32, buf, 0, ops["fetch_add"], None, buf, None, 1, None, None, None) == E  # count
    assert ref(_lib.HTAB_ATOMIC_U32, buf, 4, 13, None, buf, None, 1, None, None, None) == E  # op
    assert ref(_lib.HTAB_ATOMIC_U32, buf, 4, ops["fetch_add"], None, None, None, 1, None, None, None) == E  # operands
    assert ref(_lib.HTAB_ATOMIC_U32, buf, 4, ops["compare_exchange"], None, buf, None, 1, None, None, None) == E
    # operations the reference does not define for the type: bitwise ops on float, the wrapping
    # increment / decrement on anything but unsigned int
    assert ref(_lib.HTAB_ATOMIC_F32, buf, 4, ops["fetch_and"], None, buf, None, 1, None, None, None) == E
    assert ref(_lib.HTAB_ATOMIC_I32, buf, 4, ops["fetch_inc_mod"], None, buf, None, 1, None, None, None) == E
    assert ref(_lib.HTAB_ATOMIC_U64, buf, 4, ops["fetch_dec_mod"], None, buf, None, 1, None, None, None) == E
    # values must be naturally aligned for the device atomics
    odd = ctypes.c_void_p(0x1004)
    assert ref(_lib.HTAB_ATOMIC_U64, odd, 4, ops["fetch_add"], None, buf, None, 1, None, None, None) == E
    assert ref(_lib.HTAB_ATOMIC_U64, buf, 4, ops["fetch_add"], None, odd, None, 1, None, None, None) == E  # operands
    assert ref(_lib.HTAB_ATOMIC_U32, buf, 4, ops["fetch_add"], ctypes.c_void_p(0x1002), buf, None, 1, None, None,
               None) == E  # int64 targets
