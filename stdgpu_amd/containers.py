"""Host-side mirror of stdgpu's unordered_map / unordered_set (Python), over the C ABI.

Names, argument meaning and return values follow the reference's host API
(src/stdgpu/unordered_map.cuh:70-538, src/stdgpu/unordered_set.cuh:61-525):
createDeviceObject / destroyDeviceObject, bulk insert(begin, end) / erase(begin, end),
size / empty / full / max_size / bucket_count / load_factor / max_load_factor, bucket(key),
device_range(), clear(), valid().  The reference has no bulk find; find() / contains() here
are its device find/contains applied to a range of keys.

Device buffers are torch tensors on the GPU (PyTorch is used for device memory and streams
only).  Unsigned keys travel in signed tensors of the same width (int32 for uint32_t, int64
for uint64_t); the bits are unchanged.  Calls are asynchronous on the current torch stream
unless a `stream` is given; size()/device_range()/valid() synchronise, like the reference's
host-side size() (a device-to-host copy).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib
from ._lib import check

_KINDS = {
    ("u32", None): _lib.HTAB_SET_U32,
    ("u64", None): _lib.HTAB_SET_U64,
    ("u64", "u32"): _lib.HTAB_MAP_U64_U32,
    ("u32", "u32"): _lib.HTAB_MAP_U32_U32,
    ("u64", "u64"): _lib.HTAB_MAP_U64_U64,
}

_TORCH_DT = {"u32": "int32", "u64": "int64"}
_BYTES = {"u32": 4, "u64": 8}


def _torch():
    import torch

    return torch


def _stream_ptr(stream) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


@dataclass
class Tuning:
    bucket_lanes: int = 8
    block_size: int = 64
    slots_per_capacity: float = 1.0
    probe_batch: int = 0  # 0 = library default
    insert_path: int = 0  # 0 and 1 = per-key CAS, 2 = partitioned by table region

    def to_c(self) -> _lib.htab_tuning:
        return _lib.htab_tuning(self.bucket_lanes, self.block_size, self.slots_per_capacity, self.probe_batch,
                                self.insert_path)


class _unordered_base:
    key: str = "u64"
    mapped: str | None = None

    def __init__(self, handle: int, capacity: int):
        self._h = ctypes.c_void_p(handle)
        self._capacity = capacity
        info = _lib.htab_info()
        check(_lib.load().htab_get_info(self._h, ctypes.byref(info)), "htab_get_info")
        self.info = info

    # ---- lifetime (unordered_base_detail.cuh:1254-1348) ----
    @classmethod
    def createDeviceObject(cls, capacity: int, tuning: Tuning | None = None):
        lib = _lib.load()
        kind = _KINDS[(cls.key, cls.mapped)]
        h = ctypes.c_void_p()
        tc = (tuning or Tuning()).to_c()
        check(lib.htab_create(kind, int(capacity), ctypes.byref(tc), ctypes.byref(h)), "htab_create")
        return cls(h.value, capacity)

    @staticmethod
    def destroyDeviceObject(obj) -> None:
        if obj._h:
            check(_lib.load().htab_destroy(obj._h), "htab_destroy")
            obj._h = ctypes.c_void_p()

    # ---- snapshot / restore (no reference counterpart) ----
    def snapshot(self, stream=None):
        """The table as bytes (numpy uint8, layout of htab_snapshot_header in htab.h): save it anywhere."""
        import numpy as np

        n = ctypes.c_int64()
        check(_lib.load().htab_snapshot_bytes(self._h, ctypes.byref(n)), "htab_snapshot_bytes")
        buf = np.empty(n.value, dtype=np.uint8)
        check(_lib.load().htab_snapshot(self._h, buf.ctypes.data_as(ctypes.c_void_p), n.value, _stream_ptr(stream)),
              "htab_snapshot")
        return buf

    @classmethod
    def restore(cls, buf):
        """A new table from snapshot() bytes (same geometry, contents and size)."""
        import numpy as np

        b = np.ascontiguousarray(buf, dtype=np.uint8)
        hd = _lib.htab_snapshot_header.from_buffer_copy(b[: ctypes.sizeof(_lib.htab_snapshot_header)].tobytes())
        if hd.kind != _KINDS[(cls.key, cls.mapped)]:
            raise TypeError("snapshot holds a different key/value kind")
        h = ctypes.c_void_p()
        check(_lib.load().htab_restore(b.ctypes.data_as(ctypes.c_void_p), b.size, ctypes.byref(h)), "htab_restore")
        return cls(h.value, hd.capacity)

    # ---- dtype helpers ----
    @classmethod
    def key_dtype(cls):
        return getattr(_torch(), _TORCH_DT[cls.key])

    @classmethod
    def mapped_dtype(cls):
        return None if cls.mapped is None else getattr(_torch(), _TORCH_DT[cls.mapped])

    def _check_keys(self, keys):
        if keys.dtype != self.key_dtype() or not keys.is_cuda or not keys.is_contiguous():
            raise TypeError(f"keys must be a contiguous CUDA tensor of {self.key_dtype()}")

    # ---- bulk operations ----
    def insert(self, keys, values=None, inserted=None, stream=None):
        """insert(begin, end): never overwrites; `inserted` (uint8, optional) receives the
        per-element bool of device insert()."""
        self._check_keys(keys)
        n = keys.numel()
        if self.mapped is not None:
            if values is None or values.numel() != n or values.dtype != self.mapped_dtype():
                raise TypeError("values must match keys and the mapped type")
        check(
            _lib.load().htab_insert_kv(self._h, _ptr(keys), _ptr(values), n, _ptr(inserted), _stream_ptr(stream)),
            "htab_insert_kv",
        )
        return inserted

    def insert_values(self, values_aos, n: int, inserted=None, stream=None):
        """insert(begin, end) over value_type entries (AoS bytes, e.g. a uint8 tensor)."""
        check(
            _lib.load().htab_insert(self._h, _ptr(values_aos), int(n), _ptr(inserted), _stream_ptr(stream)),
            "htab_insert",
        )
        return inserted

    def find(self, keys, hits=None, values_out=None, stream=None):
        """contains()/find() over a range: returns (hits uint8, values or None)."""
        torch = _torch()
        self._check_keys(keys)
        n = keys.numel()
        if hits is None:
            hits = torch.empty(n, dtype=torch.uint8, device=keys.device)
        if self.mapped is not None and values_out is None:
            values_out = torch.zeros(n, dtype=self.mapped_dtype(), device=keys.device)
        check(
            _lib.load().htab_find(self._h, _ptr(keys), n, _ptr(hits), _ptr(values_out), _stream_ptr(stream)),
            "htab_find",
        )
        return hits, values_out

    def contains(self, keys, stream=None):
        return self.find(keys, stream=stream)[0]

    def erase(self, keys, erased=None, stream=None):
        """erase(begin, end); `erased` (uint8, optional) receives device erase()'s 0/1."""
        self._check_keys(keys)
        check(
            _lib.load().htab_erase(self._h, _ptr(keys), keys.numel(), _ptr(erased), _stream_ptr(stream)), "htab_erase"
        )
        return erased

    def assign(self, keys, values, found=None, stream=None):
        """find(key)->second = value for every present key (maps; each key once per call)."""
        self._check_keys(keys)
        check(_lib.load().htab_assign(self._h, _ptr(keys), _ptr(values), keys.numel(), _ptr(found),
                                      _stream_ptr(stream)), "htab_assign")
        return found

    def assign_from_deque(self, keys, flags, deque, count=None, stream=None):
        """Free-list take (maps): every keys[i] with flags[i] == 1 (the flags a bulk insert wrote) gets
        the front element of `deque` as its value; flags[i] becomes 2 when the deque ran empty first, 3
        when keys[i] is absent.  `count` (int64 device tensor of one element, optional) receives the
        number popped.  No host synchronisation."""
        self._check_keys(keys)
        check(_lib.load().htab_assign_from_deque(self._h, _ptr(keys), _ptr(flags), keys.numel(), deque._h, _ptr(count),
                                                 _stream_ptr(stream)), "htab_assign_from_deque")
        return flags

    def erase_far_blocks(self, center, radius: float, block_size: float, deque, bits: int = 21,
                         offset: int = 1 << 20, count=None, stream=None):
        """Free-list evict (maps): erases every entry whose key, read as a packed 3-D block coordinate
        (`bits` bits per axis, x highest, each stored + `offset`), has its block centre farther than
        `radius` from `center`, and pushes its value onto the back of `deque` (stopping when the deque
        is full).  `count` (int64 device tensor of one element, optional) receives the number erased."""
        r = _lib.htab_block_region()
        r.bits, r.reserved, r.offset, r.block_size = int(bits), 0, int(offset), float(block_size)
        for a in range(3):
            r.center[a] = float(center[a])
        r.radius = float(radius)
        check(_lib.load().htab_erase_far_blocks(self._h, ctypes.byref(r), deque._h, _ptr(count), _stream_ptr(stream)),
              "htab_erase_far_blocks")

    def clear(self, stream=None):
        check(_lib.load().htab_clear(self._h, _stream_ptr(stream)), "htab_clear")

    # ---- queries ----
    def size(self, stream=None) -> int:
        out = ctypes.c_int64()
        check(_lib.load().htab_size(self._h, ctypes.byref(out), _stream_ptr(stream)), "htab_size")
        return out.value

    def empty(self) -> bool:
        return self.size() == 0

    def full(self) -> bool:
        return self.size() == self.max_size()

    def max_size(self) -> int:
        return self.info.max_size

    def bucket_count(self) -> int:
        return self.info.bucket_count

    def load_factor(self) -> float:
        # float(size) / float(bucket_count), as unordered_base_detail.cuh:1152-1158
        import numpy as np

        return float(np.float32(self.size()) / np.float32(self.bucket_count()))

    def max_load_factor(self) -> float:
        return 1.0

    def bucket(self, key: int) -> int:
        out = ctypes.c_int64()
        check(_lib.load().htab_bucket(self._h, ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF), ctypes.byref(out)), "htab_bucket")
        return out.value

    def valid(self, stream=None) -> bool:
        out = ctypes.c_int32()
        check(_lib.load().htab_valid(self._h, ctypes.byref(out), _stream_ptr(stream)), "htab_valid")
        return bool(out.value)

    def device_range(self, stream=None):
        """Slot indices of all stored entries, ascending (int64 device tensor)."""
        torch = _torch()
        cap = self.info.limit + 2
        idx = torch.empty(max(cap, 1), dtype=torch.int64, device="cuda")
        n = ctypes.c_int64()
        check(_lib.load().htab_range(self._h, _ptr(idx), cap, ctypes.byref(n), _stream_ptr(stream)), "htab_range")
        return idx[: n.value]

    def contents(self, stream=None):
        """(keys, values) of all stored entries in slot order (device tensors)."""
        torch = _torch()
        cap = self.info.limit + 2
        keys = torch.empty(max(cap, 1), dtype=self.key_dtype(), device="cuda")
        vals = None if self.mapped is None else torch.empty(max(cap, 1), dtype=self.mapped_dtype(), device="cuda")
        n = ctypes.c_int64()
        check(
            _lib.load().htab_contents(self._h, _ptr(keys), _ptr(vals), cap, ctypes.byref(n), _stream_ptr(stream)),
            "htab_contents",
        )
        return keys[: n.value], (None if vals is None else vals[: n.value])

    def view(self) -> _lib.htab_view:
        v = _lib.htab_view()
        check(_lib.load().htab_get_view(self._h, ctypes.byref(v)), "htab_get_view")
        return v


def _make(name: str, key: str, mapped: str | None):
    return type(name, (_unordered_base,), {"key": key, "mapped": mapped})


unordered_set_u32 = _make("unordered_set_u32", "u32", None)
unordered_set_u64 = _make("unordered_set_u64", "u64", None)
unordered_map_u64_u32 = _make("unordered_map_u64_u32", "u64", "u32")
unordered_map_u32_u32 = _make("unordered_map_u32_u32", "u32", "u32")
unordered_map_u64_u64 = _make("unordered_map_u64_u64", "u64", "u64")


def unordered_map(key: str = "u64", mapped: str = "u32"):
    return {("u64", "u32"): unordered_map_u64_u32, ("u32", "u32"): unordered_map_u32_u32,
            ("u64", "u64"): unordered_map_u64_u64}[(key, mapped)]


def unordered_set(key: str = "u32"):
    return {"u32": unordered_set_u32, "u64": unordered_set_u64}[key]


class _ring:
    """Shared host side of vector and deque: bulk calls over a ring of elem_bytes-byte elements.
    Elements travel as tensors of `dtype` (one element per entry) or, for records that no dtype
    covers, as uint8 tensors of shape (n, elem_bytes)."""

    _prefix = ""

    def __init__(self, capacity: int, dtype=None, elem_bytes: int | None = None):
        torch = _torch()
        self.dtype = dtype if dtype is not None else (torch.int32 if elem_bytes in (None, 4) else torch.uint8)
        item = torch.empty((), dtype=self.dtype).element_size()
        self.elem_bytes = int(elem_bytes) if elem_bytes is not None else item
        if self.elem_bytes != item and self.dtype != torch.uint8:
            raise ValueError("elem_bytes must match the dtype's size (or use uint8 records)")
        h = ctypes.c_void_p()
        fn = getattr(_lib.load(), f"htab_{self._prefix}_create_sized")
        check(fn(int(capacity), self.elem_bytes, ctypes.byref(h)), f"htab_{self._prefix}_create_sized")
        self._h = h
        self.capacity = capacity

    def _call(self, name, *args):
        check(getattr(_lib.load(), f"htab_{self._prefix}_{name}")(self._h, *args), f"htab_{self._prefix}_{name}")

    def _records(self, n):
        torch = _torch()
        item = torch.empty((), dtype=self.dtype).element_size()
        shape = (n,) if item == self.elem_bytes else (n, self.elem_bytes)
        return torch.empty(shape, dtype=self.dtype, device="cuda")

    def destroy(self):
        if self._h:
            check(getattr(_lib.load(), f"htab_{self._prefix}_destroy")(self._h), f"htab_{self._prefix}_destroy")
            self._h = ctypes.c_void_p()

    def clear(self, stream=None):
        self._call("clear", _stream_ptr(stream))

    def view(self) -> _lib.htab_ring_view:
        """device pointers for kernels that push and pop per thread (include/stdgpu_amd/ring_device.hpp)"""
        v = _lib.htab_ring_view()
        self._call("get_view", ctypes.byref(v))
        return v

    def _push(self, name, values, pushed, stream):
        n = values.shape[0] if values.dim() > 0 else 0
        self._call(name, _ptr(values), int(n), _ptr(pushed), _stream_ptr(stream))
        return pushed

    def _pop(self, name, n, out, popped, stream):
        torch = _torch()
        if out is None:
            out = self._records(n)
        if popped is None:
            popped = torch.empty(n, dtype=torch.uint8, device="cuda")
        self._call(name, _ptr(out), int(n), _ptr(popped), _stream_ptr(stream))
        return out, popped

    def size(self, stream=None) -> int:
        out = ctypes.c_int64()
        self._call("size", ctypes.byref(out), _stream_ptr(stream))
        return out.value

    def valid(self, stream=None) -> bool:
        out = ctypes.c_int32()
        self._call("valid", ctypes.byref(out), _stream_ptr(stream))
        return out.value != 0

    def contents(self, stream=None):
        """the elements front (bottom) to back (top)"""
        out = self._records(max(self.capacity, 1))
        n = ctypes.c_int64()
        self._call("contents", _ptr(out), ctypes.byref(n), _stream_ptr(stream))
        return out[: n.value]


class deque(_ring):
    """deque<T> (deque.cuh) with bulk push_back / push_front / pop_back / pop_front; a bulk call
    reserves its positions once and returns what a loop of single calls would."""

    _prefix = "deque"

    def fill_iota(self, stream=None):
        self._call("fill_iota", _stream_ptr(stream))

    def push_back(self, values, pushed=None, stream=None):
        return self._push("push_back", values, pushed, stream)

    def push_front(self, values, pushed=None, stream=None):
        return self._push("push_front", values, pushed, stream)

    def pop_front(self, n: int, out=None, popped=None, stream=None):
        return self._pop("pop_front", n, out, popped, stream)

    def pop_back(self, n: int, out=None, popped=None, stream=None):
        return self._pop("pop_back", n, out, popped, stream)


class vector(_ring):
    """vector<T> (vector.cuh) with bulk push_back / pop_back; as a stack it is the excess free list's
    container."""

    _prefix = "vector"

    def fill_iota(self, first: int = 0, stream=None):
        self._call("fill_iota", ctypes.c_uint32(int(first)), _stream_ptr(stream))

    def push_back(self, values, pushed=None, stream=None):
        return self._push("push_back", values, pushed, stream)

    def pop_back(self, n: int, out=None, popped=None, stream=None):
        return self._pop("pop_back", n, out, popped, stream)


class deque_u32(deque):
    """deque<uint32_t>, the voxel pipeline's free list of slot numbers"""

    def __init__(self, capacity: int):
        super().__init__(capacity, elem_bytes=4)


class vector_u32(vector):
    """vector<uint32_t> used as a stack"""

    def __init__(self, capacity: int):
        super().__init__(capacity, elem_bytes=4)


class bitset:
    """bitset over 32-bit blocks (bitset.cuh): bulk set / reset / flip / test of bit positions, and
    set() / reset() / flip() / count() / all() / any() / none() of the whole set."""

    _OPS = {"set": _lib.HTAB_BIT_SET, "reset": _lib.HTAB_BIT_RESET, "flip": _lib.HTAB_BIT_FLIP}

    def __init__(self, nbits: int):
        h = ctypes.c_void_p()
        check(_lib.load().htab_bitset_create(int(nbits), ctypes.byref(h)), "htab_bitset_create")
        self._h = h
        self.nbits = nbits

    @classmethod
    def createDeviceObject(cls, size: int):
        return cls(size)

    @staticmethod
    def destroyDeviceObject(obj) -> None:
        obj.destroy()

    def destroy(self):
        if self._h:
            check(_lib.load().htab_bitset_destroy(self._h), "htab_bitset_destroy")
            self._h = ctypes.c_void_p()

    def size(self) -> int:
        return self.nbits

    def update(self, op: str, positions, stream=None):
        """op in set / reset / flip for each position (int64 device tensor); returns the previous bits."""
        torch = _torch()
        old = torch.empty(positions.numel(), dtype=torch.uint8, device=positions.device)
        check(_lib.load().htab_bitset_update(self._h, self._OPS[op], _ptr(positions), positions.numel(), _ptr(old),
                                             None, _stream_ptr(stream)), "htab_bitset_update")
        return old

    def test(self, positions, stream=None):
        torch = _torch()
        out = torch.empty(positions.numel(), dtype=torch.uint8, device=positions.device)
        check(_lib.load().htab_bitset_test(self._h, _ptr(positions), positions.numel(), _ptr(out),
                                           _stream_ptr(stream)), "htab_bitset_test")
        return out

    def set(self, stream=None):
        check(_lib.load().htab_bitset_all_bits(self._h, _lib.HTAB_BIT_SET, _stream_ptr(stream)), "htab_bitset_all_bits")

    def reset(self, stream=None):
        check(_lib.load().htab_bitset_all_bits(self._h, _lib.HTAB_BIT_RESET, _stream_ptr(stream)),
              "htab_bitset_all_bits")

    def flip(self, stream=None):
        check(_lib.load().htab_bitset_all_bits(self._h, _lib.HTAB_BIT_FLIP, _stream_ptr(stream)),
              "htab_bitset_all_bits")

    def count(self, stream=None) -> int:
        out = ctypes.c_int64()
        check(_lib.load().htab_bitset_count(self._h, ctypes.byref(out), _stream_ptr(stream)), "htab_bitset_count")
        return out.value

    def all(self) -> bool:
        return self.count() == self.nbits

    def any(self) -> bool:
        return self.count() > 0

    def none(self) -> bool:
        return self.count() == 0


_ATOMIC_TYPES = {
    "int": _lib.HTAB_ATOMIC_I32, "i32": _lib.HTAB_ATOMIC_I32,
    "unsigned int": _lib.HTAB_ATOMIC_U32, "u32": _lib.HTAB_ATOMIC_U32,
    "unsigned long long int": _lib.HTAB_ATOMIC_U64, "u64": _lib.HTAB_ATOMIC_U64,
    "float": _lib.HTAB_ATOMIC_F32, "f32": _lib.HTAB_ATOMIC_F32,
}
_ATOMIC_CTYPES = {_lib.HTAB_ATOMIC_I32: ctypes.c_int32, _lib.HTAB_ATOMIC_U32: ctypes.c_uint32,
                  _lib.HTAB_ATOMIC_U64: ctypes.c_uint64, _lib.HTAB_ATOMIC_F32: ctypes.c_float}


def _atomic_torch_dtype(code: int):
    """device tensors of the values: int32 (also for unsigned int bits), int64 (unsigned long long bits), float32"""
    torch = _torch()
    return {_lib.HTAB_ATOMIC_I32: torch.int32, _lib.HTAB_ATOMIC_U32: torch.int32,
            _lib.HTAB_ATOMIC_U64: torch.int64, _lib.HTAB_ATOMIC_F32: torch.float32}[code]


def _atomic_bulk(call, name, code, op, operands, targets, expected, stream):
    torch = _torch()
    opc = _lib.HTAB_ATOMIC_OPS[op]
    ref = operands if operands is not None else targets
    n = ref.numel() if ref is not None else 0
    dev = ref.device if ref is not None else "cuda"
    old = None if op == "store" else torch.empty(n, dtype=_atomic_torch_dtype(code), device=dev)
    done = torch.empty(n, dtype=torch.uint8, device=dev)
    check(call(opc, _ptr(targets), _ptr(operands), _ptr(expected), n, _ptr(old), _ptr(done), _stream_ptr(stream)),
          name)
    return old, done


def atomic_ref_apply(type: str, values, op: str, operands=None, targets=None, expected=None, stream=None):
    """atomic_ref<T> over the elements of a device tensor: the bulk form of op (see atomic.apply)."""
    code = _ATOMIC_TYPES[type]
    lib = _lib.load()
    return _atomic_bulk(lambda *a: lib.htab_atomic_ref_apply(code, _ptr(values), values.numel(), *a),
                        "htab_atomic_ref_apply", code, op, operands, targets, expected, stream)


class atomic:
    """atomic<T> (atomic.cuh) for T in int / unsigned int / unsigned long long int / float: count device
    values (count = 1 is one atomic<T>), bulk device operations, host load() / store().

    apply(op, operands, targets, expected) applies operation i to element targets[i] (element 0 when
    targets is None) and returns (old, done): the values replaced, and 1 per operation that applied
    (target in range; for compare_exchange, the value equalled expected[i] and was replaced).  Operands
    are device tensors of the value type's bits (int32 for int and unsigned int, int64 for unsigned
    long long, float32 for float)."""

    def __init__(self, type: str = "unsigned int", count: int = 1):
        self.type_code = _ATOMIC_TYPES[type]
        h = ctypes.c_void_p()
        check(_lib.load().htab_atomic_create(self.type_code, int(count), ctypes.byref(h)), "htab_atomic_create")
        self._h = h
        self.count = count

    @classmethod
    def createDeviceObject(cls, type: str = "unsigned int", count: int = 1):
        return cls(type, count)

    @staticmethod
    def destroyDeviceObject(obj) -> None:
        obj.destroy()

    def destroy(self):
        if self._h:
            check(_lib.load().htab_atomic_destroy(self._h), "htab_atomic_destroy")
            self._h = ctypes.c_void_p()

    @property
    def dtype(self):
        return _atomic_torch_dtype(self.type_code)

    def data_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(_lib.load().htab_atomic_data(self._h, ctypes.byref(p)), "htab_atomic_data")
        return int(p.value)

    def load(self, index: int = 0, stream=None):
        v = _ATOMIC_CTYPES[self.type_code]()
        check(_lib.load().htab_atomic_load(self._h, int(index), ctypes.byref(v), _stream_ptr(stream)),
              "htab_atomic_load")
        return v.value

    def store(self, value, index: int = 0, stream=None):
        v = _ATOMIC_CTYPES[self.type_code](value)
        check(_lib.load().htab_atomic_store(self._h, int(index), ctypes.byref(v), _stream_ptr(stream)),
              "htab_atomic_store")

    def apply(self, op: str, operands=None, targets=None, expected=None, stream=None):
        lib = _lib.load()
        return _atomic_bulk(lambda *a: lib.htab_atomic_apply(self._h, *a), "htab_atomic_apply", self.type_code, op,
                            operands, targets, expected, stream)

    def bulk_load(self, targets, stream=None):
        return self.apply("load", None, targets, stream=stream)[0]

    def exchange(self, operands, targets=None, stream=None):
        return self.apply("exchange", operands, targets, stream=stream)

    def compare_exchange(self, expected, desired, targets=None, stream=None):
        return self.apply("compare_exchange", desired, targets, expected, stream=stream)

    def fetch_add(self, operands, targets=None, stream=None):
        return self.apply("fetch_add", operands, targets, stream=stream)

    def fetch_sub(self, operands, targets=None, stream=None):
        return self.apply("fetch_sub", operands, targets, stream=stream)

    def fetch_min(self, operands, targets=None, stream=None):
        return self.apply("fetch_min", operands, targets, stream=stream)

    def fetch_max(self, operands, targets=None, stream=None):
        return self.apply("fetch_max", operands, targets, stream=stream)

    def fetch_and(self, operands, targets=None, stream=None):
        return self.apply("fetch_and", operands, targets, stream=stream)

    def fetch_or(self, operands, targets=None, stream=None):
        return self.apply("fetch_or", operands, targets, stream=stream)

    def fetch_xor(self, operands, targets=None, stream=None):
        return self.apply("fetch_xor", operands, targets, stream=stream)

    def fetch_inc_mod(self, operands, targets=None, stream=None):
        return self.apply("fetch_inc_mod", operands, targets, stream=stream)

    def fetch_dec_mod(self, operands, targets=None, stream=None):
        return self.apply("fetch_dec_mod", operands, targets, stream=stream)
