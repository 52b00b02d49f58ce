"""ctypes binding of oracle/liboracle.so, the CPU restatement of stdgpu's hash-table engine
(oracle/htab_oracle.cpp).  Test infrastructure only: the parity checker and the timed CPU
baseline; the product never calls it.

Pieces and what pins each (tests/test_oracle.py):
  * Table (sizing, bucket(), insert / find / erase, valid): reference KATs, exact closed forms,
    the reference tests' colliding keys, brute force against Python sets and dicts, tiny-table
    capacity cases, the examples' closed-form sums;
  * Deque, Vector (the free lists): multiset conservation, FIFO / LIFO order, full and empty;
  * voxel_insert (the reference's caller pattern, not library code): brute force against a
    Python deque and dict, exact key -> slot map and free-list order;
  * atomic_step / atomic_sequence / atomic_linearize (pure Python): the reference atomic tests'
    closed forms, wrap-around of the modular increment / decrement, and a planted tamper that
    the linearisation check must reject.
No function here is left without such a check.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KIND = {("u32", None): 0, ("u64", None): 1, ("u64", "u32"): 2, ("u32", "u32"): 3, ("u64", "u64"): 4}
NP = {"u32": np.uint32, "u64": np.uint64}

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = ctypes.CDLL(LIB_PATH)
    P, I64, I32, U64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
    sig = {
        "orc_create": (P, [I32, I64]),
        "orc_destroy": (None, [P]),
        "orc_insert": (None, [P, P, P, I64, P, I32]),
        "orc_find": (None, [P, P, I64, P, P, I32]),
        "orc_erase": (None, [P, P, I64, P, I32]),
        "orc_size": (I64, [P]),
        "orc_contents": (I64, [P, P, P]),
        "orc_valid": (I32, [P]),
        "orc_clear": (None, [P]),
        "orc_corrupt": (I32, [P, I32]),
        "orc_bucket_count": (I64, [P]),
        "orc_max_size": (I64, [P]),
        "orc_excess_free": (I64, [P]),
        "orc_bucket": (I64, [P, U64]),
        "orc_fibonacci_hashing": (I64, [U64, I64]),
        "orc_sizing": (None, [I64, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "orc_vector_create": (P, [I64]),
        "orc_vector_destroy": (None, [P]),
        "orc_vector_size": (I64, [P]),
        "orc_vector_fill_iota": (None, [P, ctypes.c_uint32]),
        "orc_vector_push_back": (None, [P, P, I64, P, I32]),
        "orc_vector_pop_back": (None, [P, P, I64, P, I32]),
        "orc_vector_contents": (I64, [P, P]),
        "orc_deque_create": (P, [I64]),
        "orc_deque_destroy": (None, [P]),
        "orc_deque_size": (I64, [P]),
        "orc_deque_push_back": (None, [P, P, I64, P, I32]),
        "orc_deque_pop_front": (None, [P, P, I64, P, I32]),
        "orc_deque_push_front": (None, [P, P, I64, P, I32]),
        "orc_deque_pop_back": (None, [P, P, I64, P, I32]),
        "orc_deque_contents": (I64, [P, P]),
        "orc_voxel_insert": (I64, [P, P, P, I64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def fibonacci_hashing(h: int, bucket_count: int) -> int:
    return load().orc_fibonacci_hashing(ctypes.c_uint64(h & 0xFFFFFFFFFFFFFFFF), bucket_count)


def sizing(capacity: int):
    bc, ex = ctypes.c_int64(), ctypes.c_int64()
    load().orc_sizing(capacity, ctypes.byref(bc), ctypes.byref(ex))
    return bc.value, ex.value


class Table:
    """unordered_set<K> / unordered_map<K, M> on the CPU, reference algorithm."""

    def __init__(self, key: str, mapped: str | None, capacity: int):
        self.key, self.mapped = key, mapped
        self._t = load().orc_create(KIND[(key, mapped)], capacity)
        if not self._t:
            raise ValueError("oracle: bad kind or capacity")

    def close(self):
        if self._t:
            load().orc_destroy(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert(self, keys, vals=None, threads: int = 1):
        keys = np.ascontiguousarray(keys, dtype=NP[self.key])
        if self.mapped is not None:
            vals = np.ascontiguousarray(vals, dtype=NP[self.mapped])
        ok = np.zeros(len(keys), dtype=np.uint8)
        load().orc_insert(self._t, _p(keys), _p(vals), len(keys), _p(ok), threads)
        return ok

    def find(self, keys, threads: int = 1):
        keys = np.ascontiguousarray(keys, dtype=NP[self.key])
        hits = np.zeros(len(keys), dtype=np.uint8)
        vals = None if self.mapped is None else np.zeros(len(keys), dtype=NP[self.mapped])
        load().orc_find(self._t, _p(keys), len(keys), _p(hits), _p(vals), threads)
        return hits, vals

    def erase(self, keys, threads: int = 1):
        keys = np.ascontiguousarray(keys, dtype=NP[self.key])
        ok = np.zeros(len(keys), dtype=np.uint8)
        load().orc_erase(self._t, _p(keys), len(keys), _p(ok), threads)
        return ok

    def size(self) -> int:
        return load().orc_size(self._t)

    def contents(self):
        n = self.size()
        keys = np.zeros(max(n, 1), dtype=NP[self.key])
        vals = None if self.mapped is None else np.zeros(max(n, 1), dtype=NP[self.mapped])
        c = load().orc_contents(self._t, _p(keys), _p(vals))
        assert c == n, (c, n)
        return keys[:c], (None if vals is None else vals[:c])

    def sorted_contents(self):
        k, v = self.contents()
        o = np.argsort(k, kind="stable")
        return k[o], (None if v is None else v[o])

    def valid(self) -> bool:
        return bool(load().orc_valid(self._t))

    def clear(self):
        load().orc_clear(self._t)

    # test-only: break one valid() invariant (0 looping chain, 1 offset out of range,
    # 2 unreachable entry, 3 duplicate key, 4 size above the occupied count, 5 lock held,
    # 6 excess list size above its occupancy); False when the table lacks the structure needed
    CORRUPTIONS = ("loop", "offset_out_of_range", "unreachable", "duplicate", "count", "lock", "excess_list")

    def corrupt(self, what: str) -> bool:
        return bool(load().orc_corrupt(self._t, self.CORRUPTIONS.index(what)))

    def bucket_count(self) -> int:
        return load().orc_bucket_count(self._t)

    def max_size(self) -> int:
        return load().orc_max_size(self._t)

    def excess_free(self) -> int:
        return load().orc_excess_free(self._t)

    def bucket(self, key: int) -> int:
        return load().orc_bucket(self._t, ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF))


def voxel_insert(table: "Table", deque: "Deque", keys) -> int:
    """The reference caller pattern: per key pop a slot, insert (key, slot), push back if present."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    return load().orc_voxel_insert(table._t, deque._d, _p(k), len(k))


class Vector:
    """vector<uint32_t> used as a stack on the CPU, reference algorithm (push_back / pop_back)."""

    def __init__(self, capacity: int):
        self._v = load().orc_vector_create(capacity)
        self.capacity = capacity

    def close(self):
        if self._v:
            load().orc_vector_destroy(self._v)
            self._v = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push_back(self, values, threads: int = 1):
        v = np.ascontiguousarray(values, dtype=np.uint32)
        ok = np.zeros(len(v), dtype=np.uint8)
        load().orc_vector_push_back(self._v, _p(v), len(v), _p(ok), threads)
        return ok

    def pop_back(self, n: int, threads: int = 1):
        out = np.zeros(n, dtype=np.uint32)
        ok = np.zeros(n, dtype=np.uint8)
        load().orc_vector_pop_back(self._v, _p(out), n, _p(ok), threads)
        return out, ok

    def fill_iota(self, first: int = 0):
        load().orc_vector_fill_iota(self._v, first)

    def size(self) -> int:
        return load().orc_vector_size(self._v)

    def contents(self):
        out = np.zeros(max(self.capacity, 1), dtype=np.uint32)
        n = load().orc_vector_contents(self._v, _p(out))
        return out[:n]


class Deque:
    """deque<uint32_t> on the CPU, reference algorithm (push_back / pop_front / push_front / pop_back)."""

    def __init__(self, capacity: int):
        self._d = load().orc_deque_create(capacity)
        self.capacity = capacity

    def close(self):
        if self._d:
            load().orc_deque_destroy(self._d)
            self._d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push_back(self, values, threads: int = 1):
        v = np.ascontiguousarray(values, dtype=np.uint32)
        ok = np.zeros(len(v), dtype=np.uint8)
        load().orc_deque_push_back(self._d, _p(v), len(v), _p(ok), threads)
        return ok

    def pop_front(self, n: int, threads: int = 1):
        out = np.zeros(n, dtype=np.uint32)
        ok = np.zeros(n, dtype=np.uint8)
        load().orc_deque_pop_front(self._d, _p(out), n, _p(ok), threads)
        return out, ok

    def push_front(self, values, threads: int = 1):
        v = np.ascontiguousarray(values, dtype=np.uint32)
        ok = np.zeros(len(v), dtype=np.uint8)
        load().orc_deque_push_front(self._d, _p(v), len(v), _p(ok), threads)
        return ok

    def pop_back(self, n: int, threads: int = 1):
        out = np.zeros(n, dtype=np.uint32)
        ok = np.zeros(n, dtype=np.uint8)
        load().orc_deque_pop_back(self._d, _p(out), n, _p(ok), threads)
        return out, ok

    def size(self) -> int:
        return load().orc_deque_size(self._d)

    def contents(self):
        out = np.zeros(max(self.capacity, 1), dtype=np.uint32)
        n = load().orc_deque_contents(self._d, _p(out))
        return out[:n]


# ------------------------------------------------------------------ atomic<T>
# Sequential statement of the reference's atomic operations: the OpenMP backend writes each as a
# critical section over the value (src/stdgpu/openmp/impl/atomic_detail.h:40-217), and the public
# fetch_inc_mod(arg) / fetch_dec_mod(arg) pass arg - 1 as the backend's limit
# (src/stdgpu/impl/atomic_detail.cuh:679-702).  Pure Python on small inputs.

ATOMIC_BITS = {"int": 32, "unsigned int": 32, "unsigned long long int": 64}


def _wrap(type_: str, v):
    if type_ == "float":
        return float(np.float32(v))
    bits = ATOMIC_BITS[type_]
    v &= (1 << bits) - 1
    if type_ == "int" and v >= 1 << 31:
        v -= 1 << 32
    return v


def atomic_step(type_: str, op: str, state, arg=0, expected=0):
    """(old, new state, replaced) of one operation on a value holding state."""
    old = state
    if op == "load":
        return old, state, True
    if op in ("store", "exchange"):
        return old, _wrap(type_, arg), True
    if op == "compare_exchange":  # atomic_detail.h:71-81
        hit = old == expected
        return old, (_wrap(type_, arg) if hit else old), hit
    if type_ == "float":
        f = np.float32
        new = {"fetch_add": lambda: f(f(old) + f(arg)), "fetch_sub": lambda: f(f(old) - f(arg)),
               "fetch_min": lambda: f(arg) if f(arg) < f(old) else f(old),  # std::min(old, arg)
               "fetch_max": lambda: f(arg) if f(old) < f(arg) else f(old)}[op]()
        return old, float(new), True
    if op == "fetch_add":
        new = old + arg
    elif op == "fetch_sub":
        new = old - arg
    elif op == "fetch_min":
        new = arg if arg < old else old
    elif op == "fetch_max":
        new = arg if old < arg else old
    elif op == "fetch_and":
        new = old & arg
    elif op == "fetch_or":
        new = old | arg
    elif op == "fetch_xor":
        new = old ^ arg
    elif op == "fetch_inc_mod":  # atomic_detail.h:195-204 with limit = arg - 1
        limit = _wrap(type_, arg - 1)
        new = 0 if old >= limit else old + 1
    elif op == "fetch_dec_mod":  # atomic_detail.h:208-217 with limit = arg - 1
        limit = _wrap(type_, arg - 1)
        new = limit if (old == 0 or old > limit) else old - 1
    else:
        raise ValueError(op)
    return old, _wrap(type_, new), True


def atomic_sequence(type_: str, op: str, initial, operands, expected=None):
    """final value and the old values of the operations applied one after another, in index order"""
    s = initial
    olds = []
    for i, a in enumerate(operands):
        old, s, _ = atomic_step(type_, op, s, a, 0 if expected is None else expected[i])
        olds.append(old)
    return s, olds


def atomic_linearize(type_: str, op: str, initial, operands, olds, expected=None):
    """The final value of some one-at-a-time order of the operations on one value in which operation i
    sees olds[i]; raises AssertionError when no order is found.  Greedy: from the current state, take
    an operation that saw it, preferring one that leaves the state unchanged.  Exact when at most one
    operation that changes the state sees each state (distinct addends, distinct exchange operands,
    monotone min / max / and / or), or when the operations that see one state are equal."""
    by_old = {}
    for i, o in enumerate(olds):
        by_old.setdefault(o, []).append(i)
    s = initial
    for _ in range(len(olds)):
        cand = by_old.get(s)
        assert cand, f"no operation saw the value {s}"
        pick = None
        for j in range(len(cand) - 1, -1, -1):
            i = cand[j]
            _, ns, _ = atomic_step(type_, op, s, operands[i], 0 if expected is None else expected[i])
            if ns == s:
                pick = j
                break
        i = cand.pop(len(cand) - 1 if pick is None else pick)
        _, s, _ = atomic_step(type_, op, s, operands[i], 0 if expected is None else expected[i])
    return s
