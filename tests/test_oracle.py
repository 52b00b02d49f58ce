"""Pins the CPU oracle (oracle/htab_oracle.cpp) before it is trusted as the parity checker.

Each check is chosen so that a plausible slip in the restatement fails it:
  * sizing and bucket() values printed by a run of the compiled reference (OpenMP backend),
    and an independent Python restatement of the closed forms;
  * the reference tests' pinned colliding / non-colliding key pairs
    (tests/stdgpu/unordered_datastructure.inc:899-934, vec_hash unordered_map.inc:181-206);
  * brute-force set semantics against Python sets, single- and multi-threaded;
  * the reference's valid() invariants after every step, and valid() false for a table with each
    invariant broken on its own (a test-only corruption hook); clear() against a fresh table;
  * the tiny-table capacity cases (unordered_datastructure.inc:1375-1739);
  * the examples' closed-form sums (examples/openmp/unordered_map.cpp:99-108,
    unordered_set.cpp:82-87);
  * multiset conservation of the deque free list, stack order of the vector;
  * the voxel caller pattern (pop, insert, push back) against a pure-Python statement of it;
  * the reference atomic tests' closed forms (tests/stdgpu/atomic.inc) for the sequential
    atomic model.
"""
from decimal import Decimal, getcontext
import math

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = 11400714819323198485


def py_fibonacci(h: int, bc: int) -> int:
    """unordered_base_detail.cuh:56-85, restated with Python integers."""
    w = bc.bit_length() - 1
    if w <= 0:
        return 0
    d = 64 - w
    h2 = h ^ (h >> d)
    return ((GOLD * h2) & ((1 << 64) - 1)) >> d


def py_sizing(capacity: int):
    """unordered_base_detail.cuh:34-48, 1263-1273 with exact decimal arithmetic."""
    want = math.ceil(float(np.float32(capacity)))
    bc = 1
    while bc < want:
        bc <<= 1
    getcontext().prec = 60
    k, n = Decimal(bc), Decimal(capacity)
    p = Decimal(1) if capacity == 1 else (1 - 1 / k) ** (int(capacity) - 1)
    ec = n * (1 - p)
    return bc, int(ec.to_integral_value(rounding="ROUND_CEILING")) + 1


# values printed by a run of the compiled reference (stdgpu OpenMP backend)
REFERENCE_KATS = [
    # kind, capacity, bucket_count, max_size, {key: bucket}
    (("u32", None), 1 << 21, 2097152, 3422806, {1: 1296111, 2: 495070, 3: 1791181, 0xFFFFFFFE: 547697}),
    (("u64", "u32"), 100000, 131072, 184444,
     {1: 81006, 2: 30941, 3: 111948, 0xFFFFFFFE: 34231, 0x0123456789ABCDEF: 28855}),
    (("u64", "u32"), 1, 1, 2, {}),
]


@pytest.mark.parametrize("kind,cap,bc,mx,buckets", REFERENCE_KATS)
def test_reference_kats(kind, cap, bc, mx, buckets):
    t = po.Table(kind[0], kind[1], cap)
    assert t.bucket_count() == bc
    assert t.max_size() == mx
    for k, b in buckets.items():
        assert t.bucket(k) == b
        assert py_fibonacci(k, bc) == b


@pytest.mark.parametrize("cap", [1, 2, 3, 7, 100, 1000, 100000, 1 << 20, 1 << 21, (1 << 21) + 1, 83886080,
                                 5_000_000, 12_345_678])
def test_sizing_closed_form(cap):
    bc, ex = po.sizing(cap)
    pbc, pex = py_sizing(cap) if cap <= 1 << 21 else (None, None)
    assert bc == 1 << (cap - 1).bit_length()
    if pbc is not None:
        assert (bc, ex) == (pbc, pex)


def test_sizing_config_capacities():
    # capacity for 2^26 keys at load 0.8 -> the reference's power-of-two buckets give load 0.5
    bc, ex = po.sizing(83886080)
    assert bc == 1 << 27
    assert bc + ex == 173202826


def test_fibonacci_random_against_python():
    rng = np.random.default_rng(5)
    for bc in [1, 2, 4, 1 << 10, 1 << 17, 1 << 21, 1 << 27, 1 << 30]:
        for h in rng.integers(0, 2**63, size=200, dtype=np.int64):
            hh = int(h) * 2 + 1
            assert po.fibonacci_hashing(hh, bc) == py_fibonacci(hh, bc)


def vec_hash(x: int, y: int, z: int) -> int:
    """unordered_map.inc:181-206: int16 components sign-extended to size_t."""
    m = (1 << 64) - 1
    return ((x & m) * 73856093 ^ (y & m) * 19349669 ^ (z & m) * 83492791) & m


def test_reference_test_collision_pairs():
    bc, _ = po.sizing(100000)  # the test fixture's table (unordered_datastructure.inc:49-70)
    b = lambda v: po.fibonacci_hashing(vec_hash(*v), bc)  # noqa: E731
    # insert_collision asserts these pairs share a bucket
    assert b((-7, -3, 15)) == b((7, 3, 15))
    assert b((-5, -15, 13)) == b((5, 15, 13))
    # insert_no_collision relies on these two not colliding
    assert b((-7, -3, 15)) != b((-5, -15, 13))


KINDS = [("u32", None), ("u64", None), ("u64", "u32"), ("u32", "u32"), ("u64", "u64")]


def _rand_keys(rng, key, n, universe=None):
    if key == "u32":
        hi = universe or (1 << 32)
        return rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    if universe:
        return rng.integers(0, universe, size=n, dtype=np.uint64)
    return rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n, dtype=np.uint64)


def _val(key, keys, mapped):
    if mapped is None:
        return None
    v = (keys.astype(np.uint64) * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)
    return v.astype(np.uint32 if mapped == "u32" else np.uint64)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("threads", [1, 4])
def test_set_semantics_bruteforce(kind, threads):
    key, mapped = kind
    rng = np.random.default_rng(11 + threads)
    cap = 3000
    t = po.Table(key, mapped, cap)
    ref = {}
    for rnd in range(4):
        ins = _rand_keys(rng, key, 1500, universe=4000)
        vals = _val(key, ins, mapped)
        ok = t.insert(ins, vals, threads=threads)
        # exactly one success per distinct new key
        new = {}
        for i, k in enumerate(ins.tolist()):
            if k not in ref and k not in new:
                new[k] = None if vals is None else int(vals[i])
        for k in set(ins.tolist()):
            idx = np.nonzero(ins == k)[0]
            assert ok[idx].sum() == (1 if k in new else 0)
        ref.update(new)
        assert t.size() == len(ref)
        assert t.valid()
        q = _rand_keys(rng, key, 2000, universe=4000)
        hits, got = t.find(q, threads=threads)
        assert hits.tolist() == [1 if k in ref else 0 for k in q.tolist()]
        if mapped is not None:
            for k, h, v in zip(q.tolist(), hits.tolist(), got.tolist()):
                if h:
                    assert v == ref[k]
        er = _rand_keys(rng, key, 700, universe=4000)
        eok = t.erase(er, threads=threads)
        gone = set()
        for k in set(er.tolist()):
            idx = np.nonzero(er == k)[0]
            assert eok[idx].sum() == (1 if k in ref else 0)
            if k in ref:
                gone.add(k)
        for k in gone:
            del ref[k]
        assert t.size() == len(ref)
        assert t.valid()
        sk, sv = t.sorted_contents()
        assert sk.tolist() == sorted(ref)
        if mapped is not None:
            assert sv.tolist() == [ref[k] for k in sorted(ref)]


def _filled_with_chains(kind, cap=64, n=60, seed=7):
    """a small table whose buckets carry chains into the excess area"""
    key, mapped = kind
    t = po.Table(key, mapped, cap)
    rng = np.random.default_rng(seed)
    keys = np.unique(_rand_keys(rng, key, 4 * n))[:n]
    assert t.insert(keys, _val(key, keys, mapped)).sum() == n
    assert t.valid()
    return t, keys


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("what", po.Table.CORRUPTIONS)
def test_valid_rejects_each_broken_invariant(kind, what):
    """unordered_base_detail.cuh:1192-1218: each of the reference's checks, broken on its own
    (a looping chain, an offset out of range, an unreachable entry, a duplicate key, a size above
    the occupied count, a held lock, an inconsistent excess list), makes valid() false."""
    t, _ = _filled_with_chains(kind)
    assert t.corrupt(what), "table lacks the structure this corruption needs"
    assert not t.valid()


@pytest.mark.parametrize("what", ["unreachable", "duplicate"])
def test_valid_rejects_corruption_the_earlier_checks_pass(what):
    """An unreachable entry keeps size == the bit count, and a duplicate keeps every entry
    reachable: only the reachability and uniqueness checks can reject them.  (Reachability is
    implied by uniqueness in the reference too: find() of an unreachable key returns end(), whose
    position differs from the entry's, so no corruption fails one check and passes the other.)"""
    t, keys = _filled_with_chains(("u64", "u32"))
    assert t.corrupt(what)
    k, _ = t.contents()
    assert len(k) == t.size()  # the occupied count still equals the number of occupied entries
    if what == "duplicate":
        h, _ = t.find(k)
        assert h.all()  # every stored key is still reachable
    assert not t.valid()


@pytest.mark.parametrize("kind", KINDS)
def test_clear_restores_a_fresh_table(kind):
    """unordered_base_detail.cuh:1220-1252: clear() empties the table, resets the excess list to
    all of [bucket_count, max_size), and the table behaves as a fresh one afterwards."""
    key, mapped = kind
    t, keys = _filled_with_chains(kind, cap=256, n=200, seed=9)
    t.erase(keys[::3])
    assert t.excess_free() < t.max_size() - t.bucket_count()
    t.clear()
    assert t.size() == 0 and t.valid()
    assert len(t.contents()[0]) == 0
    assert t.excess_free() == t.max_size() - t.bucket_count()
    h, _ = t.find(keys)
    assert h.sum() == 0
    # a later insert equals a fresh table's (identical layout for a sequential insert)
    fresh = po.Table(key, mapped, 256)
    rng = np.random.default_rng(10)
    more = np.unique(_rand_keys(rng, key, 400))[:220]
    vals = _val(key, more, mapped)
    assert t.insert(more, vals).tolist() == fresh.insert(more, vals).tolist()
    a, av = t.sorted_contents()
    b, bv = fresh.sorted_contents()
    np.testing.assert_array_equal(a, b)
    if mapped is not None:
        np.testing.assert_array_equal(av, bv)
    assert t.size() == fresh.size() and t.excess_free() == fresh.excess_free() and t.valid()


def test_tiny_table_capacity_cases():
    # unordered_datastructure.inc:1375-1443: createDeviceObject(1) holds exactly two keys
    t = po.Table("u64", "u32", 1)
    assert (t.bucket_count(), t.max_size()) == (1, 2)
    assert t.insert(np.array([10], np.uint64), np.array([1], np.uint32)).tolist() == [1]
    assert t.insert(np.array([20], np.uint64), np.array([2], np.uint32)).tolist() == [1]
    assert t.size() == 2 and t.valid()
    assert t.insert(np.array([30], np.uint64), np.array([3], np.uint32)).tolist() == [0]
    assert t.size() == 2 and t.valid()
    # insert_multiple_while_full: many threads, same key, none inserted
    ok = t.insert(np.full(1000, 30, np.uint64), np.full(1000, 3, np.uint32), threads=8)
    assert ok.sum() == 0 and t.size() == 2


def test_tiny_table_parallel_one_free():
    # unordered_datastructure.inc:1549-1603: one free place, many random keys -> exactly one wins
    t = po.Table("u64", "u32", 1)
    t.insert(np.array([123], np.uint64), np.array([1], np.uint32))
    rng = np.random.default_rng(3)
    keys = rng.integers(1, 2**40, size=5000, dtype=np.uint64)
    ok = t.insert(keys, np.ones(5000, np.uint32), threads=8)
    assert ok.sum() == 1
    assert t.size() == 2 and t.valid()


def test_excess_list_exhaustion_refuses_free_head():
    # unordered_base_detail.cuh:980-981: an empty excess list refuses every insert, even one
    # whose bucket head is free
    t = po.Table("u64", None, 4)  # bc 4, excess = ceil(4(1-(3/4)^3)) + 1 = 4
    bc = t.bucket_count()
    ex = t.max_size() - bc
    same = [k for k in range(1, 100000) if t.bucket(k) == 0][: ex + 1]
    ok = t.insert(np.array(same, np.uint64))
    assert ok.sum() == ex + 1  # one at the head, ex in the excess list
    assert t.excess_free() == 0
    other = next(k for k in range(1, 100000) if t.bucket(k) == 1)
    assert t.insert(np.array([other], np.uint64)).tolist() == [0]


def test_examples_closed_form_sums():
    # examples/openmp/unordered_set.cpp:60-87 and unordered_map.cpp:70-108
    n = 100
    odd = np.arange(1, n + 1)[np.arange(1, n + 1) % 2 == 1]
    neigh = np.concatenate([odd - 1, odd, odd + 1]).astype(np.uint32)
    s = po.Table("u32", None, n)
    s.insert(neigh, threads=4)
    k, _ = s.contents()
    assert s.size() == n + 1 and int(k.astype(np.int64).sum()) == n * (n + 1) // 2
    m = po.Table("u32", "u32", n)
    m.insert(neigh, (neigh.astype(np.int64) ** 2).astype(np.uint32), threads=4)
    k, v = m.contents()
    assert m.size() == n + 1
    assert int(k.astype(np.int64).sum()) == n * (n + 1) // 2
    assert int(v.astype(np.int64).sum()) == n * (n + 1) * (2 * n + 1) // 6


def test_deque_multiset_conservation():
    cap = 1 << 12
    q = po.Deque(cap)
    q.push_back(np.arange(cap, dtype=np.uint32))
    assert q.size() == cap
    assert q.push_back(np.array([7], np.uint32)).tolist() == [0]  # full
    out, ok = q.pop_front(1000, threads=4)
    assert ok.all()
    rest = q.contents()
    assert sorted(np.concatenate([out, rest]).tolist()) == list(range(cap))
    # FIFO order single-threaded
    out2, _ = q.pop_front(10)
    assert out2.tolist() == rest[:10].tolist()
    q.push_back(out)
    assert q.size() == cap - 10
    got, ok = q.pop_front(cap)
    assert ok.sum() == cap - 10
    assert sorted(np.concatenate([got[ok == 1], out2]).tolist()) == list(range(cap))


def test_deque_both_ends_equal_a_plain_deque():
    """push_front / pop_back (deque_detail.cuh:262-388, positions from atomicDec-style wraps of begin
    and end) with push_back / pop_front, single-threaded: every result, refusal and the contents front
    to back equal a bounded double-ended queue (collections.deque), also across the ring's ends"""
    import collections

    rng = np.random.default_rng(5)
    cap = 37
    q = po.Deque(cap)
    model = collections.deque()
    for step in range(400):
        op = int(rng.integers(0, 4))
        n = int(rng.integers(0, 25))
        if op < 2:
            vals = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
            ok = (q.push_back if op == 0 else q.push_front)(vals)
            want = []
            for x in vals:
                if len(model) < cap:
                    (model.append if op == 0 else model.appendleft)(int(x))
                    want.append(1)
                else:
                    want.append(0)
            assert ok.tolist() == want, step
        else:
            out, ok = (q.pop_back if op == 2 else q.pop_front)(n)
            for i in range(n):
                if model:
                    assert ok[i] == 1 and int(out[i]) == (model.pop() if op == 2 else model.popleft()), step
                else:
                    assert ok[i] == 0 and out[i] == 0xFFFFFFFF
        assert q.size() == len(model)
        assert q.contents().tolist() == list(model), step


def test_vector_stack_semantics():
    """vector<uint32_t> as the excess stack (vector_detail.cuh:181-302, 361-368): LIFO order single-
    threaded, refusal when full / empty, clamped size, iota fill, multiset conservation with threads."""
    cap = 1 << 12
    v = po.Vector(cap)
    assert v.size() == 0
    out, ok = v.pop_back(3)
    assert ok.tolist() == [0, 0, 0] and out.tolist() == [0xFFFFFFFF] * 3  # empty
    assert v.push_back(np.arange(10, dtype=np.uint32)).all()
    assert v.contents().tolist() == list(range(10))
    out, ok = v.pop_back(4)
    assert ok.all() and out.tolist() == [9, 8, 7, 6]  # last in, first out
    assert v.size() == 6
    pushed = v.push_back(np.arange(100, 100 + cap, dtype=np.uint32))
    assert pushed.sum() == cap - 6 and v.size() == cap  # full: the rest is refused
    assert v.push_back(np.array([5], np.uint32)).tolist() == [0]
    v.fill_iota(1000)
    assert v.size() == cap and v.contents().tolist() == list(range(1000, 1000 + cap))
    # threads: the popped values and the rest are the whole set, each once
    out, ok = v.pop_back(1500, threads=4)
    assert ok.all()
    rest = v.contents()
    assert sorted(np.concatenate([out, rest]).tolist()) == list(range(1000, 1000 + cap))
    assert rest.tolist() == list(range(1000, 1000 + cap - 1500))  # the bottom stays in place
    assert v.push_back(out, threads=4).all()
    assert sorted(v.contents().tolist()) == list(range(1000, 1000 + cap))
    v.close()


# ---------------------------------------------------------------------------------------------
# committed golden fixtures (tests/golden/make_golden.py)
import json  # noqa: E402
import os  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_golden_reference_kats_match_oracle():
    kats = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
    for name, k in kats.items():
        kind = ("u32", None) if name.startswith("set") else ("u64", "u32")
        t = po.Table(kind[0], kind[1], k["capacity"])
        assert t.bucket_count() == k["bucket_count"] and t.max_size() == k["max_size"]
        for key, b in k["bucket"].items():
            assert t.bucket(int(key)) == b


def test_golden_small_workloads_regenerate_identically():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fresh = mg.small_workloads()
    stored = np.load(os.path.join(GOLDEN, "small_workloads.npz"))
    assert sorted(stored.files) == sorted(fresh)
    for k in stored.files:
        np.testing.assert_array_equal(stored[k], fresh[k], err_msg=k)


# ------------------------------------------------------------------ atomic<T>
# Known answers of the reference's own atomic tests (tests/stdgpu/atomic.inc): for_each_index over
# N = 40000 with sequence = 1..N (iota from 1), and N = 50000 with modulus N / 10 = 5000 for the
# wrapping increment / decrement, each from the stated start value.

ATOMIC_INT_TYPES = ["int", "unsigned int", "unsigned long long int"]


@pytest.mark.parametrize("t", ATOMIC_INT_TYPES)
def test_atomic_sequence_known_answers(t):
    n = 40000
    seq = list(range(1, n + 1))
    total = n * (n + 1) // 2
    assert po.atomic_sequence(t, "fetch_add", 0, seq)[0] == total  # atomic.inc:850-866
    assert po.atomic_sequence(t, "fetch_sub", total, seq)[0] == 0  # atomic.inc:1030-1051
    top = (1 << 31) - 1 if t == "int" else (1 << po.ATOMIC_BITS[t]) - 1
    assert po.atomic_sequence(t, "fetch_min", top, seq)[0] == 1  # atomic.inc:1197-1213
    lowest = -(1 << 31) if t == "int" else 0
    assert po.atomic_sequence(t, "fetch_max", lowest, seq)[0] == n  # atomic.inc:1324-1340
    bits = po.ATOMIC_BITS[t]
    pat = [po._wrap(t, 1 << i) for i in range(bits)]
    ones = po._wrap(t, (1 << bits) - 1)
    assert po.atomic_sequence(t, "fetch_or", 0, pat)[0] == ones  # one bit per thread (atomic.inc:1479-1495, 1871-1887)
    assert po.atomic_sequence(t, "fetch_xor", 0, pat)[0] == ones
    assert po.atomic_sequence(t, "fetch_and", ones, [po._wrap(t, ~(1 << i)) for i in range(bits)])[0] == 0
    # the reference's compare-exchange loop adds the sequence (atomic.inc:584-703, 670-703)
    s = 0
    for a in seq[:2000]:
        while True:
            seen = s
            _, s, hit = po.atomic_step(t, "compare_exchange", s, seen + a, seen)
            if hit:
                break
    assert s == 2000 * 2001 // 2


def test_atomic_wrapping_increment_decrement():
    n, mod = 50000, 5000
    assert po.atomic_sequence("unsigned int", "fetch_inc_mod", 42, [mod] * n)[0] == 42  # atomic.inc:1990-2007
    assert po.atomic_sequence("unsigned int", "fetch_dec_mod", 42, [mod] * n)[0] == 42  # atomic.inc:2046-2063
    # the backend's limit = arg - 1: inc wraps after limit to 0; dec wraps from 0 (or above limit) to limit
    assert po.atomic_step("unsigned int", "fetch_inc_mod", 4999, 5000)[1] == 0
    assert po.atomic_step("unsigned int", "fetch_inc_mod", 7000, 5000)[1] == 0
    assert po.atomic_step("unsigned int", "fetch_inc_mod", 4998, 5000)[1] == 4999
    assert po.atomic_step("unsigned int", "fetch_dec_mod", 0, 5000)[1] == 4999
    assert po.atomic_step("unsigned int", "fetch_dec_mod", 7000, 5000)[1] == 4999
    assert po.atomic_step("unsigned int", "fetch_dec_mod", 1, 5000)[1] == 0
    # wrapping increment of every start value below the modulus is (s + 1) mod m; decrement (s - 1) mod m
    for m in (1, 2, 7):
        for s0 in range(m):
            assert po.atomic_step("unsigned int", "fetch_inc_mod", s0, m)[1] == (s0 + 1) % m
            assert po.atomic_step("unsigned int", "fetch_dec_mod", s0, m)[1] == (s0 - 1) % m


def test_atomic_integer_wrap_and_float():
    assert po.atomic_step("int", "fetch_add", (1 << 31) - 1, 1)[1] == -(1 << 31)
    assert po.atomic_step("unsigned int", "fetch_sub", 0, 1)[1] == (1 << 32) - 1
    assert po.atomic_step("unsigned long long int", "fetch_add", (1 << 64) - 1, 2)[1] == 1
    assert po.atomic_step("float", "fetch_add", 1.0, 2.0 ** -30)[1] == 1.0  # float32 rounding
    assert po.atomic_step("float", "fetch_min", 3.0, -1.5)[1] == -1.5
    assert po.atomic_step("float", "compare_exchange", -0.0, 5.0, 0.0)[2]  # value comparison: -0 == +0
    nan = float("nan")
    assert not po.atomic_step("float", "compare_exchange", nan, 5.0, nan)[2]


def test_atomic_linearize_accepts_orders_and_rejects_tampering():
    rng = np.random.default_rng(3)
    n = 3000
    ops = [int(x) for x in rng.integers(1, 1000, size=n)]
    for op, init in (("fetch_add", 0), ("fetch_max", 0), ("fetch_min", 1 << 20), ("exchange", -1)):
        args = ops if op != "exchange" else list(range(n))
        order = rng.permutation(n)
        s, olds = init, [None] * n
        for i in order:
            olds[i], s, _ = po.atomic_step("int", op, s, args[i])
        assert po.atomic_linearize("int", op, init, args, olds) == s
    # an add history with two old values swapped (between ops with different addends) has no order
    s, olds = 0, [None] * n
    for i in range(n):
        olds[i], s, _ = po.atomic_step("int", "fetch_add", s, ops[i])
    j = next(k for k in range(1, n) if ops[k] != ops[0])
    bad = list(olds)
    bad[0], bad[j] = bad[j], bad[0]
    with pytest.raises(AssertionError):
        po.atomic_linearize("int", "fetch_add", 0, ops, bad)
    # an exchange history where one operation saw a value nobody wrote
    s, olds = -1, [None] * 100
    for i in range(100):
        olds[i], s, _ = po.atomic_step("int", "exchange", s, i)
    olds[50] = 12345
    with pytest.raises(AssertionError):
        po.atomic_linearize("int", "exchange", -1, list(range(100)), olds)


def test_voxel_caller_pattern_bruteforce():
    """voxel_insert against a pure-Python statement of the caller pattern: per key pop the front slot,
    insert (key, slot), push the slot back when the key is present; a key that finds the free list
    empty is skipped.  Exact key -> slot map and free-list order, including the empty-list case."""
    import collections

    rng = np.random.default_rng(5)
    for cap, nkeys, universe in ((64, 200, 40), (64, 200, 100), (1000, 5000, 700)):
        keys = rng.integers(1, universe + 1, size=nkeys, dtype=np.uint64)
        tab = po.Table("u64", "u32", capacity=4 * universe)
        q = po.Deque(cap)
        q.push_back(np.arange(cap, dtype=np.uint32))
        added = po.voxel_insert(tab, q, keys)
        free = collections.deque(range(cap))
        want = {}
        for k in keys.tolist():
            if not free:
                continue
            s = free.popleft()
            if k in want:
                free.append(s)
            else:
                want[k] = s
        assert added == len(want)
        k, v = tab.contents()
        assert dict(zip(k.tolist(), v.tolist())) == want
        assert q.contents().tolist() == list(free)
        assert tab.valid()
