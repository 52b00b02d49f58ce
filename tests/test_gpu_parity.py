"""Parity of the HIP library (through its C ABI) with the CPU oracle on identical seeded inputs.

The bar is bit-exact: sorted (key, value) contents, find/contains hit vectors, mapped values
and size() after every step.  Where inputs carry duplicates, which duplicate reports
`inserted` depends on thread order in the reference too, so those flags are compared per
distinct key (exactly one winner per new key).  Full-size workloads are checked through
size-independent properties (membership computed on the host, counts, a checksum of the
contents, valid()).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from stdgpu_amd import workloads as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

NPT = {"u32": (np.uint32, np.int32), "u64": (np.uint64, np.int64)}
CLS = {
    ("u32", None): sa.unordered_set_u32,
    ("u64", None): sa.unordered_set_u64,
    ("u64", "u32"): sa.unordered_map_u64_u32,
    ("u32", "u32"): sa.unordered_map_u32_u32,
    ("u64", "u64"): sa.unordered_map_u64_u64,
}


def dev(a, kind):
    u, s = NPT[kind]
    return torch.from_numpy(np.ascontiguousarray(a, dtype=u).view(s)).cuda()


def host(t, kind):
    u, _ = NPT[kind]
    return t.cpu().numpy().view(u)


def u8(t):
    return t.cpu().numpy().astype(np.uint8)


def sorted_contents(tab, key, mapped):
    k, v = tab.contents()
    k = host(k, key)
    o = np.argsort(k, kind="stable")
    return k[o], (None if v is None else host(v, mapped)[o])


def assert_same_table(tab, ref, key, mapped):
    gk, gv = sorted_contents(tab, key, mapped)
    rk, rv = ref.sorted_contents()
    assert tab.size() == ref.size()
    np.testing.assert_array_equal(gk, rk)
    if mapped is not None:
        np.testing.assert_array_equal(gv, rv)
    assert tab.valid()


def per_key_wins(keys, flags):
    """number of reported successes per distinct key"""
    u, inv = np.unique(keys, return_inverse=True)
    return u, np.bincount(inv, weights=flags.astype(np.int64), minlength=len(u)).astype(np.int64)


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("cap", [1, 2, 7, 100, 100000, 1 << 21, 83886080])
def test_logical_sizes_follow_reference(cap):
    t = sa.unordered_map_u64_u32.createDeviceObject(cap)
    bc, ex = po.sizing(cap)
    assert t.bucket_count() == bc
    assert t.max_size() == bc + ex
    assert t.size() == 0 and t.empty()
    for k in [1, 2, 3, 0xFFFFFFFE, 0x0123456789ABCDEF]:
        assert t.bucket(k) == po.fibonacci_hashing(k, bc)
    sa.unordered_map_u64_u32.destroyDeviceObject(t)


def test_config_set_u32_full_size():
    """set<uint32_t>, capacity 2^21, 2^20 unique keys from seed 0: insert, find (half present),
    erase a seeded half, find again; contents / hits / size after every step."""
    w = wl.set_u32()
    ref = po.Table("u32", None, w.capacity)
    tab = sa.unordered_set_u32.createDeviceObject(w.capacity)
    keys = dev(w.keys, "u32")
    ins = torch.empty(len(w.keys), dtype=torch.uint8, device="cuda")
    tab.insert(keys, inserted=ins)
    rins = ref.insert(w.keys)
    np.testing.assert_array_equal(u8(ins), rins)
    assert tab.size() == len(w.keys)
    assert_same_table(tab, ref, "u32", None)

    q = dev(w.queries, "u32")
    hits, _ = tab.find(q)
    rh, _ = ref.find(w.queries)
    np.testing.assert_array_equal(u8(hits), rh)
    np.testing.assert_array_equal(rh, w.expect_hit)

    er = torch.empty(len(w.erase), dtype=torch.uint8, device="cuda")
    tab.erase(dev(w.erase, "u32"), erased=er)
    rer = ref.erase(w.erase)
    np.testing.assert_array_equal(u8(er), rer)
    assert_same_table(tab, ref, "u32", None)
    hits2, _ = tab.find(q)
    rh2, _ = ref.find(w.queries)
    np.testing.assert_array_equal(u8(hits2), rh2)
    sa.unordered_set_u32.destroyDeviceObject(tab)


def _map_sequence(w, lanes=8, check_ins_flags=True):
    ref = po.Table("u64", "u32", w.capacity)
    tab = sa.unordered_map_u64_u32.createDeviceObject(w.capacity, sa.Tuning(bucket_lanes=lanes))
    keys, vals = dev(w.keys, "u64"), dev(w.values, "u32")
    ins = torch.empty(len(w.keys), dtype=torch.uint8, device="cuda")
    tab.insert(keys, vals, inserted=ins)
    rins = ref.insert(w.keys, w.values)
    if check_ins_flags:
        np.testing.assert_array_equal(u8(ins), rins)
    else:
        uk, gw = per_key_wins(w.keys, u8(ins))
        _, rw = per_key_wins(w.keys, rins)
        np.testing.assert_array_equal(gw, rw)
        assert (gw == 1).all()
    assert_same_table(tab, ref, "u64", "u32")

    q = dev(w.queries, "u64")
    hits, got = tab.find(q)
    rh, rv = ref.find(w.queries)
    np.testing.assert_array_equal(u8(hits), rh)
    gv = host(got, "u32")
    np.testing.assert_array_equal(gv[rh == 1], rv[rh == 1])

    er = torch.empty(len(w.erase), dtype=torch.uint8, device="cuda")
    tab.erase(dev(w.erase, "u64"), erased=er)
    np.testing.assert_array_equal(u8(er), ref.erase(w.erase))
    assert_same_table(tab, ref, "u64", "u32")

    rk = w.reinsert
    rvv = wl.value_of(rk)
    ins2 = torch.empty(len(rk), dtype=torch.uint8, device="cuda")
    tab.insert(dev(rk, "u64"), dev(rvv, "u32"), inserted=ins2)
    np.testing.assert_array_equal(u8(ins2), ref.insert(rk, rvv))
    assert_same_table(tab, ref, "u64", "u32")
    hits3, _ = tab.find(q)
    np.testing.assert_array_equal(u8(hits3), ref.find(w.queries)[0])
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_config_map_u64_uniform_reduced(lanes):
    """map<uint64_t,uint32_t> at load 0.8: insert, find (90% hit), erase 25%, re-insert."""
    _map_sequence(wl.map_u64_uniform(n=1 << 18), lanes=lanes)


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_config_map_u64_zipf_reduced(lanes):
    """the same sequence with Zipf(1.1) draws: many lanes race on the same key and bucket."""
    _map_sequence(wl.map_u64_zipf(n=1 << 18), lanes=lanes, check_ins_flags=False)


KINDS = list(CLS)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_all_kinds_random_rounds(kind, lanes):
    key, mapped = kind
    rng = np.random.default_rng(1234 + lanes)
    # sized so neither table overflows (parity is defined where neither refuses an insert):
    # at most 40000 distinct keys in 60000 slots
    cap = 60000
    ref = po.Table(key, mapped, cap)
    tab = CLS[kind].createDeviceObject(cap, sa.Tuning(bucket_lanes=lanes))
    uni = 40000
    for rnd in range(4):
        ks = rng.integers(0, uni, size=12000, dtype=np.uint64).astype(NPT[key][0])
        vs = None if mapped is None else (ks.astype(np.uint64) * np.uint64(2654435761) % np.uint64(1 << 32)).astype(NPT[mapped][0])
        ins = torch.empty(len(ks), dtype=torch.uint8, device="cuda")
        tab.insert(dev(ks, key), None if vs is None else dev(vs, mapped), inserted=ins)
        rins = ref.insert(ks, vs)
        _, gw = per_key_wins(ks, u8(ins))
        _, rw = per_key_wins(ks, rins)
        np.testing.assert_array_equal(gw, rw)
        assert_same_table(tab, ref, key, mapped)
        q = rng.integers(0, uni, size=8000, dtype=np.uint64).astype(NPT[key][0])
        h, v = tab.find(dev(q, key))
        rh, rv = ref.find(q)
        np.testing.assert_array_equal(u8(h), rh)
        if mapped is not None:
            np.testing.assert_array_equal(host(v, mapped)[rh == 1], rv[rh == 1])
        e = rng.integers(0, uni, size=5000, dtype=np.uint64).astype(NPT[key][0])
        er = torch.empty(len(e), dtype=torch.uint8, device="cuda")
        tab.erase(dev(e, key), erased=er)
        _, ge = per_key_wins(e, u8(er))
        _, re_ = per_key_wins(e, ref.erase(e))
        np.testing.assert_array_equal(ge, re_)
        assert_same_table(tab, ref, key, mapped)
    CLS[kind].destroyDeviceObject(tab)


@pytest.mark.parametrize("kind", [("u32", None), ("u64", "u32"), ("u32", "u32")])
def test_reserved_key_values_are_storable(kind):
    """all-ones and all-ones-minus-one are the layout's EMPTY/TOMB markers; as user keys they
    live in side slots and behave like any other key (the reference accepts every key)."""
    key, mapped = kind
    u = NPT[key][0]
    top = np.iinfo(u).max
    ks = np.array([top, top - 1, 0, 5, top, top - 1], dtype=u)
    vs = None if mapped is None else np.array([1, 2, 3, 4, 9, 9], dtype=NPT[mapped][0])
    ref = po.Table(key, mapped, 100)
    tab = CLS[kind].createDeviceObject(100)
    ins = torch.empty(6, dtype=torch.uint8, device="cuda")
    tab.insert(dev(ks, key), None if vs is None else dev(vs, mapped), inserted=ins)
    np.testing.assert_array_equal(u8(ins), ref.insert(ks, vs))
    assert_same_table(tab, ref, key, mapped)
    h, v = tab.find(dev(ks, key))
    rh, rv = ref.find(ks)
    np.testing.assert_array_equal(u8(h), rh)
    if mapped is not None:
        np.testing.assert_array_equal(host(v, mapped), rv)
    er = torch.empty(2, dtype=torch.uint8, device="cuda")
    tab.erase(dev(np.array([top, 7], dtype=u), key), erased=er)
    np.testing.assert_array_equal(u8(er), ref.erase(np.array([top, 7], dtype=u)))
    assert_same_table(tab, ref, key, mapped)


def test_tiny_table_holds_exactly_two():
    # unordered_datastructure.inc:1375-1443 (createDeviceObject(1): two keys, the third fails)
    tab = sa.unordered_map_u64_u32.createDeviceObject(1)
    assert tab.max_size() == 2
    for k, expect in [(10, 1), (20, 1), (30, 0)]:
        ins = torch.empty(1, dtype=torch.uint8, device="cuda")
        tab.insert(dev(np.array([k]), "u64"), dev(np.array([k]), "u32"), inserted=ins)
        assert u8(ins).tolist() == [expect]
    assert tab.size() == 2 and tab.full() and tab.valid()
    ins = torch.empty(100000, dtype=torch.uint8, device="cuda")
    tab.insert(dev(np.full(100000, 30), "u64"), dev(np.full(100000, 3), "u32"), inserted=ins)
    assert int(u8(ins).sum()) == 0 and tab.size() == 2


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_tiny_table_parallel_one_free(lanes):
    # unordered_datastructure.inc:1549-1603: one free place, 100000 random keys -> exactly one wins
    tab = sa.unordered_map_u64_u32.createDeviceObject(1, sa.Tuning(bucket_lanes=lanes))
    tab.insert(dev(np.array([123]), "u64"), dev(np.array([1]), "u32"))
    rng = np.random.default_rng(9)
    ks = rng.integers(1, 2**40, size=100000, dtype=np.uint64)
    ins = torch.empty(len(ks), dtype=torch.uint8, device="cuda")
    tab.insert(dev(ks, "u64"), dev(np.ones(len(ks)), "u32"), inserted=ins)
    assert int(u8(ins).sum()) == 1
    assert tab.size() == 2 and tab.full() and tab.valid()
    won = ks[u8(ins) == 1][0]
    h, _ = tab.find(dev(np.array([123, won]), "u64"))
    assert u8(h).tolist() == [1, 1]


@pytest.mark.parametrize("block", [64, 256])
def test_one_stripe_table_plans_its_launches(block):
    """A table whose limit lies below its slots (capacity 2^16 at 2 slots per capacity: the limit is
    the reference's max_size) counts its size on one stripe.  Bulk launches that cannot reach the
    limit count on all stripes and fold back into it; a launch that can is ranked exactly, so the
    table stops at its limit.  A fold that went missing would show as a wrong rank base below."""
    tab = sa.unordered_map_u64_u32.createDeviceObject(1 << 16, sa.Tuning(block_size=block, slots_per_capacity=2.0))
    limit = tab.info.limit
    assert limit == tab.max_size() < tab.info.physical_slots
    rng = np.random.default_rng(61)
    keys = np.unique(rng.integers(1, 2**62, size=300000, dtype=np.uint64))[:200000]
    keys = keys[rng.permutation(len(keys))]

    def ins(k):
        flags = torch.empty(len(k), dtype=torch.uint8, device="cuda")
        tab.insert(dev(k, "u64"), dev(k & np.uint64(0xFFFF), "u32"), inserted=flags)
        return int(u8(flags).sum())

    assert ins(keys[:50000]) == 50000 and tab.size() == 50000       # cannot reach the limit
    assert ins(keys[50000:90000]) == 40000 and tab.size() == 90000  # cannot reach it either
    er = torch.empty(20000, dtype=torch.uint8, device="cuda")
    tab.erase(dev(keys[:20000], "u64"), erased=er)
    assert int(u8(er).sum()) == 20000 and tab.size() == 70000 and tab.valid()
    # this one can: exactly limit - 70000 of 60000 new keys go in
    assert ins(keys[90000:150000]) == limit - 70000
    assert tab.size() == limit and tab.full() and tab.valid()
    assert ins(keys[150000:151000]) == 0 and tab.size() == limit
    tab.erase(dev(keys[20000:20010], "u64"))
    assert tab.size() == limit - 10
    assert ins(keys[151000:152000]) == 10 and tab.size() == limit and tab.valid()
    h, _ = tab.find(dev(keys[20010:90000], "u64"))
    assert int(u8(h).sum()) == 90000 - 20010


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_contention_same_key(lanes):
    # insert_multiple_* / erase_multiple_* (unordered_datastructure.inc:1167-1373)
    tab = sa.unordered_map_u64_u32.createDeviceObject(100000, sa.Tuning(bucket_lanes=lanes))
    n = 100000
    ins = torch.empty(n, dtype=torch.uint8, device="cuda")
    tab.insert(dev(np.full(n, 77), "u64"), dev(np.full(n, 5), "u32"), inserted=ins)
    assert int(u8(ins).sum()) == 1 and tab.size() == 1
    er = torch.empty(n, dtype=torch.uint8, device="cuda")
    tab.erase(dev(np.full(n, 77), "u64"), erased=er)
    assert int(u8(er).sum()) == 1 and tab.size() == 0 and tab.valid()


def test_empty_inputs_and_clear():
    tab = sa.unordered_set_u64.createDeviceObject(1000)
    z = torch.empty(0, dtype=torch.int64, device="cuda")
    tab.insert(z)
    tab.erase(z)
    h, _ = tab.find(z)
    assert h.numel() == 0 and tab.size() == 0
    tab.insert(dev(np.arange(1, 501), "u64"))
    assert tab.size() == 500
    tab.clear()
    assert tab.size() == 0 and tab.valid()
    h, _ = tab.find(dev(np.arange(1, 501), "u64"))
    assert int(u8(h).sum()) == 0
    assert tab.device_range().numel() == 0


def test_device_range_ascending_and_complete():
    w = wl.map_u64_uniform(n=1 << 16)
    tab = sa.unordered_map_u64_u32.createDeviceObject(w.capacity)
    tab.insert(dev(w.keys, "u64"), dev(w.values, "u32"))
    idx = tab.device_range().cpu().numpy()
    assert len(idx) == len(w.keys) == tab.size()
    assert (np.diff(idx) > 0).all()
    k, v = tab.contents()
    np.testing.assert_array_equal(np.sort(host(k, "u64")), np.sort(w.keys))
    np.testing.assert_array_equal(host(v, "u32"), wl.value_of(host(k, "u64")))


def test_examples_closed_form_sums():
    n = 100
    odd = np.arange(1, n + 1)[np.arange(1, n + 1) % 2 == 1]
    neigh = np.concatenate([odd - 1, odd, odd + 1]).astype(np.uint32)
    s = sa.unordered_set_u32.createDeviceObject(n)
    s.insert(dev(neigh, "u32"))
    k, _ = s.contents()
    assert s.size() == n + 1 and int(host(k, "u32").astype(np.int64).sum()) == n * (n + 1) // 2
    m = sa.unordered_map_u32_u32.createDeviceObject(n)
    m.insert(dev(neigh, "u32"), dev((neigh.astype(np.int64) ** 2), "u32"))
    k, v = m.contents()
    assert int(host(k, "u32").astype(np.int64).sum()) == n * (n + 1) // 2
    assert int(host(v, "u32").astype(np.int64).sum()) == n * (n + 1) * (2 * n + 1) // 6


def test_config_map_u64_full_size_properties():
    """2^26 keys at load 0.8: insert, find (90% hit), erase 25%, re-insert, checked through
    membership computed on the host, counts, a checksum of the contents and valid()."""
    w = wl.map_u64_uniform()
    n = len(w.keys)
    tab = sa.unordered_map_u64_u32.createDeviceObject(w.capacity)
    keys, vals = dev(w.keys, "u64"), dev(w.values, "u32")
    ins = torch.empty(n, dtype=torch.uint8, device="cuda")
    tab.insert(keys, vals, inserted=ins)
    assert int(ins.sum()) == n and tab.size() == n
    q = dev(w.queries, "u64")
    hits, got = tab.find(q)
    srt = np.sort(w.keys)
    pos = np.searchsorted(srt, w.queries)
    pos[pos == n] = 0
    member = srt[pos] == w.queries
    h = u8(hits)
    np.testing.assert_array_equal(h, member.astype(np.uint8))
    np.testing.assert_array_equal(host(got, "u32")[member], wl.value_of(w.queries[member]))
    er = torch.empty(len(w.erase), dtype=torch.uint8, device="cuda")
    tab.erase(dev(w.erase, "u64"), erased=er)
    assert int(er.sum()) == len(w.erase) and tab.size() == n - len(w.erase)
    k, v = tab.contents()
    kk = host(k, "u64")
    remaining = np.setdiff1d(w.keys, w.erase, assume_unique=True)
    assert len(kk) == len(remaining)
    with np.errstate(over="ignore"):
        assert kk.sum(dtype=np.uint64) == remaining.sum(dtype=np.uint64)
    np.testing.assert_array_equal(host(v, "u32"), wl.value_of(kk))
    ins2 = torch.empty(len(w.reinsert), dtype=torch.uint8, device="cuda")
    tab.insert(dev(w.reinsert, "u64"), dev(wl.value_of(w.reinsert), "u32"), inserted=ins2)
    assert int(ins2.sum()) == len(w.reinsert) and tab.size() == n
    hits3, _ = tab.find(q)
    np.testing.assert_array_equal(u8(hits3), member.astype(np.uint8))
    assert tab.valid()
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


def test_deque_free_list_multiset():
    cap = 1 << 16
    q = sa.deque_u32(cap)
    q.fill_iota()
    assert q.size() == cap
    out, ok = q.pop_front(40000)
    assert int(ok.sum()) == 40000 and q.size() == cap - 40000
    rest = q.contents().cpu().numpy()
    got = out.cpu().numpy()
    assert sorted(np.concatenate([got, rest]).view(np.uint32).tolist()) == list(range(cap))
    pushed = torch.empty(40000, dtype=torch.uint8, device="cuda")
    q.push_back(out, pushed=pushed)
    assert int(pushed.sum()) == 40000 and q.size() == cap
    more = torch.zeros(10, dtype=torch.int32, device="cuda")
    p2 = torch.empty(10, dtype=torch.uint8, device="cuda")
    q.push_back(more, pushed=p2)
    assert int(p2.sum()) == 0  # full
    out2, ok2 = q.pop_front(cap + 100)
    assert int(ok2.sum()) == cap and q.size() == 0
    vals = out2.cpu().numpy().view(np.uint32)[u8(ok2) == 1]
    assert sorted(vals.tolist()) == list(range(cap))
    q.destroy()


def test_against_committed_golden_fixtures():
    """the GPU path against the committed fixture vectors (tests/golden/small_workloads.npz)"""
    import os

    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_workloads.npz"))
    s = sa.unordered_set_u32.createDeviceObject(int(g["set_capacity"][0]))
    ins = torch.empty(len(g["set_keys"]), dtype=torch.uint8, device="cuda")
    s.insert(dev(g["set_keys"], "u32"), inserted=ins)
    np.testing.assert_array_equal(u8(ins), g["set_ins"])
    np.testing.assert_array_equal(u8(s.find(dev(g["set_queries"], "u32"))[0]), g["set_hits1"])
    er = torch.empty(len(g["set_erase"]), dtype=torch.uint8, device="cuda")
    s.erase(dev(g["set_erase"], "u32"), erased=er)
    np.testing.assert_array_equal(u8(er), g["set_erased"])
    np.testing.assert_array_equal(u8(s.find(dev(g["set_queries"], "u32"))[0]), g["set_hits2"])
    k, _ = s.contents()
    np.testing.assert_array_equal(np.sort(host(k, "u32")), g["set_contents"])
    m = sa.unordered_map_u64_u32.createDeviceObject(int(g["map_capacity"][0]))
    ins = torch.empty(len(g["map_keys"]), dtype=torch.uint8, device="cuda")
    m.insert(dev(g["map_keys"], "u64"), dev(g["map_vals"], "u32"), inserted=ins)
    np.testing.assert_array_equal(u8(ins), g["map_ins"])
    h, v = m.find(dev(g["map_queries"], "u64"))
    np.testing.assert_array_equal(u8(h), g["map_hits"])
    np.testing.assert_array_equal(np.where(u8(h) == 1, host(v, "u32"), 0), g["map_found_vals"])
    er = torch.empty(len(g["map_erase"]), dtype=torch.uint8, device="cuda")
    m.erase(dev(g["map_erase"], "u64"), erased=er)
    np.testing.assert_array_equal(u8(er), g["map_erased"])
    k, v = m.contents()
    o = np.argsort(host(k, "u64"))
    np.testing.assert_array_equal(host(k, "u64")[o], g["map_contents_keys"])
    np.testing.assert_array_equal(host(v, "u32")[o], g["map_contents_vals"])


def test_config_map_u64_zipf_full_size_properties():
    """The 2^26 Zipf(1.1) sequence (heavy duplicates, generated on the device): exactly one inserted
    flag per distinct key, size = distinct keys, every present query hits with value_of(key), erase
    and re-insert counts, valid() at the end."""
    from stdgpu_amd import workloads_device as wd

    n = 1 << 26
    w = wd.map_u64_zipf("cuda", n=n)
    tab = sa.unordered_map_u64_u32.createDeviceObject(w["capacity"])
    ins = torch.empty(n, dtype=torch.uint8, device="cuda")
    tab.insert(w["keys"], w["values"], inserted=ins)
    assert int(ins.sum()) == w["distinct"] and tab.size() == w["distinct"]
    # one winner per distinct key: the winners' keys are exactly the distinct keys
    winners = torch.sort(w["keys"][ins.bool()])[0]
    assert winners.numel() == w["distinct"] and bool((winners[1:] != winners[:-1]).all())
    hits, got = tab.find(w["queries"])
    nhit = int(round(n * 0.9))
    present = torch.isin(w["queries"], w["keys"])
    assert torch.equal(hits.bool(), present)
    assert int(present.sum()) >= nhit  # 90% of the queries are drawn from the keys
    assert torch.equal(got[present], wd.value_of(w["queries"][present]))
    er = torch.empty(w["erase"].numel(), dtype=torch.uint8, device="cuda")
    tab.erase(w["erase"], erased=er)
    assert int(er.sum()) == w["erase"].numel()
    assert tab.size() == w["distinct"] - w["erase"].numel()
    h2, _ = tab.find(w["erase"])
    assert int(h2.sum()) == 0
    ins2 = torch.empty(w["reinsert"].numel(), dtype=torch.uint8, device="cuda")
    tab.insert(w["reinsert"], w["reinsert_values"], inserted=ins2)
    assert int(ins2.sum()) == w["reinsert"].numel() and tab.size() == w["distinct"]
    assert tab.valid()
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


def test_misaligned_inputs_are_refused():
    """keys, values and pairs must be aligned to their size; a misaligned pointer is refused before any
    device work (HTAB_E_ARG) instead of faulting"""
    import ctypes

    from stdgpu_amd import _lib

    lib = _lib.load()
    t = sa.unordered_map_u64_u32.createDeviceObject(1000)
    buf = torch.zeros(64, dtype=torch.int64, device="cuda")
    base = buf.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    h = t._h
    E = _lib.HTAB_E_ARG
    assert lib.htab_insert_kv(h, base + 4, base, 4, None, s) == E  # u64 keys at a 4-byte offset
    assert lib.htab_insert_kv(h, base, base + 2, 4, None, s) == E  # u32 values at a 2-byte offset
    assert lib.htab_insert(h, base + 4, 4, None, s) == E  # 16-byte pairs need 8-byte alignment
    hits = torch.zeros(4, dtype=torch.uint8, device="cuda")
    assert lib.htab_find(h, base + 4, 4, hits.data_ptr(), None, s) == E
    assert lib.htab_find(h, base, 4, hits.data_ptr(), base + 1, s) == E
    assert lib.htab_erase(h, base + 4, 4, None, s) == E
    assert lib.htab_assign(h, base, base + 2, 4, None, s) == E
    # aligned calls on the same buffers go through, and the table is untouched by the refused ones
    assert lib.htab_insert_kv(h, base, base + 256, 4, None, s) == 0
    torch.cuda.synchronize()
    assert t.size() == 1 and t.valid()  # four zero keys: one distinct key
    sa.unordered_map_u64_u32.destroyDeviceObject(t)
