"""The partitioned bulk insert (keys partitioned by table region, each region inserted in LDS, the
keys whose probe sequence leaves their region inserted afterwards by the per-key kernel) must give
exactly what the per-key insert and the oracle give: the same sorted contents, values and size(),
valid() true, finds that hit exactly the stored keys.

Tables are made with insert_path = 2, which takes the partitioned path whenever it applies (any
number of keys and regions), against twins with insert_path = 1 (one CAS per key).  The cases
cover every kind and bucket width; uniform keys, heavy duplicates, the two reserved key values;
inserts into tables holding keys and tombstones; near-full tables (many keys leave their region,
some find no free slot at all); a table of one region (probe sequences wrap around the end).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

KINDS = {
    "set_u32": (sa.unordered_set_u32, "u32", None, np.uint32, np.int32, None),
    "set_u64": (sa.unordered_set_u64, "u64", None, np.uint64, np.int64, None),
    "map_u64_u32": (sa.unordered_map_u64_u32, "u64", "u32", np.uint64, np.int64, np.int32),
    "map_u32_u32": (sa.unordered_map_u32_u32, "u32", "u32", np.uint32, np.int32, np.int32),
    "map_u64_u64": (sa.unordered_map_u64_u64, "u64", "u64", np.uint64, np.int64, np.int64),
}


def keys_of(rng, kdt, n, universe=None, zipf=False):
    hi = universe or (2**32 if kdt == np.uint32 else 2**64)
    if zipf:
        ranks = np.minimum(rng.zipf(1.1, size=n), 1 << 20).astype(np.uint64)
        with np.errstate(over="ignore"):
            k = (ranks * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(1)
        return (k % np.uint64(hi)).astype(kdt) if kdt == np.uint32 else k.astype(kdt)
    if kdt == np.uint32 or universe:
        return rng.integers(0, hi, size=n, dtype=np.uint64).astype(kdt)
    return rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n, dtype=np.uint64)


def vals_of(keys, vdt):
    if vdt is None:
        return None
    with np.errstate(over="ignore"):
        v = (keys.astype(np.uint64) * np.uint64(0xBF58476D1CE4E5B9)) >> np.uint64(7)
    return v.view(np.int64).astype(vdt)


class Twin:
    """the same operations on a partitioned-path table, a per-key-path table and the oracle"""

    def __init__(self, name, cap, lanes):
        cls, key, mapped, kdt, sdt, vdt = KINDS[name]
        self.kdt, self.sdt, self.vdt = kdt, sdt, vdt
        self.a = cls.createDeviceObject(cap, sa.Tuning(bucket_lanes=lanes, insert_path=2))
        self.b = cls.createDeviceObject(cap, sa.Tuning(bucket_lanes=lanes, insert_path=1))
        self.ref = po.Table(key, mapped, cap)
        self.cls = cls

    def d(self, a, dt):
        return torch.from_numpy(np.ascontiguousarray(a).view(dt)).cuda()

    def insert(self, keys):
        keys = np.ascontiguousarray(keys, dtype=self.kdt)
        v = vals_of(keys, self.vdt)
        for t in (self.a, self.b):
            if v is None:
                t.insert(self.d(keys, self.sdt))
            else:
                t.insert(self.d(keys, self.sdt), self.d(v, self.vdt))
        self.ref.insert(keys, None if v is None else v.view(np.uint32 if self.vdt == np.int32 else np.uint64))

    def erase(self, keys):
        keys = np.ascontiguousarray(keys, dtype=self.kdt)
        for t in (self.a, self.b):
            t.erase(self.d(keys, self.sdt))
        self.ref.erase(keys)

    def check(self, probe):
        torch.cuda.synchronize()
        assert self.a.valid() and self.b.valid()
        for t in (self.a, self.b):
            k, v = t.contents()
            k = k.cpu().numpy().view(self.kdt)
            o = np.argsort(k, kind="stable")
            rk, rv = self.ref.sorted_contents()
            np.testing.assert_array_equal(k[o], rk)
            if v is not None:
                np.testing.assert_array_equal(v.cpu().numpy()[o].view(rv.dtype), rv)
            assert t.size() == self.ref.size()
        probe = np.ascontiguousarray(probe, dtype=self.kdt)
        h, _ = self.a.find(self.d(probe, self.sdt))
        rh, _ = self.ref.find(probe)
        np.testing.assert_array_equal(h.cpu().numpy(), rh)

    def close(self):
        self.cls.destroyDeviceObject(self.a)
        self.cls.destroyDeviceObject(self.b)


@pytest.mark.parametrize("name", list(KINDS))
@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_partitioned_insert_equals_per_key_and_oracle(name, lanes):
    rng = np.random.default_rng(lanes * 7 + len(name))
    cap = 1 << 15
    tw = Twin(name, cap, lanes)
    kdt = tw.kdt
    first = np.unique(keys_of(rng, kdt, int(cap * 0.9)))[: int(cap * 0.6)]
    tw.insert(first)                                     # into an empty table
    tw.check(np.concatenate([first[:500], keys_of(rng, kdt, 500)]))
    tw.erase(first[::3])                                 # tombstones and EMPTY slots
    more = np.concatenate([first[::3], keys_of(rng, kdt, int(cap * 0.2), zipf=True),
                           first[1::7]])                 # re-inserts, duplicates, present keys
    rng.shuffle(more)
    tw.insert(more)
    tw.check(np.concatenate([more[:500], first[:500]]))
    tw.close()


@pytest.mark.parametrize("name", ["map_u64_u32", "set_u32"])
def test_partitioned_insert_near_full_and_reserved_keys(name):
    rng = np.random.default_rng(5)
    cap = 1 << 14
    tw = Twin(name, cap, 8)
    kdt = tw.kdt
    allk = np.unique(keys_of(rng, kdt, 2 * cap))
    rng.shuffle(allk)
    tw.insert(allk[: int(cap * 0.9)])
    tw.check(allk[:300])
    # the two reserved key values travel the per-key path; then more keys than free slots
    resv = np.array([np.iinfo(kdt).max, np.iinfo(kdt).max - 1] * 5, dtype=kdt)
    tw.insert(np.concatenate([resv, allk[int(cap * 0.9): int(cap * 1.5)]]))
    # full in both: every slot, and the two reserved values in their side slots
    assert tw.a.size() == tw.b.size() == tw.a.info.limit + 2
    torch.cuda.synchronize()
    assert tw.a.valid() and tw.b.valid()
    # which of the surplus keys got in is decided by the race, so only the pre-existing keys, the
    # reserved values and the counts are compared once the table is full
    k = tw.a.contents()[0].cpu().numpy().view(kdt)
    assert set(allk[: int(cap * 0.9)].tolist()) <= set(k.tolist())
    assert {int(np.iinfo(kdt).max), int(np.iinfo(kdt).max) - 1} <= set(k.tolist())
    tw.close()


@pytest.mark.parametrize("lanes,cap", [(1, 1 << 20), (8, 1 << 22)])
def test_partitioned_insert_many_super_bins(lanes, cap):
    """tables of 512 and 2048 regions (4 and 16 super-bins of the first pass), filled in two launches"""
    rng = np.random.default_rng(lanes)
    tw = Twin("map_u64_u32", cap, lanes)
    first = np.unique(keys_of(rng, np.uint64, int(cap * 0.5)))
    tw.insert(first)
    tw.check(np.concatenate([first[:1000], keys_of(rng, np.uint64, 1000)]))
    tw.erase(first[::4])
    more = np.concatenate([first[::4], keys_of(rng, np.uint64, int(cap * 0.25))])
    tw.insert(more)
    tw.check(np.concatenate([more[:1000], first[:1000]]))
    tw.close()


def test_partitioned_insert_one_region_wraps():
    """a table of one region: keys homed at its last buckets run on to bucket 0"""
    rng = np.random.default_rng(9)
    tw = Twin("map_u64_u32", 400, 8)  # 50 buckets: one region
    for rnd in range(4):
        ks = keys_of(rng, np.uint64, 120, universe=1000)
        tw.insert(ks)
        tw.check(keys_of(rng, np.uint64, 200, universe=1000))
        tw.erase(keys_of(rng, np.uint64, 60, universe=1000))
        tw.check(keys_of(rng, np.uint64, 200, universe=1000))
    tw.close()


def test_partitioned_full_size_sequence():
    """2^26 keys at load 0.8 through the partitioned path: size, valid(), every key found, erased
    keys absent, re-inserted found"""
    from stdgpu_amd import workloads_device as wd

    n = 1 << 26
    keys = wd.stream(77, n, "cuda")  # a bijective mix of 0..n-1: distinct keys
    assert not bool((keys == -1).any()) and not bool((keys == -2).any())
    perm = torch.randperm(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    erase = keys[perm[: n // 4]].contiguous()
    w = {"capacity": int(np.ceil(n / 0.8)), "keys": keys, "values": wd.value_of(keys), "erase": erase,
         "reinsert": erase, "reinsert_values": wd.value_of(erase)}
    tab = sa.unordered_map_u64_u32.createDeviceObject(w["capacity"], sa.Tuning(insert_path=2))
    tab.insert(w["keys"], w["values"])
    assert tab.size() == n and tab.valid()
    hits = torch.empty(n, dtype=torch.uint8, device="cuda")
    got = torch.empty(n, dtype=torch.int32, device="cuda")
    tab.find(w["keys"], hits=hits, values_out=got)
    assert bool(hits.all()) and torch.equal(got, w["values"])
    tab.erase(w["erase"])
    assert tab.size() == n - w["erase"].numel()
    h2 = torch.empty(w["erase"].numel(), dtype=torch.uint8, device="cuda")
    tab.find(w["erase"], hits=h2)
    assert int(h2.sum()) == 0
    tab.insert(w["reinsert"], w["reinsert_values"])
    assert tab.size() == n and tab.valid()
    tab.find(w["keys"], hits=hits, values_out=got)
    assert bool(hits.all()) and torch.equal(got, w["values"])
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


def _sorted_contents(t):
    k, v = t.contents()
    o = torch.argsort(k)
    return k[o], v[o]


@pytest.mark.parametrize("hot", [False, True])
def test_default_insert_chooses_by_size_and_sample(hot):
    """Default tuning (insert_path 0) on a table past 256 MB (2^25 slots of 16 B): a launch of 0.6 of the
    slots in distinct keys takes the partitioned path; the same launch with every other key replaced by
    one hot key has 1 in 2 of its sampled keys repeat, so the sample hands it to the per-key kernel (the
    partitioned kernels exit).  Either way the contents, values and size equal a twin table that
    inserts with one compare-and-swap per key (insert_path 1), and a later re-insert of a fifth of the
    slots (below the size rule: per-key) and finds agree too."""
    from stdgpu_amd import workloads_device as wd

    cap = 1 << 25
    n = int(cap * 0.6)
    keys = wd.stream(91, n, "cuda")
    if hot:
        keys[::2] = keys[1]
    vals = wd.value_of(keys)
    a = sa.unordered_map_u64_u32.createDeviceObject(cap)
    b = sa.unordered_map_u64_u32.createDeviceObject(cap, sa.Tuning(insert_path=1))
    assert a.info.physical_slots * 16 > (256 << 20)
    for t in (a, b):
        t.insert(keys, vals)
    distinct = int(torch.unique(keys).numel())
    assert a.size() == b.size() == distinct
    ka, va = _sorted_contents(a)
    kb, vb = _sorted_contents(b)
    assert torch.equal(ka, kb) and torch.equal(va, vb)
    assert a.valid() and b.valid()
    more = wd.stream(92, cap // 5, "cuda")
    for t in (a, b):
        t.erase(keys[: n // 4])
        t.insert(more, wd.value_of(more))
    assert a.size() == b.size()
    ka, va = _sorted_contents(a)
    kb, vb = _sorted_contents(b)
    assert torch.equal(ka, kb) and torch.equal(va, vb)
    h, _ = a.find(keys)
    h2, _ = b.find(keys)
    assert torch.equal(h, h2)
    assert a.valid()
    for t in (a, b):
        sa.unordered_map_u64_u32.destroyDeviceObject(t)
