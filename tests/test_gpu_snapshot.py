"""Snapshot / restore of tables: a restored table has the original's geometry, contents, size and
probe behaviour, and the same later operations give the same results on both."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402

KINDS = {
    "set_u32": (sa.unordered_set_u32, np.uint32, np.int32, None),
    "set_u64": (sa.unordered_set_u64, np.uint64, np.int64, None),
    "map_u64_u32": (sa.unordered_map_u64_u32, np.uint64, np.int64, np.int32),
    "map_u32_u32": (sa.unordered_map_u32_u32, np.uint32, np.int32, np.int32),
    "map_u64_u64": (sa.unordered_map_u64_u64, np.uint64, np.int64, np.int64),
}


def dev(a, dt):
    return torch.from_numpy(np.array(a, copy=True).view(dt)).cuda()


def keys_of(rng, n, kdt):
    hi = 2**32 if kdt == np.uint32 else 2**64
    k = np.unique(rng.integers(0, hi, size=n + 64, dtype=np.uint64).astype(kdt))[:n]
    rng.shuffle(k)
    # both reserved key values (all ones, all ones - 1) are ordinary keys too
    k[:2] = np.array([~np.uint64(0), ~np.uint64(0) - np.uint64(1)], dtype=np.uint64).astype(kdt)
    return k


def same(a, b, kind, probe, slot_order=True):
    cls, kdt, sdt, vdt = KINDS[kind]
    assert a.size() == b.size()
    assert a.info.physical_buckets == b.info.physical_buckets and a.info.limit == b.info.limit
    ka, va = a.contents()
    kb, vb = b.contents()
    if not slot_order:
        # after concurrent inserts the slots a key lands in may differ: compare as sets
        oa, ob = torch.argsort(ka), torch.argsort(kb)
        ka, kb = ka[oa], kb[ob]
        va, vb = (None, None) if vdt is None else (va[oa], vb[ob])
    assert torch.equal(ka, kb)  # with slot_order: the layout is restored as it was
    if vdt is not None:
        assert torch.equal(va, vb)
    ha, xa = a.find(dev(probe, sdt))
    hb, xb = b.find(dev(probe, sdt))
    assert torch.equal(ha, hb)
    if vdt is not None:
        assert torch.equal(xa[ha.bool()], xb[hb.bool()])
    assert a.valid() and b.valid()


@pytest.mark.parametrize("lanes", [1, 8])
@pytest.mark.parametrize("kind", list(KINDS))
def test_restored_table_equals_original(kind, lanes):
    cls, kdt, sdt, vdt = KINDS[kind]
    rng = np.random.default_rng(lanes * 7 + len(kind))
    t = cls.createDeviceObject(50000, sa.Tuning(bucket_lanes=lanes))
    k = keys_of(rng, 36000, kdt)
    vals = None if vdt is None else dev(rng.integers(0, 2**31, size=len(k)).astype(vdt), vdt)
    t.insert(dev(k, sdt), vals)
    t.erase(dev(k[1000:6000], sdt))  # tombstones and lowered overflow counts travel too
    buf = t.snapshot()
    r = cls.restore(buf)
    probe = np.concatenate([k[:8000], keys_of(rng, 2000, kdt)[2:]])
    same(t, r, kind, probe)
    # a snapshot of the restored table is byte-identical to the original's
    np.testing.assert_array_equal(r.snapshot(), buf)
    # the same later operations give the same results
    more = keys_of(rng, 9000, kdt)[2:]
    mv = None if vdt is None else dev(rng.integers(0, 2**31, size=len(more)).astype(vdt), vdt)
    fa = torch.zeros(len(more), dtype=torch.uint8, device="cuda")
    fb = torch.zeros(len(more), dtype=torch.uint8, device="cuda")
    t.insert(dev(more, sdt), mv, inserted=fa)
    r.insert(dev(more, sdt), mv, inserted=fb)
    assert torch.equal(fa, fb)
    ea = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    eb = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    t.erase(dev(k[6000:10000], sdt), erased=ea)
    r.erase(dev(k[6000:10000], sdt), erased=eb)
    assert torch.equal(ea, eb)
    same(t, r, kind, probe, slot_order=False)
    cls.destroyDeviceObject(t)
    cls.destroyDeviceObject(r)


def test_empty_and_exact_limit_tables():
    for cap in (1, 3, 1000):
        t = sa.unordered_map_u64_u32.createDeviceObject(cap)
        r = sa.unordered_map_u64_u32.restore(t.snapshot())
        assert r.size() == 0 and r.valid() and r.max_size() == t.max_size()
        k = dev(np.arange(1, 6, dtype=np.uint64), np.int64)
        v = dev(np.arange(5, dtype=np.int32), np.int32)
        fa = torch.zeros(5, dtype=torch.uint8, device="cuda")
        fb = torch.zeros(5, dtype=torch.uint8, device="cuda")
        t.insert(k, v, inserted=fa)
        r.insert(k, v, inserted=fb)
        assert int(fa.sum()) == int(fb.sum()) == min(5, t.info.limit)
        r2 = sa.unordered_map_u64_u32.restore(t.snapshot())
        assert r2.size() == t.size() and r2.full() == t.full() and r2.valid()
        for x in (t, r, r2):
            sa.unordered_map_u64_u32.destroyDeviceObject(x)
    with pytest.raises(TypeError):
        s = sa.unordered_set_u32.createDeviceObject(10)
        try:
            sa.unordered_map_u64_u32.restore(s.snapshot())
        finally:
            sa.unordered_set_u32.destroyDeviceObject(s)
