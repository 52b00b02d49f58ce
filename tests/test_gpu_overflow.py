"""Overflow bookkeeping under adversarial and churning workloads.

Keys are chosen so that many share one home bucket (the last one, so their probe sequences wrap
around the end of the table), then erased head-first and tail-first; other tables churn at load
0.95.  After every step the contents, hit vectors, mapped values, per-element flags and size()
must equal a plain set/dict model (set semantics: insert never overwrites, erase removes), and
valid() must hold; valid() recomputes every bucket's overflow count from the stored keys and
checks that no bucket keys run through holds an EMPTY slot.
"""
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

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def fmix64(x):
    """MurmurHash3's 64-bit finaliser (the table's key mixer), vectorised over uint64."""
    x = np.asarray(x, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xC4CEB9FE1A85EC53)
        x ^= x >> np.uint64(33)
    return x


def home_bucket(keys, nbuckets):
    h = fmix64(keys) >> np.uint64(32)
    return ((h * np.uint64(nbuckets)) >> np.uint64(32)).astype(np.int64)


def value_of(keys, vdt):
    """mapped value of a key (any deterministic function; insert never overwrites)"""
    k = np.asarray(keys, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = (k * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(17)
    return v.astype(np.uint64).view(np.int64).astype(vdt) if vdt is not None else None


def keys_homed_at(bucket, nbuckets, count, kdt, rng, exclude=()):
    out, seen = [], set(int(x) for x in exclude)
    hi = 2**32 - 3 if kdt == np.uint32 else 2**63
    while len(out) < count:
        cand = rng.integers(1, hi, size=200000, dtype=np.uint64)
        sel = cand[home_bucket(cand, nbuckets) == bucket]
        for k in sel.tolist():
            if k not in seen:
                seen.add(k)
                out.append(k)
                if len(out) == count:
                    break
    return np.array(out, dtype=kdt)


class Model:
    def __init__(self, vdt):
        self.d = {}
        self.vdt = vdt

    def insert(self, keys):
        flags = np.zeros(len(keys), np.uint8)
        vals = value_of(keys, self.vdt) if self.vdt is not None else [None] * len(keys)
        for i, (k, v) in enumerate(zip(keys.tolist(), vals if self.vdt is None else vals.tolist())):
            if k not in self.d:
                self.d[k] = v
                flags[i] = 1
        return flags

    def erase(self, keys):
        flags = np.zeros(len(keys), np.uint8)
        for i, k in enumerate(keys.tolist()):
            if k in self.d:
                del self.d[k]
                flags[i] = 1
        return flags


def dev(a, sdt):
    # a fresh contiguous copy (a reversed one-element view keeps its negative stride otherwise)
    return torch.from_numpy(np.array(a, copy=True).view(sdt)).cuda()


def g_insert(tab, keys, kind):
    cls, kdt, sdt, vdt = KINDS[kind]
    ins = torch.zeros(len(keys), dtype=torch.uint8, device="cuda")
    vals = None if vdt is None else torch.from_numpy(value_of(keys, vdt)).cuda()
    tab.insert(dev(keys, sdt), vals, inserted=ins)
    return ins.cpu().numpy()


def g_erase(tab, keys, kind):
    cls, kdt, sdt, vdt = KINDS[kind]
    er = torch.zeros(len(keys), dtype=torch.uint8, device="cuda")
    tab.erase(dev(keys, sdt), erased=er)
    return er.cpu().numpy()


def check_table(tab, model, kind, probe):
    cls, kdt, sdt, vdt = KINDS[kind]
    hits, vals = tab.find(dev(probe, sdt))
    want = np.array([int(k) in model.d for k in probe.tolist()], np.uint8)
    np.testing.assert_array_equal(hits.cpu().numpy(), want)
    if vdt is not None:
        got = vals.cpu().numpy()[want == 1]
        exp = np.array([model.d[int(k)] for k in probe[want == 1].tolist()], dtype=vdt)
        np.testing.assert_array_equal(got, exp)
    k, _ = tab.contents()
    got_keys = np.sort(k.cpu().numpy().view(kdt))
    np.testing.assert_array_equal(got_keys, np.sort(np.array(list(model.d.keys()), dtype=kdt)))
    assert tab.size() == len(model.d)
    assert tab.valid()


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
@pytest.mark.parametrize("kind", list(KINDS))
def test_collision_chain_wraps_and_unwinds(kind, lanes):
    cls, kdt, sdt, vdt = KINDS[kind]
    rng = np.random.default_rng(100 + lanes)
    tab = cls.createDeviceObject(4096, sa.Tuning(bucket_lanes=lanes))
    nb = tab.info.physical_buckets
    W = tab.info.physical_slots // nb
    chain = keys_homed_at(nb - 1, nb, 4 * W, kdt, rng)          # runs over the table's end
    absent = keys_homed_at(nb - 1, nb, 16, kdt, rng, exclude=chain)
    background = np.unique(rng.integers(1, 2**31, size=nb * W // 4, dtype=np.uint64).astype(kdt))
    background = background[~np.isin(background, chain) & ~np.isin(background, absent)]
    model = Model(vdt)
    probe = np.concatenate([chain, absent, background[:256]])

    np.testing.assert_array_equal(g_insert(tab, background, kind), model.insert(background))
    np.testing.assert_array_equal(g_insert(tab, chain, kind), model.insert(chain))
    check_table(tab, model, kind, probe)
    # head-first: the keys closest to the home bucket go first
    for part in np.array_split(chain, 4):
        np.testing.assert_array_equal(g_erase(tab, part, kind), model.erase(part))
        check_table(tab, model, kind, probe)
    np.testing.assert_array_equal(g_insert(tab, chain, kind), model.insert(chain))
    check_table(tab, model, kind, probe)
    # tail-first
    for part in np.array_split(chain[::-1], 4):
        np.testing.assert_array_equal(g_erase(tab, part, kind), model.erase(part))
        check_table(tab, model, kind, probe)
    # the chain refills buckets that now hold tombstones
    np.testing.assert_array_equal(g_insert(tab, chain[::2], kind), model.insert(chain[::2]))
    check_table(tab, model, kind, probe)
    cls.destroyDeviceObject(tab)


@pytest.mark.parametrize("kind", list(KINDS))
def test_churn_at_load_095(kind):
    cls, kdt, sdt, vdt = KINDS[kind]
    rng = np.random.default_rng(7)
    cap = 20000
    tab = cls.createDeviceObject(cap)
    universe = np.unique(rng.integers(1, 2**31, size=int(cap * 0.95) + 500, dtype=np.uint64).astype(kdt))
    universe = universe[: int(cap * 0.95)]
    model = Model(vdt)
    for rnd in range(24):
        ins = rng.choice(universe, size=int(cap * 0.3), replace=False)
        np.testing.assert_array_equal(g_insert(tab, ins, kind), model.insert(ins))
        er = rng.choice(universe, size=int(cap * 0.2), replace=False)
        np.testing.assert_array_equal(g_erase(tab, er, kind), model.erase(er))
        if rnd % 6 == 5:
            check_table(tab, model, kind, rng.choice(universe, size=4000, replace=False))
    cls.destroyDeviceObject(tab)
