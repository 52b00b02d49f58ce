"""valid() rejects broken tables (the reference's valid(), unordered_base_detail.cuh:1192-1218).

A filled table is corrupted by plain host writes into its device memory (the slot bytes, the
overflow counts, the size stripes, found through htab_get_view), one invariant at a time:
  * duplicate   - a stored key copied into a free slot of its own bucket (size raised to match);
  * ovf_count   - one bucket's overflow count raised by one;
  * empty_in_overflowed_bucket - a TOMB turned into EMPTY in a bucket that keys run through;
  * size        - the size counter one above the number of stored keys;
  * unreachable - a key moved to an EMPTY slot of a bucket off its probe sequence;
  * away_count  - (map<u64, u32> with buckets of several slots) a nonzero away counter lowered by
    one, so a key stored past its home bucket is no longer announced there.
Each must make htab_valid report 0, and writing the original bytes back must make it report 1
again, so the corruption and not the harness made the difference.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from test_gpu_overflow import home_bucket, value_of  # noqa: E402

KINDS = {
    # name: (class, key dtype, signed key dtype, mapped signed dtype, slot bytes)
    "set_u32": (sa.unordered_set_u32, np.uint32, np.int32, None, 4),
    "set_u64": (sa.unordered_set_u64, np.uint64, np.int64, None, 8),
    "map_u64_u32": (sa.unordered_map_u64_u32, np.uint64, np.int64, np.int32, 16),
    "map_u32_u32": (sa.unordered_map_u32_u32, np.uint32, np.int32, np.int32, 8),
    "map_u64_u64": (sa.unordered_map_u64_u64, np.uint64, np.int64, np.int64, 16),
}

_hip = None


def hip():
    global _hip
    if _hip is None:
        torch.cuda.init()
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip


def d2h(ptr, nbytes):
    out = np.empty(nbytes, dtype=np.uint8)
    assert hip().hipDeviceSynchronize() == 0
    assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), nbytes, 2) == 0
    return out


def h2d(ptr, buf):
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    assert hip().hipMemcpy(ctypes.c_void_p(ptr), buf.ctypes.data_as(ctypes.c_void_p), buf.size, 1) == 0
    assert hip().hipDeviceSynchronize() == 0


class Image:
    """host copy of a table's slots, overflow counts and first size stripe, written back on demand"""

    def __init__(self, tab, kdt, slot_bytes):
        v = tab.view()
        self.v, self.kdt, self.sb = v, kdt, slot_bytes
        self.W = 16 * v.bucket_lanes // slot_bytes
        self.nb = v.nbuckets
        self.nslots = self.nb * self.W
        self.slot_area = (self.nslots + 2) * slot_bytes
        self.slots = d2h(v.slots, self.slot_area)
        self.ovf_arr = None if not v.ovf else d2h(v.ovf, 4 * self.nb)
        self.stripe = d2h(v.stripes, 8)

    def key_words(self):
        kb = np.dtype(self.kdt).itemsize
        return self.slots.reshape(-1, self.sb)[: self.nslots, :kb].copy().view(self.kdt).reshape(-1)

    def set_key(self, slot, key):
        kb = np.dtype(self.kdt).itemsize
        assert 0 <= slot < self.nslots
        self.slots[slot * self.sb: slot * self.sb + kb] = np.array([key], dtype=self.kdt).view(np.uint8)

    def ovf(self):
        if self.ovf_arr is not None:
            return self.ovf_arr.view(np.uint32).copy()
        # in the padding (bytes 12-15) of each bucket's first slot
        rows = self.slots[: self.nslots * self.sb].reshape(self.nb, self.W * self.sb)
        return rows[:, 12:16].copy().view(np.uint32).reshape(-1)

    def set_ovf(self, b, value):
        assert 0 <= b < self.nb
        word = np.array([value], dtype=np.uint32).view(np.uint8)
        if self.ovf_arr is not None:
            self.ovf_arr[4 * b: 4 * b + 4] = word
        else:
            off = b * self.W * self.sb + 12
            self.slots[off: off + 4] = word

    def away_words(self):
        """padding words of every slot (bytes 12-15), shape (nbuckets, W); in-bucket layouts only"""
        rows = self.slots[: self.nslots * self.sb].reshape(self.nb, self.W, self.sb)
        return rows[:, :, 12:16].copy().view(np.uint32).reshape(self.nb, self.W)

    def set_away_field(self, b, f, value):
        off = (b * self.W + 1 + f // 2) * self.sb + 12 + 2 * (f % 2)
        self.slots[off: off + 2] = np.array([value], dtype=np.uint16).view(np.uint8)

    def add_size(self, d):
        self.stripe = (self.stripe.view(np.int64) + d).view(np.uint8)

    def write(self):
        h2d(self.v.slots, self.slots)
        if self.ovf_arr is not None:
            h2d(self.v.ovf, self.ovf_arr)
        h2d(self.v.stripes, self.stripe)


def filled(name, lanes, cap=4096, seed=1):
    cls, kdt, sdt, vdt, sb = KINDS[name]
    rng = np.random.default_rng(seed)
    hi = 2**32 - 3 if kdt == np.uint32 else 2**62
    keys = np.unique(rng.integers(1, hi, size=cap, dtype=np.uint64).astype(kdt))[: int(cap * 0.8)]
    tab = cls.createDeviceObject(cap, sa.Tuning(bucket_lanes=lanes))
    k = torch.from_numpy(keys.view(sdt)).cuda()
    if vdt is None:
        tab.insert(k)
    else:
        tab.insert(k, torch.from_numpy(value_of(keys, vdt)).cuda())
    tab.erase(torch.from_numpy(keys[::4].copy().view(sdt)).cuda())
    torch.cuda.synchronize()
    assert tab.valid()
    return tab, kdt, sb


EMPTY = {np.uint32: np.uint32(0xFFFFFFFF), np.uint64: np.uint64(0xFFFFFFFFFFFFFFFF)}
TOMB = {np.uint32: np.uint32(0xFFFFFFFE), np.uint64: np.uint64(0xFFFFFFFFFFFFFFFE)}


def corrupt(img, what):
    kdt = img.kdt
    keys = img.key_words()
    stored = (keys != EMPTY[kdt]) & (keys != TOMB[kdt])
    ovf = img.ovf()
    bucket_of_slot = np.arange(img.nslots) // img.W
    if what == "duplicate":
        for s in np.nonzero(stored)[0]:
            b = bucket_of_slot[s]
            free = [x for x in range(b * img.W, (b + 1) * img.W) if not stored[x]]
            if free:
                img.set_key(free[0], keys[s])
                img.add_size(1)
                return True
        return False
    if what == "ovf_count":
        b = np.nonzero(ovf > 0)[0]
        if len(b) == 0:
            return False
        img.set_ovf(int(b[0]), int(ovf[b[0]]) + 1)
        return True
    if what == "empty_in_overflowed_bucket":
        tomb = keys == TOMB[kdt]
        for s in np.nonzero(tomb)[0]:
            if ovf[bucket_of_slot[s]] > 0:
                img.set_key(int(s), EMPTY[kdt])
                return True
        return False
    if what == "away_count":
        if img.ovf_arr is not None or img.W < 2:
            return False  # counters live in slot padding: map<u64, u32> with buckets of several slots
        words = img.away_words()
        nctr = 2 * (img.W - 1)
        for b in np.nonzero(ovf > 0)[0]:
            for f in range(nctr):
                c = int((words[b, 1 + f // 2] >> (16 * (f % 2))) & 0xFFFF)
                if 0 < c < 0x8000:
                    img.set_away_field(int(b), f, c - 1)
                    return True
        return False
    if what == "size":
        img.add_size(1)
        return True
    if what == "unreachable":
        homes = home_bucket(keys.astype(np.uint64), img.nb)
        empty_slots = np.nonzero(keys == EMPTY[kdt])[0]
        for s in np.nonzero(stored & (bucket_of_slot == homes))[0][:64]:
            far = [e for e in empty_slots
                   if (bucket_of_slot[e] - homes[s]) % img.nb > img.nb // 2 and ovf[bucket_of_slot[e]] == 0]
            if far:
                img.set_key(int(far[0]), keys[s])
                img.set_key(int(s), TOMB[kdt] if ovf[bucket_of_slot[s]] > 0 else EMPTY[kdt])
                return True
        return False
    raise ValueError(what)


CASES = ["duplicate", "ovf_count", "empty_in_overflowed_bucket", "size", "unreachable", "away_count"]


@pytest.mark.parametrize("name", list(KINDS))
@pytest.mark.parametrize("lanes", [1, 8])
@pytest.mark.parametrize("what", CASES)
def test_valid_rejects_corrupted_table(name, lanes, what):
    tab, kdt, sb = filled(name, lanes)
    img = Image(tab, kdt, sb)
    original = (img.slots.copy(), None if img.ovf_arr is None else img.ovf_arr.copy(), img.stripe.copy())
    if not corrupt(img, what):
        pytest.skip(f"no bucket of this table has the structure {what} needs")
    img.write()
    assert not tab.valid(), what
    img.slots, img.ovf_arr, img.stripe = original
    img.write()
    assert tab.valid()
    type(tab).destroyDeviceObject(tab)


def test_corruption_cases_apply_to_the_benchmarked_kind():
    """the default map<u64, u32> table at load 0.8 has every structure the cases need"""
    tab, kdt, sb = filled("map_u64_u32", 8)
    for what in CASES:
        img = Image(tab, kdt, sb)
        assert corrupt(img, what), what
    type(tab).destroyDeviceObject(tab)


def away_recount(img):
    """away counters recomputed on the host from the stored keys: keys homed at bucket h and stored in
    another bucket, by fingerprint (bits 8-31 of (lo32 ^ hi32) * 0x9E3779B1 mod 2^32, scaled to 2(W-1)
    counters)"""
    keys = img.key_words()
    stored = (keys != EMPTY[img.kdt]) & (keys != TOMB[img.kdt])
    slots = np.nonzero(stored)[0]
    k = keys[slots].astype(np.uint64)
    homes = home_bucket(k, img.nb)
    away = (slots // img.W) != homes
    nctr = 2 * (img.W - 1)
    ka = k[away]
    x = ((ka & np.uint64(0xFFFFFFFF)) ^ (ka >> np.uint64(32))) * np.uint64(0x9E3779B1) & np.uint64(0xFFFFFFFF)
    f = ((x >> np.uint64(8)) * np.uint64(nctr)) >> np.uint64(24)
    cnt = np.zeros((img.nb, nctr), dtype=np.int64)
    np.add.at(cnt, (homes[away], f.astype(np.int64)), 1)
    return cnt, int(away.sum())


def away_fields(img):
    words = np.ascontiguousarray(img.away_words()[:, 1:])  # slots 1..W-1
    return words.view(np.uint16).reshape(img.nb, -1).astype(np.int64)


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_away_counters_equal_a_host_recount(lanes):
    """after inserts, erases of a quarter and re-inserts (bulk), every away counter equals the count of
    stored keys it stands for, recomputed on the host from the slot bytes alone"""
    cls, kdt, sdt, vdt, sb = KINDS["map_u64_u32"]
    cap = 1 << 15
    rng = np.random.default_rng(11)
    keys = np.unique(rng.integers(1, 2**62, size=cap, dtype=np.uint64))[: int(cap * 0.9)]
    rng.shuffle(keys)
    tab = cls.createDeviceObject(cap, sa.Tuning(bucket_lanes=lanes))
    k = torch.from_numpy(keys.view(sdt)).cuda()
    v = torch.from_numpy(value_of(keys, vdt)).cuda()
    tab.insert(k, v)
    quarter = torch.from_numpy(keys[::4].copy().view(sdt)).cuda()
    tab.erase(quarter)
    img = Image(tab, kdt, sb)
    want, n_away = away_recount(img)
    assert n_away > 0
    assert (away_fields(img) == want).all()
    tab.insert(quarter, torch.from_numpy(value_of(keys[::4], vdt)).cuda())
    img = Image(tab, kdt, sb)
    want, n_away = away_recount(img)
    assert (away_fields(img) == want).all()
    assert tab.size() == len(keys) and tab.valid()
    hits, got = tab.find(k)
    assert bool(hits.all())
    type(tab).destroyDeviceObject(tab)
