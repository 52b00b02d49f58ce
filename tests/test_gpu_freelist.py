"""Free-list calls between a map and a deque<uint32_t> of its values (htab_assign_from_deque,
htab_erase_far_blocks): the take must hand out the deque's front elements, each once, to exactly the
keys the insert added; the evict must erase exactly the keys numpy's float64 selection calls far (the
oracle's table after the same erase holds the same keys), push exactly their values, and leave every
overflow and away count as valid() recomputes them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from stdgpu_amd import voxel as vx  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def d64(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


def deque_of(values, cap):
    q = sa.deque_u32(cap)
    if len(values):
        q.push_back(torch.from_numpy(np.asarray(values, dtype=np.uint32).view(np.int32)).cuda())
    return q


def deque_list(q):
    return [int(x) & 0xFFFFFFFF for x in q.contents().cpu().tolist()]


def table_dict(t):
    k, v = t.contents()
    kn = k.cpu().numpy()
    kn = kn.view(np.uint64) if kn.dtype.itemsize == 8 else kn.view(np.uint32)
    return dict(zip(kn.tolist(), (v.cpu().numpy().view(np.uint32)).tolist()))


@pytest.mark.parametrize("deque_len", [1000, 5000])
def test_take_hands_out_the_front_elements_to_the_added_keys(deque_len):
    rng = np.random.default_rng(5)
    tab = sa.unordered_map_u64_u32.createDeviceObject(1 << 14)
    old = rng.integers(0, 1 << 62, size=1500, dtype=np.int64)
    tab.insert(d64(old), torch.full((1500,), 7, dtype=torch.int32, device="cuda"))
    # 3000 keys: half already present, a quarter new, a quarter repeats of the new ones
    new = rng.integers(1 << 62, 1 << 63, size=750, dtype=np.int64)
    keys = np.concatenate([old, new, new])
    rng.shuffle(keys)
    keys = d64(keys)
    flags = torch.empty(keys.numel(), dtype=torch.uint8, device="cuda")
    tab.insert(keys, torch.full((keys.numel(),), -1, dtype=torch.int32, device="cuda"), inserted=flags)
    f0 = flags.cpu().numpy().copy()
    assert int((f0 == 1).sum()) == 750
    front = list(range(100, 100 + deque_len))
    q = deque_of(front, 8192)
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    tab.assign_from_deque(keys, flags, q, count=count)
    f1 = flags.cpu().numpy()
    popped = min(750, deque_len)
    assert int(count.item()) == popped
    assert np.array_equal(f1[f0 != 1], f0[f0 != 1])
    assert int((f1 == 1).sum()) == popped and int((f1 == 2).sum()) == 750 - popped
    got = table_dict(tab)
    kk = keys.cpu().numpy().view(np.uint64)
    given = [got[int(k)] for k in kk[f1 == 1]]
    # the first `popped` elements of the deque, each to one added key
    assert sorted(given) == front[:popped]
    for k in kk[f1 == 2]:
        assert got[int(k)] == 0xFFFFFFFF
    for k in old.view(np.uint64):
        assert got[int(k)] == 7
    assert deque_list(q) == front[popped:]
    assert q.valid() and tab.valid()
    # a flag of 1 on an absent key: 3, nothing popped
    absent = d64(np.array([12345], dtype=np.int64))
    fl = torch.ones(1, dtype=torch.uint8, device="cuda")
    tab.assign_from_deque(absent, fl, q, count=count)
    assert int(fl.item()) == 3 and int(count.item()) == 0 and deque_list(q) == front[popped:]
    q.destroy()
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


def packed_blocks(rng, n, span):
    c = rng.integers(-span, span, size=(n, 3), dtype=np.int64)
    k = np.unique(vx.pack(c[:, 0], c[:, 1], c[:, 2]))
    rng.shuffle(k)
    return k


@pytest.mark.parametrize("room_short", [False, True])
def test_evict_erases_exactly_the_far_blocks_and_pushes_their_values(room_short):
    rng = np.random.default_rng(9)
    # a full table (load 0.9 of the capacity, then erases and re-inserts, so keys sit past their home
    # buckets behind tombstones), blocks within +-400 of the origin (+-25.6 m)
    keys = packed_blocks(rng, 30000, 400)
    n = len(keys)
    cap = int(n / 0.9)
    tab = sa.unordered_map_u64_u32.createDeviceObject(cap)
    ref = po.Table("u64", "u32", cap)
    vals = np.arange(n, dtype=np.uint32)
    tab.insert(d64(keys), torch.from_numpy(vals.view(np.int32)).cuda())
    ref.insert(keys.view(np.uint64), vals)
    gone = keys[: n // 5]
    tab.erase(d64(gone))
    ref.erase(gone.view(np.uint64))
    tab.insert(d64(gone), torch.from_numpy(vals[: n // 5].view(np.int32)).cuda())
    ref.insert(gone.view(np.uint64), vals[: n // 5])
    # the two reserved key values are blocks too (all fields 2^21 - 1 and the last one 2^21 - 2)
    side = np.array([-1, -2], dtype=np.int64)
    tab.insert(d64(side), torch.tensor([n, n + 1], dtype=torch.int32, device="cuda"))
    ref.insert(side.view(np.uint64), np.array([n, n + 1], dtype=np.uint32))
    allk = np.concatenate([keys, side])
    allv = np.concatenate([vals, np.array([n, n + 1], dtype=np.uint32)])
    cam = (3.1, -2.0, 7.7)
    far = vx.far_mask(allk, cam)
    nfar = int(far.sum())
    assert 1000 < nfar < n - 1000 and far[-2:].all()
    room = nfar - 777 if room_short else nfar + 10
    q = sa.deque_u32(room + 50)
    q.push_back(torch.arange(50, dtype=torch.int32, device="cuda") + 900000)
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    tab.erase_far_blocks(cam, vx.RADIUS, vx.BLOCK, q, count=count)
    erased = int(count.item())
    assert erased == min(nfar, room)
    after = table_dict(tab)
    assert tab.size() == len(allk) - erased == len(after)
    # near blocks all stay, with their values; far blocks stay only past the deque's room
    near_k = allk[~far].view(np.uint64)
    assert all(after[int(k)] == int(v) for k, v in zip(near_k, allv[~far]))
    stayed = [int(k) for k in allk[far].view(np.uint64) if int(k) in after]
    assert len(stayed) == nfar - erased
    pushed = deque_list(q)
    assert pushed[:50] == list(range(900000, 900050))
    gone_vals = sorted(int(v) for k, v in zip(allk[far].view(np.uint64), allv[far]) if int(k) not in after)
    assert sorted(pushed[50:]) == gone_vals
    # the oracle's table after erasing the same keys holds the same entries
    ref.erase(np.array([int(k) for k in allk[far].view(np.uint64) if int(k) not in after], dtype=np.uint64))
    rk, rv = ref.sorted_contents()
    assert dict(zip(rk.tolist(), rv.tolist())) == after
    # every overflow and away count recomputed, and every remaining key still found
    assert tab.valid() and q.valid()
    h, v = tab.find(d64(np.array(list(after.keys()), dtype=np.uint64).view(np.int64)))
    assert bool(h.all())
    q.destroy()
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)


def test_free_list_calls_refuse_sets_and_mismatched_deques():
    from stdgpu_amd import _lib

    s = sa.unordered_set_u64.createDeviceObject(100)
    m = sa.unordered_map_u64_u64.createDeviceObject(100)
    q = sa.deque_u32(10)
    k = torch.zeros(1, dtype=torch.int64, device="cuda")
    f = torch.ones(1, dtype=torch.uint8, device="cuda")
    for t in (s, m):
        with pytest.raises(_lib.HtabError):
            t.assign_from_deque(k, f, q)
        with pytest.raises(_lib.HtabError):
            t.erase_far_blocks((0, 0, 0), 1.0, 1.0, q)
    mm = sa.unordered_map_u64_u32.createDeviceObject(100)
    with pytest.raises(_lib.HtabError):
        mm.erase_far_blocks((0, 0, 0), 1.0, 1.0, q, bits=22)
    assert _lib.load().htab_last_error()
    q.destroy()


def test_free_list_calls_on_u32_keys():
    """unordered_map<u32,u32> (8-byte slots, 16-slot buckets): blocks of 10 bits per axis; the take
    gives every added key one of the deque's front elements, the evict erases exactly numpy's far
    selection and pushes its values, and the oracle's table after the same calls holds the same
    entries."""
    rng = np.random.default_rng(21)
    c = rng.integers(-500, 500, size=(6000, 3), dtype=np.int64)
    off = 1 << 9
    k = np.unique(((c[:, 0] + off) << 20) | ((c[:, 1] + off) << 10) | (c[:, 2] + off)).astype(np.uint32)
    n = len(k)
    cap = int(n / 0.85)
    tab = sa.unordered_map_u32_u32.createDeviceObject(cap)
    ref = po.Table("u32", "u32", cap)
    keys = torch.from_numpy(k.view(np.int32)).cuda()
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")
    tab.insert(keys, torch.full((n,), -1, dtype=torch.int32, device="cuda"), inserted=flags)
    assert bool((flags == 1).all())
    q = sa.deque_u32(n + 100)
    q.fill_iota()
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    tab.assign_from_deque(keys, flags, q, count=count)
    assert int(count.item()) == n and deque_list(q) == list(range(n, n + 100))
    got = table_dict(tab)
    assert sorted(got.values()) == list(range(n))
    ref.insert(k, np.array([got[int(x)] for x in k], dtype=np.uint32))
    # the far test on 10-bit fields, block size 1.0, centre (3, -2, 7), radius 300
    fx = ((k >> 20) & 1023).astype(np.int64) - off
    fy = ((k >> 10) & 1023).astype(np.int64) - off
    fz = (k & 1023).astype(np.int64) - off
    d2 = (fx + 0.5 - 3.0) ** 2 + (fy + 0.5 + 2.0) ** 2 + (fz + 0.5 - 7.0) ** 2
    far = d2 > 300.0 * 300.0
    assert 500 < int(far.sum()) < n - 500
    tab.erase_far_blocks((3.0, -2.0, 7.0), 300.0, 1.0, q, bits=10, offset=off, count=count)
    assert int(count.item()) == int(far.sum())
    after = table_dict(tab)
    assert set(after) == set(int(x) for x in k[~far])
    pushed = deque_list(q)[100:]
    assert sorted(pushed) == sorted(got[int(x)] for x in k[far])
    ref.erase(k[far])
    rk, rv = ref.sorted_contents()
    assert dict(zip(rk.tolist(), rv.tolist())) == after
    assert tab.valid() and q.valid()
    q.destroy()
    sa.unordered_map_u32_u32.destroyDeviceObject(tab)
