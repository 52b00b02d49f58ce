"""htab_range_noalloc (device_range that allocates nothing, scratch in a static per-GPU array) gives
exactly what htab_range gives, on two calls in a row, the second over a smaller table; and it
refuses, before any device work, a table with more slots than its static scratch serves."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from stdgpu_amd import _lib  # noqa: E402


def noalloc_range(tab, cap):
    idx = torch.full((max(cap, 1),), -7, dtype=torch.int64, device="cuda")
    n_dev = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    rc = _lib.load().htab_range_noalloc(tab._h, ctypes.c_void_p(idx.data_ptr()), cap, ctypes.c_void_p(n_dev.data_ptr()),
                                        ctypes.c_void_p(s.cuda_stream))
    assert rc == 0, _lib.load().htab_last_error()
    torch.cuda.synchronize()
    n = int(n_dev.item())
    return idx[:n], n


def filled(cls, cap, n, seed, kdt, sdt):
    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(1, 2**62, size=n, dtype=np.uint64).astype(kdt))
    tab = cls.createDeviceObject(cap)
    k = torch.from_numpy(keys.view(sdt)).cuda()
    if cls.mapped is None:
        tab.insert(k)
    else:
        tab.insert(k, (k & 0x7FFFFFFF).to(cls.mapped_dtype()))
    return tab, keys


def test_noalloc_range_equals_range_on_two_calls_second_smaller():
    big, keys = filled(sa.unordered_map_u64_u32, 1 << 20, 600000, 1, np.uint64, np.int64)
    big.erase(torch.from_numpy(keys[::3].copy().view(np.int64)).cuda())
    small, _ = filled(sa.unordered_set_u32, 5000, 3000, 2, np.uint32, np.int32)
    for tab in (big, small):  # the second call covers fewer slots and fewer entries
        ref = tab.device_range()
        got, n = noalloc_range(tab, tab.info.limit + 2)
        assert n == ref.numel() == tab.size()
        assert torch.equal(got, ref)
    # a capacity below the count: only cap indices are written, the count is the full one
    ref = small.device_range()
    got, n = noalloc_range(small, 100)
    assert n == ref.numel()
    assert torch.equal(got[:100], ref[:100])
    sa.unordered_map_u64_u32.destroyDeviceObject(big)
    sa.unordered_set_u32.destroyDeviceObject(small)


def test_noalloc_range_refuses_tables_beyond_its_scratch():
    slots = _lib.HTAB_NOALLOC_RANGE_TILES * 4096
    tab = sa.unordered_set_u32.createDeviceObject(slots + 8192)  # 4-byte slots: about 0.5 GB
    idx = torch.empty(1, dtype=torch.int64, device="cuda")
    n_dev = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    rc = _lib.load().htab_range_noalloc(tab._h, ctypes.c_void_p(idx.data_ptr()), 1, ctypes.c_void_p(n_dev.data_ptr()),
                                        None)
    assert rc == _lib.HTAB_E_LIMIT
    torch.cuda.synchronize()
    assert int(n_dev.item()) == -1  # nothing ran
    sa.unordered_set_u32.destroyDeviceObject(tab)
