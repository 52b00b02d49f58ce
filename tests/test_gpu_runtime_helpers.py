"""The C ABI's runtime helpers that the C++ facade's host code goes through (htab_rt_*): allocation,
copies both ways with a stream wait, and the zero fill kernel over ranges whose ends are not 16-byte
aligned (its head, 16-byte body and tail)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd  # noqa: E402,F401
from stdgpu_amd import _lib  # noqa: E402


@pytest.mark.parametrize("start,length", [(0, 4096), (3, 1), (5, 17), (1, 300), (13, 70000), (16, 32)])
def test_zero_fill_and_copies(start, length):
    lib = _lib.load()
    total = 80000
    p = ctypes.c_void_p()
    assert lib.htab_rt_alloc(total, ctypes.byref(p)) == 0 and p.value
    host = np.arange(total, dtype=np.uint32).astype(np.uint8) | 1  # no zero byte anywhere
    assert lib.htab_rt_copy(p, host.ctypes.data_as(ctypes.c_void_p), total, 1, None) == 0
    assert lib.htab_rt_zero(ctypes.c_void_p(p.value + start), length, None) == 0
    assert lib.htab_rt_sync(None) == 0 and lib.htab_rt_launch_status() == 0
    back = np.zeros(total, dtype=np.uint8)
    assert lib.htab_rt_copy(back.ctypes.data_as(ctypes.c_void_p), p, total, 0, None) == 0
    want = host.copy()
    want[start:start + length] = 0
    np.testing.assert_array_equal(back, want)
    # device to device
    q = ctypes.c_void_p()
    assert lib.htab_rt_alloc(total, ctypes.byref(q)) == 0
    assert lib.htab_rt_copy(q, p, total, 2, None) == 0
    back2 = np.zeros(total, dtype=np.uint8)
    assert lib.htab_rt_copy(back2.ctypes.data_as(ctypes.c_void_p), q, total, 0, None) == 0
    np.testing.assert_array_equal(back2, want)
    assert lib.htab_rt_free(p) == 0 and lib.htab_rt_free(q) == 0
