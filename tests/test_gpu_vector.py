"""vector<uint32_t> stack (wave-aggregated push_back / pop_back) against the oracle's restatement of
the reference stack (vector_detail.cuh:181-302).  Element order within one bulk push depends on
thread order in both, so the checks are the order-free facts: the bottom of the stack is untouched,
a bulk pop returns exactly the values the last pushes put on top (as a multiset), refusals happen
exactly at the capacity and at empty, sizes match, and nothing is lost or duplicated."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def dev(a):
    return torch.from_numpy(np.array(a, dtype=np.uint32, copy=True).view(np.int32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("cap", [1, 100, 1 << 16, (1 << 20) + 3])
def test_vector_matches_reference_stack(cap):
    rng = np.random.default_rng(cap)
    g, r = sa.vector_u32(cap), po.Vector(cap)
    a = rng.integers(0, 2**32 - 1, size=cap // 2 + 1, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32 - 1, size=cap // 3 + 1, dtype=np.uint64).astype(np.uint32)
    pa = torch.zeros(len(a), dtype=torch.uint8, device="cuda")
    g.push_back(dev(a), pushed=pa)
    ra = r.push_back(a, threads=4)
    assert int(pa.sum()) == int(ra.sum()) and g.size() == r.size()
    bottom = np.sort(host(g.contents()))
    np.testing.assert_array_equal(bottom, np.sort(r.contents()))
    pb = torch.zeros(len(b), dtype=torch.uint8, device="cuda")
    g.push_back(dev(b), pushed=pb)
    rb = r.push_back(b, threads=4)
    assert int(pb.sum()) == int(rb.sum()) and g.size() == r.size() == min(cap, len(a) + len(b))
    # the first push's values stay at the bottom, in place
    np.testing.assert_array_equal(np.sort(host(g.contents())[: int(pa.sum())]), bottom)
    # popping what the second push added returns exactly those values
    k = int(pb.sum())
    out, popped = g.pop_back(k)
    rout, rpopped = r.pop_back(k, threads=4)
    assert int(popped.sum()) == int(rpopped.sum()) == k
    np.testing.assert_array_equal(np.sort(host(out)), np.sort(b[host(pb.to(torch.int32)).astype(bool)] if k else b[:0]))
    np.testing.assert_array_equal(np.sort(host(out)), np.sort(rout))
    # pop past empty: exactly the size comes back, the rest is refused with the all-ones marker
    n = g.size()
    out2, popped2 = g.pop_back(n + 77)
    assert int(popped2.sum()) == n and g.size() == 0
    o2 = host(out2)
    assert (o2[host(popped2.to(torch.int32)) == 0] == 0xFFFFFFFF).all()
    np.testing.assert_array_equal(np.sort(o2[host(popped2.to(torch.int32)) == 1]), bottom)
    # push past full
    big = np.arange(cap + 50, dtype=np.uint32)
    pf = torch.zeros(len(big), dtype=torch.uint8, device="cuda")
    g.push_back(dev(big), pushed=pf)
    assert int(pf.sum()) == cap and g.size() == cap
    np.testing.assert_array_equal(np.sort(host(g.contents())), np.sort(big[host(pf.to(torch.int32)).astype(bool)]))
    # iota
    g.fill_iota(12345)
    assert g.size() == cap
    np.testing.assert_array_equal(host(g.contents()), np.arange(12345, 12345 + cap, dtype=np.uint32))
    g.destroy()
    r.close()


def test_bulk_calls_equal_a_sequential_loop():
    """A bulk call reserves its positions at once, so it equals a loop of single push_back / pop_back
    (pop_front) calls, element by element: checked against the oracle run on one thread."""
    rng = np.random.default_rng(3)
    cap = 5000
    gv, rv = sa.vector_u32(cap), po.Vector(cap)
    gq, rq = sa.deque_u32(cap), po.Deque(cap)
    for step in range(12):
        n = int(rng.integers(0, 3000))
        if step % 3 != 2:
            a = rng.integers(0, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)
            for g, r in ((gv, rv), (gq, rq)):
                pg = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")[:n]
                g.push_back(dev(a), pushed=pg)
                np.testing.assert_array_equal(pg.cpu().numpy(), r.push_back(a))
        else:
            for g, r, pop in ((gv, rv, "pop_back"), (gq, rq, "pop_front")):
                out, ok = getattr(g, pop)(n)
                rout, rok = getattr(r, pop)(n)
                np.testing.assert_array_equal(ok.cpu().numpy(), rok)
                np.testing.assert_array_equal(host(out), rout)
        for g, r in ((gv, rv), (gq, rq)):
            assert g.size() == r.size()
            np.testing.assert_array_equal(host(g.contents()), r.contents())
    for g in (gv, gq):
        g.destroy()
    rv.close()
    rq.close()


def test_clear_view_and_dropping_pops():
    """clear() empties the stack in one kernel; the view names the elements, the size word and the
    capacity; a pop with no output buffer drops the top values (the C++ facade's erase)"""
    import ctypes

    from stdgpu_amd import _lib

    v = sa.vector_u32(1000)
    v.fill_iota(5)
    assert v.size() == 1000
    view = v.view()
    assert view.capacity == 1000 and view.data and view.size
    assert ctypes.sizeof(_lib.htab_vector_view) == 24
    flags = torch.zeros(300, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert _lib.load().htab_vector_pop_back(v._h, None, 300, flags.data_ptr(), s) == 0
    torch.cuda.synchronize()
    assert int(flags.sum()) == 300 and v.size() == 700
    assert (host(v.contents()) == np.arange(5, 705, dtype=np.uint32)).all()
    v.clear()
    torch.cuda.synchronize()
    assert v.size() == 0 and v.contents().numel() == 0
    pushed = torch.zeros(3, dtype=torch.uint8, device="cuda")
    v.push_back(dev([7, 8, 9]), pushed=pushed)
    assert int(pushed.sum()) == 3 and v.size() == 3
    v.destroy()
