"""vector<uint32_t> stack (wave-aggregated push_back / pop_back) against the oracle's restatement of
the reference stack (vector_detail.cuh:181-302).  Element order within one bulk push depends on
thread order in both, so the checks are the order-free facts: the bottom of the stack is untouched,
a bulk pop returns exactly the values the last pushes put on top (as a multiset), refusals happen
exactly at the capacity and at empty, sizes match, and nothing is lost or duplicated."""
import collections

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
    assert view.capacity == 1000 and view.data and view.state and view.end and view.elem_bytes == 4
    assert ctypes.sizeof(_lib.htab_vector_view) == 64
    assert v.valid()
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


# ---------------------------------------------------------------- vector<T> / deque<T> of any size
class RingModel:
    """The reference's single calls in sequence (vector_detail.cuh:181-302, deque_detail.cuh:202-448):
    push_back / push_front append at an end while size < capacity, pop_back / pop_front remove at an
    end while size > 0; a call that finds the container full or empty changes nothing."""

    def __init__(self, cap):
        self.cap, self.items = cap, collections.deque()

    def push(self, rows, back):
        ok = []
        for r in rows:
            if len(self.items) < self.cap:
                (self.items.append if back else self.items.appendleft)(bytes(r))
                ok.append(1)
            else:
                ok.append(0)
        return np.array(ok, dtype=np.uint8)

    def pop(self, n, back):
        out, ok = [], []
        for _ in range(n):
            if self.items:
                out.append(self.items.pop() if back else self.items.popleft())
                ok.append(1)
            else:
                out.append(None)
                ok.append(0)
        return out, np.array(ok, dtype=np.uint8)


def records(rng, n, eb):
    return rng.integers(0, 256, size=(n, eb), dtype=np.uint8)


@pytest.mark.parametrize("eb", [1, 2, 6, 8, 16, 24])
def test_deque_records_at_both_ends_equal_a_sequential_loop(eb):
    """deque<T> with T of eb bytes: bulk push_back / push_front / pop_back / pop_front (sizes that
    overrun and underrun the capacity) give exactly the reference's single calls in sequence, and
    the ring's contents front to back equal the model after every call"""
    rng = np.random.default_rng(100 + eb)
    cap = 1000
    q = sa.deque(cap, elem_bytes=eb, dtype=torch.uint8)
    m = RingModel(cap)
    ops = ["push_back", "push_front", "pop_back", "pop_front"]
    for step in range(40):
        op = ops[int(rng.integers(0, 4))]
        n = int(rng.integers(0, 700))
        if op.startswith("push"):
            rows = records(rng, n, eb)
            pushed = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")[:n]
            getattr(q, op)(torch.from_numpy(rows).cuda(), pushed=pushed)
            want = m.push(rows, op == "push_back")
            assert (pushed.cpu().numpy() == want).all(), (step, op)
        else:
            out, popped = getattr(q, op)(n)
            want_rows, want_ok = m.pop(n, op == "pop_back")
            got_ok = popped.cpu().numpy()
            assert (got_ok == want_ok).all(), (step, op)
            got = out.cpu().numpy().reshape(n, eb)
            for i in range(n):
                if want_ok[i]:
                    assert bytes(got[i]) == want_rows[i], (step, op, i)
                else:
                    assert (got[i] == 255).all()
        assert q.size() == len(m.items)
        c = q.contents().cpu().numpy().reshape(-1, eb)
        assert [bytes(r) for r in c] == list(m.items), (step, op)
        assert q.valid()
    q.clear()
    assert q.size() == 0 and q.valid()
    q.destroy()


@pytest.mark.parametrize("dtype,eb", [(torch.int64, 8), (torch.int16, 2), (torch.uint8, 12)])
def test_vector_of_other_sizes_equals_a_sequential_loop(dtype, eb):
    """vector<T> for 8-, 2- and 12-byte T: bulk push_back / pop_back against the model, contents
    bottom to top, valid() after every call"""
    rng = np.random.default_rng(7 + eb)
    cap = 3000
    v = sa.vector(cap, dtype=dtype, elem_bytes=eb)
    m = RingModel(cap)
    for step in range(30):
        n = int(rng.integers(0, 1500))
        if step % 3 != 2:
            rows = records(rng, n, eb)
            t = torch.from_numpy(rows.copy()).cuda()
            if dtype != torch.uint8:
                t = t.view(dtype).reshape(n)
            pushed = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")[:n]
            v.push_back(t, pushed=pushed)
            assert (pushed.cpu().numpy() == m.push(rows, True)).all()
        else:
            out, popped = v.pop_back(n)
            want_rows, want_ok = m.pop(n, True)
            assert (popped.cpu().numpy() == want_ok).all()
            got = out.cpu().contiguous().view(torch.uint8).numpy().reshape(n, eb)
            for i in range(n):
                if want_ok[i]:
                    assert bytes(got[i]) == want_rows[i]
        c = v.contents().cpu().contiguous().view(torch.uint8).numpy().reshape(-1, eb)
        assert [bytes(r) for r in c] == list(m.items)
        assert v.size() == len(m.items) and v.valid()
    v.destroy()


def test_ring_arguments_are_checked_before_device_work():
    from stdgpu_amd import _lib

    lib = _lib.load()
    import ctypes

    h = ctypes.c_void_p()
    assert lib.htab_deque_create_sized(100, 0, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert lib.htab_deque_create_sized(100, 257, ctypes.byref(h)) == _lib.HTAB_E_ARG
    assert lib.htab_vector_create_sized(0, 8, ctypes.byref(h)) == _lib.HTAB_E_ARG
    q = sa.deque(10, dtype=torch.int64)
    # 8-byte elements need 8-byte aligned arrays
    assert lib.htab_deque_push_back(q._h, ctypes.c_void_p(0x1004), 1, None, None) == _lib.HTAB_E_ARG
    # fill_iota is defined for 4-byte elements
    assert lib.htab_deque_fill_iota(q._h, None) == _lib.HTAB_E_ARG
    q.destroy()


def test_deque_u32_both_ends_equal_the_oracle():
    """deque<uint32_t>: the same sequence of bulk calls at both ends on the GPU and through the
    oracle's single calls (deque_detail.cuh:202-448, one thread): every returned element, flag, the
    size and the contents front to back agree"""
    rng = np.random.default_rng(23)
    cap = 4096
    g, r = sa.deque_u32(cap), po.Deque(cap)
    for step in range(60):
        op = int(rng.integers(0, 4))
        n = int(rng.integers(0, 3000))
        if op < 2:
            vals = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
            pushed = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")[:n]
            (g.push_back if op == 0 else g.push_front)(dev(vals), pushed=pushed)
            want = (r.push_back if op == 0 else r.push_front)(vals)
            assert (pushed.cpu().numpy() == want).all(), step
        else:
            out, popped = (g.pop_back if op == 2 else g.pop_front)(n)
            want_out, want_ok = (r.pop_back if op == 2 else r.pop_front)(n)
            assert (popped.cpu().numpy() == want_ok).all(), step
            assert (host(out) == want_out).all(), step
        assert g.size() == r.size()
        assert (host(g.contents()) == r.contents()).all(), step
    assert g.valid()
    g.destroy()
    r.close()
