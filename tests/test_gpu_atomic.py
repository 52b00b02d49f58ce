"""atomic<T> / atomic_ref<T> (atomic.cuh) on the GPU against the oracle's sequential statement of each
operation.

The reference's own tests (tests/stdgpu/atomic.inc) run one operation per thread on one value and
check the final value against a closed form; the same sequences run here as bulk calls, and beyond the
final value every returned old value is checked: the oracle must find an order of the operations,
applied one at a time, in which each operation sees the value the GPU returned for it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

NP = {"int": np.int32, "unsigned int": np.uint32, "unsigned long long int": np.uint64, "float": np.float32}
BITS_DT = {"int": np.int32, "unsigned int": np.int32, "unsigned long long int": np.int64, "float": np.float32}
INT_TYPES = ["int", "unsigned int", "unsigned long long int"]


def dev(vals, t):
    if t == "unsigned long long int":
        a = np.array([int(x) for x in vals], dtype=np.uint64)
    else:
        a = np.asarray(vals, dtype=NP[t])
    return torch.from_numpy(np.ascontiguousarray(a).view(BITS_DT[t]).copy()).cuda()


def host(x, t):
    a = x.cpu().numpy().view(NP[t])
    return [float(v) for v in a] if t == "float" else [int(v) for v in a]


def top(t):
    return (1 << 31) - 1 if t == "int" else (1 << po.ATOMIC_BITS[t]) - 1


def lowest(t):
    return -(1 << 31) if t == "int" else 0


def run(a, t, op, operands, initial, targets=None, expected=None):
    a.store(initial)
    old, done = a.apply(op, None if operands is None else dev(operands, t), targets,
                        None if expected is None else dev(expected, t))
    torch.cuda.synchronize()
    return (None if old is None else host(old, t)), done.cpu().numpy()


@pytest.mark.parametrize("t", INT_TYPES)
def test_reference_sequences(t):
    """the closed forms of atomic.inc's sequence tests, and an order for every returned old value"""
    a = sa.atomic.createDeviceObject(t)
    assert a.load() == 0  # createDeviceObject fills 0
    n = 40000
    seq = list(range(1, n + 1))
    total = n * (n + 1) // 2
    cases = [("fetch_add", 0, seq, total), ("fetch_sub", total, seq, 0), ("fetch_min", top(t), seq, 1),
             ("fetch_max", lowest(t), seq, n)]
    bits = po.ATOMIC_BITS[t]
    pat = [po._wrap(t, 1 << i) for i in range(bits)]
    ones = po._wrap(t, (1 << bits) - 1)
    cases += [("fetch_or", 0, pat, ones), ("fetch_xor", 0, pat, ones),
              ("fetch_and", ones, [po._wrap(t, ~(1 << i)) for i in range(bits)], 0)]
    for op, init, ops, want in cases:
        old, done = run(a, t, op, ops, init)
        assert a.load() == want, op
        assert done.all()
        assert po.atomic_linearize(t, op, init, ops, old) == want, op
    # exchange of distinct values: the old values chain from the start value to the final one
    ex = [po._wrap(t, 7 * i + 3) for i in range(n)]
    old, done = run(a, t, "exchange", ex, 5)
    assert po.atomic_linearize(t, "exchange", 5, ex, old) == a.load()
    # store: the value is one of the stored operands
    run(a, t, "store", ex[:1000], 5)
    assert a.load() in set(ex[:1000])
    # load sees the value
    old, _ = run(a, t, "load", None, 77, targets=torch.zeros(300, dtype=torch.int64, device="cuda"))
    assert old == [77] * 300
    sa.atomic.destroyDeviceObject(a)


@pytest.mark.parametrize("t", INT_TYPES + ["float"])
def test_compare_exchange(t):
    a = sa.atomic.createDeviceObject(t)
    n = 5000
    # every operation expects the start value: exactly one replaces it, the others see its value
    desired = [po._wrap(t, i + 1) if t != "float" else float(i + 1) for i in range(n)]
    old, done = run(a, t, "compare_exchange", desired, 0, expected=[0] * n)
    assert done.sum() == 1
    w = int(np.flatnonzero(done)[0])
    assert a.load() == desired[w]
    assert all(o == 0 for i, o in enumerate(old) if i == w) and all(o == desired[w] for i, o in enumerate(old) if i != w)
    # the reference's compare-exchange loop (atomic.inc:584-703) driven in rounds: each pending
    # operation retries with the value it saw until it replaces it; the sum of the sequence results
    m = 256
    seq = np.arange(1, m + 1)
    a.store(0)
    seen = np.zeros(m)
    pending = np.ones(m, dtype=bool)
    rounds = 0
    while pending.any():
        idx = np.flatnonzero(pending)
        exp = seen[idx]
        old, done = a.apply("compare_exchange", dev(exp + seq[idx], t), None, dev(exp, t))
        old = np.array(host(old, t), dtype=np.float64)
        done = done.cpu().numpy().astype(bool)
        assert done.sum() >= 1  # some operation always succeeds
        seen[idx] = old
        pending[idx[done]] = False
        rounds += 1
        assert rounds <= m
    assert a.load() == m * (m + 1) // 2
    sa.atomic.destroyDeviceObject(a)


def test_wrapping_increment_decrement():
    """atomic.inc:1990-2063: 50000 wrapping increments (then decrements) modulo 5000 from 42 give 42"""
    a = sa.atomic.createDeviceObject("unsigned int")
    n, mod = 50000, 5000
    for op in ("fetch_inc_mod", "fetch_dec_mod"):
        old, done = run(a, "unsigned int", op, [mod] * n, 42)
        assert a.load() == 42 and done.all()
        assert po.atomic_linearize("unsigned int", op, 42, [mod] * n, old) == 42
    # the edges of the wrap: above the limit
    old, _ = run(a, "unsigned int", "fetch_inc_mod", [5000], 7000)
    assert old == [7000] and a.load() == 0
    run(a, "unsigned int", "fetch_dec_mod", [5000], 0)
    assert a.load() == 4999
    sa.atomic.destroyDeviceObject(a)


@pytest.mark.parametrize("count", [1, 3, 1000])
@pytest.mark.parametrize("t", INT_TYPES)
@pytest.mark.parametrize("op", ["fetch_add", "fetch_min", "fetch_max", "fetch_or", "exchange"])
def test_random_targets(t, op, count):
    """operations on many elements: waves whose lanes name one element and waves that mix elements,
    out-of-range targets skipped; per element the final value and an order of its operations"""
    rng = np.random.default_rng(count * 31 + len(op) + len(t))
    n = 30000
    a = sa.atomic.createDeviceObject(t, count)
    init = [po._wrap(t, int(v)) for v in rng.integers(0, 1 << 20, size=count)]
    if op == "fetch_min":
        init = [top(t)] * count
    for i, v in enumerate(init):
        a.store(v, i)
    # runs of one target (uniform waves) mixed with random targets and a few out of range
    tg = np.repeat(rng.integers(0, count, size=n // 64 + 1), 64)[:n]
    mix = rng.random(n) < 0.3
    tg[mix] = rng.integers(0, count, size=int(mix.sum()))
    tg[rng.random(n) < 0.01] = count + 5
    tg[rng.random(n) < 0.01] = -1
    if op == "exchange":
        ops = [po._wrap(t, 11 * i + 1) for i in range(n)]
    elif op == "fetch_or":
        ops = [po._wrap(t, 1 << int(b)) for b in rng.integers(0, po.ATOMIC_BITS[t], size=n)]
    else:
        ops = [int(v) for v in rng.integers(1, 1 << 16, size=n)]
    old, done = a.apply(op, dev(ops, t), torch.from_numpy(tg).cuda())
    torch.cuda.synchronize()
    old, done = host(old, t), done.cpu().numpy()
    inr = (tg >= 0) & (tg < count)
    np.testing.assert_array_equal(done, inr.astype(np.uint8))
    for e in range(count):
        idx = np.flatnonzero(tg == e)
        final = po.atomic_linearize(t, op, init[e], [ops[i] for i in idx], [old[i] for i in idx])
        assert a.load(e) == final
        if op != "exchange":  # commutative: the final value is the sequential one
            assert final == po.atomic_sequence(t, op, init[e], [ops[i] for i in idx])[0]
    sa.atomic.destroyDeviceObject(a)


def test_float_operations():
    a = sa.atomic.createDeviceObject("float")
    n = 4000
    seq = [float(i) for i in range(1, n + 1)]  # integer sums below 2^24 are exact in float
    old, done = run(a, "float", "fetch_add", seq, 0.0)
    assert a.load() == n * (n + 1) / 2 and done.all()
    assert po.atomic_linearize("float", "fetch_add", 0.0, seq, old) == n * (n + 1) / 2
    run(a, "float", "fetch_sub", seq, n * (n + 1) / 2)
    assert a.load() == 0.0
    big = float(np.float32(1e30))  # the float value the device holds
    old, _ = run(a, "float", "fetch_min", seq, big)
    assert a.load() == 1.0 and po.atomic_linearize("float", "fetch_min", big, seq, old) == 1.0
    old, _ = run(a, "float", "fetch_max", [-x for x in seq], -big)
    assert a.load() == -1.0 and po.atomic_linearize("float", "fetch_max", -big, [-x for x in seq], old) == -1.0
    # value comparison: -0 equals +0
    old, done = run(a, "float", "compare_exchange", [5.0], -0.0, expected=[0.0])
    assert done.tolist() == [1] and a.load() == 5.0
    with pytest.raises(sa.HtabError):
        a.apply("fetch_and", dev([1.0], "float"))
    sa.atomic.destroyDeviceObject(a)


def test_atomic_ref_and_errors():
    # atomic_ref over a caller's device array
    vals = torch.zeros(10, dtype=torch.int64, device="cuda")
    tg = torch.arange(10000, device="cuda", dtype=torch.int64) % 10
    ops = torch.arange(1, 10001, device="cuda", dtype=torch.int64)
    old, done = sa.atomic_ref_apply("unsigned long long int", vals, "fetch_add", ops, tg)
    want = np.zeros(10, dtype=np.int64)
    np.add.at(want, np.arange(10000) % 10, np.arange(1, 10001))
    np.testing.assert_array_equal(vals.cpu().numpy(), want)
    assert bool(done.all())
    # host load / store of the top bits of a 64-bit value
    a = sa.atomic.createDeviceObject("unsigned long long int", 2)
    a.store((1 << 64) - 3, 1)
    assert a.load(1) == (1 << 64) - 3 and a.load(0) == 0
    old, _ = a.apply("fetch_add", dev([5], "unsigned long long int"), torch.ones(1, dtype=torch.int64, device="cuda"))
    assert host(old, "unsigned long long int") == [(1 << 64) - 3] and a.load(1) == 2
    with pytest.raises(sa.HtabError):
        a.apply("fetch_inc_mod", dev([5], "unsigned long long int"))  # unsigned int only
    with pytest.raises(sa.HtabError):
        a.load(2)
    sa.atomic.destroyDeviceObject(a)
    i = sa.atomic.createDeviceObject("int")
    run(i, "int", "fetch_add", [1], (1 << 31) - 1)
    assert i.load() == -(1 << 31)  # two's complement wrap
    sa.atomic.destroyDeviceObject(i)
