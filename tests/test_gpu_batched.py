"""Batched form: many tables in one persistent launch (and as a replayed HIP graph) must give every
table exactly what a plain per-table call gives it (contents, flags, hits, values, size), and what
the oracle gives it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import stdgpu_amd as sa  # noqa: E402
from stdgpu_amd import workloads as wl  # noqa: E402
from stdgpu_amd.batched import TableBatch  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def d64(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


def d32(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def sorted_contents(t):
    k, v = t.contents()
    k = k.cpu().numpy().view(np.uint64)
    v = v.cpu().numpy().view(np.uint32)
    o = np.argsort(k)
    return k[o], v[o]


def make_tables(w, lanes):
    return [sa.unordered_map_u64_u32.createDeviceObject(int(c), sa.Tuning(bucket_lanes=lanes)) for c in w.capacities]


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_batched_phases_equal_plain_calls(lanes):
    w = wl.batched_maps(ntables=64, log2_lo=6, log2_hi=12, seed=11)
    bt = make_tables(w, lanes)
    pt = make_tables(w, lanes)
    batch = TableBatch(bt)
    n = len(w.ins_keys)
    keys, vals = d64(w.ins_keys), d32(w.ins_vals)
    flags = torch.zeros(n, dtype=torch.uint8, device="cuda")
    batch.run("insert", d64(w.ins_offsets), keys, vals, flags=flags)
    bflags = flags.cpu().numpy()
    for t, tab in enumerate(pt):
        a, b = w.ins_offsets[t], w.ins_offsets[t + 1]
        f = torch.zeros(b - a, dtype=torch.uint8, device="cuda")
        tab.insert(keys[a:b], vals[a:b], inserted=f)
        np.testing.assert_array_equal(bflags[a:b], f.cpu().numpy())
        assert bt[t].size() == tab.size() == b - a
    # find
    fk = d64(w.find_keys)
    hits = torch.zeros(len(w.find_keys), dtype=torch.uint8, device="cuda")
    got = torch.zeros(len(w.find_keys), dtype=torch.int32, device="cuda")
    batch.run("find", d64(w.find_offsets), fk, flags=hits, vals_out=got)
    bh, bv = hits.cpu().numpy(), got.cpu().numpy().view(np.uint32)
    for t, tab in enumerate(pt):
        a, b = w.find_offsets[t], w.find_offsets[t + 1]
        h, v = tab.find(fk[a:b])
        np.testing.assert_array_equal(bh[a:b], h.cpu().numpy())
        hv = h.cpu().numpy() == 1
        np.testing.assert_array_equal(bv[a:b][hv], v.cpu().numpy().view(np.uint32)[hv])
    # erase
    ek = d64(w.erase_keys)
    ef = torch.zeros(len(w.erase_keys), dtype=torch.uint8, device="cuda")
    batch.run("erase", d64(w.erase_offsets), ek, flags=ef)
    be = ef.cpu().numpy()
    for t, tab in enumerate(pt):
        a, b = w.erase_offsets[t], w.erase_offsets[t + 1]
        f = torch.zeros(b - a, dtype=torch.uint8, device="cuda")
        tab.erase(ek[a:b], erased=f)
        np.testing.assert_array_equal(be[a:b], f.cpu().numpy())
        ka, va = sorted_contents(bt[t])
        kb, vb = sorted_contents(tab)
        np.testing.assert_array_equal(ka, kb)
        np.testing.assert_array_equal(va, vb)
        assert bt[t].valid()
    # oracle on a few tables
    for t in [0, 17, 63]:
        ref = po.Table("u64", "u32", int(w.capacities[t]))
        a, b = w.ins_offsets[t], w.ins_offsets[t + 1]
        ref.insert(w.ins_keys[a:b], w.ins_vals[a:b])
        a, b = w.erase_offsets[t], w.erase_offsets[t + 1]
        ref.erase(w.erase_keys[a:b])
        rk, rv = ref.sorted_contents()
        ka, va = sorted_contents(bt[t])
        np.testing.assert_array_equal(ka, rk)
        np.testing.assert_array_equal(va, rv)
    batch.destroy()


def test_batched_graph_replays_match_eager():
    w = wl.batched_maps(ntables=128, log2_lo=6, log2_hi=11, seed=12)
    tabs = make_tables(w, 8)
    batch = TableBatch(tabs)
    batch.run("insert", d64(w.ins_offsets), d64(w.ins_keys), d32(w.ins_vals))
    torch.cuda.synchronize()
    fk, fo = d64(w.find_keys), d64(w.find_offsets)
    eager_h = torch.zeros(len(w.find_keys), dtype=torch.uint8, device="cuda")
    eager_v = torch.zeros(len(w.find_keys), dtype=torch.int32, device="cuda")
    batch.run("find", fo, fk, flags=eager_h, vals_out=eager_v)
    torch.cuda.synchronize()
    gh = torch.zeros_like(eager_h)
    gv = torch.zeros_like(eager_v)
    s = torch.cuda.Stream()
    g = batch.capture("find", fo, fk, flags=gh, vals_out=gv, stream=s)
    other = sa.unordered_set_u64.createDeviceObject(1000)
    for rep in range(3):
        gh.zero_()
        gv.zero_()
        torch.cuda.synchronize()
        g.launch(s)
        s.synchronize()
        assert torch.equal(gh, eager_h), rep
        assert torch.equal(gv[eager_h == 1], eager_v[eager_h == 1]), rep
        other.insert(torch.arange(1, 100, dtype=torch.int64, device="cuda") + rep)  # eager work between replays
        torch.cuda.synchronize()
    g.destroy()
    batch.destroy()


def test_batched_mixed_ops_one_launch():
    """one launch: tables 0,3,6.. insert, 1,4,7.. find, 2,5,8.. erase (the serving form's launch)"""
    w = wl.batched_maps(ntables=48, log2_lo=6, log2_hi=11, seed=13)
    tabs = make_tables(w, 8)
    plain = make_tables(w, 8)
    batch = TableBatch(tabs)
    # pre-load every table so finds and erases have something to hit
    batch.run("insert", d64(w.ins_offsets), d64(w.ins_keys), d32(w.ins_vals))
    for t, tab in enumerate(plain):
        a, b = w.ins_offsets[t], w.ins_offsets[t + 1]
        tab.insert(d64(w.ins_keys[a:b]), d32(w.ins_vals[a:b]))
    ntab = len(tabs)
    ops = np.array([t % 3 for t in range(ntab)], np.int8)
    # per table: inserts use fresh keys, finds use the find keys, erases use the erase keys
    parts_k, parts_v, offs = [], [], [0]
    fresh = wl.stream(99, int(w.sizes.sum()))
    for t in range(ntab):
        if ops[t] == 0:
            k = fresh[w.ins_offsets[t]: w.ins_offsets[t + 1]][: max(1, w.sizes[t] // 4)]
        elif ops[t] == 1:
            k = w.find_keys[w.find_offsets[t]: w.find_offsets[t + 1]]
        else:
            k = w.erase_keys[w.erase_offsets[t]: w.erase_offsets[t + 1]]
        parts_k.append(k)
        parts_v.append(wl.value_of(k))
        offs.append(offs[-1] + len(k))
    keys = np.concatenate(parts_k)
    vals = np.concatenate(parts_v)
    offs = np.array(offs, np.int64)
    flags = torch.zeros(len(keys), dtype=torch.uint8, device="cuda")
    vout = torch.zeros(len(keys), dtype=torch.int32, device="cuda")
    dk = d64(keys)
    batch.run(torch.from_numpy(ops).cuda(), d64(offs), dk, d32(vals), flags=flags, vals_out=vout)
    fl = flags.cpu().numpy()
    for t in range(ntab):
        a, b = offs[t], offs[t + 1]
        tab = plain[t]
        if ops[t] == 0:
            f = torch.zeros(b - a, dtype=torch.uint8, device="cuda")
            tab.insert(dk[a:b], d32(vals[a:b]), inserted=f)
            np.testing.assert_array_equal(fl[a:b], f.cpu().numpy())
        elif ops[t] == 1:
            h, _ = tab.find(dk[a:b])
            np.testing.assert_array_equal(fl[a:b], h.cpu().numpy())
        else:
            f = torch.zeros(b - a, dtype=torch.uint8, device="cuda")
            tab.erase(dk[a:b], erased=f)
            np.testing.assert_array_equal(fl[a:b], f.cpu().numpy())
        ka, va = sorted_contents(tabs[t])
        kb, vb = sorted_contents(tab)
        np.testing.assert_array_equal(ka, kb)
        np.testing.assert_array_equal(va, vb)
    batch.destroy()


def test_segments_coalesce_requests_like_plain_calls():
    """Several requests per table in one launch (same operation within a table), each with its own
    buffers, against the same requests issued as plain calls in order."""
    rng = np.random.default_rng(21)
    ntab = 12
    caps = [4000 + 500 * t for t in range(ntab)]
    tabs = [sa.unordered_map_u64_u32.createDeviceObject(c) for c in caps]
    plain = [sa.unordered_map_u64_u32.createDeviceObject(c) for c in caps]
    batch = TableBatch(tabs)
    # phase 1: two insert requests per table (overlapping keys); phase 2: finds; phase 3: erases
    for rnd, op in enumerate(["insert", "find", "erase", "find"]):
        segs, expect = [], []
        for t in range(ntab):
            for r in range(2):
                k = rng.integers(0, 3000, size=int(rng.integers(1, 900)), dtype=np.uint64) + np.uint64(t << 32)
                v = wl.value_of(k)
                dk, dv = d64(k), d32(v)
                f = torch.zeros(len(k), dtype=torch.uint8, device="cuda")
                vo = torch.zeros(len(k), dtype=torch.int32, device="cuda")
                segs.append((t, op, dk, dv, f, vo))
                expect.append((t, dk, dv))
        batch.run_segments(segs)
        torch.cuda.synchronize()
        for (t, op_, dk, dv, f, vo), (_, ek, ev) in zip(segs, expect):
            if op == "find":
                h, v = plain[t].find(ek)
                np.testing.assert_array_equal(f.cpu().numpy(), h.cpu().numpy())
                hv = h.cpu().numpy() == 1
                np.testing.assert_array_equal(vo.cpu().numpy()[hv], v.cpu().numpy()[hv])
        # inserts / erases: compare contents after applying the same requests in order
        if op in ("insert", "erase"):
            for (t, op_, dk, dv, f, vo) in segs:
                if op == "insert":
                    plain[t].insert(dk, dv)
                else:
                    plain[t].erase(dk)
            for t in range(ntab):
                ka, va = sorted_contents(tabs[t])
                kb, vb = sorted_contents(plain[t])
                np.testing.assert_array_equal(ka, kb)
                np.testing.assert_array_equal(va, vb)
    batch.destroy()


def test_per_table_baseline_matches_batched():
    w = wl.batched_maps(ntables=32, log2_lo=6, log2_hi=11, seed=14)
    bt = make_tables(w, 8)
    pt = make_tables(w, 8)
    b1, b2 = TableBatch(bt), TableBatch(pt)
    n = len(w.ins_keys)
    f1 = torch.zeros(n, dtype=torch.uint8, device="cuda")
    f2 = torch.zeros(n, dtype=torch.uint8, device="cuda")
    b1.run("insert", d64(w.ins_offsets), d64(w.ins_keys), d32(w.ins_vals), flags=f1)
    b2.run_per_table("insert", w.ins_offsets, d64(w.ins_keys), d32(w.ins_vals), flags=f2)
    assert torch.equal(f1, f2)
    for t in range(len(bt)):
        assert bt[t].size() == pt[t].size()


def test_misaligned_segments_and_requests_are_refused():
    """the batched and served forms take per-segment pointers: keys, values and value outputs must be
    aligned to their size, else the whole call is refused (HTAB_E_ARG) before any device work"""
    import ctypes

    from stdgpu_amd import _lib
    from stdgpu_amd import serving as sv

    lib = _lib.load()
    tabs = [sa.unordered_map_u64_u32.createDeviceObject(1000) for _ in range(2)]
    batch = TableBatch(tabs)
    srv = sv.Server(tabs, nstreams=1)
    buf = torch.zeros(256, dtype=torch.int64, device="cuda")
    base = buf.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    E = _lib.HTAB_E_ARG
    INS = _lib.HTAB_OP_INSERT
    bad = [
        _lib.htab_segment(base + 4, base + 1024, None, None, 4, 1, INS),  # u64 keys at a 4-byte offset
        _lib.htab_segment(base, base + 1026, None, None, 4, 1, INS),  # u32 values at a 2-byte offset
        _lib.htab_segment(base, None, base + 1500, base + 1026, 4, 1, _lib.HTAB_OP_FIND),  # value outputs
    ]
    good = _lib.htab_segment(base, base + 1024, None, None, 4, 0, INS)
    arr_ns = np.zeros(2, np.int64)
    sub, done = np.zeros(2, np.int64), np.zeros(2, np.int64)
    st = np.zeros(2, np.int32)
    stats = _lib.htab_server_stats()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    for b in bad:
        segs = (_lib.htab_segment * 2)(good, b)
        assert lib.htab_batch_run_segments(batch._b, segs, 2, s) == E
        assert lib.htab_server_run(srv._s, segs, p(arr_ns), 2, p(sub), p(done), p(st), ctypes.byref(stats)) == E
    torch.cuda.synchronize()
    assert tabs[0].size() == 0 and tabs[1].size() == 0  # the good segment of a refused call did not run
    segs = (_lib.htab_segment * 1)(good)
    assert lib.htab_batch_run_segments(batch._b, segs, 1, s) == 0
    torch.cuda.synchronize()
    assert tabs[0].size() == 1 and tabs[1].size() == 0
    srv.destroy()
    batch.destroy()


@pytest.mark.parametrize("block", [64, 512, 1024])
def test_batched_set_kind_any_workgroup_size(block):
    """unordered_set<u64> tables (a kind whose persistent kernel exists for 256-thread workgroups
    only) of any workgroup size: the batch cuts its work into chunks of the workgroup size it
    launches, so every element of every table is processed exactly as by plain calls."""
    rng = np.random.default_rng(block)
    ntab = 24
    caps = [int(c) for c in rng.integers(500, 20000, size=ntab)]
    bt = [sa.unordered_set_u64.createDeviceObject(c, sa.Tuning(block_size=block)) for c in caps]
    pt = [sa.unordered_set_u64.createDeviceObject(c, sa.Tuning(block_size=block)) for c in caps]
    batch = TableBatch(bt)
    sizes = [int(0.8 * c) for c in caps]
    keys = [rng.integers(1, 2**60, size=s, dtype=np.uint64) for s in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    allk = d64(np.concatenate(keys))
    flags = torch.zeros(int(off[-1]), dtype=torch.uint8, device="cuda")
    batch.run("insert", d64(off), allk, flags=flags)
    bflags = flags.cpu().numpy()
    q = d64(np.concatenate([k[::2] for k in keys]))
    qoff = np.concatenate([[0], np.cumsum([len(k[::2]) for k in keys])]).astype(np.int64)
    hits = torch.zeros(int(qoff[-1]), dtype=torch.uint8, device="cuda")
    batch.run("erase", d64(qoff), q, flags=hits)
    bh = hits.cpu().numpy()
    for t, tab in enumerate(pt):
        a, b = off[t], off[t + 1]
        f = torch.zeros(int(b - a), dtype=torch.uint8, device="cuda")
        tab.insert(allk[a:b], inserted=f)
        np.testing.assert_array_equal(bflags[a:b], f.cpu().numpy())
        assert int(f.sum()) == len(np.unique(keys[t]))
        a, b = qoff[t], qoff[t + 1]
        e = torch.zeros(int(b - a), dtype=torch.uint8, device="cuda")
        tab.erase(q[a:b], erased=e)
        np.testing.assert_array_equal(bh[a:b], e.cpu().numpy())
        ka, _ = bt[t].contents()
        kb, _ = tab.contents()
        np.testing.assert_array_equal(np.sort(ka.cpu().numpy().view(np.uint64)), np.sort(kb.cpu().numpy().view(np.uint64)))
        assert bt[t].size() == tab.size() and bt[t].valid()
    batch.destroy()


def test_mixed_operations_on_one_table_and_find_without_hits_are_refused():
    """segments of one table naming different operations, and a per-table find without a hit array
    (table 0 empty, so the first table with work starts past element 0), are refused with
    HTAB_E_ARG before any device work"""
    import ctypes

    from stdgpu_amd import _lib

    lib = _lib.load()
    tabs = [sa.unordered_map_u64_u32.createDeviceObject(1000) for _ in range(2)]
    batch = TableBatch(tabs)
    buf = torch.zeros(256, dtype=torch.int64, device="cuda")
    base = buf.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    E = _lib.HTAB_E_ARG
    ins = _lib.htab_segment(base, base + 1024, None, None, 4, 1, _lib.HTAB_OP_INSERT)
    era = _lib.htab_segment(base + 64, None, None, None, 4, 1, _lib.HTAB_OP_ERASE)
    other = _lib.htab_segment(base + 64, None, None, None, 4, 0, _lib.HTAB_OP_ERASE)
    segs = (_lib.htab_segment * 3)(ins, other, era)
    assert lib.htab_batch_run_segments(batch._b, segs, 3, s) == E
    torch.cuda.synchronize()
    assert tabs[0].size() == 0 and tabs[1].size() == 0
    # one op per table (different across tables) is accepted
    segs = (_lib.htab_segment * 2)(ins, other)
    assert lib.htab_batch_run_segments(batch._b, segs, 2, s) == 0
    torch.cuda.synchronize()
    assert tabs[1].size() == 1  # the buffer holds four copies of key 0
    off = (ctypes.c_int64 * 3)(0, 0, 4)
    assert lib.htab_batch_run_per_table(batch._b, _lib.HTAB_OP_FIND, off, ctypes.c_void_p(base), None, None, None,
                                        ctypes.c_void_p(s)) == E
    batch.destroy()


@pytest.mark.parametrize("load", [0.5, 0.8])
def test_tuned_persistent_launch_over_4096_tables_equals_plain_calls(load):
    """The benchmark's full batched shape: 4096 unordered_map<u64,u32> tables of 2^8..2^18 keys
    (log-uniform), insert all, find (70 % hits), erase 30 %, each phase one graph-captured persistent
    launch with the tuning the autotuner picks on this workload (8-slot buckets, probe batch 8,
    64-thread workgroups), at that load and at 0.8.  Every flag, hit, value, table size and the
    contents of sampled tables equal one plain call per table on tables of the default tuning."""
    from stdgpu_amd import autotune as at
    from stdgpu_amd import workloads_device as wd

    cfg = {"bucket_lanes": 8, "probe_batch": 8, "probe_group": 64, "block_size": 64, "load": load}
    w = wd.batched_maps("cuda", ntables=4096)
    tabs = at.build_tables(w["sizes"], cfg)
    batch = TableBatch(tabs)
    _, out = at.run_phases(batch, w)
    plain = at.build_tables(w["sizes"], {"bucket_lanes": 8, "probe_batch": 0, "block_size": sa.Tuning().block_size,
                                         "load": load})
    pbatch = TableBatch(plain)
    res = {}
    for op, name, flag in (("insert", "ins", "inserted"), ("find", "find", "hits"), ("erase", "erase", "erased")):
        n = w[f"{name}_keys"].numel()
        f = torch.zeros(n, dtype=torch.uint8, device="cuda")
        vo = torch.zeros(n, dtype=torch.int32, device="cuda") if op == "find" else None
        pbatch.run_per_table(op, w[f"{name}_offsets"].cpu().numpy(), w[f"{name}_keys"],
                             w["ins_vals"] if op == "insert" else None, f, vo)
        res[flag], res["values"] = f, vo if vo is not None else res.get("values")
    torch.cuda.synchronize()
    for flag in ("inserted", "hits", "erased"):
        assert torch.equal(out[flag], res[flag]), flag
    hit = res["hits"] == 1
    assert torch.equal(out["values"][hit], res["values"][hit])
    assert int(hit.sum()) >= int(0.69 * w["find_keys"].numel())
    e_off = w["erase_offsets"].cpu().numpy()
    want = w["sizes"] - np.diff(e_off)
    for t in range(len(tabs)):
        assert tabs[t].size() == plain[t].size() == want[t], t
    for t in (0, 1, 777, 2048, 4095):
        ka, va = sorted_contents(tabs[t])
        kb, vb = sorted_contents(plain[t])
        np.testing.assert_array_equal(ka, kb)
        np.testing.assert_array_equal(va, vb)
        assert tabs[t].valid() and plain[t].valid()
    batch.destroy()
    pbatch.destroy()
    for t in tabs + plain:
        sa.unordered_map_u64_u32.destroyDeviceObject(t)
