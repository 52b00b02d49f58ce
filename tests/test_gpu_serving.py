"""Served form: continuous batching over resident tables must give every request the result it gets
alone, i.e. every find's hit vector and values and every table's final contents equal an in-order
replay of the same request log through the plain API, at several arrival rates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from stdgpu_amd import serving as sv  # noqa: E402


@pytest.mark.parametrize("rate,nstreams", [(None, 4), (20000.0, 4), (None, 1), (None, 8)])
def test_served_results_equal_in_order_replay(rate, nstreams):
    caps = sv.capacities(ntables=16, log2_lo=10, log2_hi=16, seed=7)
    pools = sv.Pools(caps, seed=77)
    log = sv.make_log(pools.sizes, nreq=600, rate_per_s=rate, seed=8, log2_lo=4, log2_hi=13)
    tables, (flags, vals, status), res = sv.served_run(caps, pools, log, nstreams=nstreams,
                                                       max_inflight_elems=1 << 14)
    assert res["requests"] == 600
    assert res["launches"] > 0
    chk = sv.verify_against_replay(caps, pools, log, tables, status, flags, vals)
    sv.destroy_tables(tables)
    assert chk["finds_identical"], chk
    assert chk["contents_identical"], chk


def test_inserts_reaching_the_limit_are_served_not_refused():
    """table 0's pool is four times its capacity, table 1's fits: no request is refused; table 0 ends
    exactly at its limit with valid() holding (which keys of the request that crosses the limit get
    in is decided by the race, as inside one plain bulk call), and every request of table 1 gives the
    in-order replay's result."""
    caps = np.array([300, 5000], np.int64)
    pools = sv.Pools(caps, seed=3, fill=4.0)  # 1200 keys for a 300-slot table
    pools.sizes[1] = 4000  # ... and 4000 for the 5000-slot one
    pools.keys[1], pools.vals[1] = pools.keys[1][:4000], pools.vals[1][:4000]
    log = sv.make_log(pools.sizes, nreq=200, rate_per_s=None, seed=9, log2_lo=5, log2_hi=9, mix=(0.7, 0.3, 0.0))
    tables, (flags, vals, status), res = sv.served_run(caps, pools, log, nstreams=2)
    assert res["refused"] == 0 and (status == 0).all()
    assert tables[0].size() == tables[0].info.limit and tables[0].valid()
    assert tables[1].size() <= tables[1].info.limit and tables[1].valid()
    chk = sv.verify_against_replay(caps, pools, log, tables, status, flags, vals, only_tables=[1])
    sv.destroy_tables(tables)
    assert chk["finds_identical"] and chk["contents_identical"], chk


def test_mostly_present_inserts_near_the_limit_are_applied():
    """a table filled close to its limit receives inserts whose keys are mostly present already: n
    exceeds the free slots, the new keys do not; every request is applied and the result equals the
    in-order replay"""
    caps = np.array([4096, 4096], np.int64)
    pools = sv.Pools(caps, seed=5, fill=0.95)
    log = sv.make_log(pools.sizes, nreq=300, rate_per_s=None, seed=12, log2_lo=8, log2_hi=11, mix=(0.8, 0.2, 0.0))
    tables, (flags, vals, status), res = sv.served_run(caps, pools, log, nstreams=2)
    assert res["refused"] == 0 and (status == 0).all()
    chk = sv.verify_against_replay(caps, pools, log, tables, status, flags, vals)
    sv.destroy_tables(tables)
    assert chk["finds_identical"] and chk["contents_identical"], chk


def test_served_long_runs_of_small_requests():
    """long same-operation runs of tiny requests (some empty): launches carry more segments than one
    kernel's argument block holds, so a launch is split over several kernels on its stream"""
    caps = sv.capacities(ntables=24, log2_lo=9, log2_hi=13, seed=3)
    pools = sv.Pools(caps, seed=31)
    log = sv.make_log(pools.sizes, nreq=4000, rate_per_s=None, seed=4, log2_lo=0, log2_hi=5, mix=(0.1, 0.85, 0.05))
    log.n[::37] = 0
    tables, (flags, vals, status), res = sv.served_run(caps, pools, log, nstreams=2, max_inflight_elems=1 << 22)
    print(res, flush=True)
    assert (status == 0).all()
    chk = sv.verify_against_replay(caps, pools, log, tables, status, flags, vals)
    sv.destroy_tables(tables)
    assert chk["finds_identical"] and chk["contents_identical"], chk
