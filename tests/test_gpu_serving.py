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


def test_admission_refuses_inserts_that_cannot_fit():
    """a table whose pool is larger than its capacity: inserts beyond the free slots are refused,
    and the refused requests never take effect (the replay skips them and still matches)."""
    caps = np.array([300, 5000], np.int64)
    pools = sv.Pools(caps, seed=3, fill=4.0)  # 1200 keys for a 300-slot table
    log = sv.make_log(pools.sizes, nreq=200, rate_per_s=None, seed=9, log2_lo=5, log2_hi=9, mix=(0.7, 0.3, 0.0))
    tables, (flags, vals, status), res = sv.served_run(caps, pools, log, nstreams=2)
    assert res["refused"] > 0
    for t in tables:
        assert t.size() <= t.info.limit
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
