"""Served form driver: resident tables, an open-loop request log, latency percentiles, and the
check that every result equals an in-order replay of the same log through the plain API.

Workload (per table t): a pool of distinct keys (a permutation of the splitmix64 stream, 0.8 x
capacity of them, so no table can exceed load 0.8).  Request i picks its table by Zipf(1.0) over the
tables, its operation insert / find / erase with probabilities 0.3 / 0.6 / 0.1, its batch size n
log-uniform in [2^lo, 2^hi], and its keys as a window of n consecutive pool keys at a random offset;
mapped values are value_of(key).  Arrivals are a Poisson process at the requested rate.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import containers as C
from . import workloads_device as wd
from ._lib import check

OPS = (_lib.HTAB_OP_INSERT, _lib.HTAB_OP_FIND, _lib.HTAB_OP_ERASE)


@dataclass
class RequestLog:
    table: np.ndarray     # int32
    op: np.ndarray        # int32
    n: np.ndarray         # int64
    start: np.ndarray     # int64, offset into the table's key pool
    arrival_ns: np.ndarray


def zipf_tables(k: int, ntables: int, rng) -> np.ndarray:
    w = 1.0 / np.arange(1, ntables + 1, dtype=np.float64)
    return rng.choice(ntables, size=k, p=w / w.sum()).astype(np.int32)


def make_log(pool_sizes, nreq: int, rate_per_s: float | None, seed: int = 5, log2_lo: float = 6,
             log2_hi: float = 20, mix=(0.3, 0.6, 0.1)) -> RequestLog:
    rng = np.random.default_rng(seed)
    ntables = len(pool_sizes)
    table = zipf_tables(nreq, ntables, rng)
    op = rng.choice(3, size=nreq, p=np.asarray(mix) / sum(mix)).astype(np.int32)
    n = np.floor(2.0 ** rng.uniform(log2_lo, log2_hi, size=nreq)).astype(np.int64)
    n = np.minimum(n, np.asarray(pool_sizes, np.int64)[table])
    start = (rng.random(nreq) * (np.asarray(pool_sizes, np.int64)[table] - n + 1)).astype(np.int64)
    if rate_per_s is None or rate_per_s <= 0:
        arrival = np.zeros(nreq, np.int64)  # everything at once: saturation
    else:
        arrival = np.cumsum(rng.exponential(1e9 / rate_per_s, size=nreq)).astype(np.int64)
    return RequestLog(table, op, n, start, arrival)


class Pools:
    """Per-table key pools and their values, resident in device memory."""

    def __init__(self, capacities, seed: int = 50, fill: float = 0.8, device="cuda"):
        import torch

        self.sizes = [max(1, int(c * fill)) for c in capacities]
        self.keys, self.vals = [], []
        for t, m in enumerate(self.sizes):
            k = wd.stream(seed + 1000 * t, m, device)
            k = torch.where(k == -1, torch.ones_like(k), k)  # all-ones is a legal key, kept out for clarity
            self.keys.append(k)
            self.vals.append(wd.value_of(k))


def build_requests(log: RequestLog, pools: Pools, out_flags, out_vals):
    """htab_segment records; request i writes its flags / values at the running offset of the log."""
    reqs = (_lib.htab_segment * max(len(log.n), 1))()
    off = 0
    for i in range(len(log.n)):
        t, s, n = int(log.table[i]), int(log.start[i]), int(log.n[i])
        k = pools.keys[t][s: s + n]
        v = pools.vals[t][s: s + n]
        op = OPS[int(log.op[i])]
        reqs[i] = _lib.htab_segment(k.data_ptr(), v.data_ptr() if op == _lib.HTAB_OP_INSERT else None,
                                    out_flags.data_ptr() + off,
                                    (out_vals.data_ptr() + 4 * off) if op == _lib.HTAB_OP_FIND else None,
                                    n, t, op)
        off += n
    return reqs


class Server:
    def __init__(self, tables, nstreams: int = 1, max_inflight_elems: int = 1 << 24):
        arr = (ctypes.c_void_p * len(tables))(*[t._h.value for t in tables])
        h = ctypes.c_void_p()
        check(_lib.load().htab_server_create(arr, len(tables), nstreams, max_inflight_elems, ctypes.byref(h)),
              "htab_server_create")
        self._s = h
        self.tables = tables

    def destroy(self):
        if self._s:
            check(_lib.load().htab_server_destroy(self._s), "htab_server_destroy")
            self._s = ctypes.c_void_p()

    def run(self, reqs, arrival_ns: np.ndarray):
        nreq = len(arrival_ns)
        arr = np.ascontiguousarray(arrival_ns, np.int64)
        sub = np.zeros(nreq, np.int64)
        done = np.zeros(nreq, np.int64)
        st = np.zeros(nreq, np.int32)
        stats = _lib.htab_server_stats()
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(_lib.load().htab_server_run(self._s, reqs, p(arr), nreq, p(sub), p(done), p(st), ctypes.byref(stats)),
              "htab_server_run")
        return sub, done, st, stats

    def replay_plain(self, reqs, nreq: int, status=None, blocking: bool = True, stream=None):
        import torch

        s = stream or torch.cuda.current_stream()
        stp = None if status is None else status.ctypes.data_as(ctypes.c_void_p)
        check(_lib.load().htab_server_replay_plain(self._s, reqs, stp, nreq, 1 if blocking else 0, int(s.cuda_stream)),
              "htab_server_replay_plain")


def latency_stats(sub, done, status) -> dict:
    ok = status == 0
    lat = (done[ok] - sub[ok]) / 1e6
    if len(lat) == 0:
        return {"p50_ms": None, "p95_ms": None, "p99_ms": None}
    return {"p50_ms": round(float(np.percentile(lat, 50)), 4), "p95_ms": round(float(np.percentile(lat, 95)), 4),
            "p99_ms": round(float(np.percentile(lat, 99)), 4), "max_ms": round(float(lat.max()), 4)}


def capacities(ntables: int = 64, log2_lo: float = 16, log2_hi: float = 26, seed: int = 6):
    rng = np.random.default_rng(seed)
    return np.floor(2.0 ** rng.uniform(log2_lo, log2_hi, size=ntables)).astype(np.int64)


def make_tables(caps, tuning=None):
    return [C.unordered_map_u64_u32.createDeviceObject(int(c), tuning) for c in caps]


def sorted_contents(t):
    """(keys, values) of a table sorted by key, on the device (key bits compared as int64)."""
    import torch

    k, v = t.contents()
    ks, o = torch.sort(k)
    return ks, v[o]


def same_contents(a, b) -> bool:
    import torch

    ka, va = sorted_contents(a)
    kb, vb = sorted_contents(b)
    return bool(torch.equal(ka, kb) and torch.equal(va, vb))


def verify_against_replay(caps, pools: Pools, log: RequestLog, served_tables, status, flags, vals,
                          only_tables=None) -> dict:
    """Replays the log in order through plain calls on fresh tables and compares every find's
    hit vector and values and every table's final contents (of the tables in only_tables, when
    given)."""
    import torch

    fresh = make_tables(caps)
    rflags = torch.zeros_like(flags)
    rvals = torch.zeros_like(vals)
    reqs = build_requests(log, pools, rflags, rvals)
    srv = Server(fresh, nstreams=1)
    srv.replay_plain(reqs, len(log.n), status=status, blocking=False)
    torch.cuda.synchronize()
    srv.destroy()
    finds_ok, off = True, 0
    ok = status == 0
    for i in range(len(log.n)):
        n = int(log.n[i])
        if ok[i] and OPS[int(log.op[i])] == _lib.HTAB_OP_FIND and (only_tables is None or int(log.table[i]) in only_tables):
            a, b = flags[off: off + n], rflags[off: off + n]
            if not torch.equal(a, b) or not torch.equal(vals[off: off + n][a == 1], rvals[off: off + n][b == 1]):
                finds_ok = False
        off += n
    check_tables = range(len(caps)) if only_tables is None else only_tables
    contents_ok = all(same_contents(served_tables[t], fresh[t]) for t in check_tables)
    for t in fresh:
        C.unordered_map_u64_u32.destroyDeviceObject(t)
    return {"finds_identical": finds_ok, "contents_identical": contents_ok}


def served_run(caps, pools, log, nstreams=1, max_inflight_elems=1 << 24, keep_outputs=True):
    """Runs the log through a fresh server over fresh tables; returns (tables, outputs, results)."""
    import torch

    tables = make_tables(caps)
    total = int(log.n.sum())
    flags = torch.zeros(max(total, 1), dtype=torch.uint8, device="cuda")
    vals = torch.zeros(max(total, 1), dtype=torch.int32, device="cuda")
    reqs = build_requests(log, pools, flags, vals)
    srv = Server(tables, nstreams=nstreams, max_inflight_elems=max_inflight_elems)
    sub, done, st, stats = srv.run(reqs, log.arrival_ns)
    srv.destroy()
    ok = st == 0
    keys_done = int(log.n[ok].sum())
    wall = max(int(done.max()) if len(done) else 0, 1)
    res = {"requests": len(log.n), "refused": int((st != 0).sum()), "keys": keys_done,
           "wall_s": round(wall / 1e9, 5), "req_per_s": round(int(ok.sum()) / (wall / 1e9), 1),
           "gops": round(keys_done / (wall / 1e9) / 1e9, 3), "launches": stats.launches,
           "segments_per_launch": round(stats.segments / max(stats.launches, 1), 2),
           "max_concurrent_launches": stats.max_concurrent_launches, **latency_stats(sub, done, st)}
    return tables, (flags, vals, st), res


def plain_run(caps, pools, log, blocking: bool):
    """The same log as plain calls in arrival order (one call per request) on fresh tables."""
    import time

    import torch

    tables = make_tables(caps)
    total = int(log.n.sum())
    flags = torch.zeros(max(total, 1), dtype=torch.uint8, device="cuda")
    vals = torch.zeros(max(total, 1), dtype=torch.int32, device="cuda")
    reqs = build_requests(log, pools, flags, vals)
    srv = Server(tables, nstreams=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    srv.replay_plain(reqs, len(log.n), blocking=blocking)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    srv.destroy()
    for t in tables:
        C.unordered_map_u64_u32.destroyDeviceObject(t)
    return {"gops": round(total / dt / 1e9, 3), "seconds": round(dt, 4)}


def destroy_tables(tables):
    for t in tables:
        C.unordered_map_u64_u32.destroyDeviceObject(t)


def main(argv=None) -> None:
    """`python -m stdgpu_amd.serving served|plain [nreq]`: the 64-table saturation log (every request
    at once) through the server, or through one plain call per request; prints one JSON line."""
    import json
    import sys

    import torch

    args = sys.argv[1:] if argv is None else argv
    mode = args[0] if args else "served"
    nreq = int(args[1]) if len(args) > 1 else 3000
    caps = capacities(ntables=64)
    pools = Pools(caps)
    log = make_log(pools.sizes, nreq=nreq, rate_per_s=None, seed=8)
    if mode == "served":
        tabs, outs, res = served_run(caps, pools, log)
        destroy_tables(tabs)
        del outs
    elif mode == "plain":
        res = plain_run(caps, pools, log, blocking=False)
    else:
        raise SystemExit("mode must be served or plain")
    torch.cuda.synchronize()
    print(json.dumps({"mode": mode, "requests": nreq, "keys": int(log.n.sum()), **res}), flush=True)


if __name__ == "__main__":
    main()
