"""Benchmark of the hashed-container hot path on one MI355X.

Workload (one "step"): unordered_map<uint64_t, uint32_t> created for 2^26 keys at load 0.8
(capacity ceil(2^26 / 0.8)); clear, bulk insert of 2^26 uniform keys (value = uint32 of
splitmix64(key)), bulk find of 2^26 queries at a 90% hit rate, bulk erase of 25% of the keys,
re-insert of the erased keys.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line: whole-step throughput in G operations/s plus per-operation rates, the
roofline of the dominant kernel (algorithmic bytes per launch / measured launch time vs the
8 TB/s HBM peak) and a CPU baseline (the oracle, a restatement of the reference's OpenMP
algorithm, timed on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per operation (u64 -> u32 map): streamed input + streamed output + one
# 64 B random sector per bucket probe (+ the 16 B slot write for insert)
BYTES_PER_OP = {"insert": 16 + 64 + 16 + 1, "find": 8 + 1 + 4 + 64, "erase": 8 + 64 + 8 + 1, "reinsert": 16 + 64 + 16 + 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the oracle on a bounded sample")
    ap.add_argument("--cpu-log2n", type=int, default=20)
    return ap.parse_args()


def cpu_baseline(log2n: int) -> dict:
    """The oracle (reference OpenMP algorithm restated) on the same workload at 2^log2n keys."""
    from oracle import pyoracle as po
    from stdgpu_amd import workloads as wl

    w = wl.map_u64_uniform(n=1 << log2n, seed=1)
    ncpu = os.cpu_count() or 1
    cand = sorted({1, min(ncpu, 16)})
    best = None
    for th in cand:
        t = po.Table("u64", "u32", w.capacity)
        t0 = time.perf_counter()
        t.insert(w.keys, w.values, threads=th)
        t.find(w.queries, threads=th)
        t.erase(w.erase, threads=th)
        t.insert(w.reinsert, wl.value_of(w.reinsert), threads=th)
        dt = time.perf_counter() - t0
        t.close()
        ops = len(w.keys) + len(w.queries) + len(w.erase) + len(w.reinsert)
        rate = ops / dt / 1e9
        if best is None or rate > best[0]:
            best = (rate, th, dt)
    return {
        "value": round(best[0], 6),
        "unit": "Gops/s",
        "cores": best[1],
        "kind": "port",
        "sample": f"same op sequence at 2^{log2n} keys (insert, find, erase 25%, re-insert), "
                  f"threads tried {cand}, best {best[1]} ({best[2]:.2f} s)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the hot path "
              f"shards by table or by hash range, see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch

    import stdgpu_amd as sa
    from stdgpu_amd import workloads as wl

    n = 1 << args.log2n
    w = wl.map_u64_uniform(n=n, seed=1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    def d(a, dt):
        return torch.from_numpy(np.ascontiguousarray(a).view(dt)).to(dev)

    keys = d(w.keys, np.int64)
    vals = d(w.values, np.int32)
    queries = d(w.queries, np.int64)
    erase = d(w.erase, np.int64)
    reins = d(w.reinsert, np.int64)
    reins_v = d(wl.value_of(w.reinsert), np.int32)
    hits = torch.empty(n, dtype=torch.uint8, device=dev)
    got = torch.empty(n, dtype=torch.int32, device=dev)
    ins_flags = torch.empty(n, dtype=torch.uint8, device=dev)
    er_flags = torch.empty(len(w.erase), dtype=torch.uint8, device=dev)

    tab = sa.unordered_map_u64_u32.createDeviceObject(
        w.capacity, sa.Tuning(bucket_lanes=args.lanes, block_size=args.block))
    stream = torch.cuda.current_stream()
    names = ["clear", "insert", "find", "erase", "reinsert"]

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        tab.clear()
        if ev is not None:
            ev[1].record(stream)
        tab.insert(keys, vals, inserted=ins_flags)
        if ev is not None:
            ev[2].record(stream)
        tab.find(queries, hits=hits, values_out=got)
        if ev is not None:
            ev[3].record(stream)
        tab.erase(erase, erased=er_flags)
        if ev is not None:
            ev[4].record(stream)
        tab.insert(reins, reins_v)
        if ev is not None:
            ev[5].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # correctness guard on the measured configuration (not timed)
    assert tab.size() == n, tab.size()
    assert abs(int(hits.sum()) - int(round(n * 0.9))) < 16, int(hits.sum())

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(events[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per = {k: [] for k in names}
    for ev in events:
        for j, k in enumerate(names):
            per[k].append(ev[j].elapsed_time(ev[j + 1]))
    step_ms = [sum(per[k][s] for k in names) for s in range(args.steps)]
    ms_step = wall * 1e3 / args.steps
    ops = {"insert": n, "find": n, "erase": len(w.erase), "reinsert": len(w.reinsert)}
    total_ops = sum(ops.values())
    value = total_ops / (ms_step / 1e3) / 1e9

    rates = {k: round(ops[k] / (np.mean(per[k]) / 1e3) / 1e9, 3) for k in ops}
    lat = {k: {"mean_ms": round(float(np.mean(per[k])), 4), "p50_ms": round(float(np.percentile(per[k], 50)), 4),
               "p99_ms": round(float(np.percentile(per[k], 99)), 4)} for k in names}
    # dominant kernel = the op with the largest share of the step
    dom = max(ops, key=lambda k: np.mean(per[k]))
    dom_ms = float(np.mean(per[dom]))
    achieved = BYTES_PER_OP[dom] * ops[dom] / (dom_ms / 1e3) / 1e9

    out = {
        "metric": "Gops/s bulk insert/find/erase (u64->u32, LF 0.8)",
        "value": round(value, 4),
        "unit": "Gops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (splitmix64 keys, value = uint32 splitmix64(key))",
        "config": {
            "workload": f"unordered_map<u64,u32> 2^{args.log2n} keys, capacity ceil(n/0.8): clear, insert, "
                        f"find 2^{args.log2n} at 90% hit, erase 25%, re-insert",
            "keys": n,
            "capacity": w.capacity,
            "bucket_lanes": args.lanes,
            "block_size": args.block,
            "physical_slots": tab.info.physical_slots,
        },
        "ops_per_step": total_ops,
        "per_op_gops": rates,
        "per_op_latency": lat,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_op": BYTES_PER_OP[dom],
            "avg_launch_ms": round(dom_ms, 4),
        },
    }
    sa.unordered_map_u64_u32.destroyDeviceObject(tab)
    del keys, vals, queries, erase, reins, reins_v, hits, got, ins_flags, er_flags
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_log2n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
