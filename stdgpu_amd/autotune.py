"""Autotuner of the batched form (many tables, one persistent launch per phase).

Sweeps, on the workload itself:
  * bucket width: 1, 4, 8 or 16 slots of 16 bytes (= lanes that load one bucket together);
  * probe group: lanes whose bucket loads are issued together = bucket lanes x probe batch,
    16, 32 or 64 (probe batch 2, 4 or 8, at most the bucket lanes); with one-slot buckets every
    lane probes its own element (probe group 1);
  * workgroup size: 64, 256, 512 or 1024 (also the chunk of elements a workgroup takes at a time);
  * load factor: 0.5, 0.7, 0.8 or 0.9 (physical slots = keys / load).
Each candidate builds its tables, runs the insert, find and erase phases once as captured HIP
graphs and is scored by the sum of the three phase times (HIP events on the launch stream).
"""
from __future__ import annotations

import itertools
import math

from . import containers as C
from .batched import TableBatch

BUCKET_LANES = (1, 4, 8, 16)
PROBE_GROUPS = (16, 32, 64)
BLOCKS = (64, 256, 512, 1024)
LOADS = (0.5, 0.7, 0.8, 0.9)


def candidates(lanes=BUCKET_LANES, groups=PROBE_GROUPS, blocks=BLOCKS, loads=LOADS):
    for g, blk, lf in itertools.product(lanes, blocks, loads):
        if g == 1:
            # one lane per bucket: the probe batch is 1 (probe_batch 0 selects it)
            yield {"bucket_lanes": 1, "probe_batch": 0, "probe_group": 1, "block_size": blk, "load": lf}
            continue
        for pg in groups:
            rb = pg // g
            if rb not in (2, 4, 8) or rb > g:
                continue
            yield {"bucket_lanes": g, "probe_batch": rb, "probe_group": pg, "block_size": blk, "load": lf}


def build_tables(sizes, cfg, cls=C.unordered_map_u64_u32):
    tun = C.Tuning(bucket_lanes=cfg["bucket_lanes"], block_size=cfg["block_size"], probe_batch=cfg["probe_batch"])
    return [cls.createDeviceObject(max(1, int(math.ceil(int(n) / cfg["load"]))), tun) for n in sizes]


def run_phases(batch: TableBatch, w: dict, graphs: bool = True, stream=None):
    """Insert, find and erase phases over the device workload dict (workloads_device.batched_maps);
    returns (ms per phase, outputs)."""
    import torch

    s = stream or torch.cuda.Stream()
    n_ins, n_find, n_er = w["ins_keys"].numel(), w["find_keys"].numel(), w["erase_keys"].numel()
    dev = w["ins_keys"].device
    out = {
        "inserted": torch.zeros(n_ins, dtype=torch.uint8, device=dev),
        "hits": torch.zeros(n_find, dtype=torch.uint8, device=dev),
        "values": torch.zeros(n_find, dtype=torch.int32, device=dev),
        "erased": torch.zeros(n_er, dtype=torch.uint8, device=dev),
    }
    phases = [
        ("insert", w["ins_offsets"], w["ins_keys"], w["ins_vals"], out["inserted"], None),
        ("find", w["find_offsets"], w["find_keys"], None, out["hits"], out["values"]),
        ("erase", w["erase_offsets"], w["erase_keys"], None, out["erased"], None),
    ]
    ms = {}
    torch.cuda.synchronize()
    for op, off, k, v, f, vo in phases:
        g = batch.capture(op, off, k, v, f, vo, stream=s) if graphs else None
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        if g is not None:
            g.launch(s)
        else:
            batch.run(op, off, k, v, f, vo, stream=s)
        b.record(s)
        b.synchronize()
        ms[op] = a.elapsed_time(b)
        if g is not None:
            g.destroy()
    return ms, out


def tune(w: dict, grid=None, sample_tables: int | None = None, verbose: bool = False) -> dict:
    """Returns {"best": cfg, "results": [cfg + phase times]} on (a sample of) the workload `w`
    (a dict from workloads_device.batched_maps)."""
    import torch

    sizes = w["sizes"]
    sub = w
    if sample_tables and sample_tables < len(sizes):
        sub = _subset(w, sample_tables)
        sizes = sub["sizes"]
    results = []
    for cfg in (grid or list(candidates())):
        tabs = build_tables(sizes, cfg)
        batch = TableBatch(tabs)
        ms, _ = run_phases(batch, sub)
        batch.destroy()
        for t in tabs:
            C.unordered_map_u64_u32.destroyDeviceObject(t)
        r = dict(cfg, **{k: round(v, 4) for k, v in ms.items()})
        r["total_ms"] = round(sum(ms.values()), 4)
        results.append(r)
        if verbose:
            print(r, flush=True)
        torch.cuda.synchronize()
    best = min(results, key=lambda r: r["total_ms"])
    return {"best": best, "results": sorted(results, key=lambda r: r["total_ms"])}


def _subset(w: dict, k: int) -> dict:
    """The first k tables of a device workload (their offsets are unchanged prefixes)."""
    sub = {"sizes": w["sizes"][:k], "capacities": w["capacities"][:k]}
    for name in ("ins", "find", "erase"):
        off = w[f"{name}_offsets"]
        end = int(off[k].item())
        sub[f"{name}_offsets"] = off[: k + 1].clone()
        sub[f"{name}_keys"] = w[f"{name}_keys"][:end]
    sub["ins_vals"] = w["ins_vals"][: int(w["ins_offsets"][k].item())]
    return sub


def describe(cfg: dict) -> str:
    return (f"bucket {cfg['bucket_lanes']} x 16 B, probe group {cfg['probe_group']} lanes "
            f"(batch {cfg['probe_batch'] or 1}), workgroup {cfg['block_size']}, load {cfg['load']}")
