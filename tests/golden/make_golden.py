"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

* reference_kats.json: values printed by the compiled reference (stdgpu 1.3.0, OpenMP backend) for
  sizing and bucket(), recorded from a run of it; they are data, not generated here.
* small_workloads.npz: seeded inputs and the oracle's outputs (sorted contents, hit vectors, values,
  per-element flags, sizes) for small instances of the set<u32> and map<u64,u32> workloads, so the
  GPU tests can be checked against fixed vectors as well as against a live oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from stdgpu_amd import workloads as wl  # noqa: E402

REFERENCE_KATS = {
    "set_u32_capacity_2^21": {"capacity": 1 << 21, "bucket_count": 2097152, "max_size": 3422806,
                              "bucket": {"1": 1296111, "2": 495070, "3": 1791181, "4294967294": 547697}},
    "map_u64_u32_capacity_100000": {"capacity": 100000, "bucket_count": 131072, "max_size": 184444,
                                    "bucket": {"1": 81006, "2": 30941, "3": 111948, "4294967294": 34231,
                                               "81985529216486895": 28855}},
    "map_u64_u32_capacity_1": {"capacity": 1, "bucket_count": 1, "max_size": 2, "bucket": {}},
}


def small_workloads() -> dict:
    out = {}
    w = wl.set_u32(n=1 << 12, capacity=1 << 13, seed=0)
    t = po.Table("u32", None, w.capacity)
    out["set_keys"], out["set_queries"], out["set_erase"] = w.keys, w.queries, w.erase
    out["set_ins"] = t.insert(w.keys)
    out["set_hits1"] = t.find(w.queries)[0]
    out["set_erased"] = t.erase(w.erase)
    out["set_hits2"] = t.find(w.queries)[0]
    out["set_contents"] = t.sorted_contents()[0]
    out["set_capacity"] = np.array([w.capacity])
    m = wl.map_u64_uniform(n=1 << 12, seed=1)
    t = po.Table("u64", "u32", m.capacity)
    out["map_keys"], out["map_vals"], out["map_queries"], out["map_erase"] = m.keys, m.values, m.queries, m.erase
    out["map_ins"] = t.insert(m.keys, m.values)
    h, v = t.find(m.queries)
    out["map_hits"], out["map_found_vals"] = h, np.where(h == 1, v, 0).astype(np.uint32)
    out["map_erased"] = t.erase(m.erase)
    k, vv = t.sorted_contents()
    out["map_contents_keys"], out["map_contents_vals"] = k, vv
    out["map_capacity"] = np.array([m.capacity])
    return out


def main():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(REFERENCE_KATS, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "small_workloads.npz"), **small_workloads())
    print("wrote", HERE)


if __name__ == "__main__":
    main()
