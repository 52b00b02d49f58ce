"""Seeded synthetic inputs for the hashed-container workloads (numpy, host side).

Every generator is a pure function of its seed, so the GPU path and the CPU oracle see
identical inputs.  Random 64-bit words come from the splitmix64 sequence: with
GOLDEN = 0x9E3779B97F4A7C15, word i (i = 1, 2, ...) of the stream `seed` is
mix(seed + i * GOLDEN), where mix is splitmix64's finaliser (shifts 30/27/31, multipliers
0xBF58476D1CE4E5B9 and 0x94D049BB133111EB).  The map value of a key is
value(key) = uint32(mix(key + GOLDEN)), i.e. the first splitmix64 output with state = key.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
U64_MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def stream(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Words offset+1 .. offset+n of the splitmix64 stream `seed`."""
    i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix(np.uint64(seed) + i * GOLDEN)


def value_of(keys: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        return (mix(np.asarray(keys, dtype=np.uint64) + GOLDEN) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def _unique_in_order(a: np.ndarray) -> np.ndarray:
    _, first = np.unique(a, return_index=True)
    return a[np.sort(first)]


# ---------------------------------------------------------------------------------------------
@dataclass
class SetU32Workload:
    """unordered_set<uint32_t>: capacity, unique keys, finds (half present), erase half."""

    capacity: int
    keys: np.ndarray       # unique, uint32, in [1, 2^32-2]
    queries: np.ndarray    # uint32
    expect_hit: np.ndarray  # uint8, membership of queries in keys
    erase: np.ndarray      # uint32, subset of keys


def set_u32(n: int = 1 << 20, capacity: int = 1 << 21, seed: int = 0) -> SetU32Workload:
    """Config: set<uint32_t>, capacity 2^21, 2^20 unique keys uniform in [1, 2^32-2] from
    seed 0; 2^20 finds, half sampled from the keys and half guaranteed absent; erase a
    seeded random half; the finds run again after the erase."""
    out = np.empty(0, dtype=np.uint32)
    off = 0
    while len(out) < n:
        need = (n - len(out)) + (n >> 4) + 64
        w = (stream(seed, need, off) >> np.uint64(32)).astype(np.uint32)
        off += need
        w = w[(w != 0) & (w != 0xFFFFFFFF)]
        out = _unique_in_order(np.concatenate([out, w]))
    keys = out[:n]
    rng = np.random.default_rng(seed + 1)
    half = n // 2
    present = keys[rng.integers(0, n, size=half)]
    absent = np.empty(0, dtype=np.uint32)
    keyset = np.sort(keys)
    off2 = 0
    while len(absent) < n - half:
        w = (stream(seed + 2, n, off2) >> np.uint64(32)).astype(np.uint32)
        off2 += n
        w = w[(w != 0) & (w != 0xFFFFFFFF)]
        pos = np.searchsorted(keyset, w)
        pos[pos == len(keyset)] = 0
        w = w[keyset[pos] != w]
        absent = np.concatenate([absent, w])
    absent = absent[: n - half]
    queries = np.concatenate([present, absent])
    perm = rng.permutation(n)
    queries = queries[perm]
    expect = np.concatenate([np.ones(half, np.uint8), np.zeros(n - half, np.uint8)])[perm]
    erase = keys[rng.permutation(n)[:half]]
    return SetU32Workload(capacity, keys, queries, expect, erase)


# ---------------------------------------------------------------------------------------------
@dataclass
class MapU64Workload:
    """unordered_map<uint64_t, uint32_t>: insert, find at a hit rate, erase a fraction, re-insert."""

    capacity: int
    keys: np.ndarray        # uint64 (unique for the uniform draw; duplicates for Zipf)
    values: np.ndarray      # uint32, value_of(keys)
    queries: np.ndarray     # uint64
    erase: np.ndarray       # uint64, subset of the distinct keys
    reinsert: np.ndarray    # uint64 (the erased keys)


def _capacity_for(n: int, load: float) -> int:
    return int(np.ceil(n / load))


def map_u64_uniform(n: int = 1 << 26, load: float = 0.8, hit: float = 0.9, erase_frac: float = 0.25,
                    seed: int = 1) -> MapU64Workload:
    """Config: map<uint64_t,uint32_t>, capacity for load 0.8 holding n keys; uniform 64-bit
    keys excluding the EMPTY sentinel (all ones); value = value_of(key); n finds at the hit
    rate; erase erase_frac of the keys; re-insert them."""
    keys = stream(seed, n)
    keys[keys == U64_MAX] = np.uint64(1)
    vals = value_of(keys)
    rng = np.random.default_rng(seed + 100)
    nhit = int(round(n * hit))
    q_present = keys[rng.integers(0, n, size=nhit)]
    q_absent = stream(seed + 7, n - nhit)  # a disjoint stream: absent except with probability ~n^2/2^64
    queries = np.concatenate([q_present, q_absent])
    queries = queries[rng.permutation(n)]
    ne = int(n * erase_frac)
    erase = keys[rng.permutation(n)[:ne]]
    return MapU64Workload(_capacity_for(n, load), keys, vals, queries, erase, erase.copy())


def zipf_ranks(n: int, s: float, universe: int, seed: int) -> np.ndarray:
    """n draws of a Zipf(s) rank over 1..universe (inverse-CDF on a seeded uniform stream)."""
    r = np.arange(1, universe + 1, dtype=np.float64)
    w = r ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = (stream(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return np.searchsorted(cdf, u, side="right").astype(np.uint64) + np.uint64(1)


def map_u64_zipf(n: int = 1 << 26, load: float = 0.8, s: float = 1.1, hit: float = 0.9,
                 erase_frac: float = 0.25, seed: int = 2, universe: int | None = None) -> MapU64Workload:
    """Config: the same sequence with Zipf(s=1.1) key draws carrying heavy duplicates; the key
    of rank r is mix(r) (distinct per rank), so value = value_of(key) stays deterministic."""
    universe = universe or n
    ranks = zipf_ranks(n, s, universe, seed)
    keys = mix(ranks * GOLDEN)
    keys[keys == U64_MAX] = np.uint64(1)
    vals = value_of(keys)
    distinct = np.unique(keys)
    rng = np.random.default_rng(seed + 100)
    nhit = int(round(n * hit))
    q_present = keys[rng.integers(0, n, size=nhit)]
    q_absent = stream(seed + 7, n - nhit)
    queries = np.concatenate([q_present, q_absent])[rng.permutation(n)]
    ne = int(len(distinct) * erase_frac)
    erase = distinct[rng.permutation(len(distinct))[:ne]]
    return MapU64Workload(_capacity_for(n, load), keys, vals, queries, erase, erase.copy())


# ---------------------------------------------------------------------------------------------
@dataclass
class BatchedWorkload:
    """Many independent map<uint64_t,uint32_t> tables; flattened per-phase arrays with offsets."""

    sizes: np.ndarray          # keys per table
    capacities: np.ndarray     # capacity per table
    ins_offsets: np.ndarray    # int64[ntables + 1]
    ins_keys: np.ndarray       # uint64, unique within each table
    ins_vals: np.ndarray       # uint32, value_of(key)
    find_offsets: np.ndarray
    find_keys: np.ndarray
    erase_offsets: np.ndarray
    erase_keys: np.ndarray


def batched_maps(ntables: int = 4096, log2_lo: float = 8, log2_hi: float = 18, load: float = 0.8,
                 hit: float = 0.7, erase_frac: float = 0.3, seed: int = 4) -> BatchedWorkload:
    """Config: ntables maps whose key counts are log-uniform in [2^lo, 2^hi]; uniform keys with
    value_of(key); all inserts, then one find per inserted key position (each a key of the same
    table with probability `hit`, else a fresh key), then erases of the first erase_frac of each
    table's keys (keys are uniform, so that is a random subset)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(log2_lo, log2_hi, size=ntables)).astype(np.int64)
    total = int(sizes.sum())
    keys = stream(seed, total)
    keys[keys == U64_MAX] = np.uint64(1)
    ins_off = np.zeros(ntables + 1, np.int64)
    ins_off[1:] = np.cumsum(sizes)
    tid = np.repeat(np.arange(ntables), sizes)
    pos = np.arange(total, dtype=np.int64) - ins_off[tid]
    u = (stream(seed + 5, total) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    v = (stream(seed + 6, total) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    own = keys[ins_off[tid] + np.minimum((v * sizes[tid]).astype(np.int64), sizes[tid] - 1)]
    f_keys = np.where(u < hit, own, stream(seed + 7, total))
    e_sizes = np.floor(sizes * erase_frac).astype(np.int64)
    e_keys = keys[pos < e_sizes[tid]]
    e_off = np.zeros(ntables + 1, np.int64)
    e_off[1:] = np.cumsum(e_sizes)
    caps = np.ceil(sizes / load).astype(np.int64)
    return BatchedWorkload(sizes, caps, ins_off, keys, value_of(keys), ins_off.copy(), f_keys, e_off, e_keys)
