"""The large synthetic workloads of stdgpu_amd.workloads, generated directly in device memory.

Same formulas as the numpy generators (splitmix64 stream, value_of, batched table layout), written
with torch int64 arithmetic (two's-complement wrap-around; right shifts masked to be logical), so the
benchmark does not spend tens of seconds building 10^8-element arrays on the host.  A CPU test checks
these against the numpy versions element for element.
"""
from __future__ import annotations

import numpy as np

from . import workloads as wl


def _i64(x: int) -> int:
    x &= 0xFFFFFFFFFFFFFFFF
    return x - (1 << 64) if x >= 1 << 63 else x


_GOLD = _i64(0x9E3779B97F4A7C15)
_M1 = _i64(0xBF58476D1CE4E5B9)
_M2 = _i64(0x94D049BB133111EB)


def _srl(z, k: int):
    """logical right shift of int64 bits"""
    return (z >> k) & ((1 << (64 - k)) - 1)


def mix(z):
    z = (z ^ _srl(z, 30)) * _M1
    z = (z ^ _srl(z, 27)) * _M2
    return z ^ _srl(z, 31)


def stream(seed: int, n: int, device, offset: int = 0):
    import torch

    i = torch.arange(offset + 1, offset + n + 1, dtype=torch.int64, device=device)
    return mix(_i64(seed) + i * _GOLD)


def value_of(keys):
    return _low32(mix(keys + _GOLD))


def _low32(z):
    import torch

    lo = z & 0xFFFFFFFF
    return torch.where(lo >= (1 << 31), lo - (1 << 32), lo).to(torch.int32)


def _unit(z):
    import torch

    return _srl(z, 11).to(torch.float64) * (1.0 / (1 << 53))


def batched_maps(device, ntables: int = 4096, log2_lo: float = 8, log2_hi: float = 18, load: float = 0.8,
                 hit: float = 0.7, erase_frac: float = 0.3, seed: int = 4) -> dict:
    """Device twin of workloads.batched_maps: returns int64/int32 tensors (uint64/uint32 bits)."""
    import torch

    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(log2_lo, log2_hi, size=ntables)).astype(np.int64)
    total = int(sizes.sum())
    keys = stream(seed, total, device)
    keys = torch.where(keys == -1, torch.ones_like(keys), keys)
    ins_off = np.zeros(ntables + 1, np.int64)
    ins_off[1:] = np.cumsum(sizes)
    sizes_t = torch.from_numpy(sizes).to(device)
    off_t = torch.from_numpy(ins_off).to(device)
    tid = torch.repeat_interleave(torch.arange(ntables, device=device), sizes_t)
    pos = torch.arange(total, dtype=torch.int64, device=device) - off_t[tid]
    u = _unit(stream(seed + 5, total, device))
    v = _unit(stream(seed + 6, total, device))
    sz = sizes_t[tid]
    own = keys[off_t[tid] + torch.minimum((v * sz.to(torch.float64)).to(torch.int64), sz - 1)]
    f_keys = torch.where(u < hit, own, stream(seed + 7, total, device))
    e_sizes = np.floor(sizes * erase_frac).astype(np.int64)
    e_keys = keys[pos < torch.from_numpy(e_sizes).to(device)[tid]]
    e_off = np.zeros(ntables + 1, np.int64)
    e_off[1:] = np.cumsum(e_sizes)
    return {
        "sizes": sizes,
        "capacities": np.ceil(sizes / load).astype(np.int64),
        "ins_offsets": off_t,
        "ins_keys": keys,
        "ins_vals": value_of(keys),
        "find_offsets": off_t,
        "find_keys": f_keys,
        "erase_offsets": torch.from_numpy(e_off).to(device),
        "erase_keys": e_keys,
    }


def check_against_numpy(ntables: int = 16, log2_lo: float = 4, log2_hi: float = 10) -> bool:
    """Element-for-element equality with the numpy generator (used by the CPU tests)."""
    import torch

    a = wl.batched_maps(ntables=ntables, log2_lo=log2_lo, log2_hi=log2_hi)
    b = batched_maps("cpu", ntables=ntables, log2_lo=log2_lo, log2_hi=log2_hi)
    eq = lambda x, y: np.array_equal(x.view(np.int64) if x.dtype == np.uint64 else x, y.numpy())  # noqa: E731
    return (np.array_equal(a.sizes, b["sizes"]) and np.array_equal(a.ins_offsets, b["ins_offsets"].numpy())
            and eq(a.ins_keys, b["ins_keys"]) and np.array_equal(a.ins_vals.view(np.int32), b["ins_vals"].numpy())
            and eq(a.find_keys, b["find_keys"]) and eq(a.erase_keys, b["erase_keys"])
            and np.array_equal(a.erase_offsets, b["erase_offsets"].numpy()) and bool(torch.is_tensor(b["ins_keys"])))


def map_u64_zipf(device, n: int = 1 << 26, load: float = 0.8, s: float = 1.1, hit: float = 0.9,
                 erase_frac: float = 0.25, seed: int = 2, universe: int | None = None) -> dict:
    """Device twin of workloads.map_u64_zipf (the Zipf(1.1) run of the large map sequence): keys,
    values, queries, erase and re-insert keys as int64/int32 tensors (uint64/uint32 bits)."""
    import torch

    universe = universe or n
    r = torch.arange(1, universe + 1, dtype=torch.float64, device=device)
    cdf = torch.cumsum(r ** (-s), 0)
    cdf /= cdf[-1].clone()
    ranks = torch.searchsorted(cdf, _unit(stream(seed, n, device)), right=True).to(torch.int64) + 1
    del r, cdf
    keys = mix(ranks * _GOLD)
    keys = torch.where(keys == -1, torch.ones_like(keys), keys)
    # ascending unsigned order, as numpy sorts uint64
    distinct = torch.unique(keys ^ _i64(1 << 63)) ^ _i64(1 << 63)
    rng = np.random.default_rng(seed + 100)
    nhit = int(round(n * hit))
    q_present = keys[torch.from_numpy(rng.integers(0, n, size=nhit)).to(device)]
    q_absent = stream(seed + 7, n - nhit, device)
    queries = torch.cat([q_present, q_absent])[torch.from_numpy(rng.permutation(n)).to(device)]
    ne = int(distinct.numel() * erase_frac)
    erase = distinct[torch.from_numpy(rng.permutation(distinct.numel())[:ne]).to(device)]
    return {"capacity": wl._capacity_for(n, load), "keys": keys, "values": value_of(keys), "queries": queries,
            "erase": erase, "reinsert": erase.clone(), "reinsert_values": value_of(erase),
            "distinct": int(distinct.numel())}


def check_zipf_against_numpy(n: int = 1 << 12) -> bool:
    """Element-for-element equality of the Zipf twin with the numpy generator (CPU tests)."""
    a = wl.map_u64_zipf(n=n)
    b = map_u64_zipf("cpu", n=n)
    eq = lambda x, y: np.array_equal(x.view(np.int64), y.numpy())  # noqa: E731
    return (a.capacity == b["capacity"] and eq(a.keys, b["keys"]) and eq(a.queries, b["queries"])
            and eq(a.erase, b["erase"]) and np.array_equal(a.values.view(np.int32), b["values"].numpy()))
