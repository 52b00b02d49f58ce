"""Measured random-access ceilings of the chip the hash tables run on.

Uses libstdgpu_amd_roofline.so (stdgpu_amd/csrc/roofline.hip): random aligned 64/128/256-byte
line reads over buffers of several sizes (inside and outside the 256 MB Infinity Cache), and
random 64-bit CAS / atomic-add throughput.  Rates are in G operations/s, timed with HIP events
on the stream the kernels run on.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstdgpu_amd_roofline.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m stdgpu_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        lib.roofline_fill.argtypes = [P, U64, P]
        lib.roofline_gather.argtypes = [P, U64, I, U64, U64, P, P]
        lib.roofline_cas.argtypes = [P, U64, U64, U64, P, P]
        lib.roofline_add.argtypes = [P, U64, U64, U64, P]
        lib.roofline_copy.argtypes = [P, P, U64, P]
        _lib = lib
    return _lib


def _time(fn, reps: int = 3) -> float:
    import torch

    s = torch.cuda.current_stream()
    fn()
    best = None
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        ms = a.elapsed_time(b)
        best = ms if best is None else min(best, ms)
    return best


def measure(sizes_mb=(64, 256, 4096), lines=(64, 128), n: int = 1 << 26) -> dict:
    import torch

    lib = load()
    s = int(torch.cuda.current_stream().cuda_stream)
    out = {"gather": {}, "cas": {}, "add": {}}
    sink = torch.zeros(4, dtype=torch.int32, device="cuda")
    big = max(sizes_mb) << 20
    buf = torch.empty(big, dtype=torch.uint8, device="cuda")
    assert lib.roofline_fill(buf.data_ptr(), big, s) == 0
    for mb in sizes_mb:
        nbytes = mb << 20
        for line in lines:
            ms = _time(lambda: lib.roofline_gather(buf.data_ptr(), nbytes, line, n, 12345, sink.data_ptr(), s))
            out["gather"][f"{line}B@{mb}MB"] = {"G_lines_per_s": round(n / ms / 1e6, 3),
                                                "GB_per_s": round(n * line / ms / 1e6, 1)}
        ms = _time(lambda: lib.roofline_cas(buf.data_ptr(), nbytes, n, 777, sink.data_ptr(), s))
        out["cas"][f"{mb}MB"] = {"G_ops_per_s": round(n / ms / 1e6, 3)}
        ms = _time(lambda: lib.roofline_add(buf.data_ptr(), nbytes, n, 778, s))
        out["add"][f"{mb}MB"] = {"G_ops_per_s": round(n / ms / 1e6, 3)}
    half = big // 2
    ms = _time(lambda: lib.roofline_copy(buf.data_ptr(), buf.data_ptr() + half, half, s))
    out["copy"] = {"bytes": half, "GB_per_s_read_plus_write": round(2 * half / ms / 1e6, 1)}
    del buf
    return out


if __name__ == "__main__":
    import json

    print(json.dumps(measure()), flush=True)
