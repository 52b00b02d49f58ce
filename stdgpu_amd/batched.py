"""Batched form: many tables served by one persistent launch (C ABI htab_batch_*).

A TableBatch groups tables of one kind and bucket width.  One call runs one operation per table
over flattened device arrays: table t owns elements [offsets[t], offsets[t+1]).  `capture` records
the launch into a HIP graph that replays with the same buffers.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import check
from .containers import _ptr, _stream_ptr

OPS = {"insert": _lib.HTAB_OP_INSERT, "find": _lib.HTAB_OP_FIND, "erase": _lib.HTAB_OP_ERASE}


class Graph:
    def __init__(self, handle: ctypes.c_void_p, keepalive):
        self._g = handle
        self._keep = keepalive  # buffers the graph reads and writes

    def launch(self, stream=None):
        check(_lib.load().htab_graph_launch(self._g, _stream_ptr(stream)), "htab_graph_launch")

    def destroy(self):
        if self._g:
            check(_lib.load().htab_graph_destroy(self._g), "htab_graph_destroy")
            self._g = ctypes.c_void_p()


class TableBatch:
    def __init__(self, tables):
        self.tables = list(tables)
        arr = (ctypes.c_void_p * len(self.tables))(*[t._h.value for t in self.tables])
        h = ctypes.c_void_p()
        check(_lib.load().htab_batch_create(arr, len(self.tables), ctypes.byref(h)), "htab_batch_create")
        self._b = h

    def destroy(self):
        if self._b:
            check(_lib.load().htab_batch_destroy(self._b), "htab_batch_destroy")
            self._b = ctypes.c_void_p()

    @staticmethod
    def _op_args(op):
        if isinstance(op, str):
            return None, OPS[op]
        return op, 0  # per-table int8 device tensor

    def run(self, op, offsets, keys, vals=None, flags=None, vals_out=None, stream=None):
        ops, opc = self._op_args(op)
        check(
            _lib.load().htab_batch_run(self._b, _ptr(ops), opc, _ptr(offsets), _ptr(keys), _ptr(vals), _ptr(flags),
                                       _ptr(vals_out), _stream_ptr(stream)),
            "htab_batch_run",
        )

    def run_segments(self, segments, stream=None):
        """segments: list of (table, op, keys, vals, flags, vals_out) with device tensors."""
        arr = (_lib.htab_segment * max(len(segments), 1))()
        for j, (t, op, k, v, f, vo) in enumerate(segments):
            arr[j] = _lib.htab_segment(_ptr(k), _ptr(v), _ptr(f), _ptr(vo), k.numel(), t, OPS[op])
        check(_lib.load().htab_batch_run_segments(self._b, arr, len(segments), _stream_ptr(stream)),
              "htab_batch_run_segments")

    def run_per_table(self, op: str, host_offsets, keys, vals=None, flags=None, vals_out=None, stream=None):
        """The plain baseline: one htab_* call per table for the same work."""
        off = (ctypes.c_int64 * len(host_offsets))(*[int(x) for x in host_offsets])
        check(
            _lib.load().htab_batch_run_per_table(self._b, OPS[op], off, _ptr(keys), _ptr(vals), _ptr(flags),
                                                 _ptr(vals_out), _stream_ptr(stream)),
            "htab_batch_run_per_table",
        )

    def capture(self, op, offsets, keys, vals=None, flags=None, vals_out=None, stream=None) -> Graph:
        """Record run(...) into a HIP graph on `stream` (a created torch stream)."""
        import torch

        s = stream or torch.cuda.Stream()
        ops, opc = self._op_args(op)
        g = ctypes.c_void_p()
        check(
            _lib.load().htab_batch_capture(self._b, _ptr(ops), opc, _ptr(offsets), _ptr(keys), _ptr(vals),
                                           _ptr(flags), _ptr(vals_out), int(s.cuda_stream), ctypes.byref(g)),
            "htab_batch_capture",
        )
        return Graph(g, (ops, offsets, keys, vals, flags, vals_out, s))


def main(argv=None) -> None:
    """`python -m stdgpu_amd.batched [bucket_lanes probe_batch block_size load]`: the 4096-table
    workload (sizes log-uniform in [2^8, 2^18]) through one graph-captured launch per phase, then
    through one plain call per table; prints one JSON line with the phase times."""
    import json
    import sys

    import torch

    from . import autotune as at
    from . import containers as C
    from . import workloads_device as wd

    args = sys.argv[1:] if argv is None else argv
    # default: the autotuner's pick in the benchmark runs (8-slot buckets, probe batch 8, 64 threads, load 0.5)
    g, rb, blk, load = (int(args[0]), int(args[1]), int(args[2]), float(args[3])) if len(args) >= 4 else (8, 8, 64, 0.5)
    cfg = {"bucket_lanes": g, "probe_batch": rb, "probe_group": g * rb, "block_size": blk, "load": load}
    w = wd.batched_maps("cuda", ntables=4096)
    out = {"config": cfg}
    for label in ("batched_warmup", "batched"):
        tabs = at.build_tables(w["sizes"], cfg)
        batch = TableBatch(tabs)
        ms, _ = at.run_phases(batch, w)
        out[label + "_ms"] = {k: round(v, 3) for k, v in ms.items()}
        batch.destroy()
        for t in tabs:
            C.unordered_map_u64_u32.destroyDeviceObject(t)
    # the plain calls run with their own default workgroup size (the tuned one is the persistent kernel's)
    tabs = at.build_tables(w["sizes"], {**cfg, "block_size": C.Tuning().block_size})
    batch = TableBatch(tabs)
    s = torch.cuda.Stream()
    pms = {}
    for op, name in (("insert", "ins"), ("find", "find"), ("erase", "erase")):
        n = w[f"{name}_keys"].numel()
        flags = torch.zeros(n, dtype=torch.uint8, device="cuda")
        vals_out = torch.zeros(n, dtype=torch.int32, device="cuda") if op == "find" else None
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        batch.run_per_table(op, w[f"{name}_offsets"].cpu().numpy(), w[f"{name}_keys"],
                            w["ins_vals"] if op == "insert" else None, flags, vals_out, stream=s)
        b.record(s)
        b.synchronize()
        pms[op] = round(a.elapsed_time(b), 3)
    out["plain_per_table_ms"] = pms
    batch.destroy()
    for t in tabs:
        C.unordered_map_u64_u32.destroyDeviceObject(t)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
