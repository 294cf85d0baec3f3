"""Batched multi-graph traversal (libmsbfs msbfs_batch_* API).

Many graphs, each with its own source list, are traversed by one persistent
launch; a run can be recorded into a hipGraph and replayed.  The reference
has no counterpart: it makes one engine call per graph and source set
(crates/gpu_bfs/src/engine.rs:219-297).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .engine import DeviceGraph, GpuBfsEngine


class Batch:
    def __init__(self, engine: GpuBfsEngine, graphs, sources, dist_dtype=np.uint32):
        self.engine = engine
        self.dgraphs = [g if isinstance(g, DeviceGraph) else engine._device_graph(g)
                        for g in graphs]
        self.counts = np.array([len(s) for s in sources], dtype=np.uint64)
        flat = (np.concatenate([np.asarray(s, dtype=np.uint32) for s in sources])
                if len(sources) else np.zeros(0, np.uint32))
        self.dtype = L.DIST_U16 if np.dtype(dist_dtype) == np.uint16 else L.DIST_U32
        self.np_dtype = np.uint16 if self.dtype == L.DIST_U16 else np.uint32
        n = len(self.dgraphs)
        handles = (ctypes.c_void_p * max(n, 1))(*[g.handle.value for g in self.dgraphs])
        h = ctypes.c_void_p()
        flat = np.ascontiguousarray(flat)
        L.check(L.lib().msbfs_batch_create(
            engine.handle, n, handles, self.counts.ctypes.data_as(ctypes.c_void_p),
            flat.ctypes.data_as(ctypes.c_void_p) if flat.size else None, self.dtype,
            ctypes.byref(h)))
        self._h = h
        self._finalizer = L.track("batch", self, L.lib().msbfs_batch_destroy, h)
        total = ctypes.c_uint64()
        self.offsets = np.zeros(max(n, 1), dtype=np.uint64)
        L.check(L.lib().msbfs_batch_layout(h, ctypes.byref(total),
                                           self.offsets.ctypes.data_as(ctypes.c_void_p)))
        self.offsets = self.offsets[:n]
        self.total = total.value
        self.num_vertices = [g.num_vertices for g in self.dgraphs]

    def set_sources(self, sources) -> None:
        flat = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.uint32)
                                                    for s in sources]))
        L.check(L.lib().msbfs_batch_set_sources(self._h, flat.ctypes.data_as(ctypes.c_void_p)))

    def run(self, out_ptr: int, out_len: int, stream: int | None = None, sync: bool = True):
        L.check(L.lib().msbfs_batch_run(self._h, ctypes.c_void_p(out_ptr), out_len,
                                        ctypes.c_void_p(stream) if stream else None, int(sync)))

    def capture(self, out_ptr: int, out_len: int, stream: int | None = None) -> None:
        L.check(L.lib().msbfs_batch_capture(self._h, ctypes.c_void_p(out_ptr), out_len,
                                            ctypes.c_void_p(stream) if stream else None))

    def replay(self, stream: int | None = None, sync: bool = True) -> None:
        L.check(L.lib().msbfs_batch_replay(self._h, ctypes.c_void_p(stream) if stream else None,
                                           int(sync)))

    def check(self) -> None:
        L.check(L.lib().msbfs_batch_check(self._h))

    def stats(self) -> dict:
        st = L.RunStats()
        L.check(L.lib().msbfs_batch_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def set_tuning(self, **kw) -> None:
        t = L.Tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        L.check(L.lib().msbfs_batch_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = L.Tuning()
        L.check(L.lib().msbfs_batch_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def autotune(self, out_ptr: int, out_len: int):
        t = L.Tuning()
        ms = ctypes.c_float()
        L.check(L.lib().msbfs_batch_autotune(self._h, ctypes.c_void_p(out_ptr), out_len,
                                             ctypes.byref(t), ctypes.byref(ms)))
        return {k: getattr(t, k) for k, _ in t._fields_}, ms.value

    def sweep(self) -> list:
        """The last autotune's candidates in the order they ran: dicts with
        the tuning knobs, "ms" (None when the candidate did not run) and
        "status"."""
        n = ctypes.c_uint32()
        L.check(L.lib().msbfs_batch_sweep(self._h, None, 0, ctypes.byref(n)))
        arr = (L.SweepEntry * max(n.value, 1))()
        L.check(L.lib().msbfs_batch_sweep(self._h, arr, n.value, ctypes.byref(n)))
        res = []
        for e in arr[:n.value]:
            d = {k: getattr(e.tuning, k) for k, _ in e.tuning._fields_}
            d["ms"] = float(e.ms) if e.ms >= 0 else None
            d["status"] = e.status
            res.append(d)
        return res

    def families(self) -> list:
        """Family (gpu_bfs.Family value) that traverses each graph: WIDE for
        the graphs in union traversals and for graphs whose visited bits
        exceed one workgroup's LDS (in a union, or, with batch_union 2, on a
        wide run of their own), DEEP for those of the persistent launch."""
        n = len(self.dgraphs)
        arr = (ctypes.c_int32 * max(n, 1))()
        L.check(L.lib().msbfs_batch_families(self._h, arr))
        return list(arr[:n])

    def split(self, out_host: np.ndarray) -> list:
        """Per-graph [S_g, V_g] views of a host copy of the batch output."""
        res = []
        for g, (off, S, V) in enumerate(zip(self.offsets, self.counts, self.num_vertices)):
            res.append(out_host[int(off):int(off) + int(S) * V].reshape(int(S), V))
        return res
