"""Serving loop (libmsbfs msbfs_server_* API): resident graphs and continuous
batching of BFS queries with admission by distance-output memory.

The reference has no serving path; its engine serves one call at a time
(crates/gpu_bfs/src/engine.rs:61-63).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .engine import DeviceGraph, GpuBfsMethod, GpuContext


class Server:
    def __init__(self, ctx: GpuContext, pool_bytes: int = 16 << 30, dist_dtype=np.uint16,
                 pack: bool = True, max_sources_per_launch: int = 2048,
                 max_queries_per_launch: int = 512, slots_per_lane: int = 4,
                 admit_wait_us: int = 0):
        cfg = L.ServerConfig()
        cfg.output_pool_bytes = pool_bytes
        cfg.max_sources_per_launch = max_sources_per_launch
        cfg.max_queries_per_launch = max_queries_per_launch
        self.dtype = L.DIST_U16 if np.dtype(dist_dtype) == np.uint16 else L.DIST_U32
        self.np_dtype = np.uint16 if self.dtype == L.DIST_U16 else np.uint32
        cfg.dist_type = self.dtype
        cfg.pack = int(pack)
        cfg.slots_per_lane = slots_per_lane
        cfg.admit_wait_us = admit_wait_us
        h = ctypes.c_void_p()
        L.check(L.lib().msbfs_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        self.graphs = []
        self._finalizer = L.track("server", self, L.lib().msbfs_server_destroy, h)

    def add_graph(self, graph, method: GpuBfsMethod = GpuBfsMethod.Auto) -> int:
        dg = graph if isinstance(graph, DeviceGraph) else DeviceGraph(self.ctx, graph)
        lane = ctypes.c_uint32()
        L.check(L.lib().msbfs_server_add_graph(self._h, dg.handle, int(method), ctypes.byref(lane)))
        self.graphs.append(dg)
        return lane.value

    def set_tuning(self, lane: int, **kw) -> None:
        t = L.Tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        L.check(L.lib().msbfs_server_set_tuning(self._h, lane, ctypes.byref(t)))

    def start(self) -> None:
        L.check(L.lib().msbfs_server_start(self._h))

    def submit(self, lane: int, sources) -> int:
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        qid = ctypes.c_uint64()
        L.check(L.lib().msbfs_server_submit(self._h, lane, src.ctypes.data_as(ctypes.c_void_p),
                                            len(src), ctypes.byref(qid)))
        return qid.value

    def wait(self, qid: int, timeout_s: float = 600.0) -> None:
        L.check(L.lib().msbfs_server_wait(self._h, qid, timeout_s))

    def next_completed(self, timeout_s: float = 600.0) -> int:
        """Oldest finished query not yet returned (completion order)."""
        qid = ctypes.c_uint64()
        L.check(L.lib().msbfs_server_next_completed(self._h, timeout_s, ctypes.byref(qid)))
        return qid.value

    def info(self, qid: int) -> dict:
        qi = L.QueryInfo()
        L.check(L.lib().msbfs_server_query_info(self._h, qid, ctypes.byref(qi)))
        return {k: getattr(qi, k) for k, _ in qi._fields_}

    def result(self, qid: int) -> np.ndarray:
        qi = self.info(qid)
        out = np.empty(qi["elements"], dtype=self.np_dtype)
        L.check(L.lib().msbfs_server_copy_result(self._h, qid, out.ctypes.data_as(ctypes.c_void_p),
                                                 out.size))
        return out

    def release(self, qid: int) -> None:
        L.check(L.lib().msbfs_server_release(self._h, qid))

    def stats(self) -> dict:
        st = L.ServerStats()
        L.check(L.lib().msbfs_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def reset_stats(self) -> None:
        L.check(L.lib().msbfs_server_reset_stats(self._h))

    def stop(self) -> None:
        L.check(L.lib().msbfs_server_stop(self._h))
