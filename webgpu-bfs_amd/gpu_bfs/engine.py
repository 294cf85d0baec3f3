"""Engine API: the Python face of the reference's crates/gpu_bfs/src/engine.rs.

`GpuBfsEngine` keeps the reference's method enum, constructors, Auto
decision table, argument checks and error behaviour; every call goes through
the C ABI of libmsbfs (include/msbfs.h) into the HIP kernels.
"""
from __future__ import annotations

import ctypes
import enum
import threading
import weakref

import numpy as np

from . import _lib as L
from .core import CsrGraph, ReachabilityBitmask


class GpuBfsMethod(enum.IntEnum):
    """GpuBfsMethod (crates/gpu_bfs/src/engine.rs:33-54), same order."""

    NaiveSequential = 0
    BatchedDenseSpMM = 1
    BitParallel32 = 2
    BitParallel64 = 3
    BitParallel64Pull = 4
    BooleanReachability = 5
    PushPullAdaptive = 6
    SparseQueue = 7
    SparseBitParallel64 = 8
    MultiWordBitParallel = 9
    IndirectMacroBatch = 10
    IndirectSparseQueuePush = 11
    MultiWordBitParallelPush = 12
    WorkgroupPersistentBfs = 13
    Auto = 14


class Family(enum.IntEnum):
    NONE = 0
    DEEP = 1
    WIDE = 2


class GpuContext:
    """Device context (GpuContext::init_blocking, crates/gpu_bfs/src/context.rs:190)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        L.check(L.lib().msbfs_context_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device
        self._finalizer = L.track("context", self, L.lib().msbfs_context_destroy, h)

    @classmethod
    def init_blocking(cls, device: int = 0) -> "GpuContext":
        return cls(device)

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        di = L.DeviceInfo()
        L.check(L.lib().msbfs_context_info(self._h, ctypes.byref(di)))
        return {"name": di.name.decode(), "arch": di.arch.decode(),
                "compute_units": di.compute_units, "lds_bytes_per_block": di.lds_bytes_per_block,
                "total_mem_bytes": di.total_mem_bytes,
                "cooperative_launch": bool(di.cooperative_launch)}

    def synchronize(self) -> None:
        L.check(L.lib().msbfs_context_synchronize(self._h))


class DeviceGraph:
    """A CSR graph resident in device memory (explicit handle; replaces the
    reference's size-keyed buffer pools, crates/gpu_bfs/src/buffer.rs:96-219)."""

    def __init__(self, ctx: GpuContext, graph: CsrGraph):
        h = ctypes.c_void_p()
        ro = np.ascontiguousarray(graph.row_offsets, dtype=np.uint32)
        ci = np.ascontiguousarray(graph.col_indices, dtype=np.uint32)
        if len(ro) != graph.num_vertices + 1:
            raise L.MsbfsError(L.ERR_INVALID_GRAPH, "row_offsets length != num_vertices + 1")
        if len(ci) != graph.num_edges:
            raise L.MsbfsError(L.ERR_INVALID_GRAPH, "col_indices length != num_edges")
        L.check(L.lib().msbfs_graph_create(
            ctx.handle, graph.num_vertices, graph.num_edges,
            ro.ctypes.data_as(ctypes.c_void_p),
            ci.ctypes.data_as(ctypes.c_void_p) if len(ci) else None, ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        self.num_vertices = graph.num_vertices
        self.num_edges = graph.num_edges
        self._finalizer = L.track("graph", self, L.lib().msbfs_graph_destroy, h)

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        gi = L.GraphInfo()
        L.check(L.lib().msbfs_graph_get_info(self._h, ctypes.byref(gi)))
        return {k: getattr(gi, k) for k, _ in gi._fields_}


def _src_array(sources) -> np.ndarray:
    a = np.asarray(sources)
    if a.size and (a.min() < 0 or a.max() > 0xFFFFFFFF):
        raise L.SourceOutOfBoundsError(L.ERR_SOURCE_OOB, "source id outside the u32 range")
    return np.ascontiguousarray(a, dtype=np.uint32)


class GpuBfsEngine:
    """GpuBfsEngine (crates/gpu_bfs/src/engine.rs:57-480) on libmsbfs.

    Host graphs passed to the run methods are uploaded once and cached per
    graph object (identity, not size); `DeviceGraph` handles can be passed
    directly to avoid the cache.
    """

    def __init__(self, context: GpuContext, method: GpuBfsMethod = GpuBfsMethod.Auto):
        h = ctypes.c_void_p()
        L.check(L.lib().msbfs_engine_create(context.handle, int(method), ctypes.byref(h)))
        self._h = h
        self._ctx = context
        self._method = GpuBfsMethod(method)
        self._graphs: "weakref.WeakKeyDictionary[CsrGraph, tuple]" = weakref.WeakKeyDictionary()
        self._lock = threading.Lock()
        self._finalizer = L.track("engine", self, L.lib().msbfs_engine_destroy, h)

    # constructors mirroring engine.rs:68-154
    @classmethod
    def new(cls, context: GpuContext, method: GpuBfsMethod) -> "GpuBfsEngine":
        return cls(context, method)

    @classmethod
    def default_with_method(cls, method: GpuBfsMethod) -> "GpuBfsEngine":
        return cls(GpuContext.init_blocking(), method)

    @classmethod
    def auto(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.Auto)

    @classmethod
    def workgroup_persistent_bfs(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.WorkgroupPersistentBfs)

    @classmethod
    def push_pull_adaptive(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.PushPullAdaptive)

    @classmethod
    def multi_word_bit_parallel(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.MultiWordBitParallel)

    @classmethod
    def bit_parallel_64(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.BitParallel64)

    @classmethod
    def bit_parallel_64_pull(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.BitParallel64Pull)

    @classmethod
    def indirect_sparse_queue_push(cls, context: GpuContext) -> "GpuBfsEngine":
        return cls(context, GpuBfsMethod.IndirectSparseQueuePush)

    @property
    def handle(self):
        return self._h

    def method(self) -> GpuBfsMethod:
        return self._method

    def context(self) -> GpuContext:
        return self._ctx

    @staticmethod
    def select_auto_method(graph: CsrGraph, num_sources: int) -> GpuBfsMethod:
        """engine.rs:157-198."""
        md = graph.max_degree() if graph.num_vertices else 0
        return GpuBfsMethod(L.lib().msbfs_select_auto_method(graph.num_vertices, graph.num_edges,
                                                             md, num_sources))

    def resolve_method(self, graph: CsrGraph, sources) -> GpuBfsMethod:
        """engine.rs:201-206."""
        if self._method == GpuBfsMethod.Auto:
            return self.select_auto_method(graph, len(sources))
        return self._method

    def resolve_family(self, graph, num_sources: int) -> Family:
        dg = self._device_graph(graph)
        m = ctypes.c_int()
        return Family(L.lib().msbfs_engine_resolve(self._h, dg.handle, num_sources,
                                                   ctypes.byref(m)))

    def set_push_pull_config(self, alpha: int, beta: int) -> None:
        """PushPullConfig { alpha, beta } (crates/gpu_bfs/src/dispatch/push_pull.rs:22-36)."""
        L.check(L.lib().msbfs_engine_set_push_pull(self._h, alpha, beta))

    def set_tuning(self, **kw) -> None:
        t = L.Tuning()
        L.check(L.lib().msbfs_engine_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        L.check(L.lib().msbfs_engine_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self) -> dict:
        t = L.Tuning()
        L.check(L.lib().msbfs_engine_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def last_stats(self) -> dict:
        st = L.RunStats()
        L.check(L.lib().msbfs_engine_last_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def level_trace(self) -> list:
        """Per-level records of the last wide-family run (needs
        set_tuning(trace_levels=N) before the run)."""
        cap = max(1, self.get_tuning()["trace_levels"])
        recs = (L.LevelRecord * cap)()
        n = ctypes.c_uint32()
        L.check(L.lib().msbfs_engine_level_trace(self._h, recs, cap, ctypes.byref(n)))
        return [{k: getattr(recs[i], k) for k, _ in L.LevelRecord._fields_ if k != "reserved"}
                for i in range(n.value)]

    def _device_graph(self, graph) -> DeviceGraph:
        if isinstance(graph, DeviceGraph):
            return graph
        with self._lock:
            hit = self._graphs.get(graph)
            key = (graph.num_vertices, graph.num_edges, id(graph.row_offsets),
                   id(graph.col_indices))
            if hit is not None and hit[0] == key:
                return hit[1]
            dg = DeviceGraph(self._ctx, graph)
            self._graphs[graph] = (key, dg)
            return dg

    # BfsEngine::run_ms_bfs (crates/core/src/lib.rs:21-39)
    def run_ms_bfs(self, graph, sources, dist_dtype=np.uint32) -> np.ndarray:
        src = _src_array(sources)
        S = len(src)
        nv = graph.num_vertices
        dt = L.dist_code(dist_dtype)
        out = np.empty(S * nv, dtype=L.dist_numpy(dt))
        if S == 0:
            return out
        if nv == 0:
            raise L.SourceOutOfBoundsError(
                L.ERR_SOURCE_OOB, f"Source vertex {int(src[0])} at index 0 exceeds graph vertex count 0")
        dg = self._device_graph(graph)
        L.check(L.lib().msbfs_run_ms_bfs(self._h, dg.handle, src.ctypes.data_as(ctypes.c_void_p), S,
                                         dt, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def autotune(self, graph, sources, dist_dtype=np.uint32, out_ptr: int = 0,
                 out_len: int = 0) -> dict:
        """Measured routing (msbfs_engine_autotune): times both families' and
        their knobs' candidates on this graph and these sources, keeps the
        fastest in the engine.  Returns the winner's method, tuning and mean
        kernel ms.  Output goes to `out_ptr` (device) or an internal buffer."""
        src = _src_array(sources)
        dg = self._device_graph(graph)
        t = L.Tuning()
        meth = ctypes.c_int32()
        ms = ctypes.c_float()
        L.check(L.lib().msbfs_engine_autotune(
            self._h, dg.handle, src.ctypes.data_as(ctypes.c_void_p), len(src),
            L.dist_code(dist_dtype), ctypes.c_void_p(out_ptr) if out_ptr else None, out_len,
            ctypes.byref(t), ctypes.byref(meth), ctypes.byref(ms)))
        self._method = GpuBfsMethod(meth.value)
        return {"method": GpuBfsMethod(meth.value), "ms": ms.value,
                "tuning": {k: getattr(t, k) for k, _ in L.Tuning._fields_}}

    def run_ms_bfs_with_method(self, method: GpuBfsMethod, graph, sources) -> np.ndarray:
        """engine.rs:300-404: runs `method` regardless of the engine's own."""
        other = GpuBfsEngine(self._ctx, method)
        other._graphs = self._graphs
        return other.run_ms_bfs(graph, sources)

    def run_ms_bfs_into(self, graph, sources, out: np.ndarray) -> np.ndarray:
        """engine.rs:219-297: fills (and resizes if needed) a caller buffer."""
        res = self.run_ms_bfs(graph, sources, dist_dtype=out.dtype if out.dtype in
                              (np.uint16, np.uint8) else np.uint32)
        if out.size != res.size:
            out.resize(res.size, refcheck=False)
        out[...] = res
        return out

    def run_ms_bfs_device(self, graph, sources, out_ptr: int, out_len: int,
                          dist_dtype=np.uint32, stream: int | None = None,
                          sync: bool = True) -> None:
        """Output left in device memory at `out_ptr` (e.g. a torch tensor's
        data_ptr()), work enqueued on `stream` (a hipStream_t as int)."""
        src = _src_array(sources)
        dt = L.dist_code(dist_dtype)
        if len(src) and graph.num_vertices == 0:
            raise L.SourceOutOfBoundsError(
                L.ERR_SOURCE_OOB, f"Source vertex {int(src[0])} at index 0 exceeds graph vertex count 0")
        dg = self._device_graph(graph)
        L.check(L.lib().msbfs_run_ms_bfs_device(
            self._h, dg.handle, src.ctypes.data_as(ctypes.c_void_p) if len(src) else None,
            len(src), dt, ctypes.c_void_p(out_ptr), out_len,
            ctypes.c_void_p(stream) if stream else None, int(sync)))

    def run_ms_bfs_device_noalloc(self, graph, sources, out_ptr: int, out_len: int,
                                  dist_dtype=np.uint32, stream: int | None = None) -> None:
        """msbfs_run_ms_bfs_device_noalloc: as run_ms_bfs_device(sync=True)
        with every workspace in the library's static per-device scratch
        array (no device or pinned allocation during the call)."""
        src = _src_array(sources)
        dt = L.dist_code(dist_dtype)
        if len(src) and graph.num_vertices == 0:
            raise L.SourceOutOfBoundsError(
                L.ERR_SOURCE_OOB, f"Source vertex {int(src[0])} at index 0 exceeds graph vertex count 0")
        dg = self._device_graph(graph)
        L.check(L.lib().msbfs_run_ms_bfs_device_noalloc(
            self._h, dg.handle, src.ctypes.data_as(ctypes.c_void_p) if len(src) else None,
            len(src), dt, ctypes.c_void_p(out_ptr), out_len,
            ctypes.c_void_p(stream) if stream else None))

    def check(self) -> None:
        L.check(L.lib().msbfs_engine_check(self._h))

    def run_ms_reachability(self, graph, sources) -> ReachabilityBitmask:
        """engine.rs:420-479."""
        src = _src_array(sources)
        S, nv = len(src), graph.num_vertices
        bm = ReachabilityBitmask(nv, S)
        if S == 0 or nv == 0:
            return bm
        dg = self._device_graph(graph)
        L.check(L.lib().msbfs_run_ms_reachability(
            self._h, dg.handle, src.ctypes.data_as(ctypes.c_void_p), S,
            bm.words.ctypes.data_as(ctypes.c_void_p), bm.words.size))
        return bm

    # ReachabilityEngine::run_reachability (crates/core/src/lib.rs:65-88)
    run_reachability = run_ms_reachability
