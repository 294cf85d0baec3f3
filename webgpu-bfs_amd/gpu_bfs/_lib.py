"""Loads libmsbfs.so (the HIP engine and its C ABI, include/msbfs.h) with ctypes.

The library is built in-tree by `make -C webgpu-bfs_amd/csrc` (or
`__graft_entry__.build()`).  There is no fallback: if the library is missing
every entry point raises.
"""
from __future__ import annotations

import ctypes

import numpy as np
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libmsbfs.so"))

OK = 0
ERR_INVALID_ARG = 1
ERR_SOURCE_OOB = 2
ERR_EMPTY_GRAPH = 3
ERR_OUT_TOO_SMALL = 4
ERR_HIP = 5
ERR_OOM = 6
ERR_DIAMETER_U16 = 7
ERR_INVALID_GRAPH = 8
ERR_UNSUPPORTED = 9
ERR_NO_DEVICE = 10
ERR_BUSY = 11
ERR_PARSE = 12
ERR_DIAMETER_U8 = 13

DIST_U32 = 0
DIST_U16 = 1
DIST_U8 = 2


def dist_code(dtype) -> int:
    """numpy dtype (uint32 / uint16 / uint8) -> msbfs_dist_type."""
    d = np.dtype(dtype)
    if d == np.uint16:
        return DIST_U16
    if d == np.uint8:
        return DIST_U8
    if d == np.uint32:
        return DIST_U32
    raise ValueError(f"distances are uint32, uint16 or uint8, not {d}")


def dist_numpy(code: int):
    return {DIST_U32: np.uint32, DIST_U16: np.uint16, DIST_U8: np.uint8}[code]

FAMILY_NONE, FAMILY_DEEP, FAMILY_WIDE = 0, 1, 2


class MsbfsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class SourceOutOfBoundsError(MsbfsError, ValueError):
    pass


class NoDeviceError(MsbfsError):
    pass


class MtxParseError(MsbfsError, ValueError):
    """Matrix Market text rejected (crates/core/src/error.rs:45-81 MtxError).
    `kind` names the variant; `line_number`, `row`, `col`, `max_rows`,
    `max_cols` carry its fields where the variant has them."""

    KINDS = {1: "Io", 2: "InvalidHeader", 3: "UnsupportedFormat", 4: "InvalidDimensions",
             5: "InvalidCoordinateLine", 6: "ZeroIndex", 7: "IndexOutOfBounds"}

    def __init__(self, code: int, message: str, detail):
        super().__init__(code, message)
        self.kind = self.KINDS.get(detail.kind, "Unknown")
        self.line_number = detail.line_number
        self.row, self.col = detail.row, detail.col
        self.max_rows, self.max_cols = detail.max_rows, detail.max_cols


class MtxErrorDetail(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("line_number", ctypes.c_uint64), ("row", ctypes.c_uint64),
                ("col", ctypes.c_uint64), ("max_rows", ctypes.c_uint64),
                ("max_cols", ctypes.c_uint64)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32),
                ("compute_units", ctypes.c_int32), ("lds_bytes_per_block", ctypes.c_int32),
                ("total_mem_bytes", ctypes.c_uint64), ("cooperative_launch", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class GraphInfo(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint64), ("num_edges", ctypes.c_uint64),
                ("max_degree", ctypes.c_uint32), ("symmetric", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "alpha", "beta", "tile_words", "words_per_lane", "deep_bits", "deep_threads",
        "wide_blocks_per_cu", "direction_rule", "trace_levels", "stepped", "deep_rows",
        "deep_lattice", "batch_union", "union_words", "deep_blocks_per_cu")]


class SweepEntry(ctypes.Structure):
    _fields_ = [("tuning", Tuning), ("ms", ctypes.c_float), ("status", ctypes.c_int32)]


class LevelRecord(ctypes.Structure):
    _fields_ = [("level", ctypes.c_uint32), ("direction", ctypes.c_uint32),
                ("frontier", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("new_bits", ctypes.c_uint64), ("start_us", ctypes.c_double),
                ("mid_us", ctypes.c_double), ("end_us", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("output_pool_bytes", ctypes.c_uint64), ("max_sources_per_launch", ctypes.c_uint32),
                ("max_queries_per_launch", ctypes.c_uint32), ("dist_type", ctypes.c_int32),
                ("pack", ctypes.c_int32), ("slots_per_lane", ctypes.c_uint32),
                ("admit_wait_us", ctypes.c_uint32)]


class QueryInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("lane", ctypes.c_uint32),
                ("num_sources", ctypes.c_uint32), ("launch", ctypes.c_uint64),
                ("device_ptr", ctypes.c_void_p), ("elements", ctypes.c_uint64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("completed", "failed", "launches", "sources",
                                               "pending", "pool_bytes", "pool_used")] + \
               [(n, ctypes.c_double) for n in ("span_s", "p50_ms", "p99_ms", "mean_ms", "max_ms")] + \
               [(n, ctypes.c_uint64) for n in ("overlapped_launches", "max_in_flight")]


class RunStats(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("family", ctypes.c_int32),
                ("levels", ctypes.c_uint32), ("push_levels", ctypes.c_uint32),
                ("pull_levels", ctypes.c_uint32), ("instances", ctypes.c_uint32),
                ("bits_per_instance", ctypes.c_uint32), ("grid_blocks", ctypes.c_uint32),
                ("block_threads", ctypes.c_uint32), ("lattice_bits", ctypes.c_uint32),
                ("kernel_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("workspace_bytes", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "msbfs_last_error": (ctypes.c_char_p, []),
    "msbfs_version": (ctypes.c_char_p, []),
    "msbfs_status_string": (ctypes.c_char_p, [I32]),
    "msbfs_context_create": (I32, [I32, PP]),
    "msbfs_context_destroy": (I32, [P]),
    "msbfs_context_info": (I32, [P, ctypes.POINTER(DeviceInfo)]),
    "msbfs_context_synchronize": (I32, [P]),
    "msbfs_graph_create": (I32, [P, U64, U64, P, P, PP]),
    "msbfs_graph_destroy": (I32, [P]),
    "msbfs_graph_get_info": (I32, [P, ctypes.POINTER(GraphInfo)]),
    "msbfs_engine_create": (I32, [P, I32, PP]),
    "msbfs_engine_destroy": (I32, [P]),
    "msbfs_engine_method": (I32, [P]),
    "msbfs_engine_set_push_pull": (I32, [P, U32, U32]),
    "msbfs_engine_set_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "msbfs_tuning_validate": (I32, [ctypes.POINTER(Tuning)]),
    "msbfs_engine_get_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "msbfs_engine_last_stats": (I32, [P, ctypes.POINTER(RunStats)]),
    "msbfs_engine_level_trace": (I32, [P, P, U32, ctypes.POINTER(U32)]),
    "msbfs_select_auto_method": (I32, [U64, U64, U32, U64]),
    "msbfs_engine_resolve": (I32, [P, P, U64, ctypes.POINTER(I32)]),
    "msbfs_run_ms_bfs": (I32, [P, P, P, U64, I32, P, U64]),
    "msbfs_run_ms_bfs_device": (I32, [P, P, P, U64, I32, P, U64, P, I32]),
    "msbfs_engine_check": (I32, [P]),
    "msbfs_run_ms_bfs_device_noalloc": (I32, [P, P, P, U64, I32, P, U64, P]),
    "msbfs_run_ms_reachability": (I32, [P, P, P, U64, P, U64]),
    "msbfs_batch_create": (I32, [P, U32, P, P, P, I32, PP]),
    "msbfs_batch_destroy": (I32, [P]),
    "msbfs_batch_layout": (I32, [P, ctypes.POINTER(U64), P]),
    "msbfs_batch_set_sources": (I32, [P, P]),
    "msbfs_batch_run": (I32, [P, P, U64, P, I32]),
    "msbfs_batch_capture": (I32, [P, P, U64, P]),
    "msbfs_batch_replay": (I32, [P, P, I32]),
    "msbfs_batch_check": (I32, [P]),
    "msbfs_batch_stats": (I32, [P, ctypes.POINTER(RunStats)]),
    "msbfs_batch_set_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "msbfs_batch_get_tuning": (I32, [P, ctypes.POINTER(Tuning)]),
    "msbfs_batch_autotune": (I32, [P, P, U64, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_float)]),
    "msbfs_batch_families": (I32, [P, P]),
    "msbfs_batch_sweep": (I32, [P, ctypes.POINTER(SweepEntry), U32, ctypes.POINTER(U32)]),
    "msbfs_server_create": (I32, [P, ctypes.POINTER(ServerConfig), PP]),
    "msbfs_server_add_graph": (I32, [P, P, I32, ctypes.POINTER(U32)]),
    "msbfs_server_set_tuning": (I32, [P, U32, ctypes.POINTER(Tuning)]),
    "msbfs_server_start": (I32, [P]),
    "msbfs_server_submit": (I32, [P, U32, P, U32, ctypes.POINTER(U64)]),
    "msbfs_server_wait": (I32, [P, U64, ctypes.c_double]),
    "msbfs_server_next_completed": (I32, [P, ctypes.c_double, ctypes.POINTER(U64)]),
    "msbfs_server_query_info": (I32, [P, U64, ctypes.POINTER(QueryInfo)]),
    "msbfs_server_copy_result": (I32, [P, U64, P, U64]),
    "msbfs_server_release": (I32, [P, U64]),
    "msbfs_server_get_stats": (I32, [P, ctypes.POINTER(ServerStats)]),
    "msbfs_server_reset_stats": (I32, [P]),
    "msbfs_server_stop": (I32, [P]),
    "msbfs_server_destroy": (I32, [P]),
    "msbfs_host_graph_get": (I32, [P, ctypes.POINTER(U64), ctypes.POINTER(U64), PP, PP]),
    "msbfs_host_graph_free": (None, [P]),
    "msbfs_csr_from_edge_list": (I32, [U64, U64, P, P, PP]),
    "msbfs_csr_from_undirected_pairs": (I32, [U64, U64, P, P, PP]),
    "msbfs_csr_to_csc": (I32, [U64, U64, P, P, P, P]),
    "msbfs_csr_is_symmetric": (I32, [U64, U64, P, P]),
    "msbfs_csr_lattice_offsets": (I32, [U64, U64, P, P, P]),
    "msbfs_csr_validate": (I32, [U64, U64, P, P]),
    "msbfs_gen_grid2d": (I32, [U64, U64, PP]),
    "msbfs_gen_road_like": (I32, [U64, U64, U64, U64, PP]),
    "msbfs_gen_rmat": (I32, [U32, U32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, U64, I32, PP]),
    "msbfs_gen_erdos_renyi_np": (I32, [U64, ctypes.c_double, U64, PP]),
    "msbfs_sample_sources": (I32, [P, U64, U64, U64, I32, P]),
    "msbfs_engine_autotune": (I32, [P, P, P, U64, I32, P, U64, P, ctypes.POINTER(I32),
                                    ctypes.POINTER(ctypes.c_float)]),
    "msbfs_mtx_parse": (I32, [ctypes.c_char_p, U64, PP]),
    "msbfs_mtx_read": (I32, [ctypes.c_char_p, PP]),
    "msbfs_mtx_last_error": (I32, [P]),
}

_lib = None

# Live library objects by teardown rank: servers first (they hold engines,
# graphs and a worker thread), then batches, engines, graphs, contexts.
RANK = {"server": 0, "batch": 1, "engine": 2, "graph": 3, "context": 4}
_live: list = []  # (rank, weakref.finalize) in creation order
_shutdown_registered = False


def track(kind: str, obj, destroy, handle):
    """Destroys `handle` with `destroy` when `obj` is collected or at
    shutdown(), whichever comes first; returns the finalizer."""
    import weakref

    fin = weakref.finalize(obj, destroy, handle)
    fin.atexit = False  # shutdown() orders these itself
    _live.append((RANK[kind], fin))
    if len(_live) > 4096:
        _live[:] = [x for x in _live if x[1].alive]
    return fin


def shutdown() -> None:
    """Destroys every live server, batch, engine, graph and context, in that
    order, while the HIP runtime is still up.  Registered with atexit right
    after torch is imported, so it runs before torch's own exit handlers (and
    before any profiler's, which run at C exit); safe to call more than once.
    Objects created afterwards are tracked again."""
    live = sorted(((r, i, f) for i, (r, f) in enumerate(_live) if f.alive),
                  key=lambda x: (x[0], -x[1]))
    _live.clear()
    for _, _, fin in live:
        fin()


def _bind_one_hip_runtime() -> None:
    """PyTorch-ROCm wheels ship their own libamdhip64.so.7, with the same
    SONAME as the system ROCm's.  The dynamic loader keeps whichever copy is
    loaded first for the whole process, and torch fails to initialise on the
    other copy; device pointers and hipStream_t handles are also only valid
    within one runtime.  Loading torch first (when it is installed) makes
    libmsbfs bind to torch's runtime, so tensors and streams can be shared."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Returns the loaded libmsbfs, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libmsbfs.so not found at {LIB_PATH}; build it with "
                "`make -C webgpu-bfs_amd/csrc` (hipcc, gfx950)")
        _bind_one_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        global _shutdown_registered
        if not _shutdown_registered:
            import atexit

            atexit.register(shutdown)
            _shutdown_registered = True
    return _lib


def last_error() -> str:
    return lib().msbfs_last_error().decode(errors="replace")


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = last_error() or lib().msbfs_status_string(rc).decode()
    if rc == ERR_SOURCE_OOB:
        raise SourceOutOfBoundsError(rc, msg)
    if rc == ERR_NO_DEVICE:
        raise NoDeviceError(rc, msg)
    if rc == ERR_PARSE:
        d = MtxErrorDetail()
        lib().msbfs_mtx_last_error(ctypes.byref(d))
        raise MtxParseError(rc, msg, d)
    raise MsbfsError(rc, msg)
