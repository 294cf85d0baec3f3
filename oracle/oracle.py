"""ctypes front-end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the timed CPU baseline.
The shipped library never imports it.

Each wrapper names the reference function it restates (see oracle.c for the
file:line citations).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
INF = 0xFFFFFFFF

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        Z = ctypes.c_size_t
        L.orc_csr_from_edge_list.restype = Z
        L.orc_csr_from_edge_list.argtypes = [Z, Z, P, P, P, P]
        L.orc_csr_to_csc.restype = None
        L.orc_csr_to_csc.argtypes = [Z, Z, P, P, P, P]
        for name in ("orc_ms_bfs_fifo", "orc_ms_bfs_bit_parallel"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [Z, P, P, P, Z, P]
        L.orc_ms_bfs_parallel.restype = ctypes.c_int
        L.orc_ms_bfs_parallel.argtypes = [Z, P, P, P, Z, P, ctypes.c_int]
        L.orc_traversed_edges.restype = ctypes.c_uint64
        L.orc_traversed_edges.argtypes = [Z, P, Z, P]
        L.orc_reachability_from_dist.restype = None
        L.orc_reachability_from_dist.argtypes = [Z, Z, P, P]
        L.orc_grid_manhattan.restype = None
        L.orc_grid_manhattan.argtypes = [Z, Z, P, Z, P]
        for name in ("orc_pack_u16", "orc_unpack_u16", "orc_pack_u8", "orc_unpack_u8"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [Z, P, P]
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class SourceOutOfBounds(ValueError):
    pass


def csr_from_edge_list(num_vertices: int, edges):
    """CsrGraph::from_edge_list (crates/core/src/graph.rs:51-80)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    keep = (e[:, 0] >= 0) & (e[:, 1] >= 0) & (e[:, 0] < 2**32) & (e[:, 1] < 2**32)
    e = e[keep]
    us, vs = _u32(e[:, 0]), _u32(e[:, 1])
    ro = np.zeros(num_vertices + 1, dtype=np.uint32)
    ci = np.zeros(max(len(us), 1), dtype=np.uint32)
    k = lib().orc_csr_from_edge_list(num_vertices, len(us), _p(us), _p(vs), _p(ro), _p(ci))
    return ro, ci[:k].copy()


def csr_to_csc(num_vertices: int, ro, ci):
    """CsrGraph::to_csc (crates/core/src/graph.rs:90-126)."""
    ro, ci = _u32(ro), _u32(ci)
    co = np.zeros(num_vertices + 1, dtype=np.uint32)
    ri = np.zeros(max(len(ci), 1), dtype=np.uint32)
    lib().orc_csr_to_csc(num_vertices, len(ci), _p(ro), _p(ci), _p(co), _p(ri))
    return co, ri[: len(ci)].copy()


def _run(fn, num_vertices, ro, ci, sources, *extra):
    ro, ci, src = _u32(ro), _u32(ci), _u32(sources)
    S = len(src)
    if S == 0 or num_vertices == 0:
        if S > 0:
            raise SourceOutOfBounds(
                f"Source vertex {int(src[0])} at index 0 exceeds graph vertex count 0")
        return np.zeros(0, dtype=np.uint32)
    dist = np.empty(S * num_vertices, dtype=np.uint32)
    ci_ = ci if len(ci) else np.zeros(1, dtype=np.uint32)
    rc = fn(num_vertices, _p(ro), _p(ci_), _p(src), S, _p(dist), *extra)
    if rc == 2:
        bad = int(np.argmax(src >= num_vertices))
        raise SourceOutOfBounds(
            f"Source vertex {int(src[bad])} at index {bad} exceeds graph vertex count {num_vertices}")
    return dist


def ms_bfs_fifo(num_vertices, ro, ci, sources):
    """ms_bfs_fifo (crates/cpu_bfs/src/fifo.rs:70-102)."""
    return _run(lib().orc_ms_bfs_fifo, num_vertices, ro, ci, sources)


def ms_bfs_parallel(num_vertices, ro, ci, sources, threads: int | None = None):
    """ms_bfs_rayon (crates/cpu_bfs/src/rayon_ms.rs:11-54), pthreads over sources."""
    t = threads or os.cpu_count() or 1
    return _run(lib().orc_ms_bfs_parallel, num_vertices, ro, ci, sources, int(t))


def ms_bfs_bit_parallel(num_vertices, ro, ci, sources):
    """ms_bfs_bit_parallel (crates/cpu_bfs/src/bit_parallel.rs:101-130)."""
    return _run(lib().orc_ms_bfs_bit_parallel, num_vertices, ro, ci, sources)


def traversed_edges(num_vertices, ro, num_sources, dist) -> int:
    """calculate_traversed_edges (crates/cli/src/bench.rs:478-491)."""
    ro, d = _u32(ro), _u32(dist)
    return int(lib().orc_traversed_edges(num_vertices, _p(ro), num_sources, _p(d)))


def reachability_from_dist(num_vertices, num_sources, dist) -> np.ndarray:
    """ReachabilityBitmask::from_canonical_distances (crates/core/src/graph.rs:571-645)."""
    wpv = (num_sources + 31) // 32 if num_sources else 0
    words = np.zeros(max(num_vertices * wpv, 1), dtype=np.uint32)
    d = _u32(dist) if len(dist) else np.zeros(1, dtype=np.uint32)
    lib().orc_reachability_from_dist(num_vertices, num_sources, _p(d), _p(words))
    return words[: num_vertices * wpv]


def grid_manhattan(width, height, sources) -> np.ndarray:
    """Grid2D::ms_bfs_oracle (crates/core/src/generators.rs:118-127)."""
    src = _u32(sources)
    out = np.empty(len(src) * width * height, dtype=np.uint32)
    lib().orc_grid_manhattan(width, height, _p(src), len(src), _p(out))
    return out


def pack_u16(d) -> np.ndarray:
    d = _u32(d)
    out = np.zeros(max((len(d) + 1) // 2, 1), dtype=np.uint32)
    lib().orc_pack_u16(len(d), _p(d) if len(d) else None, _p(out))
    return out[: (len(d) + 1) // 2]


def unpack_u16(n, words) -> np.ndarray:
    w = _u32(words) if len(words) else np.zeros(1, dtype=np.uint32)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    lib().orc_unpack_u16(n, _p(w), _p(out))
    return out[:n]


def pack_u8(d) -> np.ndarray:
    d = _u32(d)
    out = np.zeros(max((len(d) + 3) // 4, 1), dtype=np.uint32)
    lib().orc_pack_u8(len(d), _p(d) if len(d) else None, _p(out))
    return out[: (len(d) + 3) // 4]


def unpack_u8(n, words) -> np.ndarray:
    w = _u32(words) if len(words) else np.zeros(1, dtype=np.uint32)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    lib().orc_unpack_u8(n, _p(w), _p(out))
    return out[:n]
