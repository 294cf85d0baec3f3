"""Graph data model: the Python face of the reference's crates/core.

`CsrGraph` mirrors `webgpu_bfs_core::CsrGraph` (crates/core/src/graph.rs:8-21)
and its helpers; the generators mirror crates/core/src/generators.rs.  The
work is done in libmsbfs (C++), so the large benchmark graphs build quickly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


def _take_host_graph(h: ctypes.c_void_p) -> "CsrGraph":
    lib = L.lib()
    V, E = ctypes.c_uint64(), ctypes.c_uint64()
    ro_p, ci_p = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        L.check(lib.msbfs_host_graph_get(h, ctypes.byref(V), ctypes.byref(E),
                                         ctypes.byref(ro_p), ctypes.byref(ci_p)))
        nv, ne = V.value, E.value
        ro = np.ctypeslib.as_array(ctypes.cast(ro_p, ctypes.POINTER(ctypes.c_uint32)),
                                   shape=(nv + 1,)).copy()
        ci = (np.ctypeslib.as_array(ctypes.cast(ci_p, ctypes.POINTER(ctypes.c_uint32)),
                                    shape=(ne,)).copy()
              if ne else np.zeros(0, dtype=np.uint32))
    finally:
        lib.msbfs_host_graph_free(h)
    return CsrGraph(nv, ne, ro, ci)


@dataclass(eq=False)
class CsrGraph:
    """Compressed sparse row graph (crates/core/src/graph.rs:8-21)."""

    num_vertices: int
    num_edges: int
    row_offsets: np.ndarray
    col_indices: np.ndarray

    def __post_init__(self):
        self.row_offsets = _u32(self.row_offsets)
        self.col_indices = _u32(self.col_indices)

    __hash__ = object.__hash__

    def __eq__(self, other) -> bool:
        return (isinstance(other, CsrGraph) and self.num_vertices == other.num_vertices
                and self.num_edges == other.num_edges
                and np.array_equal(self.row_offsets, other.row_offsets)
                and np.array_equal(self.col_indices, other.col_indices))

    # CsrGraph::from_edge_list (graph.rs:51-80)
    @staticmethod
    def from_edge_list(num_vertices: int, edges) -> "CsrGraph":
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        keep = (e >= 0).all(axis=1) & (e < 2**32).all(axis=1)
        e = e[keep]
        us, vs = _u32(e[:, 0]), _u32(e[:, 1])
        h = ctypes.c_void_p()
        L.check(L.lib().msbfs_csr_from_edge_list(num_vertices, len(us), _ptr(us), _ptr(vs),
                                                 ctypes.byref(h)))
        return _take_host_graph(h)

    def to_csc(self) -> "CscGraph":
        """CsrGraph::to_csc (graph.rs:90-126)."""
        co = np.zeros(self.num_vertices + 1, dtype=np.uint32)
        ri = np.zeros(self.num_edges, dtype=np.uint32)
        L.check(L.lib().msbfs_csr_to_csc(self.num_vertices, self.num_edges, _ptr(self.row_offsets),
                                         _ptr(self.col_indices), _ptr(co), _ptr(ri)))
        return CscGraph(self.num_vertices, self.num_edges, co, ri)

    def is_symmetric(self) -> bool:
        """CsrGraph::is_symmetric (graph.rs:129-132)."""
        r = L.lib().msbfs_csr_is_symmetric(self.num_vertices, self.num_edges,
                                           _ptr(self.row_offsets), _ptr(self.col_indices))
        if r < 0:
            L.check(-r)
        return r == 1

    def lattice_offsets(self) -> list:
        """Id offsets of a lattice graph (symmetric, every edge joining ids
        that differ by one of at most four offsets), ascending; [] for any
        other graph.  The deep family keeps such a graph's adjacency in LDS
        (a few bits per vertex)."""
        d = np.zeros(4, np.uint32)
        r = L.lib().msbfs_csr_lattice_offsets(self.num_vertices, self.num_edges,
                                              _ptr(self.row_offsets), _ptr(self.col_indices),
                                              _ptr(d))
        if r < 0:
            L.check(-r)
        return [int(x) for x in d[:r]]

    def to_undirected(self) -> "CsrGraph":
        """CsrGraph::to_undirected (graph.rs:136-150)."""
        if self.is_symmetric():
            return CsrGraph(self.num_vertices, self.num_edges, self.row_offsets.copy(),
                            self.col_indices.copy())
        us = np.repeat(np.arange(self.num_vertices, dtype=np.uint32), np.diff(self.row_offsets))
        vs = self.col_indices
        h = ctypes.c_void_p()
        L.check(L.lib().msbfs_csr_from_undirected_pairs(self.num_vertices, len(us), _ptr(_u32(us)),
                                                        _ptr(vs), ctypes.byref(h)))
        return _take_host_graph(h)

    def validate(self) -> None:
        """CsrGraph::validate (graph.rs:226-278); raises MsbfsError on failure."""
        if len(self.row_offsets) != self.num_vertices + 1:
            raise L.MsbfsError(L.ERR_INVALID_GRAPH,
                               f"row_offsets has {len(self.row_offsets)} entries, expected "
                               f"{self.num_vertices + 1}")
        if len(self.col_indices) != self.num_edges:
            raise L.MsbfsError(L.ERR_INVALID_GRAPH, "col_indices length != num_edges")
        L.check(L.lib().msbfs_csr_validate(self.num_vertices, self.num_edges,
                                           _ptr(self.row_offsets), _ptr(self.col_indices)))
        for u in np.nonzero(np.diff(self.row_offsets) > 1)[0]:
            nb = self.neighbors(int(u))
            if np.any(nb[1:] <= nb[:-1]):
                raise L.MsbfsError(L.ERR_INVALID_GRAPH, f"unsorted or duplicate neighbours at {u}")

    def is_empty(self) -> bool:
        return self.num_vertices == 0

    def degree(self, u: int) -> int:
        if not 0 <= u < self.num_vertices:
            raise IndexError(f"Vertex {u} out of bounds for graph with {self.num_vertices} vertices")
        return int(self.row_offsets[u + 1]) - int(self.row_offsets[u])

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets.astype(np.int64))

    def neighbors(self, u: int) -> np.ndarray:
        if not 0 <= u < self.num_vertices:
            raise IndexError(f"Vertex {u} out of bounds for graph with {self.num_vertices} vertices")
        return self.col_indices[self.row_offsets[u]:self.row_offsets[u + 1]]

    def max_degree(self) -> int:
        return int(self.degrees().max()) if self.num_vertices else 0

    def avg_degree(self) -> float:
        return self.num_edges / self.num_vertices if self.num_vertices else 0.0


@dataclass(eq=False)
class CscGraph:
    """Compressed sparse column graph (crates/core/src/graph.rs:27-40)."""

    num_vertices: int
    num_edges: int
    col_offsets: np.ndarray
    row_indices: np.ndarray

    def in_degree(self, v: int) -> int:
        return int(self.col_offsets[v + 1]) - int(self.col_offsets[v])

    def predecessors(self, v: int) -> np.ndarray:
        return self.row_indices[self.col_offsets[v]:self.col_offsets[v + 1]]


class ReachabilityBitmask:
    """Vertex-major reachability bits (crates/core/src/graph.rs:446-646):
    bit s%32 of words[v*words_per_vertex + s//32] is set iff v is reachable
    from the s-th source."""

    def __init__(self, num_vertices: int, num_sources: int, words=None):
        self.num_vertices = num_vertices
        self.num_sources = num_sources
        self.words_per_vertex = (num_sources + 31) // 32 if num_sources else 0
        n = num_vertices * self.words_per_vertex
        self.words = np.zeros(n, dtype=np.uint32) if words is None else _u32(words)
        if len(self.words) != n:
            raise ValueError("Words buffer length must exactly equal num_vertices * words_per_vertex")

    def is_reachable(self, s_idx: int, v: int) -> bool:
        if s_idx >= self.num_sources or v >= self.num_vertices or self.words_per_vertex == 0:
            return False
        return bool((self.words[v * self.words_per_vertex + s_idx // 32] >> (s_idx % 32)) & 1)

    def total_reachable_pairs(self) -> int:
        return int(np.unpackbits(self.words.view(np.uint8)).sum())

    def __eq__(self, other) -> bool:
        return (isinstance(other, ReachabilityBitmask) and self.num_vertices == other.num_vertices
                and self.num_sources == other.num_sources
                and np.array_equal(self.words, other.words))


# ---------------------------------------------------------------- generators

def parse_mtx(text) -> CsrGraph:
    """MtxReader::parse (crates/core/src/mtx.rs:124-287) of Matrix Market text
    (str or bytes): coordinate files only, values ignored, self loops dropped,
    symmetric files mirrored, canonical CSR.  Raises MtxParseError."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_mtx_parse(data, len(data), ctypes.byref(h)))
    return _take_host_graph(h)


def read_mtx(path) -> CsrGraph:
    """CsrGraph::from_mtx_reader over a file (the CLI's `.mtx` loader)."""
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_mtx_read(os.fsencode(path), ctypes.byref(h)))
    return _take_host_graph(h)


def write_mtx(graph: CsrGraph, path, symmetric: bool = False) -> None:
    """Writes `graph` as a coordinate pattern file (1-based).  With
    `symmetric=True` only entries u > v are written and the banner says
    symmetric, so a symmetric graph reads back identical."""
    ro, ci = graph.row_offsets.astype(np.int64), graph.col_indices.astype(np.int64)
    rows = np.repeat(np.arange(graph.num_vertices, dtype=np.int64), np.diff(ro))
    if symmetric:
        keep = rows > ci
        rows, cols = rows[keep], ci[keep]
    else:
        cols = ci
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern "
                + ("symmetric" if symmetric else "general") + "\n")
        f.write(f"{graph.num_vertices} {graph.num_vertices} {len(rows)}\n")
        if len(rows):
            np.savetxt(f, np.stack([rows + 1, cols + 1], 1), fmt="%d")


def generate_grid2d(width: int, height: int) -> CsrGraph:
    """Grid2D::new(width, height).graph (crates/core/src/generators.rs:28-80)."""
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_gen_grid2d(width, height, ctypes.byref(h)))
    return _take_host_graph(h)


def generate_road_like(num_vertices: int = 114_599, width: int = 339, extra_edges: int = 5_068,
                       seed: int = 0x5EED) -> CsrGraph:
    """High-diameter road-like graph: seeded Kruskal spanning tree of a
    `width`-column grid plus `extra_edges` extra grid edges."""
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_gen_road_like(num_vertices, width, extra_edges, seed, ctypes.byref(h)))
    return _take_host_graph(h)


GRAPH500_A, GRAPH500_B, GRAPH500_C, GRAPH500_D = 0.57, 0.19, 0.19, 0.05


def generate_rmat(scale: int, edge_factor: int, a: float = GRAPH500_A, b: float = GRAPH500_B,
                  c: float = GRAPH500_C, d: float = GRAPH500_D, seed: int = 0x5EED,
                  directed: bool = False) -> CsrGraph:
    """R-MAT (crates/core/src/generators.rs:277-342) on this library's stream."""
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_gen_rmat(scale, edge_factor, a, b, c, d, seed, int(directed),
                                   ctypes.byref(h)))
    return _take_host_graph(h)


def generate_graph500_rmat(scale: int, edge_factor: int, seed: int = 0x5EED) -> CsrGraph:
    return generate_rmat(scale, edge_factor, seed=seed)


def generate_erdos_renyi_np(num_vertices: int, p: float, seed: int = 0x5EED) -> CsrGraph:
    """G(n, p) by geometric skipping (crates/core/src/generators.rs:139-191)."""
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_gen_erdos_renyi_np(num_vertices, p, seed, ctypes.byref(h)))
    return _take_host_graph(h)


def csr_from_undirected_pairs(num_vertices: int, pairs) -> CsrGraph:
    """csr_from_undirected_pairs (crates/core/src/generators.rs:547-556)."""
    e = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    us, vs = _u32(e[:, 0]), _u32(e[:, 1])
    h = ctypes.c_void_p()
    L.check(L.lib().msbfs_csr_from_undirected_pairs(num_vertices, len(us), _ptr(us), _ptr(vs),
                                                    ctypes.byref(h)))
    return _take_host_graph(h)


def sample_sources(graph: CsrGraph, count: int, seed: int = 0x5EED,
                   non_isolated: bool = False) -> np.ndarray:
    """Uniform sources without replacement, sorted ascending (the CLI's
    random strategy, crates/cli/src/graph_loader.rs:181-190, on this
    library's stream)."""
    out = np.zeros(count, dtype=np.uint32)
    L.check(L.lib().msbfs_sample_sources(_ptr(graph.row_offsets), graph.num_vertices, count, seed,
                                         int(non_isolated), _ptr(out)))
    return out


def traversed_edges(graph: CsrGraph, num_sources: int, dist: np.ndarray) -> int:
    """Traversed edges as the reference's CLI counts them
    (crates/cli/src/bench.rs:478-491): for every source, the out-degrees of
    all reached vertices."""
    deg = graph.degrees()
    d = np.asarray(dist).reshape(num_sources, graph.num_vertices)
    inf = np.iinfo(d.dtype).max
    return int(((d != inf) * deg[None, :]).sum())
