"""gpu_bfs -- MI355X multi-source BFS engine (Python face of libmsbfs).

Mirrors the reference's `gpu_bfs` and `webgpu_bfs_core` crates
(crates/gpu_bfs/src/engine.rs, crates/core/src): CSR graphs in, per-source
hop distances dist[s*|V|+v] out.  All traversal runs in hand-written HIP
kernels for gfx950 behind the C ABI in include/msbfs.h.
"""
from ._lib import (DIST_U8, DIST_U16, DIST_U32, LIB_PATH, MsbfsError, MtxParseError, NoDeviceError, shutdown,
                   SourceOutOfBoundsError, lib)
from .core import (CscGraph, CsrGraph, ReachabilityBitmask, csr_from_undirected_pairs,
                   generate_erdos_renyi_np, generate_graph500_rmat, generate_grid2d,
                   generate_rmat, generate_road_like, parse_mtx, read_mtx, sample_sources,
                   traversed_edges, write_mtx)
from .engine import DeviceGraph, Family, GpuBfsEngine, GpuBfsMethod, GpuContext
from .batch import Batch
from .server import Server

__all__ = [
    "CsrGraph", "CscGraph", "ReachabilityBitmask", "GpuContext", "GpuBfsEngine", "GpuBfsMethod",
    "DeviceGraph", "Family", "Batch", "Server", "MsbfsError", "SourceOutOfBoundsError", "NoDeviceError",
    "MtxParseError", "parse_mtx", "read_mtx", "write_mtx",
    "generate_grid2d", "generate_road_like", "generate_rmat", "generate_graph500_rmat",
    "generate_erdos_renyi_np", "csr_from_undirected_pairs", "sample_sources", "traversed_edges",
    "DIST_U32", "DIST_U16", "DIST_U8", "LIB_PATH", "lib", "shutdown",
]
