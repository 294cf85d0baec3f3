#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/.

reference_kats.json -- known-answer vectors that the reference's own tests
    hold (typed in as data, with the file:line each comes from).
scipy_bfs.npz -- small graphs, sources and hop distances computed by
    scipy.sparse.csgraph.shortest_path(unweighted=True), a library
    independent of both this repository's oracle and its kernels.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF

KATS = {
    "disconnected_5": {
        "source": "crates/cpu_bfs/src/tests.rs:125-169",
        "num_vertices": 5,
        "row_offsets": [0, 2, 4, 6, 7, 8],
        "col_indices": [1, 2, 0, 2, 0, 1, 4, 3],
        "sources": [0, 3],
        "expected": [0, 1, 1, INF, INF, INF, INF, INF, 0, 1],
    },
    "single_vertex": {
        "source": "crates/cpu_bfs/src/tests.rs:172-181",
        "num_vertices": 1, "row_offsets": [0, 0], "col_indices": [],
        "sources": [0], "expected": [0],
    },
    "zero_edge_4": {
        "source": "crates/cpu_bfs/src/tests.rs:183-203",
        "num_vertices": 4, "row_offsets": [0, 0, 0, 0, 0], "col_indices": [],
        "sources": [0, 2],
        "expected": [0, INF, INF, INF, INF, INF, 0, INF],
    },
    "line_5": {
        "source": "crates/cpu_bfs/src/tests.rs:205-215",
        "num_vertices": 5, "row_offsets": [0, 1, 3, 5, 7, 8],
        "col_indices": [1, 0, 2, 1, 3, 2, 4, 3],
        "sources": [0, 4],
        "expected": [0, 1, 2, 3, 4, 4, 3, 2, 1, 0],
    },
    "grid_4x4_sources_0_5": {
        "source": "crates/cli/src/verify.rs:226-239",
        "grid": [4, 4], "sources": [0, 5],
        "expected": [0, 1, 2, 3, 1, 2, 3, 4, 2, 3, 4, 5, 3, 4, 5, 6,
                     2, 1, 2, 3, 1, 0, 1, 2, 2, 1, 2, 3, 3, 2, 3, 4],
    },
    "grid_3x3_oracle": {
        "source": "crates/core/src/generators.rs:646-663",
        "grid": [3, 3], "sources": [0, 4, 8],
        "spot_checks": [[0, 8, 4], [1, 0, 2], [2, 8, 0]],
    },
    "grid_10x10_oracle": {
        "source": "crates/core/src/generators.rs:625-644",
        "grid": [10, 10], "sources": [0],
        "spot_checks": [[0, 99, 18], [0, 0, 0]],
        "manhattan": [[55, 5, 5], [0, 99, 18]],
    },
    "grid_4x3_csr": {
        "source": "crates/core/src/generators.rs:594-608",
        "grid": [4, 3], "num_vertices": 12, "num_edges": 34,
    },
    "csr_construction": {
        "source": "crates/core/src/graph.rs:652-672",
        "num_vertices": 3, "edges": [[0, 1], [0, 2], [1, 2], [2, 0]],
        "row_offsets": [0, 2, 3, 4], "col_indices": [1, 2, 2, 0], "max_degree": 2,
    },
    "csr_filter": {
        "source": "crates/core/src/graph.rs:674-691",
        "num_vertices": 3, "edges": [[0, 0], [0, 1], [0, 1], [1, 2], [2, 2]],
        "row_offsets": [0, 1, 2, 2], "col_indices": [1, 2],
    },
    "csc_transpose": {
        "source": "crates/core/src/graph.rs:693-714",
        "num_vertices": 3, "edges": [[0, 1], [0, 2], [1, 2]],
        "col_offsets": [0, 0, 1, 3], "row_indices": [0, 0, 1],
    },
    "symmetry": {
        "source": "crates/core/src/graph.rs:716-742",
        "directed_edges": [[0, 1], [1, 2]],
        "symmetric_edges": [[0, 1], [1, 0], [1, 2], [2, 1]],
        "undirected_row_offsets": [0, 1, 3, 4], "undirected_col_indices": [1, 0, 2, 1],
    },
    "reachability_bitmask": {
        "source": "crates/core/src/graph.rs:760-797",
        "num_vertices": 5, "num_sources": 70, "words_per_vertex": 3,
        "set": [[0, 0], [33, 0], [69, 4]],
        "dist_entries": [[0, 0, 0], [33, 0, 2], [69, 4, 5]],
        "reachable_counts": [[0, 2], [4, 1], [1, 0]], "total_pairs": 3,
    },
    "u8_pack_lengths": {
        "source": "crates/gpu_bfs/src/tests.rs:163-222",
        "lengths": [0, 1, 2, 3, 4, 5, 7, 8, 15, 16, 31, 32, 33, 100, 1024, 65537],
        "rule": "expected[i] = INF if i % 7 == 0 else i % 255",
    },
    "quantisation_boundary": {
        "source": "crates/gpu_bfs/tests/m24_adversarial_challenge.rs:397-503",
        "path_lengths": [253, 254, 255, 256],
    },
    "batch_sizes": {
        "source": "tests/tier2_boundaries/b6_batch_sizes.rs:5-54",
        "sizes": [1, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1024,
                  2048],
    },
    "auto_router": {
        "source": "crates/gpu_bfs/src/engine.rs:502-605",
        "low_degree": [0, 1, 4, 7, 8],
        "low_degree_sources": [1, 2, 16, 31, 32, 63, 64, 256, 512, 1024, 2048],
        "intermediate_degree": [9, 16, 32, 33, 50, 64, 100, 127],
        "hub_degree": [128, 256, 1024, 38625],
        "hub_sources": [1, 2, 16, 32, 63, 64, 128, 256, 384, 512, 1024, 2048],
    },
}


def _csr_from_pairs(n, pairs, symmetric):
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if symmetric:
        pairs = np.vstack([pairs, pairs[:, ::-1]])
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pairs = np.unique(pairs, axis=0)
    ro = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ro, pairs[:, 0] + 1, 1)
    ro = np.cumsum(ro)
    return ro.astype(np.uint32), pairs[:, 1].astype(np.uint32)


def _scipy_dist(n, ro, ci, sources):
    m = sp.csr_matrix((np.ones(len(ci)), ci.astype(np.int64), ro.astype(np.int64)), shape=(n, n))
    d = shortest_path(m, directed=True, unweighted=True, indices=np.asarray(sources))
    d = np.where(np.isinf(d), INF, d).astype(np.uint64)
    return d.astype(np.uint32).reshape(-1)


def main():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)

    rng = np.random.Generator(np.random.PCG64(2024))
    arrays = {}
    cases = []
    # undirected ER, directed random, two components with a tail, star + path
    n = 200
    e = rng.integers(0, n, size=(500, 2))
    cases.append(("er200", n, e, True))
    n = 150
    e = rng.integers(0, n, size=(400, 2))
    cases.append(("directed150", n, e, False))
    e = [(i, i + 1) for i in range(59)] + [(60 + i, 60 + (i + 1) % 40) for i in range(40)]
    cases.append(("path_and_cycle", 110, e, True))
    e = [(0, i) for i in range(1, 80)] + [(79, 80 + i) for i in range(20)]
    cases.append(("star_tail", 100, e, True))
    for name, n, e, sym in cases:
        ro, ci = _csr_from_pairs(n, e, sym)
        src = np.sort(rng.choice(n, size=min(n, 70), replace=False)).astype(np.uint32)
        src = np.concatenate([src, src[:3]])  # duplicates are independent sources
        arrays[f"{name}__num_vertices"] = np.array([n], dtype=np.uint32)
        arrays[f"{name}__row_offsets"] = ro
        arrays[f"{name}__col_indices"] = ci
        arrays[f"{name}__sources"] = src
        arrays[f"{name}__dist"] = _scipy_dist(n, ro, ci, src)
    np.savez_compressed(os.path.join(HERE, "scipy_bfs.npz"), **arrays)
    print("wrote", sorted({k.split("__")[0] for k in arrays}))


if __name__ == "__main__":
    main()
