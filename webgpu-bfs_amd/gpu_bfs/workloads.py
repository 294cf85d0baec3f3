"""Benchmark workloads: the synthetic graphs and source sets of the five
measured configurations, built with this library's seeded generators.

  grid256   Grid2D 256x256 (65,536 V, 261,120 directed entries), S=64
  road      road-like 114,599 V on a 339-column grid: seeded Kruskal
            spanning tree + 5,068 extra grid edges (239,332 entries),
            S in {1, 256, 2048}
  rmat20    R-MAT scale 20, edge factor 16 (0.57/0.19/0.19/0.05),
            symmetrised, no self loops or duplicates; S in {64, 2048}
            drawn from non-isolated vertices; u16 distances
  batch512  512 G(n, p) graphs, n log-uniform in [1024, 65536], average
            degree 8; S log-uniform in [1, 256] per graph
  serve     road + R-MAT scale 22 resident; Poisson query arrivals
"""
from __future__ import annotations

import math

import numpy as np

from . import core

SEED = 0x5EED

ROAD_V, ROAD_WIDTH, ROAD_EXTRA = 114_599, 339, 5_068


def grid256():
    return core.generate_grid2d(256, 256)


def road():
    return core.generate_road_like(ROAD_V, ROAD_WIDTH, ROAD_EXTRA, SEED)


def rmat(scale: int = 20):
    return core.generate_rmat(scale, 16, 0.57, 0.19, 0.19, 0.05, seed=SEED)


def sources(graph, count: int, non_isolated: bool = False, seed: int = SEED):
    return core.sample_sources(graph, count, seed=seed + count, non_isolated=non_isolated)


def batch512(n_graphs: int = 512, seed: int = SEED, v_lo: int = 1024, v_hi: int = 65536,
             s_lo: int = 1, s_hi: int = 256, avg_degree: float = 8.0):
    """Returns [(graph, sources)] for the batched uneven workload."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for g in range(n_graphs):
        nv = int(round(math.exp(rng.uniform(math.log(v_lo), math.log(v_hi)))))
        ns = int(round(math.exp(rng.uniform(math.log(s_lo), math.log(s_hi)))))
        ns = max(1, min(ns, nv))
        graph = core.generate_erdos_renyi_np(nv, avg_degree / (nv - 1), seed=seed * 1000 + g)
        src = core.sample_sources(graph, ns, seed=seed * 7919 + g)
        out.append((graph, src))
    return out
