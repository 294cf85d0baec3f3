"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks, each chosen so that a plausible mistake in the oracle fails one:
  * the reference's own known-answer vectors (tests/golden/reference_kats.json);
  * scipy.sparse.csgraph BFS distances (tests/golden/scipy_bfs.npz), an
    independent library;
  * closed forms: Grid2D Manhattan distance, path d = |u - v|, cycle
    min(k, n - k);
  * brute force (Floyd-Warshall over 0/1 adjacency) on tiny random graphs;
  * the reference's 3-way CPU equivalence (FIFO == parallel == bit-parallel,
    crates/cpu_bfs/src/tests.rs:13-31);
  * planted mistakes in a copy of an oracle output are caught.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF


def kats():
    with open(os.path.join(HERE, "golden", "reference_kats.json")) as f:
        return json.load(f)


def grid_csr(w, h):
    """Grid2D CSR built from its definition (neighbours N, W, E, S)."""
    ro, ci = [0], []
    for y in range(h):
        for x in range(w):
            v = y * w + x
            if y > 0:
                ci.append(v - w)
            if x > 0:
                ci.append(v - 1)
            if x + 1 < w:
                ci.append(v + 1)
            if y + 1 < h:
                ci.append(v + w)
            ro.append(len(ci))
    return np.array(ro, np.uint32), np.array(ci, np.uint32)


def three_way(V, ro, ci, src):
    a = orc.ms_bfs_fifo(V, ro, ci, src)
    b = orc.ms_bfs_parallel(V, ro, ci, src, 4)
    c = orc.ms_bfs_bit_parallel(V, ro, ci, src)
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)
    return a


@pytest.mark.parametrize("name", ["disconnected_5", "single_vertex", "zero_edge_4", "line_5"])
def test_reference_kat_graphs(name):
    k = kats()[name]
    d = three_way(k["num_vertices"], k["row_offsets"], k["col_indices"], k["sources"])
    assert d.tolist() == k["expected"]


def test_grid_4x4_kat_and_manhattan():
    k = kats()["grid_4x4_sources_0_5"]
    ro, ci = grid_csr(4, 4)
    d = three_way(16, ro, ci, k["sources"])
    assert d.tolist() == k["expected"]
    assert np.array_equal(orc.grid_manhattan(4, 4, k["sources"]), d)


def test_grid_spot_checks():
    k3 = kats()["grid_3x3_oracle"]
    m = orc.grid_manhattan(3, 3, k3["sources"]).reshape(3, 9)
    for s, v, want in k3["spot_checks"]:
        assert m[s, v] == want
    k10 = kats()["grid_10x10_oracle"]
    m = orc.grid_manhattan(10, 10, [55, 0]).reshape(2, 100)
    assert m[0, 5] == 5 and m[1, 99] == 18


@pytest.mark.parametrize("w,h", [(10, 10), (17, 5), (1, 30), (33, 1), (64, 48)])
def test_grid_bfs_equals_manhattan(w, h):
    ro, ci = grid_csr(w, h)
    rng = np.random.default_rng(w * 100 + h)
    src = rng.integers(0, w * h, size=9).astype(np.uint32)
    d = three_way(w * h, ro, ci, src)
    assert np.array_equal(d, orc.grid_manhattan(w, h, src))


def test_scipy_fixtures():
    z = np.load(os.path.join(HERE, "golden", "scipy_bfs.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    assert len(names) == 4
    for n in names:
        V = int(z[f"{n}__num_vertices"][0])
        d = three_way(V, z[f"{n}__row_offsets"], z[f"{n}__col_indices"], z[f"{n}__sources"])
        assert np.array_equal(d, z[f"{n}__dist"]), n


def test_path_and_cycle_closed_forms():
    n = 300
    ro, ci = orc.csr_from_edge_list(n, [(i, i + 1) for i in range(n - 1)] +
                                    [(i + 1, i) for i in range(n - 1)])
    src = [0, 17, n - 1]
    d = three_way(n, ro, ci, src).reshape(3, n)
    for k, s in enumerate(src):
        assert np.array_equal(d[k], np.abs(np.arange(n) - s))
    ro, ci = orc.csr_from_edge_list(n, [(i, (i + 1) % n) for i in range(n)] +
                                    [((i + 1) % n, i) for i in range(n)])
    d = three_way(n, ro, ci, [0]).reshape(n)
    k = np.arange(n)
    assert np.array_equal(d, np.minimum(k, n - k))


def _floyd(n, ro, ci):
    D = np.full((n, n), np.inf)
    np.fill_diagonal(D, 0)
    for u in range(n):
        for e in range(ro[u], ro[u + 1]):
            D[u, ci[e]] = min(D[u, ci[e]], 1)
    for k in range(n):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    return np.where(np.isinf(D), INF, D).astype(np.uint64).astype(np.uint32)


def test_brute_force_tiny_graphs():
    rng = np.random.default_rng(99)
    for trial in range(60):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(0, 3 * n + 1))
        edges = rng.integers(0, n, size=(m, 2))
        ro, ci = orc.csr_from_edge_list(n, edges)
        src = rng.integers(0, n, size=int(rng.integers(1, 2 * n + 1))).astype(np.uint32)
        d = three_way(n, ro, ci, src).reshape(len(src), n)
        F = _floyd(n, ro, ci)
        assert np.array_equal(d, F[src]), trial


def test_csr_builder_kats():
    k = kats()
    c = k["csr_construction"]
    ro, ci = orc.csr_from_edge_list(c["num_vertices"], c["edges"])
    assert ro.tolist() == c["row_offsets"] and ci.tolist() == c["col_indices"]
    c = k["csr_filter"]
    ro, ci = orc.csr_from_edge_list(c["num_vertices"], c["edges"])
    assert ro.tolist() == c["row_offsets"] and ci.tolist() == c["col_indices"]
    c = k["csc_transpose"]
    ro, ci = orc.csr_from_edge_list(c["num_vertices"], c["edges"])
    co, ri = orc.csr_to_csc(c["num_vertices"], ro, ci)
    assert co.tolist() == c["col_offsets"] and ri.tolist() == c["row_indices"]


def test_transpose_against_numpy():
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(1, 60))
        e = rng.integers(0, n, size=(int(rng.integers(0, 300)), 2))
        ro, ci = orc.csr_from_edge_list(n, e)
        co, ri = orc.csr_to_csc(n, ro, ci)
        us = np.repeat(np.arange(n), np.diff(ro.astype(np.int64)))
        order = np.lexsort((us, ci))  # by destination, then source
        assert np.array_equal(ri, us[order])
        assert np.array_equal(np.diff(co.astype(np.int64)), np.bincount(ci, minlength=n))


def test_traversed_edges_definition():
    ro, ci = grid_csr(8, 8)
    src = [0, 9, 63]
    d = orc.ms_bfs_fifo(64, ro, ci, src)
    assert orc.traversed_edges(64, ro, 3, d) == 3 * len(ci)
    k = kats()["disconnected_5"]
    d = orc.ms_bfs_fifo(5, k["row_offsets"], k["col_indices"], [0, 3])
    # source 0 reaches {0,1,2} (deg 2 each), source 3 reaches {3,4} (deg 1 each)
    assert orc.traversed_edges(5, k["row_offsets"], 2, d) == 6 + 2


def test_reachability_kat():
    k = kats()["reachability_bitmask"]
    nv, ns = k["num_vertices"], k["num_sources"]
    d = np.full(ns * nv, INF, dtype=np.uint32)
    for s, v, val in k["dist_entries"]:
        d[s * nv + v] = val
    w = orc.reachability_from_dist(nv, ns, d)
    assert len(w) == nv * k["words_per_vertex"]
    expect = np.zeros_like(w)
    for s, v in k["set"]:
        expect[v * 3 + s // 32] |= np.uint32(1 << (s % 32))
    assert np.array_equal(w, expect)
    assert int(np.unpackbits(w.view(np.uint8)).sum()) == k["total_pairs"]


def test_u8_u16_codec_kat():
    for n in kats()["u8_pack_lengths"]["lengths"]:
        i = np.arange(n)
        expect = np.where(i % 7 == 0, INF, i % 255).astype(np.uint32)
        assert np.array_equal(orc.unpack_u8(n, orc.pack_u8(expect)), expect)
        e16 = np.where(i % 7 == 0, INF, (i * 37) % 65535).astype(np.uint32)
        assert np.array_equal(orc.unpack_u16(n, orc.pack_u16(e16)), e16)
    # explicit layout: low half first, padding 0xFFFF / 0xFF
    assert orc.pack_u16([1, 2, 3]).tolist() == [0x00020001, 0xFFFF0003]
    assert orc.pack_u8([1, 2, 3, 4, 5]).tolist() == [0x04030201, 0xFFFFFF05]


def test_errors_and_empty():
    ro, ci = grid_csr(3, 3)
    with pytest.raises(orc.SourceOutOfBounds):
        orc.ms_bfs_fifo(9, ro, ci, [0, 9])
    assert orc.ms_bfs_fifo(9, ro, ci, []).size == 0
    with pytest.raises(orc.SourceOutOfBounds):
        orc.ms_bfs_parallel(0, [0], [], [0])


def test_planted_mistakes_are_caught():
    """A wrong distance, a dropped reachable vertex or a transposed row in
    an oracle output fails the closed-form / invariant checks."""
    ro, ci = grid_csr(12, 9)
    src = np.array([3, 50, 107], np.uint32)
    good = orc.ms_bfs_fifo(108, ro, ci, src)
    man = orc.grid_manhattan(12, 9, src)
    assert np.array_equal(good, man)
    bad = good.copy()
    bad[5] += 1
    assert not np.array_equal(bad, man)
    bad = good.copy()
    bad[108 + 7] = INF
    assert not np.array_equal(bad, man)
    bad = good.reshape(3, 108)[[1, 0, 2]].reshape(-1)
    assert not np.array_equal(bad, man)
