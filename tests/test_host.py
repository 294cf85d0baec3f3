"""CPU-side tests of libmsbfs: the C ABI loads and exports every symbol the
header declares; the host graph library (CSR build, transpose, symmetry,
validation, generators, source sampling) matches the oracle and the
reference's known answers; the Auto decision table matches the reference's
tests.  No kernel runs here."""
import json
import os
import re

import numpy as np
import pytest

import gpu_bfs as gb
from gpu_bfs import _lib
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def kats():
    with open(os.path.join(HERE, "golden", "reference_kats.json")) as f:
        return json.load(f)


def header_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(msbfs_\w+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = gb.lib()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the same set
    assert set(_lib.SIGNATURES) <= set(names)


def test_version_and_status_strings():
    lib = gb.lib()
    assert lib.msbfs_version().decode().count(".") == 2
    assert lib.msbfs_status_string(7).decode() == "distance does not fit u16"


def test_csr_from_edge_list_matches_oracle_random():
    rng = np.random.default_rng(1)
    for _ in range(30):
        n = int(rng.integers(1, 80))
        e = rng.integers(0, n + 3, size=(int(rng.integers(0, 400)), 2))  # some out of range
        g = gb.CsrGraph.from_edge_list(n, e)
        ro, ci = orc.csr_from_edge_list(n, e)
        assert np.array_equal(g.row_offsets, ro) and np.array_equal(g.col_indices, ci)
        g.validate()


def test_csr_kats():
    k = kats()
    c = k["csr_construction"]
    g = gb.CsrGraph.from_edge_list(c["num_vertices"], c["edges"])
    assert g.row_offsets.tolist() == c["row_offsets"]
    assert g.col_indices.tolist() == c["col_indices"]
    assert g.max_degree() == c["max_degree"]
    c = k["csc_transpose"]
    g = gb.CsrGraph.from_edge_list(c["num_vertices"], c["edges"])
    csc = g.to_csc()
    assert csc.col_offsets.tolist() == c["col_offsets"]
    assert csc.row_indices.tolist() == c["row_indices"]
    c = k["symmetry"]
    d = gb.CsrGraph.from_edge_list(3, c["directed_edges"])
    s = gb.CsrGraph.from_edge_list(3, c["symmetric_edges"])
    assert not d.is_symmetric() and s.is_symmetric()
    u = d.to_undirected()
    assert u.row_offsets.tolist() == c["undirected_row_offsets"]
    assert u.col_indices.tolist() == c["undirected_col_indices"]


def test_transpose_matches_oracle():
    g = gb.generate_rmat(10, 8, directed=True)
    assert not g.is_symmetric()
    csc = g.to_csc()
    co, ri = orc.csr_to_csc(g.num_vertices, g.row_offsets, g.col_indices)
    assert np.array_equal(csc.col_offsets, co) and np.array_equal(csc.row_indices, ri)


def test_validate_rejects_bad_graphs():
    bad = [
        gb.CsrGraph(3, 2, [1, 1, 2, 2], [1, 2]),     # nonzero start
        gb.CsrGraph(3, 2, [0, 2, 1, 2], [1, 2]),     # non-monotonic
        gb.CsrGraph(3, 2, [0, 1, 2, 2], [1, 3]),     # column out of range
        gb.CsrGraph(3, 3, [0, 1, 2, 2], [1, 2, 0]),  # edge count mismatch
    ]
    for g in bad:
        with pytest.raises(gb.MsbfsError):
            g.validate()


def test_grid_generator_matches_reference_construction():
    k = kats()["grid_4x3_csr"]
    g = gb.generate_grid2d(*k["grid"])
    assert (g.num_vertices, g.num_edges) == (k["num_vertices"], k["num_edges"])
    g.validate()
    from test_oracle import grid_csr

    for w, h in [(4, 3), (5, 5), (1, 7), (256, 256)]:
        g = gb.generate_grid2d(w, h)
        ro, ci = grid_csr(w, h) if w * h < 5000 else (None, None)
        if ro is not None:
            assert np.array_equal(g.row_offsets, ro) and np.array_equal(g.col_indices, ci)
        assert g.num_edges == 2 * ((w - 1) * h + w * (h - 1))
    g = gb.generate_grid2d(256, 256)
    assert (g.num_vertices, g.num_edges) == (65536, 261120)
    with pytest.raises(gb.MsbfsError):
        gb.generate_grid2d(0, 4)


def test_road_like_generator_shape():
    g = gb.generate_road_like()
    assert (g.num_vertices, g.num_edges) == (114_599, 239_332)
    assert g.is_symmetric()
    g.validate()
    # every edge joins grid neighbours of the 339-column layout
    us = np.repeat(np.arange(g.num_vertices), g.degrees())
    vs = g.col_indices.astype(np.int64)
    diff = np.abs(vs - us)
    assert np.all((diff == 339) | ((diff == 1) & (np.minimum(us, vs) % 339 != 338)))
    assert g.max_degree() <= 4
    # connected: one BFS reaches everything
    d = orc.ms_bfs_fifo(g.num_vertices, g.row_offsets, g.col_indices, [0])
    assert d.max() < 0xFFFFFFFF
    # deterministic
    assert gb.generate_road_like() == g


def test_rmat_generator_properties():
    g = gb.generate_rmat(12, 16)
    assert g.num_vertices == 4096 and g.is_symmetric()
    g.validate()
    assert g.num_edges > 0 and g.num_edges % 2 == 0
    us = np.repeat(np.arange(g.num_vertices), g.degrees())
    assert np.all(us != g.col_indices)  # no self loops
    assert gb.generate_rmat(12, 16) == g
    assert gb.generate_rmat(12, 16, seed=1) != g
    with pytest.raises(gb.MsbfsError):
        gb.generate_rmat(8, 4, 0.5, 0.5, 0.5, 0.5)


def test_erdos_renyi_generator_properties():
    n = 4000
    g = gb.generate_erdos_renyi_np(n, 8 / (n - 1), seed=3)
    assert g.is_symmetric()
    g.validate()
    assert abs(g.avg_degree() - 8.0) < 0.5
    assert gb.generate_erdos_renyi_np(1, 0.5).num_edges == 0


def test_sample_sources():
    g = gb.generate_rmat(10, 4)
    s = gb.sample_sources(g, 200, seed=9, non_isolated=True)
    assert len(np.unique(s)) == 200 and np.all(np.diff(s.astype(np.int64)) > 0)
    assert np.all(g.degrees()[s] > 0)
    with pytest.raises(gb.MsbfsError):
        gb.sample_sources(g, g.num_vertices + 1)


def test_traversed_edges_matches_oracle():
    g = gb.generate_rmat(9, 4)
    s = gb.sample_sources(g, 40, non_isolated=True)
    d = orc.ms_bfs_fifo(g.num_vertices, g.row_offsets, g.col_indices, s)
    assert gb.traversed_edges(g, 40, d) == orc.traversed_edges(g.num_vertices, g.row_offsets, 40, d)


def _star(deg):
    # make_graph_with_max_degree (crates/gpu_bfs/src/engine.rs:496-500)
    return gb.CsrGraph.from_edge_list(deg + 2, [(0, v) for v in range(1, deg + 1)])


def test_auto_router_decision_table():
    k = kats()["auto_router"]
    M = gb.GpuBfsMethod
    sel = gb.GpuBfsEngine.select_auto_method
    assert sel(gb.CsrGraph.from_edge_list(0, []), 64) == M.NaiveSequential
    assert sel(_star(5), 0) == M.NaiveSequential
    for deg in k["low_degree"]:
        g = _star(deg)
        assert g.max_degree() == deg
        for s in k["low_degree_sources"]:
            assert sel(g, s) == M.WorkgroupPersistentBfs
    for deg in k["intermediate_degree"]:
        g = _star(deg)
        assert sel(g, 1) == M.SparseQueue
        assert sel(g, 16) == M.BitParallel32
        assert sel(g, 32) == M.BitParallel64 and sel(g, 63) == M.BitParallel64
        for s in (64, 512, 2048):
            assert sel(g, s) == M.MultiWordBitParallel
    for deg in k["hub_degree"]:
        g = _star(deg)
        for s in k["hub_sources"]:
            assert sel(g, s) == M.IndirectSparseQueuePush


def test_product_path_has_no_oracle_or_fallback():
    """The shipped package never imports the oracle and has no CPU BFS."""
    pkg = os.path.join(ROOT, "webgpu-bfs_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower() or f.endswith(".md"), f


def test_context_without_device_raises():
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "import gpu_bfs as gb\n"
            "try:\n    gb.GpuContext(0)\n    print('device')\n"
            "except gb.NoDeviceError:\n    print('nodevice')\n") % os.path.join(ROOT, "webgpu-bfs_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() in ("device", "nodevice"), out.stderr


def test_lattice_offsets():
    """The deep family keeps a lattice graph's adjacency in LDS: a graph
    qualifies when it is symmetric and every edge joins ids that differ by
    one of at most four offsets.  Grids have offsets {1, width} by
    construction, the 8-neighbour grid {1, w-1, w, w+1}, a path {1}; five
    offsets, a directed graph, an edgeless graph and R-MAT do not qualify."""
    assert gb.generate_grid2d(5, 7).lattice_offsets() == [1, 5]
    assert gb.generate_grid2d(256, 256).lattice_offsets() == [1, 256]
    assert gb.generate_road_like().lattice_offsets() == [1, 339]
    assert gb.CsrGraph.from_edge_list(10, [(i, i + 1) for i in range(9)]).to_undirected() \
        .lattice_offsets() == [1]
    w, h = 9, 6
    king = []
    for y in range(h):
        for x in range(w):
            for dx, dy in ((1, 0), (-1, 1), (0, 1), (1, 1)):
                if 0 <= x + dx < w and y + dy < h:
                    king.append((y * w + x, (y + dy) * w + x + dx))
    kg = gb.CsrGraph.from_edge_list(w * h, king).to_undirected()
    assert kg.lattice_offsets() == [1, w - 1, w, w + 1]
    five = [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5)]
    assert gb.CsrGraph.from_edge_list(6, five).to_undirected().lattice_offsets() == []
    assert gb.CsrGraph.from_edge_list(6, five[:4]).to_undirected().lattice_offsets() == [1, 2, 3, 4]
    assert gb.CsrGraph.from_edge_list(3, [(0, 1), (1, 2)]).lattice_offsets() == []  # directed
    assert gb.CsrGraph.from_edge_list(4, []).lattice_offsets() == []
    assert gb.generate_rmat(10, 8).lattice_offsets() == []


def test_tuning_validation_rejects_bad_fields():
    """Every set_tuning entry point (engine, batch, server lane) shares
    msbfs_tuning_validate; a workgroup narrower than one wave would make the
    deep kernels' per-wave strides zero, so it must never reach a launch."""
    import ctypes

    lib = gb.lib()

    def rc(**kw):
        t = _lib.Tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        return lib.msbfs_tuning_validate(ctypes.byref(t))

    assert rc() == _lib.OK
    assert rc(deep_threads=64, deep_bits=1, tile_words=16, words_per_lane=2) == _lib.OK
    assert rc(batch_union=2, union_words=4, deep_lattice=1, deep_blocks_per_cu=1) == _lib.OK
    for bad in (dict(deep_threads=32), dict(deep_threads=1), dict(deep_threads=96),
                dict(deep_threads=2048), dict(deep_bits=3), dict(deep_bits=64),
                dict(tile_words=3), dict(tile_words=128), dict(words_per_lane=3),
                dict(direction_rule=2), dict(wide_blocks_per_cu=17), dict(deep_rows=3),
                dict(deep_lattice=3), dict(batch_union=3), dict(union_words=3),
                dict(union_words=128), dict(deep_blocks_per_cu=33)):
        assert rc(**bad) == _lib.ERR_INVALID_ARG, bad
        assert _lib.last_error()
    assert lib.msbfs_tuning_validate(None) == _lib.ERR_INVALID_ARG


def test_shutdown_order_and_idempotence():
    """gpu_bfs.shutdown() (also run by atexit, registered once the library is
    loaded) destroys live objects servers-first, contexts-last, each once."""
    calls = []

    class Obj:
        pass

    objs = []
    for kind in ("context", "graph", "engine", "batch", "server", "graph"):
        o = Obj()
        objs.append(o)
        _lib.track(kind, o, calls.append, kind)
    _lib.shutdown()
    assert calls == ["server", "batch", "engine", "graph", "graph", "context"]
    _lib.shutdown()
    del objs
    assert len(calls) == 6
    import atexit  # noqa: F401  (registered by lib())
    assert _lib._shutdown_registered
