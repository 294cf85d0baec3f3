"""GPU parity: every traversal family of libmsbfs, called through the C ABI,
must return the oracle's distances bit for bit.

Small cases compare whole outputs with the oracle (oracle/oracle.c, itself
pinned by tests/test_oracle.py) or with the golden fixtures; the full-size
configurations compare every row the oracle can produce in seconds and check
size-independent properties (source rows are 0 at the source, adjacent
vertices differ by at most one level on undirected graphs, traversed-edge
totals agree) on the rest.
"""
import json
import os

import numpy as np
import pytest

import gpu_bfs as gb
from gpu_bfs import workloads as wl
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INF32 = 0xFFFFFFFF
M = gb.GpuBfsMethod
METHODS = [M.Auto, M.WorkgroupPersistentBfs, M.PushPullAdaptive, M.BitParallel64,
           M.MultiWordBitParallel]
THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)


def kats():
    with open(os.path.join(HERE, "golden", "reference_kats.json")) as f:
        return json.load(f)


def oracle(g, src):
    return orc.ms_bfs_parallel(g.num_vertices, g.row_offsets, g.col_indices, src, THREADS)


def to_u16_expect(d32):
    return np.where(d32 == INF32, 0xFFFF, d32).astype(np.uint16)


@pytest.fixture(scope="module")
def engines(ctx):
    return {m: gb.GpuBfsEngine(ctx, m) for m in METHODS}


def kat_graph(k):
    if "grid" in k:
        return gb.generate_grid2d(*k["grid"])
    return gb.CsrGraph(k["num_vertices"], len(k["col_indices"]), k["row_offsets"],
                       k["col_indices"])


@pytest.mark.parametrize("name", ["disconnected_5", "single_vertex", "zero_edge_4", "line_5",
                                  "grid_4x4_sources_0_5"])
def test_reference_kats(engines, name):
    k = kats()[name]
    g = kat_graph(k)
    for m, e in engines.items():
        say(name, m.name)
        got = e.run_ms_bfs(g, k["sources"])
        assert got.tolist() == k["expected"], (name, m.name)


def test_scipy_fixtures(engines):
    z = np.load(os.path.join(HERE, "golden", "scipy_bfs.npz"))
    for n in sorted({k.split("__")[0] for k in z.files}):
        V = int(z[f"{n}__num_vertices"][0])
        ro, ci = z[f"{n}__row_offsets"], z[f"{n}__col_indices"]
        g = gb.CsrGraph(V, len(ci), ro, ci)
        for m, e in engines.items():
            say("scipy", n, m.name)
            assert np.array_equal(e.run_ms_bfs(g, z[f"{n}__sources"]), z[f"{n}__dist"]), (n, m)


def test_grid256_s64_manhattan(engines):
    """Grid2D 256x256, S=64: bit-exact against the Manhattan closed form."""
    g = wl.grid256()
    src = wl.sources(g, 64)
    expect = orc.grid_manhattan(256, 256, src)
    for m, e in engines.items():
        say("grid256 S=64", m.name)
        assert np.array_equal(e.run_ms_bfs(g, src), expect), m.name


@pytest.fixture(scope="module")
def road():
    return wl.road()


@pytest.mark.parametrize("S", [1, 256, 2048])
def test_road_like_full(engines, road, S):
    src = wl.sources(road, S)
    expect = oracle(road, src)
    ms = [M.Auto, M.WorkgroupPersistentBfs] if S == 2048 else [M.Auto, M.PushPullAdaptive]
    for m in ms:
        say("road S=%d" % S, m.name)
        got = engines[m].run_ms_bfs(road, src)
        assert np.array_equal(got, expect), (S, m.name, int(np.count_nonzero(got != expect)))


def test_road_like_deep_tile_variants(ctx, road):
    """Every task width whose visited bits fit in LDS, at three workgroup
    sizes (64 threads: the one-wave kernel; neighbour tables of width 4 and
    8); widths that do not fit fall back to the wide family."""
    small = gb.generate_road_like(20_000, 160, 900, seed=3)
    sparse = gb.generate_erdos_renyi_np(20_000, 1.6 / 19_999, seed=21)
    assert 4 < sparse.max_degree() <= 8  # a width-8 neighbour table
    for g, widths in ((road, (1, 2, 4, 8)), (small, (16, 32)), (sparse, (1, 4, 32))):
        src = wl.sources(g, 300)
        expect = oracle(g, src)
        for bits in widths:
            for threads in (64, 256, 1024):
                e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
                e.set_tuning(deep_bits=bits, deep_threads=threads)
                say("road V=%d S=300 deep bits" % g.num_vertices, bits, "threads", threads)
                assert np.array_equal(e.run_ms_bfs(g, src), expect), (bits, threads)
                assert e.last_stats()["bits_per_instance"] == bits
    e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
    e.set_tuning(deep_bits=16)  # 16 bits x 114,599 vertices exceed LDS
    src = wl.sources(road, 64)
    assert np.array_equal(e.run_ms_bfs(road, src), oracle(road, src))
    assert e.last_stats()["family"] == int(gb.Family.WIDE)
    # capped persistent launches (deep_blocks_per_cu): fewer workgroups take
    # the same tasks from the queue, lattice kernel and workgroup kernel
    src = wl.sources(road, 700)
    expect = oracle(road, src)
    cus = ctx.info()["compute_units"]
    for lattice in (0, 2):
        for cap in (1, 2):
            e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
            e.set_tuning(deep_bits=1, deep_lattice=lattice, deep_blocks_per_cu=cap)
            say("road S=700 deep_blocks_per_cu", cap, "lattice", lattice)
            assert np.array_equal(e.run_ms_bfs(road, src), expect), (cap, lattice)
            assert e.last_stats()["grid_blocks"] <= cus * cap, (cap, e.last_stats())


def test_deep_one_wave_routing(ctx, road):
    """Deep traversals take a one-task-per-wave kernel when |V| / levels <=
    64 (road-like: ~37 vertices per level) and keep the workgroup kernel
    otherwise (Grid 256x256: ~128).  Both graphs are lattices (ids joined by
    offsets 1 and the row length) small enough for LDS, whose level count a
    host BFS from vertex 0 gives when the graph is uploaded, so the first
    traversal is routed like the later ones (which record the level count
    they see); the one-task-per-wave kernel is the lattice kernel (adjacency
    in LDS), and with the lattice turned off the one-wave kernel.  Every run
    matches the oracle."""
    for g, wave in ((road, True), (gb.generate_grid2d(256, 256), False)):
        src = wl.sources(g, 512)
        expect = oracle(g, src)
        dg = gb.DeviceGraph(ctx, g)
        e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        for run in range(3):
            say("one-wave routing V=%d run" % g.num_vertices, run)
            assert np.array_equal(e.run_ms_bfs(dg, src), expect), (g.num_vertices, run)
            st = e.last_stats()
            assert st["bits_per_instance"] == 1
            assert (st["lattice_bits"] == 2) == wave, (g.num_vertices, run, st)
            assert st["lattice_bits"] or st["block_threads"] != 64
        e.set_tuning(deep_lattice=2)
        say("one-wave routing V=%d, lattice off" % g.num_vertices)
        assert np.array_equal(e.run_ms_bfs(dg, src), expect), g.num_vertices
        st = e.last_stats()
        assert st["lattice_bits"] == 0 and (st["block_threads"] == 64) == wave, (g.num_vertices, st)


@pytest.fixture(scope="module")
def rmat20():
    return wl.rmat(20)


def _undirected_invariants(g, src, d, inf):
    S, V = len(src), g.num_vertices
    d = d.reshape(S, V)
    assert np.all(d[np.arange(S), src] == 0)
    us = np.repeat(np.arange(V), g.degrees())
    vs = g.col_indices
    for r in range(S):
        a, b = d[r, us].astype(np.int64), d[r, vs].astype(np.int64)
        fin = (a != inf) & (b != inf)
        assert np.all(np.abs(a[fin] - b[fin]) <= 1)
        assert np.all((a == inf) == (b == inf))


def test_rmat20_s64_u16(engines, rmat20):
    src = wl.sources(rmat20, 64, non_isolated=True)
    expect = to_u16_expect(oracle(rmat20, src))
    for m in (M.Auto, M.PushPullAdaptive, M.MultiWordBitParallel, M.BitParallel64):
        say("rmat20 S=64 u16", m.name)
        got = engines[m].run_ms_bfs(rmat20, src, dist_dtype=np.uint16)
        assert np.array_equal(got, expect), m.name


def test_rmat20_s2048_u16(engines, rmat20):
    """Every path at S=2048 on R-MAT scale 20; all rows are compared with the
    oracle in test_gpu_configs.py, here the first and last 64 rows of each
    method plus the undirected-distance invariants."""
    g = rmat20
    src = wl.sources(g, 2048, non_isolated=True)
    rows = np.concatenate([np.arange(64), np.arange(2048 - 64, 2048)])
    expect = to_u16_expect(oracle(g, src[rows])).reshape(len(rows), -1)
    for m in (M.Auto, M.MultiWordBitParallel):
        say("rmat20 S=2048 u16", m.name)
        got = engines[m].run_ms_bfs(g, src, dist_dtype=np.uint16).reshape(2048, -1)
        assert np.array_equal(got[rows], expect), m.name
        assert np.all(got[np.arange(2048), src] == 0)
    _undirected_invariants(g, src[rows[:8]], got[rows[:8]].reshape(-1), 0xFFFF)


def test_batch_sizes_word_boundaries(engines):
    g = gb.generate_erdos_renyi_np(3000, 8 / 2999, seed=11)
    for S in kats()["batch_sizes"]["sizes"]:
        src = gb.sample_sources(g, min(S, 3000), seed=S)
        expect = oracle(g, src)
        for m in (M.Auto, M.WorkgroupPersistentBfs, M.PushPullAdaptive):
            say("ER3000 batch S=%d" % S, m.name)
            assert np.array_equal(engines[m].run_ms_bfs(g, src), expect), (S, m.name)


def test_duplicate_sources(engines):
    g = gb.generate_grid2d(5, 5)
    src = [7, 7, 12, 7]
    expect = oracle(g, src)
    for m, e in engines.items():
        assert np.array_equal(e.run_ms_bfs(g, src), expect)


def test_directed_graph_uses_true_transpose(engines):
    g = gb.generate_rmat(12, 8, directed=True)
    assert not g.is_symmetric()
    src = gb.sample_sources(g, 100, seed=4)
    expect = oracle(g, src)
    for m, e in engines.items():
        say("directed rmat12", m.name)
        assert np.array_equal(e.run_ms_bfs(g, src), expect), m.name


def test_reachability_matches_oracle(engines):
    g = gb.generate_erdos_renyi_np(2000, 1.5 / 1999, seed=5)  # many components
    src = gb.sample_sources(g, 150, seed=2)
    d = oracle(g, src)
    expect = orc.reachability_from_dist(g.num_vertices, 150, d)
    for m in (M.Auto, M.WorkgroupPersistentBfs, M.PushPullAdaptive):
        say("reachability", m.name)
        bm = engines[m].run_ms_reachability(g, src)
        assert np.array_equal(bm.words, expect), m.name


def test_u16_quantisation_boundaries(engines):
    for n in kats()["quantisation_boundary"]["path_lengths"]:
        g = gb.csr_from_undirected_pairs(n, [(i, i + 1) for i in range(n - 1)])
        expect = to_u16_expect(oracle(g, [0, n - 1]))
        for m in (M.Auto, M.PushPullAdaptive):
            assert np.array_equal(engines[m].run_ms_bfs(g, [0, n - 1], dist_dtype=np.uint16),
                                  expect)


def test_u16_overflow_is_an_error(engines):
    n = 70_000
    g = gb.csr_from_undirected_pairs(n, [(i, i + 1) for i in range(n - 1)])
    for m in (M.Auto, M.PushPullAdaptive):
        say("u16 overflow", m.name)
        with pytest.raises(gb.MsbfsError) as ei:
            engines[m].run_ms_bfs(g, [0], dist_dtype=np.uint16)
        assert ei.value.code == 7
        d = engines[m].run_ms_bfs(g, [0, 69_999])
        assert np.array_equal(d.reshape(2, n)[0], np.arange(n))
        assert np.array_equal(d.reshape(2, n)[1], n - 1 - np.arange(n))


def test_argument_errors(engines, ctx):
    g = gb.generate_grid2d(3, 3)
    e = engines[M.Auto]
    with pytest.raises(gb.SourceOutOfBoundsError) as ei:
        e.run_ms_bfs(g, [0, 9])
    assert "Source vertex 9 at index 1 exceeds graph vertex count 9" in str(ei.value)
    assert e.run_ms_bfs(g, []).size == 0
    empty = gb.CsrGraph(0, 0, [0], [])
    with pytest.raises(gb.SourceOutOfBoundsError):
        e.run_ms_bfs(empty, [0])
    assert e.run_ms_bfs(empty, []).size == 0
    import torch

    out = torch.empty(5, dtype=torch.int32, device="cuda")
    with pytest.raises(gb.MsbfsError) as ei:
        e.run_ms_bfs_device(g, [0], out.data_ptr(), 5)
    assert ei.value.code == 4
    # the engine is still usable after errors
    assert np.array_equal(e.run_ms_bfs(g, [4]), oracle(g, [4]))


def test_device_output_on_torch_stream(engines, road):
    import torch

    src = wl.sources(road, 128)
    expect = oracle(road, src)
    out = torch.empty(128 * road.num_vertices, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    for m in (M.Auto, M.PushPullAdaptive):
        out.fill_(7)
        torch.cuda.synchronize()  # the run goes on stream s
        engines[m].run_ms_bfs_device(road, src, out.data_ptr(), out.numel(),
                                     stream=s.cuda_stream, sync=False)
        engines[m].check()
        s.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint32), expect), m.name


def test_repeated_runs_identical(engines):
    g = gb.generate_rmat(14, 16)
    src = gb.sample_sources(g, 500, seed=8, non_isolated=True)
    for m in (M.Auto, M.PushPullAdaptive, M.WorkgroupPersistentBfs):
        first = engines[m].run_ms_bfs(g, src)
        for _ in range(3):
            assert np.array_equal(engines[m].run_ms_bfs(g, src), first), m.name
        assert np.array_equal(first, oracle(g, src))


def test_interleaved_graphs_same_size_do_not_alias(engines):
    """Two different graphs of identical |V| and |E| (the reference's pools
    are keyed on sizes only) give their own answers."""
    a = gb.generate_road_like(20_000, 200, 500, seed=1)
    b = gb.generate_road_like(20_000, 200, 500, seed=2)
    assert (a.num_vertices, a.num_edges) == (b.num_vertices, b.num_edges) and a != b
    src = [0, 1234, 19_999]
    ea, eb = oracle(a, src), oracle(b, src)
    for m in (M.Auto, M.PushPullAdaptive):
        for _ in range(2):
            assert np.array_equal(engines[m].run_ms_bfs(a, src), ea)
            assert np.array_equal(engines[m].run_ms_bfs(b, src), eb)


def test_matrix_market_file_end_to_end(engines, tmp_path):
    """A graph read from a Matrix Market file (the reference CLI's `.mtx`
    path, crates/core/src/mtx.rs) runs bit-exact on both families."""
    g = gb.generate_road_like(8_000, 90, 500, seed=3)
    p = tmp_path / "road.mtx"
    gb.write_mtx(g, p, symmetric=True)
    h = gb.read_mtx(p)
    src = gb.sample_sources(h, 130, seed=5)
    expect = oracle(h, src)
    for m in (M.Auto, M.WorkgroupPersistentBfs, M.PushPullAdaptive):
        say("mtx road", m.name)
        assert np.array_equal(engines[m].run_ms_bfs(h, src), expect), m.name


def to_u8_expect(d32):
    return np.where(d32 == INF32, 0xFF, d32).astype(np.uint8)


def test_u8_output_and_boundaries(engines):
    """u8 distances (pack_distances_u32_to_u8.wgsl:18-53: u8 iff every level
    < 255): paths around the boundary, and graphs with unreachable pairs on
    every family."""
    for n in kats()["quantisation_boundary"]["path_lengths"]:
        g = gb.csr_from_undirected_pairs(n, [(i, i + 1) for i in range(n - 1)])
        expect = oracle(g, [0, n - 1])
        for m in (M.Auto, M.PushPullAdaptive):
            say("u8 path", n, m.name)
            if n - 1 <= 254:
                got = engines[m].run_ms_bfs(g, [0, n - 1], dist_dtype=np.uint8)
                assert np.array_equal(got, to_u8_expect(expect)), (n, m.name)
            else:
                with pytest.raises(gb.MsbfsError) as ei:
                    engines[m].run_ms_bfs(g, [0, n - 1], dist_dtype=np.uint8)
                assert ei.value.code == 13
    g = gb.generate_erdos_renyi_np(3000, 1.5 / 2999, seed=9)  # many components
    src = gb.sample_sources(g, 150, seed=3)
    expect = to_u8_expect(oracle(g, src))
    for m in (M.Auto, M.WorkgroupPersistentBfs, M.PushPullAdaptive):
        say("u8 ER3000", m.name)
        assert np.array_equal(engines[m].run_ms_bfs(g, src, dist_dtype=np.uint8), expect), m.name


def test_compact_host_transfer(engines, ctx):
    """Host outputs cross PCIe in the narrowest width that holds every
    distance below its all-ones sentinel: 4 bits when there are at most 15
    levels (R-MAT), 8 up to 255, 12 up to 4,095 (road-like), else u16, and
    u32 after a rerun when a distance reaches 65,535; every path widens to
    the caller's type exactly, at each width boundary and with ragged last
    groups."""
    import torch
    g = gb.generate_rmat(16, 16)
    src = gb.sample_sources(g, 600, seed=6, non_isolated=True)
    expect = oracle(g, src)
    e = engines[M.Auto]
    say("compact transfer rmat16 S=600")
    assert np.array_equal(e.run_ms_bfs(g, src), expect)
    assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint16), to_u16_expect(expect))
    assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint8), to_u8_expect(expect))
    out = torch.empty(len(src) * g.num_vertices, dtype=torch.uint8, device="cuda")
    e.run_ms_bfs_device(g, src, out.data_ptr(), out.numel(), dist_dtype=np.uint8)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), to_u8_expect(expect))
    road = gb.generate_road_like(20_000, 160, 900, seed=4)  # eccentricities > 255: 12-bit wire
    rs = gb.sample_sources(road, 300, seed=8)
    say("compact transfer road S=300")
    rexp = oracle(road, rs)
    assert np.array_equal(e.run_ms_bfs(road, rs), rexp)
    assert np.array_equal(e.run_ms_bfs(road, rs, dist_dtype=np.uint16), to_u16_expect(rexp))
    # wire width boundaries: a path of n vertices from vertex 0 has n levels
    # (4 bits up to 15, 8 up to 255, 12 up to 4,095, then u16); odd sizes
    # leave a ragged last group of 8
    for n in (2, 15, 16, 255, 256, 4095, 4096, 5001):
        p = gb.CsrGraph.from_edge_list(n, [(i, i + 1) for i in range(n - 1)]).to_undirected()
        ps = np.array([0, n - 1, n // 2, 0, 1 % n], dtype=np.uint32)[: 3 if n % 2 else 5]
        pexp = oracle(p, ps)
        for m in (M.WorkgroupPersistentBfs, M.PushPullAdaptive):
            say("compact transfer path", n, m.name)
            assert np.array_equal(engines[m].run_ms_bfs(p, ps), pexp), (n, m.name)
            assert np.array_equal(engines[m].run_ms_bfs(p, ps, dist_dtype=np.uint16),
                                  to_u16_expect(pexp)), (n, m.name)
            assert engines[m].last_stats()["levels"] >= n, (n, engines[m].last_stats())


def test_engine_autotune_keeps_results(ctx):
    """Measured routing: whichever family and knobs win, the distances are the
    oracle's; an explicit method stays itself."""
    road = gb.generate_road_like(20_000, 160, 900, seed=4)
    rmat = gb.generate_rmat(14, 16)
    for g, S in ((road, 300), (rmat, 200)):
        src = gb.sample_sources(g, S, seed=1, non_isolated=True)
        expect = oracle(g, src)
        e = gb.GpuBfsEngine(ctx, M.Auto)
        r = e.autotune(g, src)
        say("autotune", g.num_vertices, r["method"].name, round(r["ms"], 3), r["tuning"])
        assert r["ms"] > 0
        assert r["method"] in (M.WorkgroupPersistentBfs, M.PushPullAdaptive)
        assert e.method() == r["method"]
        assert np.array_equal(e.run_ms_bfs(g, src), expect)
    e = gb.GpuBfsEngine(ctx, M.BitParallel64Pull)
    src = gb.sample_sources(rmat, 64, seed=2, non_isolated=True)
    r = e.autotune(rmat, src, dist_dtype=np.uint16)
    assert r["method"] == M.BitParallel64Pull
    assert np.array_equal(e.run_ms_bfs(rmat, src, dist_dtype=np.uint16),
                          to_u16_expect(oracle(rmat, src)))


def hub_and_path(n_leaves=3000, path=60):
    """A hub of degree n_leaves (a heavy vertex on both lists) with a path of
    `path` vertices hanging from one leaf: the wide family passes its deferred
    levels and stores later distances at once (block pull, heavy resolve with
    lane = vertex), and light vertices sit beside the heavy one."""
    n = 1 + n_leaves + path
    pairs = [(0, i) for i in range(1, n_leaves + 1)]
    pairs += [(1, n_leaves + 1)] + [(n_leaves + 1 + i, n_leaves + 2 + i) for i in range(path - 1)]
    return gb.csr_from_undirected_pairs(n, pairs)


def test_random_graphs_every_path(ctx):
    """Seeded sweep over the wide family's code paths and the deep family:
    hub-heavy R-MAT (undirected and directed: true transpose, heavy chunks in
    pull and push, heavy resolve), a hub with a long path (more levels than
    the final pass defers, so both pull layouts and both heavy resolves run),
    G(n, p) with many components; S at word boundaries; push-only,
    pull-only, adaptive and deep methods; default and narrow wave tiles
    (several word groups per vertex); u32, u16 and u8 outputs and the
    reachability words.  Bar: bit-exact against the oracle."""
    graphs = {
        "rmat13": gb.generate_rmat(13, 16, seed=11),
        "rmat12d": gb.generate_rmat(12, 16, seed=12, directed=True),
        "hubpath": hub_and_path(),
        "er6000": gb.generate_erdos_renyi_np(6000, 3.0 / 5999, seed=13),
    }
    cases = [(M.Auto, {}), (M.PushPullAdaptive, {}), (M.BitParallel64Pull, {}),
             (M.BitParallel64, {}), (M.PushPullAdaptive, {"stepped": 2}),
             (M.BitParallel64Pull, {"stepped": 2}), (M.BitParallel64, {"stepped": 2}),
             (M.WorkgroupPersistentBfs, {}),
             (M.WorkgroupPersistentBfs, {"deep_threads": 64}),
             (M.PushPullAdaptive, {"tile_words": 4}), (M.PushPullAdaptive, {"tile_words": 8})]
    for gname, g in graphs.items():
        assert g.max_degree() > 256 or gname == "er6000", gname
        for S in (1, 64, 65, 700):
            src = gb.sample_sources(g, S, seed=S + 1)
            d32 = oracle(g, src)
            for m, tune in cases:
                e = gb.GpuBfsEngine(ctx, m)
                if tune:
                    e.set_tuning(**tune)
                say("sweep", gname, S, m.name, tune)
                assert np.array_equal(e.run_ms_bfs(g, src), d32), (gname, S, m.name, tune)
            e = gb.GpuBfsEngine(ctx, M.PushPullAdaptive)
            say("sweep u16/u8/reach", gname, S)
            assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint16), to_u16_expect(d32))
            if int(d32[d32 != INF32].max(initial=0)) < 255:
                assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint8), to_u8_expect(d32))
            bm = e.run_ms_reachability(g, src)
            assert np.array_equal(bm.words, orc.reachability_from_dist(g.num_vertices, S, d32))


def test_one_engine_across_layouts(ctx):
    """One wide-family engine reused while the workspace layout changes
    (graph size, source words, deferred levels, reachability-only runs):
    frontier slots are not cleared between runs, so each run must write
    every row it later reads.  Bar: bit-exact against the oracle."""
    rmat = gb.generate_rmat(13, 16, seed=11)
    hub = hub_and_path()
    er = gb.generate_erdos_renyi_np(6000, 3.0 / 5999, seed=13)
    e = gb.GpuBfsEngine(ctx, M.PushPullAdaptive)
    runs = [(rmat, 2048, "u32"), (rmat, 64, "u16"), (hub, 700, "u32"), (rmat, 2048, "reach"),
            (er, 65, "u32"), (rmat, 2048, "u16"), (hub, 1, "u32"), (rmat, 2048, "u32"),
            (er, 700, "reach"), (er, 700, "u32")]
    for i, (g, S, kind) in enumerate(runs):
        src = gb.sample_sources(g, S, seed=100 + i)
        d32 = oracle(g, src)
        say("layouts", i, g.num_vertices, S, kind)
        if kind == "reach":
            bm = e.run_ms_reachability(g, src)
            assert np.array_equal(bm.words, orc.reachability_from_dist(g.num_vertices, S, d32)), i
        elif kind == "u16":
            got = e.run_ms_bfs(g, src, dist_dtype=np.uint16)
            assert np.array_equal(got, to_u16_expect(d32)), i
        else:
            assert np.array_equal(e.run_ms_bfs(g, src), d32), i
        assert e.last_stats()["family"] == int(gb.Family.WIDE), i


def test_two_words_per_lane_pull(ctx):
    """Source word counts where deferred pull levels give each lane two
    adjacent words by default (even, >= 16: 16, 18 with idle lanes, 64),
    against one word per lane (asked for, or implied by an explicit tile
    width) and two words on a wider explicit tile (several word groups at
    S=4096); hub-heavy R-MAT (heavy in-chunks) and the hub with a path
    (direct levels after the deferred ones).  Bar: bit-exact against the
    oracle."""
    graphs = {"rmat13": gb.generate_rmat(13, 16, seed=11), "hubpath": hub_and_path()}
    for gname, g in graphs.items():
        for S in (1024, 1090, 4096):
            # repeated sources where S exceeds |V| (independent rows)
            src = np.resize(gb.sample_sources(g, min(S, g.num_vertices), seed=S + 7), S)
            d32 = oracle(g, src)
            for tune in ({}, {"tile_words": 16}, {"words_per_lane": 1},
                         {"tile_words": 8, "words_per_lane": 2}):
                e = gb.GpuBfsEngine(ctx, M.PushPullAdaptive)
                if tune:
                    e.set_tuning(**tune)
                say("pair lanes", gname, S, tune)
                assert np.array_equal(e.run_ms_bfs(g, src), d32), (gname, S, tune)
            e = gb.GpuBfsEngine(ctx, M.BitParallel64Pull)
            assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint16), to_u16_expect(d32))


def test_deep_reachability_rows(ctx):
    """Deep-family reachability: each task writes its sources' rows of a
    source-major bitmap, transposed into the vertex-major words afterwards;
    both deep kernels (64 threads: the one-wave kernel), several task
    widths, S not a multiple of 32 and |V| not a multiple of 64.  Bar:
    bit-exact against the oracle's reachability."""
    g = gb.generate_road_like(20_000, 160, 900, seed=3)
    assert g.num_vertices % 64 != 0
    for S in (70, 300):
        src = gb.sample_sources(g, S, seed=S)
        expect = orc.reachability_from_dist(g.num_vertices, S, oracle(g, src))
        for bits, threads in ((1, 64), (1, 256), (4, 64), (8, 1024)):
            e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
            e.set_tuning(deep_bits=bits, deep_threads=threads)
            say("deep reach", S, bits, threads)
            bm = e.run_ms_reachability(g, src)
            assert e.last_stats()["family"] == int(gb.Family.DEEP)
            assert np.array_equal(bm.words, expect), (S, bits, threads)


def _device_u32(e, g, src, fill=0x55555555):
    import torch

    n = len(src) * g.num_vertices
    out = torch.full((max(n, 1),), fill - (1 << 32) if fill >= 1 << 31 else fill,
                     dtype=torch.int32, device="cuda")
    e.run_ms_bfs_device(g, src, out.data_ptr(), n)
    return out[:n].cpu().numpy().view(np.uint32)


def test_logged_u32_rows(ctx, road):
    """u32 device output of the one-wave kernel goes through per-bin logs
    kept in the output rows themselves (4,096-vertex bins, rewritten through
    LDS when a task ends).  Checked against the oracle on every row for:
    the road-like graph at S=2048 (28 bins, the last one partial), a sparse
    graph with many unreachable vertices (sentinels), |V| below, at and just
    above one bin, a single vertex, and duplicated sources; the output
    buffer starts with garbage each time."""
    sparse = gb.generate_erdos_renyi_np(20_000, 1.6 / 19_999, seed=21)
    rng = np.random.default_rng(4)
    n = 1100  # the log is used with four or more tasks per CU
    cases = [(road, wl.sources(road, 2048)),
             (sparse, wl.sources(sparse, n)),
             (gb.generate_grid2d(64, 64), rng.integers(0, 4096, n).astype(np.uint32)),
             (gb.generate_grid2d(17, 241),
              np.concatenate([[0, 4096, 4096, 2000, 5], rng.integers(0, 4097, n)]).astype(np.uint32)),
             (gb.generate_grid2d(9, 13), rng.integers(0, 117, n).astype(np.uint32)),
             (gb.CsrGraph.from_edge_list(1, []), np.zeros(n, np.uint32))]
    for g, src in cases:
        expect = oracle(g, src)
        e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        e.set_tuning(deep_bits=1, deep_threads=64)
        for fill in (0x55555555, 0xFFFFFFFF, 0):
            say("logged u32 V=%d S=%d fill %x" % (g.num_vertices, len(src), fill))
            got = _device_u32(e, g, src, fill)
            assert np.array_equal(got, expect), (g.num_vertices, len(src), fill,
                                                 int(np.count_nonzero(got != expect)))
            st = e.last_stats()
            assert st["block_threads"] == 64 and st["bits_per_instance"] == 1
            assert st["instances"] >= 4 * ctx.info()["compute_units"]


def _king_grid(w, h):
    edges = []
    for y in range(h):
        for x in range(w):
            for dx, dy in ((1, 0), (-1, 1), (0, 1), (1, 1)):
                if 0 <= x + dx < w and y + dy < h:
                    edges.append((y * w + x, (y + dy) * w + x + dx))
    return gb.CsrGraph.from_edge_list(w * h, edges).to_undirected()


def test_lattice_kernel_matches_oracle(ctx, road):
    """The lattice kernel (one source per wave, the graph's adjacency in LDS
    as 2 or 4 bits per vertex shared by the workgroup's waves) against the
    oracle on every row: road-like graphs (offsets 1 and the row length,
    missing grid edges), a 4-neighbour grid whose row ends have no +1 edge,
    an 8-neighbour grid (four offsets: 4 bits per vertex), a path (one
    offset), a three-offset triangular lattice; u32 stored directly (the
    default) and through the binned log (deep_rows=1, four or more tasks per
    CU), u16, u8, reachability, host output, duplicated sources, output
    buffers starting with garbage, one to eight waves per workgroup, and a
    second run on the same engine."""
    rng = np.random.default_rng(11)
    small = gb.generate_road_like(20_000, 160, 900, seed=3)
    path = gb.CsrGraph.from_edge_list(3000, [(i, i + 1) for i in range(2999)]).to_undirected()
    w = 37
    tri = gb.CsrGraph.from_edge_list(
        w * 40, [(u, u + d) for u in range(w * 40) for d in (1, w, w + 1)
                 if u + d < w * 40 and (d == w or u % w != w - 1)]).to_undirected()
    cases = [(road, [1, 339], 2048), (small, None, 1100), (small, None, 300),
             (gb.generate_grid2d(17, 241), [1, 17], 1100), (_king_grid(45, 52), [1, 44, 45, 46], 600),
             (path, [1], 70), (tri, [1, w, w + 1], 1500), (gb.generate_grid2d(64, 64), [1, 64], 5)]
    for g, offs, S in cases:
        if offs is not None:
            assert g.lattice_offsets() == offs
        lat = 2 if len(g.lattice_offsets()) <= 2 else 4
        src = rng.integers(0, g.num_vertices, S).astype(np.uint32)
        src[: min(S, 3)] = src[0]  # duplicates
        expect = oracle(g, src)
        dg = gb.DeviceGraph(ctx, g)
        e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        for rows, rep in ((0, 0), (0, 1), (1, 0), (1, 1)):
            e.set_tuning(deep_bits=1, deep_lattice=1, deep_rows=rows)
            say("lattice V=%d S=%d u32 device rows %d rep %d" % (g.num_vertices, S, rows, rep))
            got = _device_u32(e, dg, src, 0x12345678 if rep else 0xFFFFFFFF)
            assert np.array_equal(got, expect), (g.num_vertices, S, rows, rep,
                                                 int(np.count_nonzero(got != expect)))
            st = e.last_stats()
            assert st["lattice_bits"] == lat and st["bits_per_instance"] == 1, st
            assert st["block_threads"] % 64 == 0 and 64 <= st["block_threads"] <= 512, st
        e.set_tuning(deep_rows=0)
        say("lattice V=%d S=%d host u32 / u16 / reachability" % (g.num_vertices, S))
        assert np.array_equal(e.run_ms_bfs(dg, src), expect)
        assert np.array_equal(e.run_ms_bfs(dg, src, np.uint16), to_u16_expect(expect))
        bm = e.run_ms_reachability(dg, src)
        assert np.array_equal(bm.words, orc.reachability_from_dist(g.num_vertices, S, expect))
        if expect[expect != INF32].max(initial=0) < 255:
            say("lattice V=%d S=%d u8" % (g.num_vertices, S))
            got8 = e.run_ms_bfs(dg, src, np.uint8)
            assert np.array_equal(got8, np.where(expect == INF32, 0xFF, expect).astype(np.uint8))


def test_staged_rows_match_oracle(ctx):
    """One source per task on the workgroup kernel keeps each row in LDS as
    4-bit level codes (distance + 1 below 14, 15 for distances stored as
    found) and writes it whole when the task ends.  Every row against the
    oracle, staged and direct, u32 and u16, output buffers starting with
    garbage, for: a sparse G(n,p) graph with unreachable vertices, a path
    whose distances run far past the codes (escapes from level 14 on), a grid
    with a neighbour table, |V| of 1, 7, 8 and 9 (partial code words), and
    duplicated sources."""
    import torch

    rng = np.random.default_rng(8)
    path = gb.CsrGraph.from_edge_list(45, [(i, i + 1) for i in range(44)])
    cases = [(gb.generate_erdos_renyi_np(3000, 2.2 / 2999, seed=5), None),
             (path, np.array([0, 44, 22, 13, 14, 15, 0], np.uint32)),
             (gb.generate_grid2d(30, 31), None),
             (gb.generate_rmat(12, 8, directed=True), None)]
    for nv in (1, 7, 8, 9):
        cases.append((gb.CsrGraph.from_edge_list(nv, [(i, (i * 3 + 1) % nv) for i in range(nv - 1)]),
                      None))
    for g, src in cases:
        if src is None:
            src = rng.integers(0, g.num_vertices, 300).astype(np.uint32)
        expect = oracle(g, src)
        n = len(src) * g.num_vertices
        for rows in (1, 2):
            for threads in (256, 1024):
                e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
                e.set_tuning(deep_bits=1, deep_threads=threads, deep_rows=rows)
                say("staged rows=%d threads=%d V=%d S=%d" % (rows, threads, g.num_vertices, len(src)))
                got = _device_u32(e, g, src, 0x33333333)
                assert np.array_equal(got, expect), (g.num_vertices, rows, threads,
                                                     int(np.count_nonzero(got != expect)))
                st = e.last_stats()
                assert st["block_threads"] == threads and st["bits_per_instance"] == 1
                out16 = torch.full((max(n, 1),), 0x1234, dtype=torch.int16, device="cuda")
                e.run_ms_bfs_device(g, src, out16.data_ptr(), n, dist_dtype=np.uint16)
                got16 = out16[:n].cpu().numpy().view(np.uint16)
                assert np.array_equal(got16, to_u16_expect(expect)), (g.num_vertices, rows, "u16")


def test_every_method_enum_matches_oracle(ctx):
    """All fifteen GpuBfsMethod values (crates/gpu_bfs/src/engine.rs:33-54,
    dispatched at :330-404) run and return the oracle's distances on a
    hub-heavy directed R-MAT graph and a small grid, at source counts on both
    sides of a 64-bit word; explicit deep tunings check NaiveSequential's
    one-source tasks."""
    graphs = [gb.generate_rmat(11, 8, directed=True), gb.generate_grid2d(23, 19)]
    for g in graphs:
        for S in (1, 37, 130):
            src = gb.sample_sources(g, S, seed=S)
            expect = oracle(g, src)
            for m in M:
                say("method", m.name, "V=%d S=%d" % (g.num_vertices, S))
                e = gb.GpuBfsEngine(ctx, m)
                got = e.run_ms_bfs(g, src)
                assert np.array_equal(got, expect), (m.name, g.num_vertices, S)
                st = e.last_stats()
                if m == M.NaiveSequential and st["family"] == int(gb.Family.DEEP):
                    assert st["bits_per_instance"] == 1
                if m == M.BatchedDenseSpMM:
                    assert st["family"] == int(gb.Family.WIDE) and st["push_levels"] == 0


def test_stepped_wide_matches_oracle(ctx):
    """The stepped wide launch (one launch per phase of each level, no grid
    barrier; the serving loop's mode) against the oracle on the sweep's
    graphs, S at word boundaries, push-only, pull-only and adaptive, u32
    (device and host), u16 and reachability.  The hub-and-path graph needs
    more levels than the first batch of launches enqueues on a new graph, so
    the continuation runs; later runs on that graph enqueue enough at once.
    The results also match the cooperative kernel's."""
    import torch

    graphs = {
        "rmat13": gb.generate_rmat(13, 16, seed=11),
        "rmat12d": gb.generate_rmat(12, 16, seed=12, directed=True),
        "hubpath": hub_and_path(),
        "er6000": gb.generate_erdos_renyi_np(6000, 3.0 / 5999, seed=13),
    }
    for gname, g in graphs.items():
        dg = gb.DeviceGraph(ctx, g)
        for S in (1, 64, 130, 700):
            src = gb.sample_sources(g, S, seed=S + 3)
            d32 = oracle(g, src)
            for m in (M.PushPullAdaptive, M.BitParallel64Pull, M.BitParallel64):
                e = gb.GpuBfsEngine(ctx, m)
                e.set_tuning(stepped=1)
                say("stepped", gname, S, m.name)
                out = torch.full((S * g.num_vertices,), 9, dtype=torch.int32, device="cuda")
                e.run_ms_bfs_device(dg, src, out.data_ptr(), out.numel())
                assert np.array_equal(out.cpu().numpy().view(np.uint32), d32), (gname, S, m.name)
                st = e.last_stats()
                assert st["family"] == int(gb.Family.WIDE)
                assert st["levels"] == int(d32[d32 != INF32].max(initial=0)) + 1, st
                assert np.array_equal(e.run_ms_bfs(dg, src), d32)
            e = gb.GpuBfsEngine(ctx, M.PushPullAdaptive)
            e.set_tuning(stepped=1)
            assert np.array_equal(e.run_ms_bfs(dg, src, dist_dtype=np.uint16), to_u16_expect(d32))
            bm = e.run_ms_reachability(dg, src)
            assert np.array_equal(bm.words, orc.reachability_from_dist(g.num_vertices, S, d32))


def _brute_force(n, adj, src):
    """Hop distances by repeated frontier expansion over the adjacency
    matrix (independent of the oracle): row i for source src[i]."""
    out = np.full((len(src), n), INF32, dtype=np.uint32)
    for i, s in enumerate(src):
        seen = np.zeros(n, dtype=bool)
        front = np.zeros(n, dtype=bool)
        front[s] = seen[s] = True
        d = 0
        while front.any():
            out[i, front] = d
            nxt = adj[front].any(axis=0) & ~seen
            seen |= nxt
            front = nxt
            d += 1
    return out.reshape(-1)


def _tiny_graphs():
    """Every undirected simple graph on 1-5 vertices and every directed one
    on 1-3 (loops excluded), as (name, n, adjacency matrix)."""
    out = []
    for n in range(1, 6):
        pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
        for mask in range(1 << len(pairs)):
            adj = np.zeros((n, n), dtype=bool)
            for b, (i, j) in enumerate(pairs):
                if (mask >> b) & 1:
                    adj[i, j] = adj[j, i] = True
            out.append((f"u{n}:{mask}", n, adj))
    for n in range(1, 4):
        arcs = [(i, j) for i in range(n) for j in range(n) if i != j]
        for mask in range(1 << len(arcs)):
            adj = np.zeros((n, n), dtype=bool)
            for b, (i, j) in enumerate(arcs):
                if (mask >> b) & 1:
                    adj[i, j] = True
            out.append((f"d{n}:{mask}", n, adj))
    return out


def test_every_tiny_graph(ctx):
    """Brute force over all small graphs: every undirected graph on up to 5
    vertices (1,099 graphs) and every directed graph on up to 3 (69), each
    with every vertex as a source plus a duplicate of vertex 0, against
    distances from repeated frontier expansion of the adjacency matrix.
    Graphs of up to 4 vertices run every method enum; all of them run the
    deep family, the adaptive wide family, the pull-only path, u16, u8 and
    the reachability words, and together as one batch in its three forms
    (union traversals, one deep launch, split).  Bar: bit-exact."""
    import torch

    graphs, sources, expect = [], [], []
    for name, n, adj in _tiny_graphs():
        us, vs = np.nonzero(adj)
        g = gb.CsrGraph.from_edge_list(n, list(zip(us.tolist(), vs.tolist())))
        src = np.array(list(range(n)) + [0], dtype=np.uint32)
        graphs.append((name, n, g))
        sources.append(src)
        expect.append(_brute_force(n, adj, src))
    engines = {m: gb.GpuBfsEngine(ctx, m) for m in M}
    core = [M.WorkgroupPersistentBfs, M.PushPullAdaptive, M.BitParallel64Pull]
    say("tiny graphs:", len(graphs))
    for (name, n, g), src, d32 in zip(graphs, sources, expect):
        for m in (list(M) if n <= 4 else core):
            got = engines[m].run_ms_bfs(g, src)
            assert np.array_equal(got, d32), (name, m.name, got, d32)
        e = engines[M.Auto]
        assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint16), to_u16_expect(d32)), name
        assert np.array_equal(e.run_ms_bfs(g, src, dist_dtype=np.uint8), to_u8_expect(d32)), name
        bm = e.run_ms_reachability(g, src)
        assert np.array_equal(bm.words, orc.reachability_from_dist(n, len(src), d32)), name
    eng = gb.GpuBfsEngine(ctx, M.Auto)
    for mode in (dict(batch_union=1), dict(batch_union=2), dict(batch_union=0, union_words=1)):
        say("tiny graphs batch", mode)
        b = gb.Batch(eng, [g for _, _, g in graphs], sources)
        b.set_tuning(**mode)
        out = torch.full((b.total,), -7, dtype=torch.int32, device="cuda")
        b.run(out.data_ptr(), b.total)
        host = out.cpu().numpy().view(np.uint32)
        for (name, n, _), got, d32 in zip(graphs, b.split(host), expect):
            assert np.array_equal(got.reshape(-1), d32), (mode, name)


def test_random_lattices(ctx):
    """Seeded sweep of random lattices through the lattice kernel: 1-4
    random offsets (up to 700), every vertex pair at an offset joined with a
    random probability (missing edges, several components), |V| from 2 to
    60,000, 1-700 sources with duplicates, direct and logged rows, u16 and
    reachability.  Bar: bit-exact against the oracle."""
    rng = np.random.default_rng(2027)
    lat_runs = 0
    for case in range(36):
        V = int(rng.integers(2, 60_000))
        k = int(rng.integers(1, 5))
        offs = sorted(set(int(x) for x in rng.integers(1, min(700, V - 1) + 1, k)))
        keep = float(rng.choice([0.5, 0.8, 0.97, 1.0]))
        us, vs = [], []
        for d in offs:
            u = np.arange(V - d, dtype=np.int64)
            u = u[rng.random(len(u)) < keep]
            us.append(u)
            vs.append(u + d)
        us, vs = np.concatenate(us), np.concatenate(vs)
        if len(us) == 0:
            continue
        g = gb.CsrGraph.from_edge_list(V, list(zip(us.tolist(), vs.tolist()))).to_undirected()
        found = g.lattice_offsets()
        assert set(found) <= set(offs) and found, (V, offs, found)
        S = int(rng.integers(1, 701))
        src = rng.integers(0, V, S).astype(np.uint32)
        src[rng.random(S) < 0.05] = src[0]  # duplicates
        expect = oracle(g, src)
        dg = gb.DeviceGraph(ctx, g)
        e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        rows = case % 2
        e.set_tuning(deep_bits=1, deep_lattice=1, deep_rows=rows)
        say("random lattice", case, "V=%d offsets=%s keep=%.2f S=%d rows=%d" % (V, found, keep, S, rows))
        got = e.run_ms_bfs(dg, src)
        st = e.last_stats()
        # (where the adjacency or a task's region does not fit LDS, the
        # one-wave or workgroup kernel runs instead, still checked)
        lat_runs += st["lattice_bits"] == (2 if len(found) <= 2 else 4)
        assert np.array_equal(got, expect), (case, V, found, S, rows,
                                             int(np.count_nonzero(got != expect)))
        if case % 3 == 0:
            assert np.array_equal(e.run_ms_bfs(dg, src, dist_dtype=np.uint16), to_u16_expect(expect)), case
            bm = e.run_ms_reachability(dg, src)
            assert np.array_equal(bm.words, orc.reachability_from_dist(V, S, expect)), case
    assert lat_runs >= 30, lat_runs
    # near the LDS limit: a 600 x 600 grid (360,000 vertices, 90 KB of
    # adjacency and 45 KB of visited bits in one workgroup)
    g = gb.generate_grid2d(600, 600)
    src = rng.integers(0, g.num_vertices, 40).astype(np.uint32)
    e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
    e.set_tuning(deep_bits=1, deep_lattice=1)
    say("random lattices: grid 600 x 600 S=40")
    assert np.array_equal(e.run_ms_bfs(g, src), oracle(g, src))
    assert e.last_stats()["lattice_bits"] == 2, e.last_stats()


def test_concurrent_callers(ctx):
    """The reference's engine is Send + Sync and concurrent calls serialise
    on its pools (engine.rs:56-63).  Here: six host threads on one context
    and two resident graphs, four with engines of their own (deep, adaptive
    wide, pull-only, Auto; host and device outputs) and two sharing one
    engine, each running a mix of calls; every result is the oracle's.
    ctypes releases the GIL during the calls, so they overlap on the host."""
    import threading

    import torch

    road = gb.generate_road_like(20_000, 160, 900, seed=4)
    rmat = gb.generate_rmat(13, 16, seed=11)
    dgs = {"road": gb.DeviceGraph(ctx, road), "rmat": gb.DeviceGraph(ctx, rmat)}
    graphs = {"road": road, "rmat": rmat}
    cases = []
    for i in range(6):
        for name in ("road", "rmat"):
            src = gb.sample_sources(graphs[name], 40 + 37 * i, seed=100 + i, non_isolated=True)
            cases.append((name, src, oracle(graphs[name], src)))
    shared = gb.GpuBfsEngine(ctx, M.Auto)
    errors = []

    def worker(k, eng, device_out):
        try:
            stream = torch.cuda.Stream()
            for j in range(3):
                for name, src, exp in cases[k::2] if k % 2 else cases[::3]:
                    dg = dgs[name]
                    if device_out:
                        out = torch.empty(len(src) * dg.num_vertices, dtype=torch.int32, device="cuda")
                        eng.run_ms_bfs_device(dg, src, out.data_ptr(), out.numel(),
                                              stream=stream.cuda_stream)
                        stream.synchronize()
                        got = out.cpu().numpy().view(np.uint32)
                    else:
                        got = eng.run_ms_bfs(dg, src)
                    if not np.array_equal(got, exp):
                        errors.append((k, name, len(src), j))
        except Exception as ex:  # reported below, in the main thread
            errors.append((k, repr(ex)))

    plan = [(0, gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs), False),
            (1, gb.GpuBfsEngine(ctx, M.PushPullAdaptive), True),
            (2, gb.GpuBfsEngine(ctx, M.BitParallel64Pull), False),
            (3, gb.GpuBfsEngine(ctx, M.Auto), True),
            (4, shared, False), (5, shared, True)]
    say("concurrent callers:", len(plan), "threads")
    ths = [threading.Thread(target=worker, args=p) for p in plan]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ths), "a caller did not finish"
    assert not errors, errors


def test_sources_change_between_calls(ctx, road):
    """Sources change between calls on one engine: the same array changed
    in place, a different array of the same length, the first one again, a
    different size, and a device-output call in between each give their own
    sources' rows."""
    import torch

    e = gb.GpuBfsEngine(ctx, M.Auto)
    a = wl.sources(road, 64)
    b = (a + 7) % road.num_vertices
    exp_a, exp_b = oracle(road, a), oracle(road, b)
    buf = a.copy()
    say("sources cache: same array, then changed in place")
    assert np.array_equal(e.run_ms_bfs(road, buf), exp_a)
    assert np.array_equal(e.run_ms_bfs(road, buf), exp_a)
    buf[:] = b
    assert np.array_equal(e.run_ms_bfs(road, buf), exp_b)
    assert np.array_equal(e.run_ms_bfs(road, a), exp_a)
    c = wl.sources(road, 65)
    assert np.array_equal(e.run_ms_bfs(road, c), oracle(road, c))
    out = torch.empty(64 * road.num_vertices, dtype=torch.int32, device="cuda")
    e.run_ms_bfs_device(road, b, out.data_ptr(), out.numel())
    assert np.array_equal(out.cpu().numpy().view(np.uint32), exp_b)
    assert np.array_equal(e.run_ms_bfs(road, a), exp_a)
