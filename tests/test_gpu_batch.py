"""GPU parity of the batched multi-graph launch (msbfs_batch_*): every graph
of a batch must get exactly the oracle's distances, eagerly, through hipGraph
replays (several, with other device work in between), after a source update,
after autotuning, and in u16 -- in every batch mode: union traversals (the
graphs joined into disjoint unions by source words, traversed bit-parallel
by the wide family; batch_union=1), one persistent deep-family launch over
per-source tasks (batch_union=2), and the split between them (batch_union=0:
graphs of at least union_words source words in unions, the default)."""
import numpy as np
import pytest

import gpu_bfs as gb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def mixed_graphs(n=24, seed=1):
    rng = np.random.default_rng(seed)
    graphs, sources = [], []
    for g in range(n):
        nv = int(rng.integers(50, 4000))
        graphs.append(gb.generate_erdos_renyi_np(nv, 6 / (nv - 1), seed=seed * 100 + g))
        k = int(rng.integers(0, 90)) if g % 7 else 0  # some graphs get no sources
        sources.append(rng.integers(0, nv, size=k).astype(np.uint32))
    graphs.append(gb.generate_grid2d(40, 30))
    sources.append(np.array([0, 1199, 600, 600], np.uint32))
    graphs.append(gb.generate_rmat(10, 6, directed=True))
    sources.append(rng.integers(0, 1024, size=70).astype(np.uint32))
    return graphs, sources


def expect_all(graphs, sources):
    return [orc.ms_bfs_parallel(g.num_vertices, g.row_offsets, g.col_indices, s, 8)
            .reshape(len(s), g.num_vertices) if len(s) else np.zeros((0, g.num_vertices), np.uint32)
            for g, s in zip(graphs, sources)]


def check(batch, out, expect, u16=False):
    host = out.cpu().numpy().view(np.uint16 if u16 else np.uint32)
    for g, (got, exp) in enumerate(zip(batch.split(host), expect)):
        if u16:
            exp = np.where(exp == 0xFFFFFFFF, 0xFFFF, exp).astype(np.uint16)
        assert np.array_equal(got, exp), f"graph {g}"


# every msbfs_tuning field a candidate of the autotuner may set
TUNING_FIELDS = ("alpha", "beta", "tile_words", "words_per_lane", "deep_bits", "deep_threads",
                 "wide_blocks_per_cu", "direction_rule", "stepped", "deep_rows", "deep_lattice",
                 "batch_union", "union_words", "deep_blocks_per_cu")
MODES = {"union": dict(batch_union=1), "deep": dict(batch_union=2),
         "split": dict(batch_union=0, union_words=2)}


def mode_of(b):
    t = b.get_tuning()
    return {1: "union", 2: "deep"}.get(t["batch_union"], "split")


@pytest.fixture(scope="module", params=list(MODES))
def setup(ctx, request):
    import torch

    graphs, sources = mixed_graphs()
    eng = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod.Auto)
    b = gb.Batch(eng, graphs, sources)
    b.set_tuning(**MODES[request.param])
    out = torch.empty(max(b.total, 1), dtype=torch.int32, device="cuda")
    return graphs, sources, eng, b, out


def test_batch_eager(setup):
    graphs, sources, eng, b, out = setup
    mode = mode_of(b)
    say("batch eager", mode, len(graphs), "graphs", b.total, "outputs")
    out.fill_(-11)
    b.run(out.data_ptr(), b.total)
    check(b, out, expect_all(graphs, sources))
    st = b.stats()
    D, W = int(gb.Family.DEEP), int(gb.Family.WIDE)
    fams = b.families()
    assert st["instances"] > 0
    if mode == "union":
        # one union per source-word class: 1 (up to 64 sources) and 2 (65-128)
        assert st["family"] == W and fams == [W if len(s) else D for s in sources]
        assert st["instances"] == 2 and st["bits_per_instance"] == 128
    elif mode == "deep":
        assert st["family"] == D and fams == [D] * len(graphs)
    else:
        # graphs of more than 64 sources in a union, the rest in the deep launch
        assert st["family"] == D
        assert fams == [W if len(s) > 64 else D for s in sources]
        assert W in fams and D in fams


def test_batch_graph_replays(setup):
    import torch

    graphs, sources, eng, b, out = setup
    expect = expect_all(graphs, sources)
    b.capture(out.data_ptr(), b.total)
    for r in range(3):
        out.fill_(-7)  # eager device work between replays
        torch.cuda.synchronize()
        say("batch replay", r)
        b.replay()
        check(b, out, expect)


def test_batch_set_sources_then_replay(setup):
    graphs, sources, eng, b, out = setup
    rng = np.random.default_rng(5)
    new = [rng.integers(0, g.num_vertices, size=len(s)).astype(np.uint32)
           for g, s in zip(graphs, sources)]
    b.set_sources(new)
    b.capture(out.data_ptr(), b.total)
    b.replay()
    check(b, out, expect_all(graphs, new))
    b.set_sources(sources)
    b.replay()
    check(b, out, expect_all(graphs, sources))


def test_batch_autotune_keeps_results(setup):
    """The autotuner's sweep covers task widths, workgroup sizes from one
    wave to 1024 threads and staged / direct rows; the winner and every
    candidate that ran give the oracle's distances."""
    import torch

    graphs, sources, eng, b, out = setup
    mode = mode_of(b)
    best, ms = b.autotune(out.data_ptr(), b.total)
    say("autotune", best, ms)
    assert best["batch_union"] in (0, 1, 2) and ms > 0
    if best["batch_union"] == 2:
        assert best["deep_bits"] in (1, 2, 4, 8, 16, 32)
    b.run(out.data_ptr(), b.total)
    expect = expect_all(graphs, sources)
    check(b, out, expect)
    sweep = b.sweep()
    ran = [c for c in sweep if c["status"] == 0]
    assert ran and all(c["ms"] > 0 for c in ran)
    assert {c["deep_threads"] for c in ran if c["batch_union"] == 2} >= {64, 128, 256, 512, 1024}
    assert {c["deep_rows"] for c in ran if c["deep_bits"] == 1} == {1, 2}
    assert any(c["batch_union"] == 1 for c in ran)
    assert any(c["batch_union"] == 0 and c["union_words"] == 2 for c in ran)
    assert min(c["ms"] for c in ran) == pytest.approx(ms, rel=1e-5)
    for c in ran:
        say("candidate", {k: c[k] for k in TUNING_FIELDS if c[k]})
        b.set_tuning(**{k: c[k] for k in TUNING_FIELDS})
        out.fill_(-3)
        torch.cuda.synchronize()  # the batch runs on its own stream
        b.run(out.data_ptr(), b.total)
        check(b, out, expect)
    b.set_tuning(**MODES[mode])


def test_batch_wide_fallback(ctx):
    """A graph whose visited bits exceed one workgroup's LDS (1.4M vertices)
    is traversed by the wide family inside the batch: eager runs, hipGraph
    replays (the wide run launched eagerly after the captured launch), a
    source update and the autotuner's wide knobs all give the oracle's
    distances."""
    import torch

    rng = np.random.default_rng(12)
    big = gb.generate_erdos_renyi_np(1_400_000, 3.0 / 1_399_999, seed=9)
    graphs = [gb.generate_erdos_renyi_np(2000, 6 / 1999, seed=3), big, gb.generate_grid2d(40, 30),
              gb.generate_erdos_renyi_np(500, 4 / 499, seed=4),
              gb.generate_erdos_renyi_np(1_300_000, 2.0 / 1_299_999, seed=10)]
    sources = [rng.integers(0, 2000, 50).astype(np.uint32),
               rng.integers(0, big.num_vertices, 70).astype(np.uint32),
               np.array([0, 1199, 5], np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)]
    eng = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod.Auto)
    b = gb.Batch(eng, graphs, sources)
    b.set_tuning(batch_union=2)
    D, W = int(gb.Family.DEEP), int(gb.Family.WIDE)
    assert b.families() == [D, W, D, D, W]
    out = torch.empty(b.total, dtype=torch.int32, device="cuda")
    expect = expect_all(graphs, sources)
    say("wide fallback eager", b.total)
    out.fill_(-9)
    torch.cuda.synchronize()  # the batch runs on its own stream
    b.run(out.data_ptr(), b.total)
    check(b, out, expect)
    b.capture(out.data_ptr(), b.total)
    for r in range(2):
        out.fill_(-7)
        torch.cuda.synchronize()
        say("wide fallback replay", r)
        b.replay()
        check(b, out, expect)
    new = [rng.integers(0, g.num_vertices, size=len(s)).astype(np.uint32)
           for g, s in zip(graphs, sources)]
    b.set_sources(new)
    b.replay()
    check(b, out, expect_all(graphs, new))
    b.set_sources(sources)
    best, ms = b.autotune(out.data_ptr(), b.total)
    say("wide fallback autotune", best, ms)
    sweep = b.sweep()
    for knob in ("tile_words", "words_per_lane", "alpha", "beta", "stepped"):
        assert any(c[knob] != 0 for c in sweep), knob
    out.fill_(-5)
    torch.cuda.synchronize()
    b.run(out.data_ptr(), b.total)
    check(b, out, expect)
    # the same graphs split (the 1.4M-vertex graph joins a union whatever its
    # source count; the others stay in the deep launch), then all in unions
    b.set_tuning(batch_union=0)
    assert b.families() == [D, W, D, D, W]
    out.fill_(-8)
    torch.cuda.synchronize()
    say("split with big graphs")
    b.run(out.data_ptr(), b.total)
    check(b, out, expect)
    b.set_tuning(batch_union=1)
    assert b.families() == [W, W, W, D, W]  # (graph 3 has no sources)
    for r in range(2):
        out.fill_(-4)
        torch.cuda.synchronize()
        say("union with big graphs, run", r)
        b.run(out.data_ptr(), b.total)
        check(b, out, expect)
    b.capture(out.data_ptr(), b.total)
    out.fill_(-6)
    torch.cuda.synchronize()
    say("union with big graphs, replay")
    b.replay()
    check(b, out, expect)


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_u16_and_errors(ctx, mode):
    import torch

    graphs, sources = mixed_graphs(8, seed=3)
    eng = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod.Auto)
    b = gb.Batch(eng, graphs, sources, dist_dtype=np.uint16)
    b.set_tuning(**MODES[mode])
    out = torch.empty(max(b.total, 1), dtype=torch.int16, device="cuda")
    b.run(out.data_ptr(), b.total)
    check(b, out, expect_all(graphs, sources), u16=True)
    bad = [s.copy() for s in sources]
    bad[2] = np.append(bad[2], np.uint32(graphs[2].num_vertices))
    with pytest.raises(gb.SourceOutOfBoundsError) as ei:
        gb.Batch(eng, graphs, bad)
    assert "graph 2" in str(ei.value)
    with pytest.raises(gb.MsbfsError) as ei:
        b.run(out.data_ptr(), b.total - 1)
    assert ei.value.code == 4


def test_batch_rejects_sub_wave_workgroups(setup):
    """A workgroup narrower than one wave is rejected on the host before any
    launch (the deep kernels stride by whole waves); the batch keeps its
    previous plan and still runs correctly."""
    graphs, sources, eng, b, out = setup
    for bad in (32, 1, 96):
        with pytest.raises(gb.MsbfsError) as ei:
            b.set_tuning(deep_threads=bad, batch_union=2)
        assert ei.value.code == 1, bad
    b.run(out.data_ptr(), b.total)
    check(b, out, expect_all(graphs, sources))
