"""GPU tests of the serving loop: every query's distances are the oracle's,
whether it ran alone, packed with others (continuous batching, concurrent
submission from another thread), or one query per traversal; admission by
output memory makes queries wait, not fail, when the pool is full."""
import threading

import numpy as np
import pytest

import gpu_bfs as gb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def graphs():
    road = gb.generate_road_like(20_000, 160, 900, seed=4)
    rmat = gb.generate_rmat(14, 16)
    return [road, rmat]


def make_queries(graphs, n, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(n):
        lane = int(rng.integers(0, 2))
        g = graphs[lane]
        k = int(np.exp(rng.uniform(0, np.log(300))))
        pool = np.nonzero(g.degrees() > 0)[0]
        qs.append((lane, rng.choice(pool, size=k).astype(np.uint32)))
    return qs


def expect(graphs, lane, src):
    g = graphs[lane]
    d = orc.ms_bfs_parallel(g.num_vertices, g.row_offsets, g.col_indices, src, 8)
    return np.where(d == 0xFFFFFFFF, 0xFFFF, d).astype(np.uint16)


def start_server(ctx, graphs, **kw):
    s = gb.Server(ctx, dist_dtype=np.uint16, **kw)
    for g in graphs:
        s.add_graph(g)
    s.start()
    return s


def test_mix_equals_alone_equals_oracle(ctx, graphs):
    qs = make_queries(graphs, 40, seed=1)
    want = [expect(graphs, lane, src) for lane, src in qs]
    s = start_server(ctx, graphs, pool_bytes=1 << 30)
    ids = [None] * len(qs)

    def client():
        for i, (lane, src) in enumerate(qs):
            ids[i] = s.submit(lane, src)

    t = threading.Thread(target=client)
    t.start()
    t.join()
    # consume in completion order (the completion queue), release at once
    got = {}
    for _ in range(len(qs)):
        qid = s.next_completed(60.0)
        assert qid not in got
        got[qid] = s.result(qid)
        s.release(qid)
    mixed = [got[qid] for qid in ids]
    st = s.stats()
    say("mix", st)
    assert st["completed"] == len(qs) and st["failed"] == 0
    assert st["launches"] < len(qs)  # queries were packed together
    alone = []
    for lane, src in qs:
        qid = s.submit(lane, src)
        s.wait(qid)
        alone.append(s.result(qid))
        s.release(qid)
    s.stop()
    for i in range(len(qs)):
        assert np.array_equal(mixed[i], want[i]), f"query {i} in the mix"
        assert np.array_equal(alone[i], want[i]), f"query {i} alone"


def test_one_query_per_launch(ctx, graphs):
    qs = make_queries(graphs, 12, seed=2)
    s = start_server(ctx, graphs, pool_bytes=1 << 30, pack=False)
    ids = [s.submit(lane, src) for lane, src in qs]
    for qid, (lane, src) in zip(ids, qs):
        s.wait(qid)
        assert np.array_equal(s.result(qid), expect(graphs, lane, src))
        s.release(qid)
    assert s.stats()["launches"] == len(qs)
    s.stop()


def test_admission_waits_for_memory(ctx, graphs):
    road = graphs[0]
    k = 100
    per_query = k * road.num_vertices * 2
    s = start_server(ctx, [road], pool_bytes=int(per_query * 2.5))  # two fit at a time
    rng = np.random.default_rng(3)
    srcs = [rng.integers(0, road.num_vertices, size=k).astype(np.uint32) for _ in range(7)]
    ids = [s.submit(0, src) for src in srcs]
    for qid, src in zip(ids, srcs):
        s.wait(qid)
        info = s.info(qid)
        assert info["state"] == 2
        assert np.array_equal(s.result(qid), expect([road], 0, src))
        s.release(qid)  # frees room for the waiting queries
    st = s.stats()
    assert st["completed"] == 7 and st["pool_used"] == 0
    with pytest.raises(gb.MsbfsError):
        s.submit(0, np.zeros(513, np.uint32))  # more than 512 sources
    with pytest.raises(gb.SourceOutOfBoundsError):
        s.submit(0, [road.num_vertices])
    s.stop()


def test_cooperative_lane_beside_deep_lanes(ctx, graphs):
    """Lanes on the wide family share the server with deep-family lanes: one
    with stepped launches (the server's default) and one forced back to the
    cooperative grid-barrier kernel, which the library keeps alone on the
    device; every result is still the oracle's."""
    road, rmat = graphs
    s = gb.Server(ctx, dist_dtype=np.uint16, pool_bytes=1 << 30)
    s.add_graph(road)                                     # deep
    s.add_graph(rmat, method=gb.GpuBfsMethod.PushPullAdaptive)  # wide, stepped
    s.add_graph(rmat)                                     # wide (Auto), cooperative below
    s.set_tuning(2, stepped=2)
    s.start()
    rng = np.random.default_rng(7)
    qs = []
    for i in range(30):
        lane = i % 3
        g = graphs[0] if lane == 0 else rmat
        pool = np.nonzero(g.degrees() > 0)[0]
        qs.append((lane, rng.choice(pool, size=int(rng.integers(1, 200))).astype(np.uint32)))
    say("coop lane beside deep lanes:", len(qs), "queries")
    ids = [s.submit(lane, src) for lane, src in qs]
    got = {}
    for _ in range(len(qs)):
        qid = s.next_completed(60.0)
        got[qid] = s.result(qid)
        s.release(qid)
    st = s.stats()
    s.stop()
    assert st["completed"] == len(qs) and st["failed"] == 0
    for qid, (lane, src) in zip(ids, qs):
        assert np.array_equal(got[qid], expect([road, rmat, rmat], lane, src)), lane


def test_mid_flight_admission(ctx, graphs):
    """Queries submitted while a long traversal of their lane runs are
    admitted at once into free launch slots: each starts (queue_ms) before
    the long one finishes, the server reports launches started while others
    were in flight, and every result equals the oracle's, alone or mixed."""
    from gpu_bfs import workloads as wl

    big_g = wl.road()
    lanes = [big_g] + list(graphs)
    s = gb.Server(ctx, dist_dtype=np.uint16, pool_bytes=4 << 30)
    for g in lanes:
        s.add_graph(g)
    s.start()
    big_src = wl.sources(big_g, 2048)
    rng = np.random.default_rng(11)
    small = []
    for i in range(12):
        lane = i % 3
        g = lanes[lane]
        pool = np.nonzero(g.degrees() > 0)[0]
        small.append((lane, rng.choice(pool, size=int(rng.integers(1, 40))).astype(np.uint32)))
    say("mid-flight: four 512-source road queries, then", len(small), "small queries")
    bigs = [s.submit(0, big_src[i:i + 512]) for i in range(0, 2048, 512)]
    ids = [s.submit(lane, src) for lane, src in small]
    got = {}
    infos = {}
    for _ in range(len(bigs) + len(small)):
        qid = s.next_completed(120.0)
        infos[qid] = s.info(qid)
        assert infos[qid]["state"] == 2, infos[qid]
        got[qid] = s.result(qid)
        s.release(qid)
    st = s.stats()
    s.stop()
    say("mid-flight stats", st, "big", [infos[b] for b in bigs])
    assert st["overlapped_launches"] >= 1 and st["max_in_flight"] >= 2, st
    t_big = min(infos[b]["latency_ms"] for b in bigs)
    started_early = [q for q in ids if infos[q]["queue_ms"] < t_big]
    assert len(started_early) >= len(ids) // 2, ([infos[q]["queue_ms"] for q in ids], t_big)
    for i, b in enumerate(bigs):
        assert np.array_equal(got[b], expect(lanes, 0, big_src[512 * i:512 * (i + 1)]))
    for qid, (lane, src) in zip(ids, small):
        assert np.array_equal(got[qid], expect(lanes, lane, src)), lane
