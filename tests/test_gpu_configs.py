"""GPU parity at the measured configurations' full workloads.

Each test builds the same seeded workload the benchmark measures
(gpu_bfs.workloads) and compares the HIP path with the oracle:

  * R-MAT scale 20, S=2048, u16: every row, the oracle run in 256-source
    chunks;
  * road-like 114,599 V, S=2048, through the push/pull family
    (PushPullAdaptive, crates/gpu_bfs/src/dispatch/push_pull.rs), every row;
  * the batched workload: 512 G(n,p) graphs of 1,024-65,536 vertices with
    1-256 sources each, one persistent launch, eagerly and replayed from a
    hipGraph, every graph against the oracle and against one engine call per
    graph;
  * the serving workload: the road-like graph and R-MAT scale 22 resident,
    queries of 1-512 log-uniform sources arriving on a Poisson clock from a
    client thread; every road query's rows and sampled R-MAT rows against
    the oracle, and every row's source at distance 0.
"""
import math
import os
import threading
import time

import numpy as np
import pytest

import gpu_bfs as gb
from gpu_bfs import workloads as wl
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

M = gb.GpuBfsMethod
THREADS = min(16, len(os.sched_getaffinity(0)))
INF32 = 0xFFFFFFFF


def say(*a):
    print("[case]", *a, flush=True)


def oracle(g, src):
    return orc.ms_bfs_parallel(g.num_vertices, g.row_offsets, g.col_indices, src, THREADS)


def u16(d32):
    return np.where(d32 == INF32, 0xFFFF, d32).astype(np.uint16)


def test_rmat20_s2048_every_row(ctx):
    g = wl.rmat(20)
    src = wl.sources(g, 2048, non_isolated=True)
    e = gb.GpuBfsEngine(ctx, M.Auto)
    say("rmat20 S=2048 u16 Auto")
    got = e.run_ms_bfs(g, src, dist_dtype=np.uint16).reshape(2048, -1)
    assert e.last_stats()["family"] == int(gb.Family.WIDE)
    for c in range(0, 2048, 256):
        say("rmat20 oracle rows", c, c + 256)
        exp = u16(oracle(g, src[c:c + 256])).reshape(256, -1)
        bad = np.nonzero(np.any(got[c:c + 256] != exp, axis=1))[0]
        assert bad.size == 0, f"rows {(c + bad[:8]).tolist()} differ"


def test_road_s2048_push_pull_family(ctx):
    g = wl.road()
    src = wl.sources(g, 2048)
    exp = oracle(g, src)
    e = gb.GpuBfsEngine(ctx, M.PushPullAdaptive)
    say("road S=2048 PushPullAdaptive (host output)")
    got = e.run_ms_bfs(g, src)
    st = e.last_stats()
    assert st["family"] == int(gb.Family.WIDE) and st["levels"] == 3080, st
    assert np.array_equal(got, exp), int(np.count_nonzero(got != exp))
    import torch

    out = torch.full((2048 * g.num_vertices,), 7, dtype=torch.int32, device="cuda")
    say("road S=2048 PushPullAdaptive (device u32)")
    e.run_ms_bfs_device(g, src, out.data_ptr(), out.numel())
    assert np.array_equal(out.cpu().numpy().view(np.uint32), exp)


def test_batch512_full_workload(ctx):
    import torch

    t0 = time.time()
    pairs = wl.batch512()
    graphs = [g for g, _ in pairs]
    sources = [s for _, s in pairs]
    say("batch512 built in %.1fs: V=%d S=%d" % (time.time() - t0,
        sum(g.num_vertices for g in graphs), sum(len(s) for s in sources)))
    eng = gb.GpuBfsEngine(ctx, M.Auto)
    dgs = [gb.DeviceGraph(ctx, g) for g in graphs]
    b = gb.Batch(eng, dgs, sources)
    out = torch.empty(b.total, dtype=torch.int32, device="cuda")
    say("batch512 eager run")
    b.run(out.data_ptr(), b.total)
    host = out.cpu().numpy().view(np.uint32)
    parts = b.split(host)
    for gi, (g, s) in enumerate(zip(graphs, sources)):
        exp = oracle(g, s).reshape(len(s), g.num_vertices)
        assert np.array_equal(parts[gi], exp), f"graph {gi} (V={g.num_vertices}, S={len(s)})"
    say("batch512 hipGraph replay x2")
    b.capture(out.data_ptr(), b.total)
    for _ in range(2):
        out.fill_(-3)
        torch.cuda.synchronize()
        b.replay()
        assert np.array_equal(out.cpu().numpy().view(np.uint32), host)
    say("batch512 plain path, one engine call per graph")
    plain = torch.full_like(out, -3)
    for gi, dg in enumerate(dgs):
        eng.run_ms_bfs_device(dg, sources[gi], plain.data_ptr() + 4 * int(b.offsets[gi]),
                              len(sources[gi]) * graphs[gi].num_vertices)
    assert torch.equal(plain, out)
    # both batch modes: the default union traversals checked above, and the
    # persistent deep launch over per-source tasks
    mode = b.get_tuning()["batch_union"]
    b.set_tuning(batch_union=1 if mode == 2 else 2)
    say("batch512 other mode (batch_union=%d)" % b.get_tuning()["batch_union"])
    out.fill_(-3)
    torch.cuda.synchronize()
    b.run(out.data_ptr(), b.total)
    assert torch.equal(plain, out)


def test_serve_workload_road_and_rmat22(ctx):
    road, rmat = wl.road(), wl.rmat(22)
    graphs = [road, rmat]
    say("serve graphs: road V=%d, rmat22 V=%d E=%d" % (road.num_vertices, rmat.num_vertices,
                                                       rmat.num_edges))
    rng = np.random.default_rng(0x5EED)
    nonzero = [np.nonzero(g.degrees() > 0)[0].astype(np.uint32) for g in graphs]
    qs = []
    for _ in range(40):
        lane = int(rng.integers(0, 2))
        k = max(1, min(512, int(math.exp(rng.uniform(0, math.log(512))))))
        qs.append((lane, rng.choice(nonzero[lane], size=k)))
    gaps = rng.exponential(1.0 / 2000.0, size=len(qs))
    s = gb.Server(ctx, pool_bytes=24 << 30, dist_dtype=np.uint16)
    for g in graphs:
        s.add_graph(g)
    s.start()
    ids = [None] * len(qs)

    def client():
        for i, (lane, src) in enumerate(qs):
            time.sleep(gaps[i])
            ids[i] = s.submit(lane, src)

    th = threading.Thread(target=client)
    th.start()
    got = {}
    for _ in range(len(qs)):
        qid = s.next_completed(300.0)
        info = s.info(qid)
        assert info["state"] == 2, info
        got[qid] = s.result(qid)
        s.release(qid)
    th.join()
    st = s.stats()
    s.stop()
    say("serve stats", st)
    assert st["completed"] == len(qs) and st["failed"] == 0
    for i, (lane, src) in enumerate(qs):
        g = graphs[lane]
        d = got[ids[i]].reshape(len(src), g.num_vertices)
        assert np.all(d[np.arange(len(src)), src] == 0), i
        rows = np.arange(len(src)) if lane == 0 else np.unique(
            [0, len(src) - 1] + rng.integers(0, len(src), size=2).tolist())
        exp = u16(oracle(g, src[rows])).reshape(len(rows), g.num_vertices)
        assert np.array_equal(d[rows], exp), (i, lane, len(src))
