#!/usr/bin/env python3
"""Multi-source BFS benchmark on one MI355X.

Default workload: the road-like high-diameter graph (114,599 V, 239,332
directed entries) with S=2048 uniform sources, u32 distances, graph and
output resident in HBM.  One step = one full MS-BFS of all S sources.

Prints one JSON line: GTEPS (traversed edges per second, counted as the
reference's CLI does: for every source, the out-degrees of all reached
vertices, crates/cli/src/bench.rs:478-491), sources/s, per-step spread
statistics (bench.rs:376-429), the roofline of the traversal kernel, the
host-buffer path timed like the reference's CLI (wall clock around
run_ms_bfs_into, bench.rs:336-372), the CPU baseline (the oracle's
parallel-over-sources restatement of the reference's Rayon path on all of
this host's allowed cores) and, for the default workload, the secondary
configurations: R-MAT scale 20 at S=2048 with its pull-phase HBM fraction,
the 512-graph batch against one call per graph, and the serving loop.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "webgpu-bfs_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "MS-BFS GTEPS & sources/s at S=2048 on 1 GPU; served p50/p99 latency vs plain"

CONFIGS = {
    # name: (graph builder, S, dist dtype, non-isolated sources)
    "road-s2048": ("road", 2048, "u32", False),
    "road-s256": ("road", 256, "u32", False),
    "road-s1": ("road", 1, "u32", False),
    "grid-s64": ("grid256", 64, "u32", False),
    "rmat-s64": ("rmat", 64, "u16", True),
    "rmat-s2048": ("rmat", 2048, "u16", True),
    "batch512": ("batch512", 0, "u32", False),
    "serve": ("serve", 0, "u16", True),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="road-s2048", choices=sorted(CONFIGS))
    ap.add_argument("--method", default="Auto")
    ap.add_argument("--tune", action="append", default=[], help="tuning knob key=value")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: every core this process may use)")
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-host-path", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default workload only: skip the R-MAT, batch and serving results")
    ap.add_argument("--serve-queries", type=int, default=300)
    ap.add_argument("--serve-rates", default="",
                    help="Poisson arrival rates in queries/s (default: 25/50/75%% of the "
                         "measured packed capacity, then 2,000 and 20,000)")
    ap.add_argument("--serve-rmat-scale", type=int, default=22)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 256 MiB (counter calibration)")
    return ap.parse_args()


# ------------------------------------------------------------------ statistics

def run_stats(durations_ms, te, S):
    """Mean, median, min, max and sample standard deviation of step times,
    harmonic and arithmetic GTEPS and sources/s, as bench.rs:376-429."""
    d = sorted(durations_ms)
    n = len(d)
    mean = sum(d) / n
    median = d[n // 2] if n % 2 else (d[n // 2 - 1] + d[n // 2]) / 2
    var = sum((x - mean) ** 2 for x in d) / (n - 1) if n > 1 else 0.0
    total_s = sum(d) / 1e3
    return {"mean_ms": round(mean, 4), "median_ms": round(median, 4), "min_ms": round(d[0], 4),
            "max_ms": round(d[-1], 4), "stddev_ms": round(math.sqrt(var), 4),
            "harmonic_gteps": round(n * te / total_s / 1e9, 3) if total_s else 0.0,
            "arithmetic_gteps": round(te / (mean / 1e3) / 1e9, 3) if mean else 0.0,
            "sources_per_s": round(S / (mean / 1e3), 1) if mean else 0.0}


def level_bytes(dist, S, V, E, deg, dist_bytes):
    """Algorithmic bytes of one MS-BFS, from its distance matrix.

    Per level L: A_L vertices with some source at distance L, EA_L their
    out-degree sum, R_L vertices with some source farther than L (or
    unreached), N_{L+1} (source, vertex) pairs at distance L+1.  With
    b = 8*ceil(S/64) bytes of source bits per vertex:
      push(L) = 8 A_L + 4 EA_L + b (A_L + 2 EA_L) + b_d N_{L+1}
      pull(L) = 4 (V+1) + 4 E + b V + 2 b R_L + b_d N_{L+1}
      total   = b_d S V + sum_L min(push(L), pull(L)).
    Returns (total, traversed edges, max distance, [(push(L), pull(L))]).
    `dist` is a torch tensor [S, V] on the device (u32 as int32 or u16 as
    int16)."""
    import torch

    inf = 0xFFFFFFFF if dist_bytes == 4 else 0xFFFF
    mask = 0xFFFFFFFF if dist_bytes == 4 else 0xFFFF
    dev = dist.device
    maxl = 0
    chunk = max(1, (1 << 27) // max(V, 1))
    for s0 in range(0, S, chunk):
        d = dist[s0:s0 + chunk].to(torch.int64) & mask
        fin = d[d != inf]
        if fin.numel():
            maxl = max(maxl, int(fin.max()))
    nl = maxl + 2
    present = torch.zeros((nl, V), dtype=torch.bool, device=dev)
    hist = torch.zeros(nl + 1, dtype=torch.int64, device=dev)
    dmax = torch.zeros(V, dtype=torch.int64, device=dev)
    vcol = torch.arange(V, device=dev)
    te = 0
    degd = torch.as_tensor(deg, dtype=torch.int64, device=dev)
    for s0 in range(0, S, chunk):
        d = dist[s0:s0 + chunk].to(torch.int64) & mask
        reached = d != inf
        te += int((reached.to(torch.int64) * degd[None, :]).sum())
        dd = torch.where(reached, d, torch.full_like(d, nl))
        dmax = torch.maximum(dmax, dd.max(dim=0).values)
        hist += torch.bincount(dd.flatten(), minlength=nl + 1)[: nl + 1]
        r = reached.nonzero()
        present[d[r[:, 0], r[:, 1]], vcol[r[:, 1]]] = True
    A = present.sum(dim=1).to(torch.int64)
    EA = (present.to(torch.float64) @ degd.to(torch.float64)).round().to(torch.int64)
    cnt_le = torch.cumsum(torch.bincount(torch.clamp(dmax, max=nl), minlength=nl + 1), 0)
    W = (S + 63) // 64
    b = 8 * W
    total = dist_bytes * S * V
    A, EA, cnt_le, hist = A.tolist(), EA.tolist(), cnt_le.tolist(), hist.tolist()
    per_level = []
    for L in range(maxl + 1):
        R = V - cnt_le[L]
        N = hist[L + 1] if L + 1 <= maxl else 0
        push = 8 * A[L] + 4 * EA[L] + b * (A[L] + 2 * EA[L]) + dist_bytes * N
        pull = 4 * (V + 1) + 4 * E + b * V + 2 * b * R + dist_bytes * N
        per_level.append((push, pull))
        total += min(push, pull)
    return total, te, maxl, per_level


def calibrate_copy():
    import torch

    n = 256 << 20
    x = torch.ones(n // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n, "bytes_written": n,
                      "launches": 3}))


_COPY_RATE = None


def device_copy_rate(nbytes=1 << 30, reps=5):
    """Read + write rate of a plain device-to-device copy (torch's copy
    kernel, 1 GiB, timed with HIP events on the current stream) on this GPU:
    a measured practical HBM ceiling to read beside the published peak.
    Measured once per process."""
    global _COPY_RATE
    if _COPY_RATE is None:
        import torch

        x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            y.copy_(x)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del x, y
        _COPY_RATE = {"gbs": round(2 * nbytes / (ms / 1e3) / 1e9, 1), "bytes_moved": 2 * nbytes,
                      "ms": round(ms, 4), "kernel": "torch device copy, 1 GiB read + 1 GiB written"}
    return _COPY_RATE


def traffic_of(config):
    """HBM bytes per launch of this workload's dominant kernels from the
    committed rocprofv3 counter summary, with where they came from."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{config}.json")
    if not os.path.exists(tfile):
        return None, None
    with open(tfile) as f:
        t = json.load(f)
    src = (f"profiles/traffic_{config}.json: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of "
           f"{t.get('kernels', 'the traversal kernels')} at commit {t.get('commit', '?')}")
    return t.get("hbm_bytes_per_launch"), src


def cpu_threads(args):
    return args.cpu_threads or len(os.sched_getaffinity(0))


# ------------------------------------------------------------------ one workload

def measure(args, config, steps, warmup, with_cpu, with_host, with_trace=False):
    """Device-resident timing of one configuration, its roofline, and
    optionally the host-buffer path, the CPU baseline and a per-level trace
    (wide family) for the pull-phase fraction."""
    import numpy as np
    import torch

    import gpu_bfs as gb
    from gpu_bfs import _lib as L
    from gpu_bfs import workloads as wl

    gname, S, dname, non_iso = CONFIGS[config]
    t0 = time.time()
    graph = getattr(wl, gname)()
    src = wl.sources(graph, S, non_isolated=non_iso)
    V, E = graph.num_vertices, graph.num_edges
    log(f"workload {config}: V={V} E={E} S={S} dist={dname} (built in {time.time() - t0:.1f}s)")

    ctx = gb.GpuContext(0)
    eng = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod[args.method])
    if args.tune:
        eng.set_tuning(**{k: int(v) for k, v in (t.split("=") for t in args.tune)})
    dg = gb.DeviceGraph(ctx, graph)
    n = S * V
    tdt = torch.int32 if dname == "u32" else torch.int16
    npdt = np.uint32 if dname == "u32" else np.uint16
    out = torch.empty(n, dtype=tdt, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def step(sync):
        eng.run_ms_bfs_device(dg, src, out.data_ptr(), n, dist_dtype=npdt, stream=stream,
                              sync=sync)

    for _ in range(warmup):
        step(True)
    torch.cuda.synchronize()
    kernel_ms, device_ms = [], []
    t_start = time.perf_counter()
    for _ in range(steps):
        step(False)
        st = eng.last_stats()
        kernel_ms.append(st["kernel_ms"])
        device_ms.append(st["total_ms"])
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t_start) * 1e3 / steps
    stats = eng.last_stats()
    log(f"{config}: timed {steps} steps: {ms_per_step:.3f} ms/step")

    deg = graph.degrees()
    dist2d = out.view(S, V)
    b_d = 4 if dname == "u32" else 2
    model_bytes, te, maxl, per_level = level_bytes(dist2d, S, V, E, deg, b_d)
    md = int(deg.max()) if V else 0
    ell = 4 if md <= 4 else (8 if md <= 8 else 0)
    lat = int(stats["lattice_bits"])
    # Compulsory HBM bytes of one launch: every distance written once, the
    # graph in the form the kernel reads (the lattice kernel: its adjacency
    # bits; otherwise the CSR, plus the neighbour table of the one-wave
    # kernel when the maximum degree is <= 8) and the sources read once.
    graph_bytes = (V * lat + 7) // 8 if lat else 4 * (V + 1) + 4 * E + 4 * ell * V
    compulsory = b_d * S * V + graph_bytes + 4 * S
    deep = stats["family"] == 1
    # The deep family keeps all per-source state in LDS, so only the
    # compulsory bytes reach HBM; the wide family keeps source words in HBM
    # and is measured against the per-level model.
    alg_bytes = compulsory if deep else model_bytes
    kmean = float(np.mean(kernel_ms))
    achieved = alg_bytes / (kmean / 1e3) / 1e9
    traffic, traffic_src = traffic_of(config)
    if deep:
        kernel = ("deep_wave_kernel (lattice adjacency in LDS)" if lat
                  else "deep_wave_kernel" if stats["block_threads"] == 64 else "deep_bfs_kernel")
    else:
        kernel = "wide_bfs_kernel + wide_final_kernel"
    res = {
        "ms_per_step": round(ms_per_step, 4),
        "gteps": round(te / (ms_per_step / 1e3) / 1e9, 3),
        "sources_per_s": round(S / (ms_per_step / 1e3), 1),
        "traversed_edges": te,
        "kernel_ms": round(kmean, 4),
        "step_device_stats": run_stats(device_ms, te, S),
        "config": {"workload": config, "graph": gname, "num_vertices": V, "num_edges": E,
                   "num_sources": S, "dist": dname, "method": args.method,
                   "family": ["none", "deep", "wide"][stats["family"]],
                   "levels": int(stats["levels"]), "max_distance": maxl,
                   "push_levels": int(stats["push_levels"]),
                   "pull_levels": int(stats["pull_levels"]),
                   "bits_per_instance": int(stats["bits_per_instance"]),
                   "lattice_bits": lat,
                   "grid_blocks": int(stats["grid_blocks"]),
                   "block_threads": int(stats["block_threads"]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": alg_bytes,
                     "byte_model": "compulsory" if deep else "level",
                     "compulsory_bytes": compulsory, "level_model_bytes": model_bytes,
                     "kernel": kernel},
    }
    cr = device_copy_rate()
    res["roofline"]["device_copy"] = cr
    res["roofline"]["frac_of_copy"] = round(achieved / cr["gbs"], 5)
    if deep and stats["levels"]:
        # the deep family is bound by its dependent levels, not by bandwidth
        res["roofline"]["latency"] = {
            "levels": int(stats["levels"]),
            "us_per_level": round(kmean * 1e3 / stats["levels"], 4)}

    if with_trace and not deep:
        # one traced run: per-level device clock, direction and frontier
        eng.set_tuning(trace_levels=max(64, maxl + 2))
        step(True)
        recs = eng.level_trace()
        eng.set_tuning(trace_levels=0)
        pull_b = pull_us = 0.0
        rows = []
        for r in recs:
            L_ = r["level"]
            if L_ >= len(per_level):
                continue
            push_b, pull_bytes = per_level[L_]
            us = r["end_us"] - r["start_us"]
            dirn = "pull" if r["direction"] == 1 else "push"
            b = pull_bytes if dirn == "pull" else push_b
            rows.append({"level": L_, "dir": dirn, "us": round(us, 1),
                         "model_gb": round(b / 1e9, 3),
                         "frac": round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 3) if us > 0 else None})
            if dirn == "pull":
                pull_b += b
                pull_us += us
        res["levels_trace"] = rows
        if pull_us > 0:
            res["pull_phase"] = {"model_bytes": int(pull_b), "us": round(pull_us, 1),
                                 "achieved_gbs": round(pull_b / (pull_us * 1e-6) / 1e9, 1),
                                 "frac": round(pull_b / (pull_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                                 "frac_of_copy": round(pull_b / (pull_us * 1e-6) / 1e9 / cr["gbs"], 4)}

    if with_host:
        # The reference's own call shape: run_ms_bfs_into a reusable host
        # buffer, wall clock around the whole call (bench.rs:336-372).
        host = np.empty(n, dtype=np.uint32)
        srcc = np.ascontiguousarray(src, dtype=np.uint32)
        lib = L.lib()

        def host_call():
            L.check(lib.msbfs_run_ms_bfs(eng.handle, dg.handle,
                                         srcc.ctypes.data_as(ctypes.c_void_p), S, L.DIST_U32,
                                         host.ctypes.data_as(ctypes.c_void_p), n))

        host_call()
        hd = []
        for _ in range(3):
            t = time.perf_counter()
            host_call()
            hd.append((time.perf_counter() - t) * 1e3)
        res["host_path"] = dict(run_stats(hd, te, S), iters=3, warmup=1,
                                call="msbfs_run_ms_bfs into a pageable u32 host array")
        res["host_path"]["gteps"] = res["host_path"]["harmonic_gteps"]
        log(f"{config}: host path {res['host_path']['mean_ms']:.2f} ms")
        if not args.no_parity:
            dv = out.cpu().numpy().view(npdt).astype(np.uint32)
            if npdt == np.uint16:
                dv = np.where(dv == 0xFFFF, 0xFFFFFFFF, dv).astype(np.uint32)
            res["host_path"]["identical_to_device"] = bool(np.array_equal(host, dv))

    if with_cpu:
        from oracle import oracle as orc

        threads = cpu_threads(args)
        # Bounded sample: as many sources as fit the time budget, estimated
        # from a 16-source probe.
        t = time.perf_counter()
        orc.ms_bfs_parallel(V, graph.row_offsets, graph.col_indices, src[:16], threads)
        per_src = (time.perf_counter() - t) / 16
        ns = int(min(S, max(16, args.cpu_budget_s / 4 / max(per_src, 1e-9))))
        sample = src[:ns]
        times = []
        cpu_out = None
        deadline = time.perf_counter() + args.cpu_budget_s
        while len(times) < 3 and (not times or time.perf_counter() < deadline):
            t = time.perf_counter()
            cpu_out = orc.ms_bfs_parallel(V, graph.row_offsets, graph.col_indices, sample, threads)
            times.append(time.perf_counter() - t)
        cpu_s = float(np.median(times))
        cpu_te = orc.traversed_edges(V, graph.row_offsets, ns, cpu_out)
        res["cpu_baseline"] = {
            "value": round(cpu_te / cpu_s / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "kind": "port",
            "sample": f"{ns} of {S} sources, {len(times)} timed runs (median {cpu_s*1e3:.1f} ms), "
                      "parallel-over-sources FIFO BFS incl. output allocation and fill",
            "sources_per_s": round(ns / cpu_s, 1),
            "threads_used": threads, "affinity_cpus": len(os.sched_getaffinity(0)),
            "os_cpu_count": os.cpu_count()}
        if "host_path" in res:
            res["host_path"]["vs_cpu_baseline"] = round(
                res["host_path"]["harmonic_gteps"] / res["cpu_baseline"]["value"], 2)
        if not args.no_parity:
            gpu = dist2d[:ns].cpu().numpy().view(npdt).reshape(-1)
            ref = cpu_out.astype(np.uint64)
            if npdt == np.uint16:
                ref = np.where(cpu_out == 0xFFFFFFFF, 0xFFFF, cpu_out).astype(np.uint16)
            res["parity"] = bool(np.array_equal(gpu, ref.astype(npdt)))
            res["parity_sources_checked"] = ns
    del out
    torch.cuda.empty_cache()
    return res


# ------------------------------------------------------------------ batch

def sweep_summary(sweep):
    """Autotune candidates as "knob=value,... : ms" strings (knobs left at
    their defaults omitted)."""
    res = []
    for c in sweep:
        knobs = ",".join(f"{k}={v}" for k, v in c.items() if k not in ("ms", "status") and v)
        res.append(f"{knobs}: {c['ms']:.4f}" if c["ms"] is not None else f"{knobs}: error {c['status']}")
    return res


def bench_batch(args, steps, warmup):
    """512 G(n,p) graphs of uneven size, each with its own source set, in the
    batch mode the autotuner picks (union traversals: the graphs joined into
    disjoint unions by source words and traversed bit-parallel; or one
    persistent hipGraph-replayed launch over per-source tasks), against one
    engine call per graph."""
    import numpy as np
    import torch

    import gpu_bfs as gb
    from gpu_bfs import workloads as wl

    t0 = time.time()
    pairs = wl.batch512()
    graphs = [g for g, _ in pairs]
    sources = [s for _, s in pairs]
    log(f"batch512: {len(graphs)} graphs, V={sum(g.num_vertices for g in graphs)} "
        f"E={sum(g.num_edges for g in graphs)} S={sum(len(s) for s in sources)} "
        f"(built in {time.time() - t0:.1f}s)")
    ctx = gb.GpuContext(0)
    eng = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod.Auto)
    dgs = [gb.DeviceGraph(ctx, g) for g in graphs]
    b = gb.Batch(eng, dgs, sources)
    out = torch.empty(b.total, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    if args.tune:
        # the given knobs instead of the autotuner (e.g. for a profile of the
        # timed replays alone)
        b.set_tuning(**{k: int(v) for k, v in (t.split("=") for t in args.tune)})
        tuned, tuned_ms = b.get_tuning(), None
        log(f"batch512 tuning given: {tuned}")
    else:
        tuned, tuned_ms = b.autotune(out.data_ptr(), b.total)
        log(f"batch512 autotune: {tuned} {tuned_ms:.3f} ms")
    b.capture(out.data_ptr(), b.total, stream=stream)
    for _ in range(warmup):
        b.replay(stream=stream)
    torch.cuda.synchronize()
    kernel_ms = []
    t_start = time.perf_counter()
    for _ in range(steps):
        b.replay(stream=stream)
        kernel_ms.append(b.stats()["kernel_ms"])
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t_start) * 1e3 / steps
    st = b.stats()

    te = 0
    for g, (off, S, V) in enumerate(zip(b.offsets, b.counts, b.num_vertices)):
        if S == 0:
            continue
        d = out[int(off):int(off) + int(S) * V].view(int(S), V)
        deg = torch.as_tensor(graphs[g].degrees(), device="cuda")
        te += int(((d != -1).to(torch.int64) * deg[None, :]).sum())

    # plain path: one engine call per graph into the same layout
    plain = torch.full_like(out, -5)
    eng_plain = gb.GpuBfsEngine(ctx, gb.GpuBfsMethod.Auto)

    def plain_all():
        for g in range(len(dgs)):
            if len(sources[g]):
                eng_plain.run_ms_bfs_device(dgs[g], sources[g],
                                            plain.data_ptr() + 4 * int(b.offsets[g]),
                                            int(b.counts[g]) * b.num_vertices[g], stream=stream,
                                            sync=False)
        eng_plain.check()

    plain_all()  # warm every per-graph workspace once
    torch.cuda.synchronize()
    t = time.perf_counter()
    plain_all()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t) * 1e3
    same = bool(torch.equal(plain, out))
    res = {"ms_per_step": round(ms_per_step, 4),
           "gteps": round(te / (ms_per_step / 1e3) / 1e9, 3),
           "kernel_ms": round(float(np.mean(kernel_ms)), 4),
           "traversed_edges": te,
           "plain_per_graph_ms": round(plain_ms, 3),
           "plain_gteps": round(te / (plain_ms / 1e3) / 1e9, 3),
           "speedup_vs_plain": round(plain_ms / ms_per_step, 2),
           "identical_to_plain": same,
           "autotune": ({"chosen": {k: v for k, v in tuned.items() if v},
                         "ms": round(tuned_ms, 4),
                         "sweep": sweep_summary(b.sweep())} if tuned_ms is not None else
                        {"given": {k: v for k, v in tuned.items() if v}}),
           "wide_graphs": int(sum(1 for f in b.families() if f == int(gb.Family.WIDE))),
           "config": {"workload": "batch512", "graphs": len(graphs),
                      "num_vertices": int(sum(g.num_vertices for g in graphs)),
                      "num_edges": int(sum(g.num_edges for g in graphs)),
                      "num_sources": int(sum(len(s) for s in sources)),
                      "mode": {1: "union traversals", 2: "persistent deep launch"}.get(
                          tuned["batch_union"],
                          "persistent deep launch (hipGraph) beside union traversals of the graphs "
                          f"of >= {tuned['union_words'] or 2} source words"),
                      "instances": int(st["instances"]),
                      "source_bits_per_instance": int(st["bits_per_instance"]),
                      "threads": int(st["block_threads"]), "grid_blocks": int(st["grid_blocks"]),
                      "hipgraph": tuned["batch_union"] != 1}}
    del out, plain
    torch.cuda.empty_cache()
    return res


# ------------------------------------------------------------------ serving

def bench_serve(args):
    """Serving: the road-like graph and an R-MAT graph (scale 22) resident;
    Poisson arrivals, each query picks a graph 50/50 and brings k
    log-uniform in [1, 512] sources; u16 distances kept on the device.
    Packed continuous batching against one query per traversal, first at
    saturation (all queries at once) to measure capacity, then at offered
    rates below and above it."""
    import threading

    import numpy as np
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    import gpu_bfs as gb
    from gpu_bfs import workloads as wl

    t0 = time.time()
    graphs = [wl.road(), wl.rmat(args.serve_rmat_scale)]
    log(f"serve: road V={graphs[0].num_vertices} E={graphs[0].num_edges}; "
        f"rmat{args.serve_rmat_scale} V={graphs[1].num_vertices} E={graphs[1].num_edges} "
        f"(built in {time.time() - t0:.1f}s)")
    # traversed edges of a source = out-degree sum of its connected component
    comp_te = []
    for g in graphs:
        m = csr_matrix((np.ones(g.num_edges, np.int8), g.col_indices, g.row_offsets),
                       shape=(g.num_vertices, g.num_vertices))
        _, lab = connected_components(m, directed=False)
        tot = np.bincount(lab, weights=g.degrees()).astype(np.int64)
        comp_te.append(tot[lab])
    rng = np.random.default_rng(0x5EED)
    n = args.serve_queries
    queries = []
    nonzero = [np.nonzero(g.degrees() > 0)[0].astype(np.uint32) for g in graphs]
    for _ in range(n):
        lane = int(rng.integers(0, 2))
        k = int(np.exp(rng.uniform(0, np.log(512))))
        k = max(1, min(512, k))
        queries.append((lane, rng.choice(nonzero[lane], size=k)))
    te_q = [int(comp_te[lane][src].sum()) for lane, src in queries]
    ctx = gb.GpuContext(0)
    dgs = [gb.DeviceGraph(ctx, g) for g in graphs]
    pool = 2 * 2 * 2048 * max(g.num_vertices for g in graphs)

    def run(rate, pack):
        srv = gb.Server(ctx, pool_bytes=pool, dist_dtype=np.uint16, pack=pack)
        for dg in dgs:
            srv.add_graph(dg)
        srv.start()
        for lane in (0, 1):  # warm both lanes
            q = srv.submit(lane, nonzero[lane][:1])
            assert srv.next_completed(120.0) == q
            srv.release(q)
        srv.reset_stats()
        gaps = rng.exponential(1.0 / rate, size=n) if rate else np.zeros(n)
        arrive = np.cumsum(gaps)

        def client():
            t_base = time.perf_counter()
            for i, (lane, src) in enumerate(queries):
                delay = arrive[i] - (time.perf_counter() - t_base)
                if delay > 0:
                    time.sleep(delay)
                srv.submit(lane, src)

        th = threading.Thread(target=client)
        th.start()
        lat = []
        while len(lat) < n:
            qid = srv.next_completed(120.0)
            info = srv.info(qid)
            if info["state"] != 2:
                raise RuntimeError(f"query {qid} failed with status {info['status']}")
            lat.append(info["latency_ms"])
            srv.release(qid)
        th.join()
        st = srv.stats()
        srv.stop()
        lat = np.array(lat)
        span = st["span_s"]
        r = {"offered_qps": rate if rate else "saturation", "queries": n,
             "launches": int(st["launches"]),
             "overlapped_launches": int(st["overlapped_launches"]),
             "max_in_flight": int(st["max_in_flight"]),
             "p50_ms": round(float(np.percentile(lat, 50)), 3),
             "p99_ms": round(float(np.percentile(lat, 99)), 3),
             "achieved_qps": round(n / span, 1),
             "sources_per_s": round(sum(len(s) for _, s in queries) / span, 1),
             "gteps": round(sum(te_q) / span / 1e9, 3)}
        log(f"serve pack={pack} rate={rate}: {r}")
        return r

    packed_sat = run(0, True)
    single_sat = run(0, False)
    cap = packed_sat["achieved_qps"]
    if args.serve_rates:
        rates = [float(r) for r in args.serve_rates.split(",")]
    else:
        rates = [round(cap * f) for f in (0.25, 0.5, 0.75)] + [2000.0, 20000.0]
    packed = [run(r, True) for r in rates]
    single = [run(r, False) for r in rates]
    return {"capacity_qps": {"packed": cap, "one_query_per_launch": single_sat["achieved_qps"]},
            "saturation": {"packed": packed_sat, "one_query_per_launch": single_sat},
            "packed": packed, "one_query_per_launch": single,
            "config": {"workload": "serve", "graphs": ["road-like 114,599 V",
                                                       f"rmat{args.serve_rmat_scale}"],
                       "queries_per_rate": n, "k": "log-uniform [1, 512]",
                       "admission": "u16 x k x |V| output bytes", "pool_bytes": pool,
                       "slots_per_lane": 4, "wide_lanes": "stepped launches"}}


# ------------------------------------------------------------------ main

def line(args, res, workload_dtype):
    """The one JSON line of the driver's contract."""
    out = {"metric": METRIC, "value": res["gteps"], "unit": "GTEPS", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": workload_dtype,
           "data": "synthetic (seeded generators; graph and distances resident in HBM)"}
    out.update({k: v for k, v in res.items() if k not in ("gteps", "ms_per_step")})
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "gpus_requested": args.gpus}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.config == "serve":
        res = bench_serve(args)
        best = max(res["packed"], key=lambda r: r["gteps"])
        out = {"metric": "served MS-BFS: p50/p99 latency and throughput, packed vs one query "
                         "per launch", "value": best["gteps"], "unit": "GTEPS", "n_gpus": 1,
               "steps": len(res["packed"]), "warmup": 1, "ms_per_step": None,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u64 source words / u16 distances",
               "data": "synthetic (road-like + R-MAT, Poisson arrivals, seeded)"}
        out.update(res)
        print(json.dumps(out))
        return
    if args.config == "batch512":
        res = bench_batch(args, args.steps, args.warmup)
        out = {"metric": "batched MS-BFS: per-batch wall time and GTEPS vs one launch per graph",
               "value": res["gteps"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u32 distances", "data": "synthetic (512 seeded G(n,p) graphs)"}
        out.update({k: v for k, v in res.items() if k not in ("gteps", "ms_per_step")})
        print(json.dumps(out))
        return

    dname = CONFIGS[args.config][2]
    res = measure(args, args.config, args.steps, args.warmup, not args.no_cpu_baseline,
                  not args.no_host_path, with_trace=args.config.startswith("rmat"))
    out = line(args, res, f"u64 source words / {dname} distances")
    if args.config == "road-s2048" and not args.no_secondary:
        sec = {}
        t0 = time.time()
        r = measure(args, "rmat-s2048", 5, 2, with_cpu=False, with_host=False, with_trace=True)
        sec["rmat-s2048"] = {k: r[k] for k in ("ms_per_step", "gteps", "sources_per_s",
                                               "kernel_ms", "roofline", "pull_phase",
                                               "levels_trace", "config") if k in r}
        sec["batch512"] = bench_batch(args, 5, 2)
        sec["serve"] = bench_serve(args)
        out["secondary"] = sec
        out["secondary_wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
