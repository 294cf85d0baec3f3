"""GPU test of msbfs_run_ms_bfs_device_noalloc (the entry point that
allocates nothing; every workspace in the library's static per-device
scratch array).  Its output must equal msbfs_run_ms_bfs_device's, bit for
bit, on two calls in a row, the second with a smaller input (the scratch
keeps the first call's contents), and it rejects what does not fit."""
import os

import numpy as np
import pytest

import gpu_bfs as gb
from gpu_bfs import workloads as wl

pytestmark = pytest.mark.gpu

M = gb.GpuBfsMethod


def say(*a):
    print("[case]", *a, flush=True)


def device_run(e, g, src, dtype, noalloc):
    import torch

    tdt = {np.uint32: torch.int32, np.uint16: torch.int16, np.uint8: torch.uint8}[dtype]
    n = len(src) * g.num_vertices
    out = torch.full((max(n, 1),), 0x55, dtype=tdt, device="cuda")
    if noalloc:
        e.run_ms_bfs_device_noalloc(g, src, out.data_ptr(), n, dist_dtype=dtype)
    else:
        e.run_ms_bfs_device(g, src, out.data_ptr(), n, dist_dtype=dtype)
    torch.cuda.synchronize()
    return out[:n].cpu().numpy().view(dtype)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint16])
def test_noalloc_equals_device_entry(ctx, dtype):
    road = wl.road()
    small = gb.generate_road_like(20_000, 160, 900, seed=3)
    grid = gb.generate_grid2d(64, 48)
    sparse = gb.generate_erdos_renyi_np(20_000, 1.6 / 19_999, seed=21)
    # (graph, S) pairs run back to back; each second call is smaller
    seqs = [[(road, 2048), (road, 300)], [(small, 700), (grid, 70)],
            [(sparse, 1100), (sparse, 5)]]
    if dtype == np.uint16:
        seqs = seqs[1:]  # road's 3,079 levels fit u16 too; keep the u16 case short
    for seq in seqs:
        ref_e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        na_e = gb.GpuBfsEngine(ctx, M.WorkgroupPersistentBfs)
        for g, S in seq:
            src = wl.sources(g, S)
            say("noalloc V=%d S=%d %s" % (g.num_vertices, S, np.dtype(dtype).name))
            want = device_run(ref_e, g, src, dtype, noalloc=False)
            # twice: the second call finds the first's contents in the scratch
            for _ in range(2):
                got = device_run(na_e, g, src, dtype, noalloc=True)
                assert np.array_equal(got, want), (g.num_vertices, S)
            st = na_e.last_stats()
            assert st["family"] == int(gb.Family.DEEP) and st["levels"] > 0


def test_noalloc_rejections(ctx):
    import torch

    e = gb.GpuBfsEngine(ctx, M.Auto)
    rmat = gb.generate_rmat(12, 16, seed=5)
    src = gb.sample_sources(rmat, 64, seed=1, non_isolated=True)
    out = torch.empty(64 * rmat.num_vertices, dtype=torch.int32, device="cuda")
    say("noalloc rejects the wide family")
    with pytest.raises(gb.MsbfsError) as ei:
        e.run_ms_bfs_device_noalloc(rmat, src, out.data_ptr(), out.numel())
    assert ei.value.code == 9 and "wide family" in str(ei.value)
    g = gb.generate_grid2d(4, 4)
    many = np.zeros(65537, np.uint32)
    out2 = torch.empty(65537 * 16, dtype=torch.int32, device="cuda")
    say("noalloc rejects too many sources")
    with pytest.raises(gb.MsbfsError) as ei:
        e.run_ms_bfs_device_noalloc(g, many, out2.data_ptr(), out2.numel())
    assert ei.value.code == 4 and "65536" in str(ei.value)
    with pytest.raises(gb.SourceOutOfBoundsError):
        e.run_ms_bfs_device_noalloc(g, [16], out2.data_ptr(), 16)
    # still usable afterwards, and equal to the allocating entry point
    src = np.arange(16, dtype=np.uint32)
    want = device_run(e, g, src, np.uint32, noalloc=False)
    assert np.array_equal(device_run(e, g, src, np.uint32, noalloc=True), want)
