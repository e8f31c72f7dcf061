"""GPU parity: the hybrid transfer engine over pageable host graphs (graphs
too large to pin): filter copies staged through pinned bounce buffers,
compaction as for pinned graphs, no zero-copy (the cost model never picks
it, forcing it is refused), rows longer than the packed staging streamed
through the ring. Bar: SSSP distances bit-exact, PageRank / PPR the same
words as the pinned engine and the in-HBM kernels."""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)


def _say(msg):
    print(f"[gpu-hybrid-pageable] {msg}", flush=True)


def graph(abi, scale=14, seed=3, wmax=255, kind=0, memory=None):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0,
                         weight_max=wmax, memory=abi.HYT_HOST_PAGEABLE if memory is None else memory)


SMALL = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=8192 * 20 + 16 + 12 * 2048 * 8,
             calibrate=0)


@pytest.mark.parametrize("engine", [-1, 0, 1])
@pytest.mark.parametrize("cache", [0, 1])
def test_pageable_sssp_bitexact(gpu, abi, orc, engine, cache):
    hg = graph(abi)
    h = abi.HybridEngine(hg, force_engine=engine, cache=cache, **SMALL)
    for src in abi.sample_sources(hg.offsets, 5, 2):
        _say(f"engine {engine} cache {cache} src {src}")
        d, st = h.sssp(int(src))
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src)))
        assert st["run"]["converged"] == 1
        assert st["parts_zero_copy"] == 0 and st["bytes_zero_copy"] == 0
        if engine == 0:
            assert st["bytes_filter"] > 0
        if engine == 1:
            assert st["bytes_compaction"] > 0


def test_pageable_refuses_zero_copy(gpu, abi):
    hg = graph(abi, scale=10)
    with pytest.raises(abi.InvalidArgument):
        abi.HybridEngine(hg, force_engine=2, **SMALL)


def test_pageable_large_staged_copies_and_calibration(gpu, abi, orc):
    """Filter windows of 8 partitions x 3 * 2^20 edges (96 MiB of columns:
    two bounce buffers per copy) and the calibration through the staged
    path."""
    hg = graph(abi, scale=21, seed=11)
    part = 3 << 20
    pack_bytes = ((1 << 22) * 4 + 16) * 4 + ((1 << 22) + 1) * 8 + 16
    h = abi.HybridEngine(hg, partition_edges=part, pack_edges=1 << 22, force_engine=0, cache=0, calibrate=1,
                         device_budget_bytes=pack_bytes + 2 * 8 * 4 * part * 8 + (1 << 20))
    info = h.info()
    _say(f"staged link {info['link_gbs_peak']:.1f} GB/s")
    assert info["link_gbs_peak"] > 0
    src = int(abi.sample_sources(hg.offsets, 12, 1)[0])
    d, st = h.sssp(src)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
    assert st["bytes_filter"] > (64 << 20)


def test_pageable_oversized_rows(gpu, abi, orc):
    """Rows longer than the packed staging stream through the ring in windows
    (single query and K lanes)."""
    hg = graph(abi, scale=15, seed=9)
    deg = np.diff(hg.offsets.astype(np.int64))
    assert deg.max() > 2048
    h = abi.HybridEngine(hg, partition_edges=512, pack_edges=2048, force_engine=1, calibrate=0,
                         device_budget_bytes=2048 * 20 + 16 + 4 * 512 * 8 + 8 * 512 * 8)
    hub = int(np.argmax(deg))
    srcs = [hub, 0, 17, 4242, 9]
    for s in srcs[:3]:
        d, st = h.sssp(s)
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s)), s
    d, st = h.sssp_k(srcs)
    for k, s in enumerate(srcs):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s)), k


@pytest.mark.parametrize("engine", [-1, 0, 1])
def test_pageable_pagerank_and_ppr_match_pinned(gpu, abi, orc, engine):
    hp = graph(abi, scale=13, seed=2, wmax=0)
    hq = graph(abi, scale=13, seed=2, wmax=0, memory=abi.HYT_HOST_PINNED)
    assert np.array_equal(hp.columns, hq.columns)
    ep = abi.HybridEngine(hp, force_engine=engine, **SMALL)
    eq = abi.HybridEngine(hq, force_engine=engine, **SMALL)
    eps = 1e-7
    rp, sp = ep.pagerank(D, eps)
    rq, _ = eq.pagerank(D, eps)
    same, _, _ = orc.delta_push_f64(hp.offsets, hp.columns, D, f32(eps))
    assert sp["run"]["converged"] == 1 and np.abs(rp - same).sum() <= 1e-6
    assert np.array_equal(rp, rq)
    seeds = [int(x) for x in abi.sample_sources(hp.offsets, 3, 4)]
    g = abi.DeviceGraph.from_host(hp)
    rk, _ = ep.ppr_k(seeds, D, 1e-8)
    alone, _ = g.ppr(seeds, D, 1e-8)
    for k in range(len(seeds)):
        assert np.array_equal(rk[:, k], alone[:, k]), k


def test_pageable_er_and_k_lanes(gpu, abi, orc):
    hg = graph(abi, scale=14, seed=61, kind=abi.HYT_GEN_ER)
    h = abi.HybridEngine(hg, force_engine=-1, **SMALL)
    src = [int(x) for x in abi.sample_sources(hg.offsets, 62, 4)]
    dk, st = h.sssp_k(src)
    for k, s in enumerate(src):
        assert np.array_equal(dk[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s)), k
    assert st["parts_zero_copy"] == 0
