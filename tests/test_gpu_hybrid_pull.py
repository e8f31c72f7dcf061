"""GPU parity: hybrid PageRank pull rounds over the host graph's incoming CSR.

Bar: the ranks equal the in-HBM PageRank's bit for bit (both add the same
64-bit words, in any order) for push-only, pull-only and auto direction,
with every in-partition resident, none, or constant eviction; in-rows longer
than a ring window (split over windows, chunked), pageable host graphs and
u64 device offsets; runs that switch between pull and push rounds and
between PageRank and SSSP on one engine keep their results. The ranks are
within L1 1e-6 of the FP64 delta-push restatement (the original's
pagerank_cpu fixed point, src/algorithms/pagerank.cpp:89-163).
"""
import numpy as np
import pytest

from conftest import csr_from_edges, f32

pytestmark = pytest.mark.gpu
D = f32(0.85)
EPS = 1e-7


def _say(msg):
    print(f"[gpu-hybrid-pull] {msg}", flush=True)


def check_transpose(hg, V):
    # the incoming CSR the engine built (on the device when it fits): the
    # stable sort of the edges by destination, sources ascending
    in_off, in_src = hg.transpose()
    src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(hg.offsets.astype(np.int64)))
    order = np.argsort(hg.columns, kind="stable")
    want_off = np.zeros(V + 1, np.uint64)
    np.cumsum(np.bincount(hg.columns, minlength=V), out=want_off[1:])
    assert np.array_equal(in_off, want_off)
    assert np.array_equal(in_src, src[order])


def host(abi, scale=13, seed=2, wmax=0, kind=0, memory=None):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0,
                         weight_max=wmax, memory=abi.HYT_HOST_PINNED if memory is None else memory)


# partitions of 2048 edges, a ring of 8 partitions per buffer, a few units
SMALL = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=8192 * 20 + 16 + 12 * 2048 * 8,
             calibrate=0)
# room for every partition of a scale-13 graph (~130 k edges, in and out)
LARGE = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=64 << 20, calibrate=0)


@pytest.mark.parametrize("cfg", ["small", "large", "nocache"])
def test_pull_push_auto_equal_in_hbm(gpu, abi, orc, cfg):
    hg = host(abi)
    conf = {"small": SMALL, "large": LARGE, "nocache": dict(SMALL, cache=0)}[cfg]
    h = abi.HybridEngine(hg, **conf)
    g = abi.DeviceGraph.from_host(hg)
    ref, rst = g.pagerank(D, EPS)
    same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(EPS))
    for direction in (abi.HYT_DIR_PUSH, abi.HYT_DIR_PULL, abi.HYT_DIR_AUTO):
        r, st = h.pagerank(D, EPS, direction=direction)
        run = st["run"]
        _say(f"{cfg} direction {direction}: rounds {run['rounds']} pull {run['pull_rounds']} "
             f"cached {st['parts_cached']} filter {st['parts_filter']} transpose {st['transpose_ms']:.1f} ms")
        assert np.array_equal(r, ref), direction
        assert run["converged"] == 1 and run["rounds"] == rst["rounds"]
        assert np.abs(r - same).sum() <= 1e-6
        if direction == abi.HYT_DIR_PULL:
            check_transpose(hg, hg.V)
        if direction == abi.HYT_DIR_PUSH:
            assert run["pull_rounds"] == 0
        elif direction == abi.HYT_DIR_PULL:
            assert run["pull_rounds"] == run["rounds"]
        else:
            assert 0 < run["pull_rounds"] < run["rounds"]
    # a second pull run: with the large budget every in-partition that fits a
    # unit (2048 in-edges) stays resident, so only longer in-rows move
    r, st = h.pagerank(D, EPS, direction=abi.HYT_DIR_PULL)
    assert np.array_equal(r, ref) and st["transpose_ms"] == 0.0
    in_off, _ = hg.transpose()
    indeg = np.diff(in_off.astype(np.int64))
    if cfg == "large":
        assert st["bytes_filter"] == st["run"]["rounds"] * 4 * int(indeg[indeg > 2048].sum())
        assert st["parts_cached"] > 0
    if cfg == "nocache":
        assert st["parts_cached"] == 0 and st["bytes_filter"] == st["run"]["rounds"] * 4 * hg.E


def test_long_in_rows_split_over_windows(gpu, abi, orc):
    # a destination with 40000 in-edges (ring windows hold 16384), one with
    # 300 (chunked), one with 33, a chain and isolated vertices
    V = 50000
    edges = [(u, 0) for u in range(1, 40001)] + [(u, 7) for u in range(100, 400)] + \
            [(u, 9) for u in range(500, 533)] + [(u, u + 1) for u in range(40001, 49000)] + [(0, 1), (7, 0)]
    off, col, _ = csr_from_edges(V, edges)
    hg = abi.HostGraph.from_csr(off, col, memory=abi.HYT_HOST_PINNED)
    g = abi.DeviceGraph(off, col)
    ref, _ = g.pagerank(D, EPS)
    for conf in (SMALL, dict(SMALL, cache=0), LARGE):
        h = abi.HybridEngine(hg, **conf)
        for direction in (abi.HYT_DIR_PULL, abi.HYT_DIR_AUTO):
            r, st = h.pagerank(D, EPS, direction=direction)
            _say(f"long rows direction {direction}: rounds {st['run']['rounds']} pull {st['run']['pull_rounds']}")
            assert np.array_equal(r, ref), direction
        h.free()
    check_transpose(hg, V)
    same, _, _ = orc.delta_push_f64(off, col, D, f32(EPS))
    assert np.abs(ref - same).sum() <= 1e-6


def test_pageable_off64_and_er(gpu, abi, orc):
    hp = host(abi, scale=13, seed=8, memory=abi.HYT_HOST_PAGEABLE)
    g = abi.DeviceGraph.from_host(hp)
    ref, _ = g.pagerank(D, EPS)
    h = abi.HybridEngine(hp, **SMALL)
    r, st = h.pagerank(D, EPS)
    assert np.array_equal(r, ref) and st["run"]["pull_rounds"] > 0
    check_transpose(hp, hp.V)
    h64 = abi.HybridEngine(hp, flags=abi.HYT_GRAPH_OFF64, **SMALL)
    r, st = h64.pagerank(D, EPS, direction=abi.HYT_DIR_PULL)
    assert np.array_equal(r, ref)
    er = host(abi, scale=13, seed=9, kind=abi.HYT_GEN_ER)
    ge = abi.DeviceGraph.from_host(er)
    he = abi.HybridEngine(er, **SMALL)
    r, st = he.pagerank(D, EPS)
    assert np.array_equal(r, ge.pagerank(D, EPS)[0]) and st["run"]["pull_rounds"] > 0


def test_switching_runs_keep_results(gpu, abi, orc):
    # weighted graph: PageRank (pull + push rounds, in-partitions cached),
    # then SSSP (out-partitions with weights), then PageRank again
    hg = host(abi, scale=13, seed=5, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    ref, _ = g.pagerank(D, EPS)
    h = abi.HybridEngine(hg, **dict(SMALL, device_budget_bytes=8 << 20))
    r1, s1 = h.pagerank(D, EPS)
    src = int(abi.sample_sources(hg.offsets, 6, 1)[0])
    d, _ = h.sssp(src)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
    r2, s2 = h.pagerank(D, EPS)
    _say(f"switching: cached {s1['parts_cached']} -> {s2['parts_cached']}")
    assert np.array_equal(r1, ref) and np.array_equal(r2, ref)
    with pytest.raises(abi.InvalidArgument):
        h.pagerank(D, EPS, direction=3)
