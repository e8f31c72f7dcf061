"""GPU parity: the u64 row-offset instantiations on small graphs.

Graphs with E >= 2^32 - 1 (RMAT scale 28 and up) run every kernel with u64
row offsets. HYT_GRAPH_OFF64 forces those instantiations on a small graph, so
each form is checked against the oracles here: SSSP bit-exact against
Dijkstra, PPR / PageRank within L1 1e-6 of the FP64 delta-push oracle and bit
for bit equal to the u32-offset run (the offsets only locate the edges).
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)


def _say(msg):
    print(f"[gpu-off64] {msg}", flush=True)


def _pair(abi, scale, seed, wmax):
    hg = abi.HostGraph(scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0, weight_max=wmax,
                       memory=abi.HYT_HOST_PINNED)
    g32 = abi.DeviceGraph.from_host(hg)
    g64 = abi.DeviceGraph.from_host(hg, off64=True)
    # the forced upload holds (V+1) u64 offsets instead of u32
    assert g64.info()["device_bytes"] - g32.info()["device_bytes"] == 4 * (hg.V + 1)
    return hg, g32, g64


def test_off64_single_source_sssp(gpu, abi, orc):
    hg, g32, g64 = _pair(abi, 13, 51, 255)
    for src in abi.sample_sources(hg.offsets, 3, 4):
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        for mode in (abi.HYT_SSSP_SYNC, abi.HYT_SSSP_ASYNC):
            for direction in (abi.HYT_DIR_PUSH, abi.HYT_DIR_PULL):
                _say(f"sssp src {src} mode {mode} dir {direction}")
                d, st = g64.sssp(int(src), mode=mode, direction=direction)
                assert np.array_equal(d, ref), (src, mode, direction)
        # small split: light rows and chunked heavy rows both run
        d, _ = g64.sssp(int(src), mode=abi.HYT_SSSP_ASYNC, light_max_degree=4, heavy_chunk=64)
        assert np.array_equal(d, ref)


def test_off64_pagerank_and_ppr(gpu, abi, orc):
    hg, g32, g64 = _pair(abi, 13, 52, 0)
    eps = 1e-8
    same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps))
    for direction in (abi.HYT_DIR_PUSH, abi.HYT_DIR_PULL):
        r64, _ = g64.pagerank(D, eps, direction=direction)
        r32, _ = g32.pagerank(D, eps, direction=direction)
        assert np.abs(r64 - same).sum() <= 1e-6
        assert np.array_equal(r64, r32)
    for s in abi.sample_sources(hg.offsets, 4, 3):
        same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=int(s))
        r64, _ = g64.ppr(int(s), D, eps)
        assert np.abs(r64 - same).sum() <= 1e-6
        assert np.array_equal(r64, g32.ppr(int(s), D, eps)[0])


@pytest.mark.parametrize("form", [1, 2])
def test_off64_batched(gpu, abi, orc, form):
    hg, g32, g64 = _pair(abi, 13, 53, 255)
    src = abi.sample_sources(hg.offsets, 5, 8)
    d, _ = g64.sssp(src, lane_form=form, light_max_degree=8, heavy_chunk=64)
    for k, s in enumerate(src):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s))), k
    hu, u32g, u64g = _pair(abi, 13, 54, 0)
    r64, _ = u64g.ppr(src, D, 1e-8, lane_form=form)
    r32, _ = u32g.ppr(src, D, 1e-8, lane_form=form)
    assert np.array_equal(r64, r32)
    for k, s in enumerate(src):
        same, _, _ = orc.delta_push_f64(hu.offsets, hu.columns, D, f32(1e-8), seed=int(s))
        assert np.abs(r64[:, k] - same).sum() <= 1e-6, k


def test_off64_lane_sessions(gpu, abi, orc):
    hg, _, g64 = _pair(abi, 12, 55, 255)
    srcs = [int(s) for s in abi.sample_sources(hg.offsets, 6, 6)]
    S = abi.LaneSession(g64, abi.HYT_QUERY_SSSP, lanes=3)
    got = _run_lanes(abi, S, srcs, 3)
    for s in srcs:
        assert np.array_equal(got[s], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s)), s
    hu, u32g, u64g = _pair(abi, 12, 56, 0)
    P = abi.LaneSession(u64g, abi.HYT_QUERY_PPR, lanes=3, damping=D, epsilon=1e-8)
    got = _run_lanes(abi, P, srcs, 3)
    for s in srcs:
        assert np.array_equal(got[s], u32g.ppr(s, D, 1e-8)[0]), s


def _run_lanes(abi, S, srcs, lanes):
    pending = list(srcs)
    owner = [None] * lanes
    got = {}
    for lane in range(lanes):
        owner[lane] = pending.pop(0)
        S.admit(lane, owner[lane])
    while any(o is not None for o in owner):
        fin, _ = abi.LaneSession.step([S], 1)
        for lane in range(lanes):
            if (fin[0] >> lane) & 1:
                got[owner[lane]] = S.read(lane)[0]
                S.release(lane)
                owner[lane] = pending.pop(0) if pending else None
                if owner[lane] is not None:
                    S.admit(lane, owner[lane])
    return got


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
def test_off64_hybrid(gpu, abi, orc, engine):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=57, weight_min=1, weight_max=255, memory=abi.HYT_HOST_PINNED)
    cfg = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=8192 * 20 + 16 + 12 * 2048 * 8,
               calibrate=0, force_engine=engine, flags=abi.HYT_GRAPH_OFF64)
    h = abi.HybridEngine(hg, **cfg)
    src = abi.sample_sources(hg.offsets, 8, 4)
    for s in src[:2]:
        _say(f"hybrid engine {engine} sssp src {s}")
        d, _ = h.sssp(int(s))
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s)))
    dk, _ = h.sssp_k(src)
    for k, s in enumerate(src):
        assert np.array_equal(dk[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s))), k
    hu = abi.HostGraph(scale=13, edge_factor=16, seed=58, weight_min=0, weight_max=0, memory=abi.HYT_HOST_PINNED)
    hy = abi.HybridEngine(hu, **cfg)
    gu = abi.DeviceGraph.from_host(hu)
    r, _ = hy.pagerank(D, 1e-8)
    assert np.array_equal(r, gu.pagerank(D, 1e-8)[0])
    rk, _ = hy.ppr_k(src, D, 1e-8)
    for k, s in enumerate(src):
        assert np.array_equal(rk[:, k], gu.ppr(int(s), D, 1e-8)[0]), k
