"""GPU parity: continuous batching (lane sessions and the serving loop).

Bar: each query's result is the same alone and in the mix. SSSP distances are
bit-exact against the Dijkstra oracle whatever lanes run beside a query and
whenever it was admitted; PPR is within L1 1e-6 of the FP64 delta-push oracle
of the same seed (same procedure, eps 1e-8) and within 1e-4 of the exact
personalized PageRank, and bit-identical to the single-query hyt_ppr result
(residuals and ranks are fixed-point words summed by integer atomics, so the
order in which lanes' pushes land changes nothing).
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)  # the damping the C ABI carries (float)


def _say(msg):
    print(f"[gpu-serve] {msg}", flush=True)


def _graph(abi, scale, seed, wmax=255):
    hg = abi.HostGraph(scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0, weight_max=wmax)
    return hg, abi.DeviceGraph.from_host(hg)


def _drain(abi, sessions, owners, pending, results, rounds=1):
    """Steps until every lane is free and nothing is pending; re-admits
    pending sources into lanes as they free up."""
    while True:
        fin, run = abi.LaneSession.step(sessions, rounds)
        for i, S in enumerate(sessions):
            for lane in range(64):
                if (fin[i] >> lane) & 1:
                    out, reached, checksum, nr = S.read(lane)
                    results[i].append((owners[i][lane], out, reached, checksum, nr))
                    S.release(lane)
                    owners[i][lane] = None
                    if pending[i]:
                        src = pending[i].pop(0)
                        S.admit(lane, src)
                        owners[i][lane] = src
        if not any(run) and not any(pending) and not any(o is not None for ow in owners for o in ow):
            return


@pytest.mark.parametrize("form", [0, 1, 2])
def test_sssp_lanes_staggered_admission(gpu, abi, orc, form):
    hg, g = _graph(abi, 12, 11)
    S = abi.LaneSession(g, abi.HYT_QUERY_SSSP, lanes=4)
    S.set_form(form)
    srcs = [int(s) for s in abi.sample_sources(hg.offsets, 3, 11)]
    owners = [[None] * 4]
    pending = [srcs[4:]]
    results = [[]]
    S.admit(0, srcs[0])
    owners[0][0] = srcs[0]
    abi.LaneSession.step([S], 1)
    S.admit(1, srcs[1])
    owners[0][1] = srcs[1]
    abi.LaneSession.step([S], 2)
    for lane, s in ((2, srcs[2]), (3, srcs[3])):
        S.admit(lane, s)
        owners[0][lane] = s
    with pytest.raises(abi.InvalidArgument):
        S.admit(0, 5)              # occupied
    with pytest.raises(abi.InvalidArgument):
        S.read(1)                  # still running
    with pytest.raises(abi.InvalidArgument):
        S.admit(4, 5)              # no such lane
    _drain(abi, [S], owners, pending, results)
    assert sorted(r[0] for r in results[0]) == sorted(srcs)
    for src, d, reached, checksum, nr in results[0]:
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)
        assert np.array_equal(d, ref), src
        fin = ref != 0xFFFFFFFF
        assert reached == fin.sum() and checksum == int(ref[fin].astype(np.uint64).sum()) and nr > 0
    assert S.info()["occupied"] == 0


@pytest.mark.parametrize("form", [0, 1, 2])
def test_ppr_lanes_staggered_admission(gpu, abi, orc, form):
    hg, g = _graph(abi, 12, 2, wmax=0)
    eps = 1e-8
    S = abi.LaneSession(g, abi.HYT_QUERY_PPR, lanes=3, damping=D, epsilon=eps)
    S.set_form(form)
    seeds = [int(s) for s in abi.sample_sources(hg.offsets, 21, 7)] + [5, 5]
    owners = [[None] * 3]
    pending = [seeds[2:]]
    results = [[]]
    S.admit(0, seeds[0])
    owners[0][0] = seeds[0]
    abi.LaneSession.step([S], 3)
    S.admit(2, seeds[1])
    owners[0][2] = seeds[1]
    _drain(abi, [S], owners, pending, results)
    assert sorted(r[0] for r in results[0]) == sorted(seeds)
    worst = 0.0
    for seed, r, _, _, _ in results[0]:
        same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=seed)
        exact = orc.pagerank_f64(hg.offsets, hg.columns, D, seed=seed)
        alone, _ = g.ppr(seed, D, eps)
        worst = max(worst, np.abs(r - same).sum())
        assert np.abs(r - same).sum() <= 1e-6
        assert np.abs(r - exact).sum() <= 1e-4
        assert np.array_equal(r, alone)
        assert abs(r.sum() - 1) < 1e-5
    _say(f"ppr lanes: worst L1 vs same-procedure oracle {worst:.2e}")


def test_sessions_of_two_graphs_overlap(gpu, abi, orc):
    hs, gs = _graph(abi, 11, 5)
    hp, gp = _graph(abi, 12, 6, wmax=0)
    A = abi.LaneSession(gs, abi.HYT_QUERY_SSSP, lanes=8)
    B = abi.LaneSession(gp, abi.HYT_QUERY_PPR, lanes=2, damping=D, epsilon=1e-8)
    sa = [int(s) for s in abi.sample_sources(hs.offsets, 1, 12)]
    sb = [int(s) for s in abi.sample_sources(hp.offsets, 2, 4)]
    owners = [[None] * 8, [None] * 2]
    for lane in range(8):
        A.admit(lane, sa[lane])
        owners[0][lane] = sa[lane]
    for lane in range(2):
        B.admit(lane, sb[lane])
        owners[1][lane] = sb[lane]
    results = [[], []]
    _drain(abi, [A, B], owners, [sa[8:], sb[2:]], results, rounds=2)
    for src, d, _, _, _ in results[0]:
        assert np.array_equal(d, orc.sssp_dijkstra(hs.offsets, hs.columns, hs.weights, src))
    for seed, r, _, _, _ in results[1]:
        same, _, _ = orc.delta_push_f64(hp.offsets, hp.columns, D, f32(1e-8), seed=seed)
        assert np.abs(r - same).sum() <= 1e-6
    assert len(results[0]) == 12 and len(results[1]) == 4


def test_serve_trace_mixed_results_match_alone(gpu, abi, orc):
    graphs = [_graph(abi, 11, 7), _graph(abi, 12, 8)]
    rng = np.random.default_rng(4)
    n = 48
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, n))  # ~2000 queries/s offered
    reqs = []
    for i in range(n):
        gi = int(rng.integers(0, 2))
        kind = abi.HYT_QUERY_SSSP if rng.random() < 0.7 else abi.HYT_QUERY_PPR
        src = int(abi.sample_sources(graphs[gi][0].offsets, 100 + i, 1)[0])
        reqs.append((float(arrivals[i]), gi, kind, src))
    # budget: the graphs plus room for only a few lanes per session, so
    # queries queue for lanes and are admitted as lanes free up
    gbytes = sum(g.info()["device_bytes"] for _, g in graphs)
    lane_room = sum(abi.lanes_state_bytes(hg.V, hg.E, int(np.diff(hg.offsets.astype(np.int64)).max()), k, 3)
                    for hg, _ in graphs for k in (abi.HYT_QUERY_SSSP, abi.HYT_QUERY_PPR))
    comp, st, res = abi.serve_trace([g for _, g in graphs], reqs, keep_results=True,
                                    device_budget_bytes=gbytes + lane_room + 4096, epsilon=1e-8)
    _say(f"serve: {st}")
    assert 0 < st["item_rounds"] <= st["rounds"]
    # every session got a lane and the lanes fit the budget (admission by memory)
    budget = gbytes + lane_room + 4096
    assert st["sessions"] == 4 and all(k >= 1 for k in st["lanes"])
    assert st["graph_bytes"] + st["state_bytes"] <= budget
    assert st["admitted"] == n
    for i, (t, gi, kind, src) in enumerate(reqs):
        assert t - 1e-3 <= comp["admit_s"][i] <= comp["done_s"][i]
        hg = graphs[gi][0]
        if kind == abi.HYT_QUERY_SSSP:
            ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)
            assert np.array_equal(res[i], ref), i
            assert comp["reached"][i] == (ref != 0xFFFFFFFF).sum()
        else:
            same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(1e-8), seed=src)
            assert np.abs(res[i] - same).sum() <= 1e-6, i
    # queries overlapped: some were admitted before an earlier one finished
    order = np.argsort(comp["admit_s"])
    overlap = any(comp["admit_s"][order[j + 1]] < comp["done_s"][order[j]] for j in range(n - 1))
    assert overlap


def test_serve_budget_too_small(gpu, abi):
    hg, g = _graph(abi, 10, 9)
    with pytest.raises(abi.HytError) as ei:
        abi.serve_trace([g], [(0.0, 0, abi.HYT_QUERY_SSSP, 0)], device_budget_bytes=4096)
    assert ei.value.status == abi.HYT_ERR_INVALID_ARGUMENT
    # room for the graph but not for one lane: out of memory, nothing runs
    gb = g.info()["device_bytes"]
    with pytest.raises(abi.HytError) as ei:
        abi.serve_trace([g], [(0.0, 0, abi.HYT_QUERY_SSSP, 0)], device_budget_bytes=gb + 1024)
    assert ei.value.status == abi.HYT_ERR_OUT_OF_MEMORY


# ----------------------------------------------- graphs served from host memory
def _pinned(abi, scale, seed, wmax=255):
    return abi.HostGraph(scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0, weight_max=wmax,
                         memory=abi.HYT_HOST_PINNED)


# small partitions and an edge budget of a few cached partitions: every
# round mixes cached, streamed, packed and zero-copy partitions
HSMALL = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=8192 * 20 + 16 + 12 * 2048 * 8,
              calibrate=0)


@pytest.mark.parametrize("form", [0, 2])
def test_hybrid_sssp_session_staggered_admission(gpu, abi, orc, form):
    hg = _pinned(abi, 13, 21)
    eng = abi.HybridEngine(hg, **HSMALL)
    S = abi.LaneSession(eng, abi.HYT_QUERY_SSSP, lanes=4)
    S.set_form(form)
    srcs = [int(s) for s in abi.sample_sources(hg.offsets, 9, 10)] + [0]
    owners = [[None] * 4]
    results = [[]]
    S.admit(0, srcs[0])
    owners[0][0] = srcs[0]
    abi.LaneSession.step([S], 2)
    S.admit(1, srcs[1])
    owners[0][1] = srcs[1]
    abi.LaneSession.step([S], 1)
    for lane, s in ((2, srcs[2]), (3, srcs[3])):
        S.admit(lane, s)
        owners[0][lane] = s
    with pytest.raises(abi.InvalidArgument):
        S.admit(0, 5)
    _drain(abi, [S], owners, [srcs[4:]], results)
    assert sorted(r[0] for r in results[0]) == sorted(srcs)
    for src, d, reached, checksum, nr in results[0]:
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)
        assert np.array_equal(d, ref), src
        fin = ref != 0xFFFFFFFF
        assert reached == fin.sum() and checksum == int(ref[fin].astype(np.uint64).sum()) and nr > 0


def test_hybrid_ppr_session_and_hbm_session_together(gpu, abi, orc):
    hp = _pinned(abi, 12, 22, wmax=0)
    eng = abi.HybridEngine(hp, **HSMALL)
    A = abi.LaneSession(eng, abi.HYT_QUERY_PPR, lanes=3, damping=D, epsilon=1e-8)
    hs, gs = _graph(abi, 11, 23)
    B = abi.LaneSession(gs, abi.HYT_QUERY_SSSP, lanes=4)
    sa = [int(s) for s in abi.sample_sources(hp.offsets, 5, 5)]
    sb = [int(s) for s in abi.sample_sources(hs.offsets, 6, 6)]
    owners = [[None] * 3, [None] * 4]
    for lane in range(3):
        A.admit(lane, sa[lane])
        owners[0][lane] = sa[lane]
    for lane in range(4):
        B.admit(lane, sb[lane])
        owners[1][lane] = sb[lane]
    results = [[], []]
    _drain(abi, [A, B], owners, [sa[3:], sb[4:]], results)
    assert len(results[0]) == 5 and len(results[1]) == 6
    worst = 0.0
    for seed, r, _, _, _ in results[0]:
        same, _, _ = orc.delta_push_f64(hp.offsets, hp.columns, D, f32(1e-8), seed=seed)
        worst = max(worst, np.abs(r - same).sum())
        assert np.abs(r - same).sum() <= 1e-6
        assert abs(r.sum() - 1) < 1e-5
    for src, d, _, _, _ in results[1]:
        assert np.array_equal(d, orc.sssp_dijkstra(hs.offsets, hs.columns, hs.weights, src))
    _say(f"hybrid ppr session: worst L1 vs same-procedure oracle {worst:.2e}")


def test_serve_trace_host_resident_graph(gpu, abi, orc):
    hs, gs = _graph(abi, 11, 31)
    hp = _pinned(abi, 13, 32)
    eng_s = abi.HybridEngine(hp, **HSMALL)
    eng_p = abi.HybridEngine(hp, **HSMALL)
    rng = np.random.default_rng(8)
    n = 40
    arrivals = np.cumsum(rng.exponential(1.0 / 1500.0, n))
    reqs = []
    for i in range(n):
        gi = int(rng.integers(0, 2))
        kind = abi.HYT_QUERY_SSSP if rng.random() < 0.7 else abi.HYT_QUERY_PPR
        offs = (hs if gi == 0 else hp).offsets
        reqs.append((float(arrivals[i]), gi, kind, int(abi.sample_sources(offs, 200 + i, 1)[0])))
    comp, st, res = abi.serve_trace([gs, (eng_s, eng_p)], reqs, keep_results=True, epsilon=1e-8,
                                    device_budget_bytes=8 << 30, max_lanes=6)
    _say(f"serve host-resident: {st}")
    assert st["admitted"] == n and st["sessions"] == 4
    assert st["parts_cached"] + st["parts_moved"] > 0 and st["parts_moved"] > 0
    for i, (t, gi, kind, src) in enumerate(reqs):
        hg = hs if gi == 0 else hp
        assert t - 1e-3 <= comp["admit_s"][i] <= comp["done_s"][i]
        if kind == abi.HYT_QUERY_SSSP:
            assert np.array_equal(res[i], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)), i
        else:
            same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(1e-8), seed=src)
            assert np.abs(res[i] - same).sum() <= 1e-6, i
    # a host-resident graph needs an engine for every query kind it receives,
    # and an engine serves one graph and one kind
    with pytest.raises(abi.InvalidArgument):
        abi.serve_trace([gs, (eng_s, None)], [(0.0, 1, abi.HYT_QUERY_PPR, 0)])
    with pytest.raises(abi.InvalidArgument):
        abi.serve_trace([gs, (eng_s, eng_s)], [(0.0, 1, abi.HYT_QUERY_SSSP, 0)])
    with pytest.raises(abi.InvalidArgument):
        abi.serve_trace([(eng_s, None), (eng_s, None)], [(0.0, 1, abi.HYT_QUERY_SSSP, 0)])
