"""GPU parity over random small graphs and options (fixed seeds): every
entry point against the oracle on each case, not on chosen inputs.

SSSP (SYNC / ASYNC, captured / persistent / host rounds, random degree
splits, waiting-row bounds, u32 / u64 offsets) and batched K-lane SSSP
bit-exact against Dijkstra; PageRank and PPR (single and K lanes) within L1
1e-6 of the FP64 delta-push restatement, lanes equal to the single query bit
for bit; the hybrid engine (random partition size, mover, cache) bit-exact
for SSSP and K-lane SSSP, PageRank equal to the in-HBM ranks; served traces
over graphs in HBM and behind hybrid engines, each result against the
oracle."""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)


def _say(msg):
    print(f"[gpu-random-sweep] {msg}", flush=True)


@pytest.mark.parametrize("block", range(3))
def test_entry_points_on_random_graphs(gpu, abi, orc, block):
    rng = np.random.default_rng(700 + block)
    for case in range(8):
        scale = int(rng.integers(8, 14))
        ef = int(rng.choice([2, 4, 8, 16]))
        kind = int(rng.choice([0, 0, 1]))
        wmax = int(rng.choice([0, 1, 4, 64, 255, 4096]))
        seed = int(rng.integers(1, 10**6))
        hg = abi.HostGraph(kind=kind, scale=scale, edge_factor=ef, seed=seed, weight_min=1 if wmax else 0,
                           weight_max=wmax, memory=abi.HYT_HOST_PINNED)
        off, col, w = hg.offsets, hg.columns, hg.weights
        srcs = [int(s) for s in abi.sample_sources(off, seed + 1, 3)] + [int(rng.integers(0, hg.V))]
        off64 = bool(rng.integers(0, 2))
        g = abi.DeviceGraph.from_host(hg, off64=off64)
        lm = int(rng.choice([0, 0, 1, 8, 32, 300]))
        ch = int(rng.choice([0, 64, 256, 4096])) if lm else 0
        hd = int(rng.choice([0, 1, 2, 7, 0xFFFFFFFF]))
        wd = int(rng.choice([0, 0, 4, 16]))
        _say(f"block {block} case {case}: kind {kind} scale {scale} ef {ef} wmax {wmax} seed {seed} "
             f"off64 {off64} split {lm}/{ch} heavy_delta {hd} wait {wd}")
        refs = {s: orc.sssp_dijkstra(off, col, w, s) for s in srcs}
        for s in srcs:
            for mode in (0, 1):
                for pers in (0, 1):
                    d, _ = g.sssp(s, mode=mode, persistent=pers, light_max_degree=lm, heavy_chunk=ch,
                                  heavy_delta=hd, wait_degree=wd)
                    assert np.array_equal(d, refs[s]), (s, mode, pers)
            d, _ = g.sssp(s, mode=1, host_rounds=1, light_max_degree=lm, heavy_chunk=ch, heavy_delta=hd)
            assert np.array_equal(d, refs[s]), (s, "host rounds")
        K = int(rng.choice([2, 3, 8, 17, 64]))
        ks = np.array([srcs[i] if i < len(srcs) else int(rng.integers(0, hg.V)) for i in range(K)], np.uint32)
        dk, _ = g.sssp(ks, lane_form=int(rng.integers(0, 3)))
        for k in range(K):
            assert np.array_equal(dk[:, k], orc.sssp_dijkstra(off, col, w, int(ks[k]))), ("batched", k)
        eps = float(rng.choice([1e-6, 1e-7]))
        r, _ = g.pagerank(D, eps, direction=int(rng.integers(0, 3)))
        same, _, _ = orc.delta_push_f64(off, col, D, f32(eps))
        assert np.abs(r - same).sum() <= 1e-6
        pk, _ = g.ppr(ks[:4], D, 1e-6)
        for k in range(min(4, K)):
            p1, _ = g.ppr(int(ks[k]), D, 1e-6)
            assert np.array_equal(pk[:, k], p1), ("ppr lanes", k)
            ref, _, _ = orc.delta_push_f64(off, col, D, f32(1e-6), seed=int(ks[k]))
            assert np.abs(p1 - ref).sum() <= 1e-6, ("ppr", k)
        pe = int(rng.choice([256, 1024, 4096]))
        budget = 608 * pe + (1 << 16)
        eng = abi.HybridEngine(hg, partition_edges=pe, pack_edges=4 * pe, device_budget_bytes=budget, calibrate=0,
                               force_engine=int(rng.choice([-1, -1, 0, 1, 2])), cache=int(rng.integers(0, 2)))
        for s in srcs[:2]:
            d, _ = eng.sssp(s)
            assert np.array_equal(d, refs[s]), ("hybrid", s)
        rr, _ = eng.pagerank(D, eps, direction=int(rng.integers(0, 3)))
        assert np.array_equal(rr, r), "hybrid pagerank"
        dk2, _ = eng.sssp_k(ks[:8])
        for k in range(min(8, K)):
            assert np.array_equal(dk2[:, k], orc.sssp_dijkstra(off, col, w, int(ks[k]))), ("hybrid k", k)
        eng.free()
        g.free()


@pytest.mark.parametrize("block", range(2))
def test_served_traces_on_random_graphs(gpu, abi, orc, block):
    rng = np.random.default_rng(800 + block)
    for trace in range(10):
        ng = int(rng.integers(1, 4))
        graphs, hosts, engines = [], [], []
        for gi in range(ng):
            hg = abi.HostGraph(kind=int(rng.choice([0, 0, 1])), scale=int(rng.integers(9, 13)),
                               edge_factor=int(rng.choice([4, 8, 16])), seed=int(rng.integers(1, 10**6)),
                               weight_min=1, weight_max=int(rng.choice([1, 64, 255])), memory=abi.HYT_HOST_PINNED)
            hosts.append(hg)
            if rng.integers(0, 4) == 0:
                pe = int(rng.choice([512, 2048]))
                conf = dict(partition_edges=pe, pack_edges=4 * pe, device_budget_bytes=pe * 8 * 200 + (1 << 20),
                            calibrate=0)
                pair = (abi.HybridEngine(hg, **conf), abi.HybridEngine(hg, **conf))
                graphs.append(pair)
                engines += list(pair)
            else:
                graphs.append(abi.DeviceGraph.from_host(hg))
        n = int(rng.integers(10, 40))
        arrivals = np.cumsum(rng.exponential(1.0 / float(rng.choice([200, 2000, 20000])), n))
        reqs = []
        for i in range(n):
            gi = int(rng.integers(0, ng))
            kindq = abi.HYT_QUERY_SSSP if rng.random() < 0.7 else abi.HYT_QUERY_PPR
            src = (int(abi.sample_sources(hosts[gi].offsets, int(rng.integers(1, 10**6)), 1)[0])
                   if rng.random() < 0.8 else int(rng.integers(0, hosts[gi].V)))
            reqs.append((float(arrivals[i]), gi, kindq, src))
        eps = float(rng.choice([1e-6, 1e-7]))
        _say(f"block {block} trace {trace}: {ng} graphs, {n} requests, eps {eps}")
        comp, st, res = abi.serve_trace(graphs, reqs, keep_results=True, epsilon=eps)
        for i, (_, gi, kindq, src) in enumerate(reqs):
            hg = hosts[gi]
            if kindq == abi.HYT_QUERY_SSSP:
                assert np.array_equal(res[i], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)), i
            else:
                same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=src)
                assert np.abs(res[i] - same).sum() <= 1e-6, i
        for e in engines:
            e.free()
