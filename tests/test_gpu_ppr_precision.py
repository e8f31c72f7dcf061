"""GPU parity sweep: PageRank / PPR precision against the FP64 oracle.

Bar (north star): normalised L1 <= 1e-6 against the FP64 restatement of the
same delta-push procedure (oracle.delta_push_f64), given the float damping and
epsilon the C ABI carries. Residuals and ranks are 64-bit fixed-point words
summed by integer atomics; the only roundings are each row's float share
(relative 2^-23) and the float output (2^-24), a bound of
d/(1-d) * 2^-23 + 2^-24 = 7.4e-7 at d = 0.85.

The sweep covers 12 PPR seeds per graph on RMAT and Erdos-Renyi graphs, global
PageRank on several graphs, every form (single query with push and pull
rounds, batched K = 8 and 64, lane sessions, the hybrid engine single and
K-lane) and prints the worst case of each; every form must also give the
single-query words bit for bit.
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)
BAR = 1e-6


def _say(msg):
    print(f"[gpu-ppr-sweep] {msg}", flush=True)


GRAPHS = [(0, 13, 41), (0, 12, 42), (1, 12, 43)]  # (kind: 0 RMAT / 1 ER, scale, seed)


def _host(abi, kind, scale, seed, pinned=False):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=0, weight_max=0,
                         memory=abi.HYT_HOST_PINNED if pinned else abi.HYT_HOST_PAGEABLE)


@pytest.mark.parametrize("kind,scale,gseed", GRAPHS)
@pytest.mark.parametrize("eps", [1e-6, 1e-8])
def test_ppr_seed_sweep_all_forms(gpu, abi, orc, kind, scale, gseed, eps):
    hg = _host(abi, kind, scale, gseed)
    g = abi.DeviceGraph.from_host(hg)
    seeds = abi.sample_sources(hg.offsets, 100 + gseed, 12)
    same = [orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=int(s))[0] for s in seeds]
    worst = {}

    def check(form, k, r):
        l1 = float(np.abs(r - same[k]).sum())
        worst[form] = max(worst.get(form, 0.0), l1)
        assert l1 <= BAR, (form, k, l1)

    singles = []
    for k, s in enumerate(seeds):
        _say(f"graph {kind}/{scale}/{gseed} eps {eps} seed {s}")
        r, st = g.ppr(int(s), D, eps)
        assert st["converged"] == 1
        check("single", k, r)
        singles.append(r)
        rp, _ = g.ppr(int(s), D, eps, direction=abi.HYT_DIR_PULL)
        assert np.array_equal(rp, r), k  # pull rounds sum the same words
    for K in (8, 12):
        rk, _ = g.ppr(seeds[:K], D, eps)
        for k in range(K):
            check(f"K={K}", k, rk[:, k])
            assert np.array_equal(rk[:, k], singles[k]), (K, k)
    r64, _ = g.ppr(np.resize(seeds, 64), D, eps)
    for k in range(64):
        assert np.array_equal(r64[:, k], singles[k % 12]), k
    _say(f"graph {kind}/{scale}/{gseed} eps {eps}: worst L1 " +
         ", ".join(f"{f} {v:.2e}" for f, v in worst.items()))


@pytest.mark.parametrize("kind,scale,gseed", GRAPHS)
def test_pagerank_sweep(gpu, abi, orc, kind, scale, gseed):
    hg = _host(abi, kind, scale, gseed)
    g = abi.DeviceGraph.from_host(hg)
    for eps in (1e-6, 1e-7, 1e-8):
        same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps))
        for direction in (abi.HYT_DIR_AUTO, abi.HYT_DIR_PUSH, abi.HYT_DIR_PULL):
            r, st = g.pagerank(D, eps, direction=direction)
            l1 = float(np.abs(r - same).sum())
            _say(f"pagerank {kind}/{scale}/{gseed} eps {eps} dir {direction}: L1 {l1:.2e}")
            assert st["converged"] == 1 and l1 <= BAR


def test_ppr_sweep_lane_sessions_and_hybrid(gpu, abi, orc):
    hg = _host(abi, 0, 13, 41, pinned=True)
    g = abi.DeviceGraph.from_host(hg)
    eps = 1e-8
    seeds = [int(s) for s in abi.sample_sources(hg.offsets, 77, 12)]
    singles = {s: g.ppr(s, D, eps)[0] for s in seeds}
    for s in seeds:
        same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=s)
        assert float(np.abs(singles[s] - same).sum()) <= BAR, s
    # lane session: 4 lanes, 12 queries admitted as lanes free up
    S = abi.LaneSession(g, abi.HYT_QUERY_PPR, lanes=4, damping=D, epsilon=eps)
    pending = list(seeds)
    owner = [None] * 4
    got = {}
    for lane in range(4):
        owner[lane] = pending.pop(0)
        S.admit(lane, owner[lane])
    while any(o is not None for o in owner):
        fin, _ = abi.LaneSession.step([S], 1)
        for lane in range(4):
            if (fin[0] >> lane) & 1:
                got[owner[lane]] = S.read(lane)[0]
                S.release(lane)
                owner[lane] = pending.pop(0) if pending else None
                if owner[lane] is not None:
                    S.admit(lane, owner[lane])
    for s in seeds:
        assert np.array_equal(got[s], singles[s]), s
    S.free()
    # hybrid engine (edges in pinned host memory), every mover
    for engine in (-1, 0, 1, 2):
        h = abi.HybridEngine(hg, force_engine=engine, partition_edges=4096, pack_edges=16384,
                             device_budget_bytes=16384 * 20 + 16 + 16 * 4096 * 8, calibrate=0)
        rk, _ = h.ppr_k(np.array(seeds[:8], np.uint32), D, eps)
        for k in range(8):
            assert np.array_equal(rk[:, k], singles[seeds[k]]), (engine, k)
        r1, _ = h.ppr_k(np.array(seeds[8:9], np.uint32), D, eps)
        assert np.array_equal(r1[:, 0], singles[seeds[8]]), engine
        h.free()
    _say("lane sessions and hybrid engine give the single-query words")
