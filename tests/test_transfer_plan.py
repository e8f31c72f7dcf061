"""Hybrid-transfer planning on the CPU: partitioning, per-partition metrics,
cost equations, engine selection and task combination.

Pinned by the original's own known answers (tests/unit_tests.cpp:803-916,
2058-2266, 2268-2466) and, on random graphs / activities / parameters,
by exact equality with the original code compiled from source (oracle/_ref).
"""
import numpy as np
import pytest

needs_ref = pytest.mark.skipif(not __import__("oracle").ref_available(),
                               reason="oracle/_ref not built (original source tree absent)")


def _graph(abi, scale=11, seed=1, ef=16):
    g = abi.HostGraph(scale=scale, edge_factor=ef, seed=seed, weight_min=0, weight_max=0)
    return g.offsets.copy(), g.columns.copy()


# ------------------------------------------------------------- known answers
def test_cost_model_known_answer(abi):
    # unit_tests.cpp:2058-2158: graph {0->1,0->2,1->2,2->0}, active {0,1},
    # partition [0,2) edges [0,3); rtt 2, cpu 14 B/s
    off = np.array([0, 2, 3, 4, 4], np.uint64)
    p = abi.cost_params(rtt=2.0, cpu_bytes_per_second=14.0)
    m = abi.partition_metrics(off, [1, 1, 0, 0], np.array([[0, 2, 0, 3]], np.uint64), p)
    av, ae, te, fb, cb, zr, za, zt = m[0].tolist()
    assert (av, ae, te) == (2, 3, 3)
    assert fb == 12 and cb == 28 and (zr, za, zt) == (2, 1, 3)
    costs, eng = abi.cost_evaluate(m, p)
    f, c, z, rtt_zc = costs[0]
    assert f == 2.0 and c == 4.0 and rtt_zc == 2.0 and z == 2.0


def test_selector_strict_boundaries(abi):
    # unit_tests.cpp:2159-2204: strict comparisons, equality falls through
    F, C, Z = abi.HYT_ENGINE_FILTER, abi.HYT_ENGINE_COMPACTION, abi.HYT_ENGINE_ZERO_COPY
    assert abi.select_engine(10.0, 7.0, 20.0, 0.8, 0.4) == C      # 7 < 8 and 7 < 8
    assert abi.select_engine(10.0, 8.0, 30.0, 0.8, 0.4) == F      # 8 == 0.8*10: not compaction
    assert abi.select_engine(10.0, 7.0, 17.5, 0.8, 0.4) == F      # 7 == 0.4*17.5
    assert abi.select_engine(10.0, 9.0, 10.0, 0.8, 0.4) == Z      # f == z -> zero-copy
    assert abi.select_engine(9.0, 9.0, 10.0, 0.8, 0.4) == F
    with pytest.raises(abi.InvalidArgument):
        abi.select_engine(float("inf"), 1.0, 1.0, 0.8, 0.4)
    with pytest.raises(abi.InvalidArgument):
        abi.select_engine(1.0, -1.0, 1.0, 0.8, 0.4)
    with pytest.raises(abi.InvalidArgument):
        abi.select_engine(1.0, 1.0, 1.0, 0.0, 0.4)


def test_cost_params_validation(abi):
    # unit_tests.cpp:2205-2266
    m = np.zeros((1, 8), np.uint64)
    for bad in ({"alpha": 0.0}, {"beta": -1.0}, {"gamma": 1.5}, {"rtt": -1.0}, {"cpu_bytes_per_second": 0.0},
                {"request_bytes": 0}, {"requests_per_tlp": 0}, {"entry_bytes": 0}, {"index_bytes": 0},
                {"alignment_bytes": 0}):
        with pytest.raises(abi.InvalidArgument):
            abi.cost_evaluate(m, abi.cost_params(**bad))


def test_zero_copy_request_arithmetic(abi):
    # unit_tests.cpp:1835-1880: 40 edges * 4 B = 160 B at m=16 -> 10 requests;
    # MR=2 -> 5 TLPs; zero-degree vertex -> no request (unit_tests.cpp:2018-2056)
    off = np.array([0, 40, 40], np.uint64)
    p = abi.cost_params(request_bytes=16, requests_per_tlp=2, rtt=1.0, gamma=1.0)
    m = abi.partition_metrics(off, [1, 1], np.array([[0, 2, 0, 40]], np.uint64), p)
    assert m[0][5] == 10 and m[0][6] == 0
    costs, _ = abi.cost_evaluate(m, p)
    assert costs[0][2] == 5.0


def test_empty_partition_zero_copy_rtt(abi):
    off = np.array([0, 0, 0], np.uint64)
    p = abi.cost_params(rtt=4.0)
    m = abi.partition_metrics(off, [1, 0], np.array([[0, 2, 0, 0]], np.uint64), p)
    costs, eng = abi.cost_evaluate(m, p)
    assert costs[0][3] == 0.625 * 4.0


def test_partition_boundaries(abi):
    # unit_tests.cpp:803-916 style: rows stay whole, oversized rows alone
    off = np.array([0, 3, 4, 14, 15, 16], np.uint64)   # degrees 3,1,10,1,1
    parts = abi.partition_graph(off, 16)                 # 16 B = 4 columns
    assert parts.tolist() == [[0, 2, 0, 4], [2, 3, 4, 14], [3, 5, 14, 16]]
    # every edge covered once, rows contiguous
    assert parts[0][2] == 0 and parts[-1][3] == 16
    assert all(parts[i][1] == parts[i + 1][0] for i in range(len(parts) - 1))
    with pytest.raises(abi.InvalidArgument):
        abi.partition_graph(off, 0)


def test_task_combiner_groupings(abi):
    # unit_tests.cpp:2268-2466: filter runs <= k, compaction / zero-copy merged,
    # ordered by first member
    F, C, Z = 0, 1, 2
    tasks = abi.combine_tasks([F, F, F, F, F, C, Z, F, C, Z, F, F], k=4)
    assert tasks == [(F, [0, 1, 2, 3]), (F, [4]), (C, [5, 8]), (Z, [6, 9]), (F, [7]), (F, [10, 11])]
    assert abi.combine_tasks([], k=4) == []
    with pytest.raises(abi.InvalidArgument):
        abi.combine_tasks([F], k=0)
    with pytest.raises(abi.InvalidArgument):
        abi.combine_tasks([7], k=4)


# -------------------------------------------------- against the compiled original
@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("target", [64, 1000, 4 << 10, 1 << 20])
def test_partition_matches_original(abi, orc, seed, target):
    off, col = _graph(abi, 11, seed)
    assert np.array_equal(abi.partition_graph(off, target), orc.ref_partition(off, col, target))


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_costs_match_original(abi, orc, seed):
    rng = np.random.default_rng(seed)
    off, col = _graph(abi, 11, seed)
    V = len(off) - 1
    active = (rng.random(V) < rng.choice([0.01, 0.1, 0.5, 0.9])).astype(np.uint8)
    alpha, beta, gamma = rng.uniform(0.2, 1.5), rng.uniform(0.1, 1.0), rng.uniform(0, 1)
    rtt, cpu = rng.uniform(0.1, 5.0), rng.uniform(1e2, 1e6)
    m, mr = int(rng.choice([16, 64, 128])), int(rng.choice([2, 64, 256]))
    d1, d2, align = int(rng.choice([4, 8])), int(rng.choice([4, 8])), int(rng.choice([64, 128]))
    target = int(rng.choice([256, 4096, 65536]))
    rm, rc, re = orc.ref_cost_evaluate(off, col, active, target, alpha, beta, gamma, rtt, cpu, m, mr, d1, d2, align)
    parts = abi.partition_graph(off, target)
    p = abi.cost_params(alpha=alpha, beta=beta, gamma=gamma, rtt=rtt, cpu_bytes_per_second=cpu, request_bytes=m,
                        requests_per_tlp=mr, entry_bytes=d1, index_bytes=d2, alignment_bytes=align)
    mine_m = abi.partition_metrics(off, active, parts, p)
    mine_c, mine_e = abi.cost_evaluate(mine_m, p)
    assert np.array_equal(mine_m, rm)
    assert np.array_equal(mine_c, rc)          # same formulas, same double arithmetic
    assert np.array_equal(mine_e, re)
    assert len(set(re.tolist())) >= 1


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_combine_matches_original(abi, orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 60))
    engines = rng.integers(0, 3, n).astype(np.int32)
    k = int(rng.integers(1, 6))
    assert abi.combine_tasks(engines, k) == orc.ref_combine(engines, k)
