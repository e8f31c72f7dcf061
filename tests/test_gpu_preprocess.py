"""GPU parity: vertex order (hub sort, reorder) and contributions on the device.

Bar: bit-exact against the oracle (orders, hub counts, f64 score bits, CSR
arrays, f64 contribution bits), the oracle being pinned against the compiled
original. Traversals of a reordered device graph give the original graph's
results with the ids permuted (SSSP distances and PageRank ranks bit for bit:
both are independent of the vertex numbering).
"""
import numpy as np
import pytest

from conftest import csr_from_edges, f32

pytestmark = pytest.mark.gpu
D = f32(0.85)


def _say(msg):
    print(f"[gpu-preprocess] {msg}", flush=True)


def _graph(abi, kind, scale, seed, wmax=255):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0,
                         weight_max=wmax)


@pytest.mark.parametrize("kind,scale,seed", [(0, 10, 1), (0, 14, 2), (1, 13, 3), (0, 16, 4)])
def test_hub_sort_matches_oracle(gpu, abi, orc, kind, scale, seed):
    hg = _graph(abi, kind, scale, seed)
    for off64 in (False, True):
        g = abi.DeviceGraph.from_host(hg, off64=off64)
        for frac in (0.0, 0.08, 1.0 / 3.0, 0.5, 1.0):
            _say(f"kind {kind} scale {scale} off64 {off64} fraction {frac}")
            order, scores, hubs = g.hub_sort(frac)
            ro, rs, rh = orc.hub_sort(hg.offsets, hg.columns, frac)
            assert hubs == rh and np.array_equal(order, ro), frac
            assert np.array_equal(scores.view(np.uint64), rs.view(np.uint64)), frac


def test_hub_sort_known_answers_and_edge_cases(gpu, abi, orc):
    # the original's unit_tests.cpp:2468-2602 graphs
    g = abi.DeviceGraph(np.array([0, 2, 3, 4], np.uint64), np.array([1, 2, 2, 0], np.uint32))
    _, s, _ = g.hub_sort(2.0 / 3.0)
    assert np.allclose(s, [0.5, 0.25, 0.5], atol=1e-12, rtol=0)
    g = abi.DeviceGraph(np.array([0, 3, 4, 5, 6, 7], np.uint64), np.array([1, 2, 3, 0, 0, 0, 0], np.uint32))
    order, _, hubs = g.hub_sort(0.4)
    assert hubs == 2 and order.tolist() == [0, 1, 2, 3, 4]
    g = abi.DeviceGraph(np.zeros(5, np.uint64), np.zeros(0, np.uint32))
    order, s, hubs = g.hub_sort()
    assert hubs == 1 and order.tolist() == [0, 1, 2, 3] and not s.any()
    g1 = abi.DeviceGraph(np.array([0, 1], np.uint64), np.array([0], np.uint32))  # one vertex, a self-loop
    order, s, hubs = g1.hub_sort(1.0)
    assert order.tolist() == [0] and s.tolist() == [1.0] and hubs == 1
    with pytest.raises(abi.InvalidArgument):
        g1.hub_sort(1.5)
    with pytest.raises(abi.InvalidArgument):
        g1.hub_sort(float("nan"))
    # parallel edges and self-loops count in both degrees
    off, col, _ = csr_from_edges(6, [(0, 1), (0, 1), (0, 0), (1, 2), (2, 1), (3, 1), (3, 1), (5, 5), (4, 1)])
    g = abi.DeviceGraph(off, col)
    for frac in (0.2, 0.5, 1.0):
        order, s, hubs = g.hub_sort(frac)
        ro, rs, rh = orc.hub_sort(off, col, frac)
        assert np.array_equal(order, ro) and hubs == rh and np.array_equal(s, rs)


def test_hub_sort_device_outputs(gpu, abi, orc, hip):
    hg = _graph(abi, 0, 12, 5)
    g = abi.DeviceGraph.from_host(hg)
    po, ps = hip.malloc(hg.V * 4), hip.malloc(hg.V * 8)
    try:
        _, _, hubs = g.hub_sort(0.08, out_ptr=po, scores_ptr=ps)
        ro, rs, rh = orc.hub_sort(hg.offsets, hg.columns, 0.08)
        assert hubs == rh
        assert np.array_equal(hip.to_host(po, hg.V, np.uint32), ro)
        assert np.array_equal(hip.to_host(ps, hg.V, np.float64), rs)
    finally:
        hip.free(po)
        hip.free(ps)


@pytest.mark.parametrize("weights", ["u32", "f32", None])
def test_reorder_matches_oracle(gpu, abi, orc, rng, weights):
    hg = _graph(abi, 0, 13, 6, wmax=0 if weights is None else 255)
    w = None
    if weights == "u32":
        w = hg.weights
    elif weights == "f32":
        w = (hg.weights.astype(np.float32) * np.float32(0.37)).astype(np.float32)
    wbits = None if w is None else np.ascontiguousarray(w).view(np.uint32)
    hub_order = orc.hub_sort(hg.offsets, hg.columns, 0.08)[0]
    for off64 in (False, True):
        for perm in (hub_order, rng.permutation(hg.V).astype(np.uint32)):
            g = abi.DeviceGraph(hg.offsets, hg.columns, w, weights_f32=weights == "f32", off64=off64)
            g.reorder(perm)
            got = g.download()
            want = orc.reorder(hg.offsets, hg.columns, wbits, perm)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
            assert (got[2] is None) == (want[2] is None)
            if want[2] is not None:
                assert np.array_equal(got[2], want[2])


def test_reorder_device_order_and_errors(gpu, abi, orc, hip):
    hg = _graph(abi, 0, 11, 7)
    g = abi.DeviceGraph.from_host(hg)
    before = g.download()
    V = hg.V
    # the first offending position decides the message; the graph is unchanged
    bad = [(np.r_[np.arange(V - 1), V + 3], "out-of-range"), (np.r_[np.arange(V - 1), 0], "duplicate"),
           (np.r_[V, np.zeros(V - 1)], "out-of-range"), (np.r_[0, 0, np.arange(V - 3), V + 9], "duplicate")]
    for order, what in bad:
        with pytest.raises(abi.InvalidArgument, match=what):
            g.reorder(order.astype(np.uint32))
    with pytest.raises(abi.InvalidArgument):
        g.reorder(np.arange(V - 1, dtype=np.uint32))
    after = g.download()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    # order in device memory: the device hub order applied on the device
    po = hip.malloc(V * 4)
    try:
        g.hub_sort(0.08, out_ptr=po)
        g.reorder(order_ptr=po)
    finally:
        hip.free(po)
    want = orc.reorder(hg.offsets, hg.columns, hg.weights, orc.hub_sort(hg.offsets, hg.columns, 0.08)[0])
    got = g.download()
    assert all(np.array_equal(a, b) for a, b in zip(got, want))


def test_reordered_graph_traversals_permute(gpu, abi, orc):
    """SSSP and PageRank / PPR on the hub-ordered device graph equal the
    original graph's results under the id map, bit for bit."""
    hg = _graph(abi, 0, 14, 8)
    g0 = abi.DeviceGraph.from_host(hg)
    g = abi.DeviceGraph.from_host(hg)
    order, _, _ = g.hub_sort(0.08)
    g.reorder(order)
    new_of_old = np.empty(hg.V, np.uint32)
    new_of_old[order] = np.arange(hg.V, dtype=np.uint32)
    for src in abi.sample_sources(hg.offsets, 3, 3):
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        for mode in (abi.HYT_SSSP_SYNC, abi.HYT_SSSP_ASYNC):
            d, st = g.sssp(int(new_of_old[src]), mode=mode)
            assert np.array_equal(d[new_of_old], ref), (src, mode)
            if mode == abi.HYT_SSSP_SYNC:
                assert st["rounds"] == g0.sssp(int(src), mode=mode)[1]["rounds"]
    hu = _graph(abi, 0, 14, 9, wmax=0)
    gu0 = abi.DeviceGraph.from_host(hu)
    gu = abi.DeviceGraph.from_host(hu)
    order, _, _ = gu.hub_sort(0.08)
    gu.reorder(order)
    new_of_old[order] = np.arange(hu.V, dtype=np.uint32)
    r0, s0 = gu0.pagerank(D, 1e-7)
    r1, s1 = gu.pagerank(D, 1e-7)
    assert np.array_equal(r1[new_of_old], r0) and s0["rounds"] == s1["rounds"]
    p0, _ = gu0.ppr(5, D, 1e-6)
    p1, _ = gu.ppr(int(new_of_old[5]), D, 1e-6)
    assert np.array_equal(p1[new_of_old], p0)


@pytest.mark.parametrize("kind,scale,seed", [(0, 12, 11), (1, 12, 12), (0, 15, 13)])
def test_pagerank_contributions_match_oracle(gpu, abi, orc, rng, kind, scale, seed):
    hg = _graph(abi, kind, scale, seed, wmax=0)
    g = abi.DeviceGraph.from_host(hg)
    V = hg.V
    xs = [np.full(V, np.float32(1.0) / np.float32(V), np.float32), rng.random(V, np.float32) / np.float32(V),
          (rng.random(V) * 10).astype(np.float32), g.pagerank(D, 1e-7)[0]]
    for x in xs:
        for d in (D, 0.5, 0.0):
            got = g.pagerank_contributions(x, d)
            want = orc.pagerank_contributions(hg.offsets, hg.columns, x, d)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), d
    with pytest.raises(abi.InvalidArgument):
        g.pagerank_contributions(xs[0], 1.0)
    with pytest.raises(abi.InvalidArgument):
        g.pagerank_contributions(xs[0][:-1], D)


def test_pagerank_contributions_parallel_edges_and_dangling(gpu, abi, orc):
    off, col, _ = csr_from_edges(7, [(0, 1), (0, 1), (0, 0), (1, 2), (2, 1), (3, 1), (3, 6), (3, 1), (6, 6)])
    g = abi.DeviceGraph(off, col)
    x = np.array([0.3, 0.1, 0.05, 0.2, 0.15, 0.12, 0.08], np.float32)
    for d in (D, 0.3):
        got = g.pagerank_contributions(x, d)
        assert np.array_equal(got, orc.pagerank_contributions(off, col, x, d))


def test_sssp_contributions_match_oracle(gpu, abi, orc, rng):
    inf, nan = np.inf, np.nan
    cur = np.array([5.0, 5.0, inf, inf, 3.0, nan, 2.0, 1e30], np.float32)
    cand = np.array([3.0, 4.75, 7.0, inf, inf, 1.0, 2.0, 1e29], np.float32)
    got = abi.sssp_contributions(cur, cand, 0.25)
    assert np.array_equal(got, orc.sssp_contributions(cur, cand, 0.25))
    cur = rng.random(100_003).astype(np.float32) * 100
    cand = cur - rng.random(100_003).astype(np.float32) * 2
    cur[::7] = inf
    cand[::11] = inf
    for tol in (0.0, 0.5, 1.5):
        got = abi.sssp_contributions(cur, cand, tol)
        assert np.array_equal(got.view(np.uint64), orc.sssp_contributions(cur, cand, tol).view(np.uint64))
    assert abi.sssp_contributions([], [], 0.0).size == 0
    with pytest.raises(abi.InvalidArgument):
        abi.sssp_contributions([1.0], [0.0], -1.0)
    with pytest.raises(abi.InvalidArgument):
        abi.sssp_contributions([1.0], [0.0], float("inf"))
