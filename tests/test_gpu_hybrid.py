"""GPU parity: the hybrid transfer engine (edges in pinned host memory).

Bar: SSSP distances bit-exact against the oracle whatever engine moves the
edges (cost model, or forced filter / compaction / zero-copy), with caching
on or off, budgets small enough to force eviction, oversized hub partitions
and multi-batch packing; PageRank within the stated L1 tolerances of the
FP64 oracles (same bars as the in-HBM path).
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)  # the damping the C ABI carries (float)


def _say(msg):
    print(f"[gpu-hybrid] {msg}", flush=True)


def pinned(abi, scale=14, seed=3, wmax=255, kind=0):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0,
                         weight_max=wmax, memory=abi.HYT_HOST_PINNED)


# partitions of 2048 edges; budget = packed staging + streaming ring + a few
# cached partitions -> constant eviction
SMALL = dict(partition_edges=2048, pack_edges=8192, device_budget_bytes=8192 * 20 + 16 + 12 * 2048 * 8,
             calibrate=0)


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
@pytest.mark.parametrize("cache", [0, 1])
def test_hybrid_sssp_bitexact(gpu, abi, orc, engine, cache):
    hg = pinned(abi)
    h = abi.HybridEngine(hg, force_engine=engine, cache=cache, **SMALL)
    for src in abi.sample_sources(hg.offsets, 5, 2):
        _say(f"engine {engine} cache {cache} src {src}")
        d, st = h.sssp(int(src))
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src)))
        assert st["run"]["converged"] == 1
        if engine == 0:
            assert st["bytes_filter"] > 0 and st["parts_filter"] > 0
        if engine == 1:
            assert st["bytes_compaction"] > 0 and st["parts_compaction"] > 0
        if engine == 2:
            assert st["parts_zero_copy"] > 0


def test_hybrid_oversized_rows_and_cost_model_mix(gpu, abi, orc):
    # partitions of 512 edges: RMAT hubs (degree in the thousands) are oversized
    hg = pinned(abi, scale=15, seed=9)
    deg = np.diff(hg.offsets.astype(np.int64))
    assert deg.max() > 4 * 512
    # budget: packed staging + a streaming ring of 4 partitions + 8 cached ones
    h = abi.HybridEngine(hg, partition_edges=512, pack_edges=4096,
                         device_budget_bytes=4096 * 20 + 16 + 4 * 512 * 8 + 8 * 512 * 8,
                         calibrate=0, cost=abi.cost_params(entry_bytes=8, rtt=1e-6, cpu_bytes_per_second=2e9))
    d, st = h.sssp(0)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0))
    _say(f"mix: filter {st['parts_filter']} compaction {st['parts_compaction']} zc {st['parts_zero_copy']} "
         f"cached {st['parts_cached']}")


def test_hybrid_float_weights_and_unweighted(gpu, abi, orc):
    hg = pinned(abi, scale=13, seed=4, wmax=64)
    ref, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 0)
    # float bit patterns in the same pinned arrays
    hg.weights[:] = hg.weights.astype(np.float32).view(np.uint32)
    h = abi.HybridEngine(hg, weights_f32=True, force_engine=-1, **SMALL)
    d, _ = h.sssp(0)
    assert np.array_equal(d, ref.view(np.uint32))
    hu = pinned(abi, scale=13, seed=6, wmax=0)
    h = abi.HybridEngine(hu, force_engine=1, **SMALL)
    d, _ = h.sssp(3)
    assert np.array_equal(d, orc.sssp_dijkstra(hu.offsets, hu.columns, None, 3))


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
def test_hybrid_pagerank(gpu, abi, orc, engine):
    hg = pinned(abi, scale=13, seed=2, wmax=0)
    h = abi.HybridEngine(hg, force_engine=engine, **SMALL)
    eps = 1e-7
    r, st = h.pagerank(D, eps)
    same, rounds, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps))
    exact = orc.pagerank_f64(hg.offsets, hg.columns, D)
    _say(f"engine {engine}: L1 same {np.abs(r - same).sum():.2e} exact {np.abs(r - exact).sum():.2e} "
         f"rounds {st['run']['rounds']}/{rounds}")
    assert st["run"]["converged"] == 1
    assert np.abs(r - same).sum() <= 1e-6
    assert np.abs(r - exact).sum() <= 2 * D * eps / (1 - D) ** 2 + 2e-6


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
def test_hybrid_erdos_renyi(gpu, abi, orc, engine):
    # the low-skew contrast graph of the out-of-memory configuration: same V
    # and E as RMAT, uniform (src, dst); SSSP bit-exact, PageRank and K-lane
    # PPR the in-HBM words, through every mover
    hg = pinned(abi, scale=14, seed=61, kind=abi.HYT_GEN_ER)
    h = abi.HybridEngine(hg, force_engine=engine, **SMALL)
    src = abi.sample_sources(hg.offsets, 62, 3)
    for s in src:
        _say(f"ER engine {engine} sssp src {s}")
        d, st = h.sssp(int(s))
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s)))
    dk, _ = h.sssp_k(src)
    for k, s in enumerate(src):
        assert np.array_equal(dk[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s))), k
    hu = pinned(abi, scale=14, seed=63, wmax=0, kind=abi.HYT_GEN_ER)
    hy = abi.HybridEngine(hu, force_engine=engine, **SMALL)
    g = abi.DeviceGraph.from_host(hu)
    eps = 1e-7
    r, st = hy.pagerank(D, eps)
    same, _, _ = orc.delta_push_f64(hu.offsets, hu.columns, D, f32(eps))
    assert st["run"]["converged"] == 1 and np.abs(r - same).sum() <= 1e-6
    assert np.array_equal(r, g.pagerank(D, eps)[0])
    rk, _ = hy.ppr_k(src, D, 1e-8)
    for k, s in enumerate(src):
        assert np.array_equal(rk[:, k], g.ppr(int(s), D, 1e-8)[0]), k


@pytest.mark.parametrize("cache", [0, 1])
def test_hybrid_contribution_order(gpu, abi, orc, cache):
    # filter runs ordered by device-computed partition contributions (the
    # original ContributionScheduler order) against partition order: same
    # SSSP distances and PageRank words, runs actually reordered
    hg = pinned(abi, scale=14, seed=64)
    h = abi.HybridEngine(hg, force_engine=0, cache=cache, **SMALL)
    for src in abi.sample_sources(hg.offsets, 65, 2):
        d0, s0 = h.sssp(int(src))
        d1, s1 = h.sssp(int(src), task_order=1)
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        assert np.array_equal(d0, ref) and np.array_equal(d1, ref)
        assert s0["runs_reordered"] == 0 and s1["runs_reordered"] > 0
        _say(f"contribution order (cache {cache}) src {src}: edges relaxed {s1['run']['edges_total']} vs "
             f"{s0['run']['edges_total']} in partition order, runs reordered {s1['runs_reordered']}")
    hu = pinned(abi, scale=14, seed=66, wmax=0)
    hp = abi.HybridEngine(hu, force_engine=0, cache=cache, **SMALL)
    r0, p0 = hp.pagerank(D, 1e-7)
    r1, p1 = hp.pagerank(D, 1e-7, task_order=1)
    assert np.array_equal(r0, r1) and p1["runs_reordered"] > 0
    with pytest.raises(abi.InvalidArgument):
        hp.ppr_k([1, 2], D, 1e-6, task_order=1)


def test_hybrid_reuse_and_errors(gpu, abi, orc):
    hg = pinned(abi, scale=12, seed=8)
    h = abi.HybridEngine(hg, **SMALL)
    for src in (0, 5, 0):
        d, _ = h.sssp(src)
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
    with pytest.raises(abi.InvalidArgument):
        h.sssp(hg.V)
    with pytest.raises(abi.InvalidArgument):
        abi.HybridEngine(hg, device_budget_bytes=1024, calibrate=0)
    info = h.info()
    assert info["partitions"] > 10 and info["device_edge_bytes"] <= SMALL["device_budget_bytes"]


def test_hybrid_weights_move_only_when_read(gpu, abi, orc):
    # PageRank moves columns only; SSSP afterwards must not reuse a
    # column-only cached slot (its distances stay bit-exact)
    hw = pinned(abi, scale=13, seed=2, wmax=255)
    hu = pinned(abi, scale=13, seed=2, wmax=0)
    assert np.array_equal(hw.columns, hu.columns)
    kw = dict(SMALL, force_engine=0, cache=1, device_budget_bytes=8192 * 20 + 16 + 400 * 2048 * 8)
    ew, eu = abi.HybridEngine(hw, **kw), abi.HybridEngine(hu, **kw)
    rw, sw = ew.pagerank(D, 1e-6)
    ru, su = eu.pagerank(D, 1e-6)
    _say(f"pagerank filter bytes weighted {sw['bytes_filter']} unweighted {su['bytes_filter']}")
    assert sw["bytes_filter"] == su["bytes_filter"] > 0
    assert np.abs(rw - ru).sum() <= 1e-6
    for src in (0, 7):
        d, st = ew.sssp(src)
        assert np.array_equal(d, orc.sssp_dijkstra(hw.offsets, hw.columns, hw.weights, src))
    # the second SSSP finds its partitions cached with weights
    assert st["parts_cached"] > 0
    r2, _ = ew.pagerank(D, 1e-6)
    assert np.abs(r2 - rw).sum() <= 1e-6


# ------------------------------------------------------------ K queries per run
def _sources(abi, hg, n, seed):
    s = [int(x) for x in abi.sample_sources(hg.offsets, seed, n)]
    return s + [s[0]]  # a repeated source shares its row with another lane


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
def test_hybrid_sssp_k_bitexact(gpu, abi, orc, engine):
    hg = pinned(abi, scale=13, seed=5)
    h = abi.HybridEngine(hg, force_engine=engine, **SMALL)
    srcs = _sources(abi, hg, 8, 7)
    refs = {s: orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s) for s in set(srcs)}
    for form in (0, 1, 2):
        _say(f"sssp_k engine {engine} form {form} K={len(srcs)}")
        d, st = h.sssp_k(srcs, lane_form=form)
        assert d.shape == (hg.V, len(srcs))
        for k, s in enumerate(srcs):
            assert np.array_equal(d[:, k], refs[s]), (form, k)
        assert st["run"]["converged"] == 1
        if engine == 0:
            assert st["parts_filter"] > 0
        if engine == 1:
            assert st["parts_compaction"] > 0
        if engine == 2:
            assert st["parts_zero_copy"] > 0


def test_hybrid_sssp_k_oversized_rows_single_lane_and_round_limit(gpu, abi, orc):
    # hubs above the packed staging go through the listed-row (zero-copy) path
    hg = pinned(abi, scale=15, seed=9)
    deg = np.diff(hg.offsets.astype(np.int64))
    assert deg.max() > 2048
    h = abi.HybridEngine(hg, partition_edges=512, pack_edges=2048, force_engine=1, calibrate=0,
                         device_budget_bytes=2048 * 20 + 16 + 4 * 512 * 8 + 8 * 512 * 8)
    hub = int(np.argmax(deg))
    srcs = [hub, 0, 17, 4242, 9]
    d, st = h.sssp_k(srcs)
    for k, s in enumerate(srcs):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, s)), k
    # K = 1 equals the single-query engine
    d1, _ = h.sssp_k([hub])
    ds, _ = h.sssp(hub)
    assert np.array_equal(d1[:, 0], ds)
    # a run stopped by the round limit leaves the engine ready for the next one
    _, st = h.sssp_k(srcs, max_rounds=2)
    assert st["run"]["converged"] == 0 and st["run"]["rounds"] == 2
    d2, _ = h.sssp_k(srcs)
    assert np.array_equal(d2, d)
    with pytest.raises(abi.InvalidArgument):
        h.sssp_k([])
    with pytest.raises(abi.InvalidArgument):
        h.sssp_k([0] * 65)
    with pytest.raises(abi.InvalidArgument):
        h.sssp_k([0, hg.V])


@pytest.mark.parametrize("engine", [-1, 0, 1, 2])
def test_hybrid_ppr_k(gpu, abi, orc, engine):
    hg = pinned(abi, scale=12, seed=2, wmax=0)
    h = abi.HybridEngine(hg, force_engine=engine, **SMALL)
    eps = 1e-8
    seeds = _sources(abi, hg, 5, 3)
    g = abi.DeviceGraph.from_host(hg)
    alone, _ = g.ppr(seeds, D, eps)
    for form in (0, 2):
        r, st = h.ppr_k(seeds, D, eps, lane_form=form)
        assert st["run"]["converged"] == 1
        worst = 0.0
        for k, s in enumerate(seeds):
            same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=s)
            worst = max(worst, np.abs(r[:, k] - same).sum())
            assert np.abs(r[:, k] - same).sum() <= 1e-6, (form, k)
            assert np.array_equal(r[:, k], alone[:, k]), (form, k)  # same words through the engine
            assert abs(r[:, k].sum() - 1) < 1e-5
        _say(f"ppr_k engine {engine} form {form}: worst L1 vs same-procedure oracle {worst:.2e}")


def test_hybrid_from_file_and_caller_csr(gpu, abi, orc, tmp_path):
    # a binary CSR file and a caller's own CSR arrays, loaded into pinned
    # memory, feed the hybrid engine like generated graphs
    hg = abi.HostGraph(scale=12, seed=4)
    path = tmp_path / "g.hytcsr"
    hg.save(path)
    ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 3)
    for g in (abi.HostGraph.load(path, memory=abi.HYT_HOST_PINNED),
              abi.HostGraph.from_csr(hg.offsets, hg.columns, hg.weights, memory=abi.HYT_HOST_PINNED)):
        h = abi.HybridEngine(g, **SMALL)
        d, _ = h.sssp(3)
        assert np.array_equal(d, ref)
        dk, _ = h.sssp_k([3, 0])
        assert np.array_equal(dk[:, 0], ref)
