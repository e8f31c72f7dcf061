"""GPU parity: batched [V][K] SSSP and PPR through the C ABI.

Bar: every SSSP column equals the single-source oracle bit-exactly; every PPR
column is within the stated L1 tolerance of the FP64 oracle of its seed.
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)  # the damping the C ABI carries (float)


def _say(msg):
    print(f"[gpu-batched] {msg}", flush=True)


@pytest.mark.parametrize("K", [2, 5, 8, 32, 64])
def test_sssp_k_matches_single_source(gpu, abi, orc, K):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=21, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 100 + K, K)
    src[-1] = src[0]  # a repeated source shares a row
    _say(f"K={K} sources {src[:4].tolist()}...")
    d, st = g.sssp(src)
    assert d.shape == (hg.V, K)
    assert st["converged"] == 1
    for k in range(K):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[k]))), k


def test_sssp_k_float_and_unweighted(gpu, abi, orc):
    hg = abi.HostGraph(scale=12, edge_factor=8, seed=5, weight_min=1, weight_max=64)
    src = np.array([0, 7, 9, 1000], np.uint32)
    gf = abi.DeviceGraph.from_host(hg, weights_f32=True)
    d, _ = gf.sssp(src)
    for k, s in enumerate(src):
        ref, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), int(s))
        assert np.array_equal(d[:, k], ref.view(np.uint32))
    gu = abi.DeviceGraph(hg.offsets, hg.columns)
    d, _ = gu.sssp(src)
    for k, s in enumerate(src):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, None, int(s)))


@pytest.mark.parametrize("light_max,chunk", [(4, 64), (1 << 20, 1 << 16)])
def test_sssp_k_work_split(gpu, abi, orc, light_max, chunk):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=4, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 3, 16)
    d, _ = g.sssp(src, light_max_degree=light_max, heavy_chunk=chunk)
    for k in range(16):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[k])))


def test_sssp_k_repeated_calls_reuse_state(gpu, abi, orc):
    """The [V][K] workspace is reused across calls of different K: marks must
    start clean every time."""
    hg = abi.HostGraph(scale=11, edge_factor=16, seed=6, weight_min=1, weight_max=64)
    g = abi.DeviceGraph.from_host(hg)
    for K in (8, 3, 16, 8):
        src = abi.sample_sources(hg.offsets, 40 + K, K)
        d, _ = g.sssp(src)
        for k in range(K):
            assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[k])))
        d1, _ = g.sssp(int(src[0]))
        assert np.array_equal(d1, d[:, 0])


@pytest.mark.parametrize("K", [2, 8, 64])
def test_ppr_k_matches_fp64(gpu, abi, orc, K):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=2, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    seeds = abi.sample_sources(hg.offsets, 7 + K, K)
    eps = 1e-8
    r, st = g.ppr(seeds, D, eps)
    assert r.shape == (hg.V, K) and st["converged"] == 1
    worst_same = worst_exact = 0.0
    for k in range(K):
        same, _, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=int(seeds[k]))
        exact = orc.pagerank_f64(hg.offsets, hg.columns, D, seed=int(seeds[k]))
        worst_same = max(worst_same, np.abs(r[:, k] - same).sum())
        worst_exact = max(worst_exact, np.abs(r[:, k] - exact).sum())
        assert abs(r[:, k].sum() - 1) < 1e-5
    _say(f"K={K}: worst L1 vs same-procedure {worst_same:.2e}, vs exact {worst_exact:.2e}")
    assert worst_same <= 1e-6
    assert worst_exact <= 1e-4


def test_ppr_k_column_equals_single(gpu, abi):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=3, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    seeds = np.array([5, 900, 5, 17], np.uint32)
    r, _ = g.ppr(seeds, D, 1e-9)
    for k, s in enumerate(seeds):
        one, _ = g.ppr(int(s), D, 1e-9)
        # fixed-point mass words: a lane gives the single query's bits
        assert np.array_equal(r[:, k], one), k


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("K", [3, 8, 64])
def test_k_lane_forms_match_oracle(gpu, abi, orc, K, form):
    # group form (next_pow2(K) lanes per edge) and item form (one lane per
    # (edge, query) pair) give the same answers; small light_max/chunk so
    # both the light and the chunked heavy kernels of each form run
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=31, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 50 + K, K)
    src[-1] = src[0]
    for lm, ch in ((256, 1024), (4, 64)):
        d, _ = g.sssp(src, lane_form=form, light_max_degree=lm, heavy_chunk=ch)
        for k in range(K):
            assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[k]))), k
    hu = abi.HostGraph(scale=12, edge_factor=16, seed=32, weight_min=0, weight_max=0)
    gu = abi.DeviceGraph.from_host(hu)
    r, st = gu.ppr(src, D, 1e-8, lane_form=form, light_max_degree=4, heavy_chunk=64)
    for k in range(K):
        same, _, _ = orc.delta_push_f64(hu.offsets, hu.columns, D, f32(1e-8), seed=int(src[k]))
        assert np.abs(r[:, k] - same).sum() <= 1e-6, k


@pytest.mark.parametrize("form", [1, 2])
def test_k_lane_full_chunk_queue_falls_back(gpu, abi, orc, form):
    # a dense graph (1024 vertices, ~0.4 M edges) with tiny light/chunk
    # thresholds needs far more heavy chunks than the queue holds (about
    # E/256 + V/16): the rows that do not fit are walked by their own
    # wavefront, and the answers stay exact
    hg = abi.HostGraph(kind=abi.HYT_GEN_ER, scale=10, edge_factor=512, seed=17, weight_min=1, weight_max=255)
    assert hg.E > 300_000
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 9, 64)
    d, _ = g.sssp(src, lane_form=form, light_max_degree=4, heavy_chunk=64)
    for k in range(0, 64, 7):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[k]))), k
    hu = abi.HostGraph(kind=abi.HYT_GEN_ER, scale=10, edge_factor=512, seed=18, weight_min=0, weight_max=0)
    gu = abi.DeviceGraph.from_host(hu)
    r, _ = gu.ppr(src, D, 1e-8, lane_form=form, light_max_degree=4, heavy_chunk=64)
    for k in range(0, 64, 9):
        same, _, _ = orc.delta_push_f64(hu.offsets, hu.columns, D, f32(1e-8), seed=int(src[k]))
        assert np.abs(r[:, k] - same).sum() <= 1e-6, k


def test_trim_releases_and_rebuilds_workspaces(gpu, abi, orc):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=23, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 5, 8)
    g.sssp(src)                                   # batched workspace
    g.sssp(int(src[0]), direction=abi.HYT_DIR_PULL)  # transpose
    before = g.info()["device_bytes"]
    g.trim()
    after = g.info()["device_bytes"]
    assert after < before
    d1, _ = g.sssp(int(src[1]), direction=abi.HYT_DIR_PULL)
    dk, _ = g.sssp(src)
    for k, s in enumerate(src):
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s))
        assert np.array_equal(dk[:, k], ref)
    assert np.array_equal(d1, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src[1])))


def test_k_lane_round_limit_and_next_call(gpu, abi, orc):
    # device-driven rounds stop at max_rounds (not converged), and the next
    # call starts from clean marks
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=29, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 6, 8)
    _, st = g.sssp(src, max_rounds=2)
    assert st["rounds"] == 2 and st["converged"] == 0
    d, st = g.sssp(src)
    assert st["converged"] == 1 and st["rounds"] > 2
    for k, s in enumerate(src):
        assert np.array_equal(d[:, k], orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s)))
    _, pst = g.ppr(src, D, 1e-6, max_rounds=3)
    assert pst["rounds"] == 3 and pst["converged"] == 0
