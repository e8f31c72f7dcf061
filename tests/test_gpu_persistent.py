"""GPU parity: the persistent single-source SSSP kernel (one cooperative
launch, data-driven queues, grid barriers between rounds) through the C ABI.

Bar: distances bit-exact against Dijkstra / the original's float Jacobi
rounds; in SYNC mode the round count, and the state after a round limit,
equal the original's iterations and state (sssp.cpp:140-232).
"""
import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-persistent] {msg}", flush=True)


def rmat(abi, scale, seed, wmin=1, wmax=64, ef=16, kind=0):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=ef, seed=seed, weight_min=wmin, weight_max=wmax)


@pytest.mark.parametrize("seed", [1, 2])
def test_rmat16_config1_bitexact(gpu, abi, orc, seed):
    """RMAT scale 16, u32 weights [1,64], source 0: both modes equal Dijkstra,
    SYNC reproduces the original's iteration count, float weights give the
    original's float bits."""
    hg = rmat(abi, 16, seed)
    g = abi.DeviceGraph.from_host(hg)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0)
    ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 0)
    for mode in (0, 1):
        _say(f"seed {seed} mode {mode}")
        d, st = g.sssp(0, mode=mode, persistent=1)
        assert np.array_equal(d, dj), mode
        assert st["converged"] == 1
        if mode == 0:
            assert st["rounds"] == it
    gf = abi.DeviceGraph.from_host(hg, weights_f32=True)
    d, st = gf.sssp(0, mode=0, persistent=1)
    assert np.array_equal(d, ref.view(np.uint32)) and st["rounds"] == it


@pytest.mark.parametrize("mode", [0, 1])
def test_matches_captured_rounds(gpu, abi, orc, mode):
    """Against the captured-round path on the same graph and sources: the
    same distances; in SYNC mode the same rounds, frontier and update totals
    (both relax each frontier row once per round)."""
    hg = rmat(abi, 15, 81, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    for src in abi.sample_sources(hg.offsets, 9, 5):
        _say(f"mode {mode} source {int(src)}")
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        dp, sp = g.sssp(int(src), mode=mode, persistent=1)
        dd, sd = g.sssp(int(src), mode=mode)
        assert np.array_equal(dp, ref) and np.array_equal(dd, ref), int(src)
        if mode == 0:
            for k in ("rounds", "converged", "frontier_total", "updates_total", "edges_total"):
                assert sp[k] == sd[k], (k, sp[k], sd[k])
        assert sp["edges_total"] >= 1


@pytest.mark.parametrize("max_rounds", [1, 2, 3, 5])
def test_truncated_sync_rounds_match_original_state(gpu, abi, orc, max_rounds):
    """SYNC persistent rounds are round-for-round Jacobi: the state after k
    rounds equals the original's after max_iterations = k."""
    hg = rmat(abi, 12, 9)
    g = abi.DeviceGraph.from_host(hg, weights_f32=True)
    ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 1, 0.0, max_rounds)
    d, st = g.sssp(1, mode=0, max_rounds=max_rounds, persistent=1)
    assert np.array_equal(d, ref.view(np.uint32))
    assert st["rounds"] == it and bool(st["converged"]) == conv


@pytest.mark.parametrize("light_max,chunk", [(0, 0), (4, 64), (64, 128), (1 << 20, 1 << 16)])
def test_degree_binning_thresholds(gpu, abi, orc, light_max, chunk):
    """Light rows only, tiny chunks (hubs over many chunks and rounds), huge
    chunks: the same distances in both modes."""
    hg = rmat(abi, 14, 6, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0)
    ref, it, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 0)
    for mode in (0, 1):
        d, st = g.sssp(0, mode=mode, light_max_degree=light_max, heavy_chunk=chunk, persistent=1)
        assert np.array_equal(d, dj), mode
        if mode == 0:
            assert st["rounds"] == it


def test_weight_kinds_offsets_and_er(gpu, abi, orc, rng):
    """Unweighted, fractional float weights, u64 device offsets and an ER
    graph."""
    hg = rmat(abi, 13, 3, wmin=0, wmax=0)
    g = abi.DeviceGraph.from_host(hg)
    for mode in (0, 1):
        d, _ = g.sssp(0, mode=mode, persistent=1)
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, None, 0))
    hw = rmat(abi, 12, 5)
    w = (hw.weights.astype(np.float32) * rng.choice(np.array([0.1, 0.37, 1.75], np.float32),
                                                    size=hw.E)).astype(np.float32)
    gf = abi.DeviceGraph(hw.offsets, hw.columns, w, weights_f32=True)
    for src in (0, 3, 777):
        ref, it, _ = orc.sssp_jacobi_f32(hw.offsets, hw.columns, w, src)
        for mode in (0, 1):
            d, st = gf.sssp(src, mode=mode, persistent=1)
            assert np.array_equal(d, ref.view(np.uint32)), (src, mode)
            if mode == 0:
                assert st["rounds"] == it
    h64 = rmat(abi, 14, 21, wmax=255)
    g64 = abi.DeviceGraph.from_host(h64, off64=True)
    for src in abi.sample_sources(h64.offsets, 3, 3):
        dj = orc.sssp_dijkstra(h64.offsets, h64.columns, h64.weights, int(src))
        for mode in (0, 1):
            d, _ = g64.sssp(int(src), mode=mode, persistent=1, light_max_degree=16, heavy_chunk=64)
            assert np.array_equal(d, dj), (int(src), mode)
    er = rmat(abi, 13, 4, kind=1, wmax=255)
    ge = abi.DeviceGraph.from_host(er)
    for mode in (0, 1):
        d, _ = ge.sssp(11, mode=mode, persistent=1)
        assert np.array_equal(d, orc.sssp_dijkstra(er.offsets, er.columns, er.weights, 11))


def test_star_hub_and_edge_cases(gpu, abi, orc):
    """A 5000-edge hub (chunked) with a chain behind it, a source with no
    out-edges, a one-vertex graph, and repeated calls on one workspace."""
    V = 6000
    edges = [(0, v) for v in range(1, 5001)] + [(5000 + i, 5001 + i) for i in range(998)]
    weights = [((v * 7919) % 50) + 1 for v in range(1, 5001)] + [3] * 998
    off, col, w = csr_from_edges(V, edges, weights)
    g = abi.DeviceGraph(off, col, w.astype(np.uint32))
    dj = orc.sssp_dijkstra(off, col, w.astype(np.uint32), 0)
    ref, it, _ = orc.sssp_jacobi_f32(off, col, w.astype(np.float32), 0)
    for mode in (0, 1):
        d, st = g.sssp(0, mode=mode, persistent=1)
        assert np.array_equal(d, dj), mode
        if mode == 0:
            assert st["rounds"] == it
    d, st = g.sssp(5999, persistent=1)
    assert d[5999] == 0 and np.all(np.delete(d, 5999) == 0xFFFFFFFF) and st["rounds"] == 1
    d, st = g.sssp(0, mode=1, persistent=1)
    assert np.array_equal(d, dj)
    g1 = abi.DeviceGraph(np.array([0, 0], np.uint64), np.array([], np.uint32))
    for mode in (0, 1):
        d, st = g1.sssp(0, mode=mode, persistent=1)
        assert d.tolist() == [0] and st["converged"] == 1 and st["rounds"] == 1


@pytest.mark.parametrize("heavy_delta,wait_degree", [(2, 0), (1, 8), (40, 0), (0xFFFFFFFF, 0)])
def test_waiting_rows_same_distances(gpu, abi, orc, heavy_delta, wait_degree):
    """ASYNC waiting rows (rows above the wait degree stay queued while their
    distance exceeds the queue's smallest by more than heavy_delta): the same
    distances for u32 weights with small chunks (hubs over many chunks),
    float weights, and a star hub with a chain behind it."""
    hg = rmat(abi, 15, 33, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    for src in abi.sample_sources(hg.offsets, 4, 3):
        _say(f"delta {heavy_delta} wait {wait_degree} source {int(src)}")
        dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        for lm, ch in ((0, 0), (16, 64)):
            d, st = g.sssp(int(src), mode=1, persistent=1, heavy_delta=heavy_delta, wait_degree=wait_degree,
                           light_max_degree=lm, heavy_chunk=ch)
            assert np.array_equal(d, dj), (int(src), lm)
            assert st["converged"] == 1
    gf = abi.DeviceGraph.from_host(hg, weights_f32=True)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0)
    d, _ = gf.sssp(0, mode=1, persistent=1, heavy_delta=heavy_delta, wait_degree=wait_degree)
    want = np.where(dj == 0xFFFFFFFF, np.inf, dj.astype(np.float64)).astype(np.float32)
    assert np.array_equal(d.view(np.float32), want)
    V = 6000
    edges = [(0, v) for v in range(1, 5001)] + [(5000 + i, 5001 + i) for i in range(998)] + [(3, 0), (4999, 2)]
    weights = [((v * 7919) % 50) + 1 for v in range(1, 5001)] + [3] * 998 + [1, 1]
    off, col, w = csr_from_edges(V, edges, weights)
    gs = abi.DeviceGraph(off, col, w.astype(np.uint32))
    for src in (0, 3, 4999):
        d, _ = gs.sssp(src, mode=1, persistent=1, heavy_delta=heavy_delta, wait_degree=wait_degree,
                       light_max_degree=32, heavy_chunk=256)
        assert np.array_equal(d, orc.sssp_dijkstra(off, col, w.astype(np.uint32), src)), src


def test_options_and_overflow(gpu, abi):
    """Bucketed or pull-direction runs are rejected with the persistent kernel;
    a u32 distance overflow is reported."""
    hg = rmat(abi, 10, 2)
    g = abi.DeviceGraph.from_host(hg)
    with pytest.raises(abi.InvalidArgument):
        g.sssp(0, mode=1, persistent=1, delta=16)
    with pytest.raises(abi.InvalidArgument):
        g.sssp(0, mode=1, persistent=1, direction=2)
    off, col, w = csr_from_edges(3, [(0, 1), (1, 2)], [0xFFFFFFF0, 0x20])
    go = abi.DeviceGraph(off, col, np.asarray(w, np.uint32))
    with pytest.raises(abi.HytError) as ei:
        go.sssp(0, persistent=1)
    assert ei.value.status == abi.HYT_ERR_OVERFLOW


def test_scale20_invariants(gpu, abi, orc):
    """Size-independent certificate at scale 20 (d[src] = 0, no edge relaxes,
    a tight in-edge per reached vertex) and agreement with Dijkstra."""
    hg = rmat(abi, 20, 11, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    src = int(abi.sample_sources(hg.offsets, 12, 1)[0])
    d, st = g.sssp(src, mode=1, persistent=1)
    off, col, w = hg.offsets, hg.columns, hg.weights
    assert d[src] == 0
    u = np.repeat(np.arange(hg.V, dtype=np.int64), np.diff(off.astype(np.int64)))
    du = d[u].astype(np.uint64)
    fin = du != 0xFFFFFFFF
    cand = du[fin] + w[fin]
    assert np.all(d[col[fin]].astype(np.uint64) <= cand)
    tight = np.zeros(hg.V, bool)
    tight[col[fin][d[col[fin]].astype(np.uint64) == cand]] = True
    reached = d != 0xFFFFFFFF
    reached[src] = False
    assert np.all(tight[reached])
    assert np.array_equal(d, orc.sssp_dijkstra(off, col, w, src))
