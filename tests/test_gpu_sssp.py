"""GPU parity: frontier SSSP through the C ABI vs the oracle.

Bar: distances bit-exact (u32 for integer weights; float bit patterns for
float weights), and in SYNC mode the round count equals the original
sssp_cpu iteration count.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import csr_from_edges

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _say(msg):
    print(f"[gpu-sssp] {msg}", flush=True)


def rmat(abi, scale, seed, wmin=1, wmax=64, ef=16, kind=0):
    return abi.HostGraph(kind=kind, scale=scale, edge_factor=ef, seed=seed, weight_min=wmin, weight_max=wmax)


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_cuda_test_graph(gpu, abi, orc, mode):
    # cuda_algorithm_tests.cu:58-87 (same graph, float weights)
    off, col, w = np.array([0, 2, 4, 5, 5], np.uint64), np.array([1, 2, 2, 3, 3], np.uint32), \
        np.array([1, 4, 2, 6, 1], np.float32)
    g = abi.DeviceGraph(off, col, w, weights_f32=True)
    d, st = g.sssp(0, mode=mode)
    ref, it, conv = orc.sssp_jacobi_f32(off, col, w, 0)
    assert d.view(np.float32).tolist() == [0.0, 1.0, 3.0, 4.0]
    assert d.tolist() == ref.view(np.uint32).tolist()
    assert st["converged"] == 1
    if mode == 0:
        assert st["rounds"] == it


@pytest.mark.parametrize("direction", [1, 2])
def test_golden_sssp_cases(gpu, abi, direction):
    data = json.loads((GOLDEN / "reference_cases.json").read_text())
    for case in data["sssp"]:
        _say(case["name"])
        w = None if case["weights"] is None else np.array(case["weights"], np.float32)
        g = abi.DeviceGraph(np.array(case["offsets"], np.uint64), np.array(case["columns"], np.uint32), w,
                            weights_f32=True)
        if case["tolerance"] != 0.0:
            # the order-dependent tolerance rule runs as ordered pull rounds
            if direction == 1:
                with pytest.raises(abi.InvalidArgument):
                    g.sssp(case["source"], mode=0, direction=direction, tolerance=case["tolerance"])
                continue
            d, st = g.sssp(case["source"], mode=0, max_rounds=case["max_iterations"], direction=direction,
                           tolerance=case["tolerance"])
        else:
            d, st = g.sssp(case["source"], mode=0, max_rounds=case["max_iterations"], direction=direction)
        assert d.tolist() == case["distance_bits"], case["name"]
        assert st["rounds"] == case["iterations"] and bool(st["converged"]) == case["converged"], case["name"]


DIRECTIONS = [(0, 0), (1, 0), (2, 0), (0, 1), (0, 50)]  # (direction, pull_percent)


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("direction,pct", DIRECTIONS)
def test_rmat16_config1_bitexact(gpu, abi, orc, seed, direction, pct):
    """RMAT scale 16, ef 16, u32 weights in [1,64], source 0 (the CPU-runnable
    case), for push-only, pull-only and mixed rounds."""
    hg = rmat(abi, 16, seed)
    g = abi.DeviceGraph.from_host(hg)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0)
    ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 0)
    assert np.array_equal(orc.f32_to_u32_dist(ref), dj)
    for mode in (0, 1):
        _say(f"seed {seed} mode {mode} direction {direction}/{pct}")
        d, st = g.sssp(0, mode=mode, direction=direction, pull_percent=pct)
        assert np.array_equal(d, dj)
        assert st["converged"] == 1
        if direction == 2:
            assert st["pull_rounds"] == st["rounds"]
        if mode == 0:
            assert st["rounds"] == it
    # float-weight graph of the same topology: float bit patterns equal the original's
    gf = abi.DeviceGraph.from_host(hg, weights_f32=True)
    d, st = gf.sssp(0, mode=0, direction=direction, pull_percent=pct)
    assert np.array_equal(d, ref.view(np.uint32)) and st["rounds"] == it


@pytest.mark.parametrize("delta", [1, 16, 100, 1000])
def test_delta_stepping_bitexact(gpu, abi, orc, delta):
    """ASYNC rounds with delta-stepping buckets: same distances, any width."""
    hg = rmat(abi, 14, 7, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    for src in (0, 4242):
        dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src)
        d, st = g.sssp(src, mode=1, delta=delta)
        assert np.array_equal(d, dj), (src, delta)
        assert st["converged"] == 1
        # a truncated bucketed run leaves the workspace clean for the next call
        g.sssp(src, mode=1, delta=delta, max_rounds=2)
        d2, _ = g.sssp(src, mode=1)
        assert np.array_equal(d2, dj)
    gf = abi.DeviceGraph.from_host(hg, weights_f32=True)
    ref, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 0)
    d, _ = gf.sssp(0, mode=1, delta=delta)
    assert np.array_equal(d, ref.view(np.uint32))


def test_fractional_float_weights_bitexact(gpu, abi, orc, rng):
    hg = rmat(abi, 12, 5)
    w = (hg.weights.astype(np.float32) * rng.choice(np.array([0.1, 0.37, 1.75], np.float32),
                                                    size=hg.E)).astype(np.float32)
    g = abi.DeviceGraph(hg.offsets, hg.columns, w, weights_f32=True)
    for src in [0, 3, 777]:
        ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src)
        for mode in (0, 1):
            d, st = g.sssp(src, mode=mode)
            assert np.array_equal(d, ref.view(np.uint32)), (src, mode)
            if mode == 0:
                assert st["rounds"] == it


@pytest.mark.parametrize("mode", [0, 1])
def test_device_driven_rounds_match_host_rounds(gpu, abi, orc, mode):
    """Push rounds replayed from the captured round graph (device-side round
    control, the default) against one host read-back per round: the same
    distances; in SYNC mode the same round count, frontier, edge and update
    totals, also when truncated by max_rounds; the cached graph serves other
    sources and is rebuilt after a trim or a split change (back and forth)."""
    hg = rmat(abi, 15, 81, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    keys = ("rounds", "converged", "frontier_total", "edges_total", "updates_total")
    for src in abi.sample_sources(hg.offsets, 9, 5):
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        dd, sd = g.sssp(int(src), mode=mode)
        dh, sh = g.sssp(int(src), mode=mode, host_rounds=1)
        assert np.array_equal(dd, ref) and np.array_equal(dh, ref), src
        assert sd["converged"] == sh["converged"] == 1
        if mode == 0:
            assert all(sd[k] == sh[k] for k in keys), (src, sd, sh)
            for mr in (1, 2, 3, 4, 5, 7, 9):
                d1, s1 = g.sssp(int(src), mode=0, max_rounds=mr)
                d2, s2 = g.sssp(int(src), mode=0, max_rounds=mr, host_rounds=1)
                assert np.array_equal(d1, d2) and all(s1[k] == s2[k] for k in keys), (src, mr)
    # timed replays: one timed expansion per round with work
    _, st = g.sssp(0, mode=mode, time_kernels=1)
    assert st["expand_launches"] == st["rounds"] and st["expand_ms"] > 0
    d, _ = g.sssp(0, mode=mode, light_max_degree=8, heavy_chunk=64)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0))
    # back to the default split without a trim (ASYNC: the rows-above-split
    # bitmap is rebuilt for the split in use)
    src = int(abi.sample_sources(hg.offsets, 10, 1)[0])
    d, _ = g.sssp(src, mode=mode)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
    d, _ = g.sssp(src, mode=mode, light_max_degree=8, heavy_chunk=64)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
    g.trim()
    d, _ = g.sssp(0, mode=mode)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0))


@pytest.mark.parametrize("device_rounds", [True, False])
def test_waiting_hubs_same_distances(gpu, abi, orc, device_rounds):
    """ASYNC rounds in which rows above light_max wait while their distance
    exceeds the queue's smallest by more than heavy_delta give Dijkstra's
    distances, for any margin (1, the most waiting; 0, the default; UINT32_MAX,
    no waiting), on integer and float weights."""
    hg = rmat(abi, 14, 91, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    wf = (hg.weights.astype(np.float32) * np.float32(0.37)).astype(np.float32)
    gf = abi.DeviceGraph(hg.offsets, hg.columns, wf, weights_f32=True)
    hr = 0 if device_rounds else 1
    for src in abi.sample_sources(hg.offsets, 92, 3):
        ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        reff, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, wf, int(src))
        base = g.sssp(int(src), mode=abi.HYT_SSSP_ASYNC, light_max_degree=16, heavy_chunk=64, host_rounds=hr,
                      heavy_delta=0xFFFFFFFF)[1]
        for hd in (0, 1, 64, 300, 100000, 0xFFFFFFFF):
            d, st = g.sssp(int(src), mode=abi.HYT_SSSP_ASYNC, light_max_degree=16, heavy_chunk=64, heavy_delta=hd,
                           host_rounds=hr)
            _say(f"src {src} heavy_delta {hd}: rounds {st['rounds']} (no waiting {base['rounds']}), "
                 f"chunk edges {st['heavy_edges']} ({base['heavy_edges']})")
            assert np.array_equal(d, ref), (src, hd)
            assert st["converged"] == 1
            df, _ = gf.sssp(int(src), mode=abi.HYT_SSSP_ASYNC, light_max_degree=16, heavy_chunk=64,
                            heavy_delta=hd, host_rounds=hr)
            assert np.array_equal(df, reff.view(np.uint32)), (src, hd)
        # SYNC mode ignores the option: the original's rounds
        d, st = g.sssp(int(src), mode=abi.HYT_SSSP_SYNC, heavy_delta=5, host_rounds=hr)
        assert np.array_equal(d, ref) and st["rounds"] == g.sssp(int(src), mode=abi.HYT_SSSP_SYNC)[1]["rounds"]


@pytest.mark.parametrize("device_rounds", [True, False])
def test_sync_edge_counts_and_chunk_split(gpu, abi, orc, device_rounds):
    """SYNC push rounds relax exactly the out-edges of the rows that improved
    in the previous round: edges_total is the degree sum of those rows over
    all rounds, and heavy_edges (edges walked by the chunk kernel) the part
    in rows above light_max. The per-round frontiers come from the oracle's
    distances after 0, 1, 2, ... iterations."""
    hg = rmat(abi, 12, 71, wmax=64)
    g = abi.DeviceGraph.from_host(hg)
    deg = np.diff(hg.offsets.astype(np.int64))
    w = hg.weights.astype(np.float32)
    for src in (0, int(abi.sample_sources(hg.offsets, 3, 1)[0])):
        ref, it, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src)
        prev = np.full(hg.V, np.inf, np.float32)
        prev[src] = 0
        frontier = np.zeros(hg.V, bool)
        frontier[src] = True
        rows = []
        for k in range(1, it + 1):
            rows.append(frontier)
            cur, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src, 0.0, k)
            frontier = cur != prev
            prev = cur
        assert np.array_equal(prev, ref) and not frontier.any()
        for light_max in (8, 64, 1 << 20):
            d, st = g.sssp(src, mode=abi.HYT_SSSP_SYNC, direction=abi.HYT_DIR_PUSH, light_max_degree=light_max,
                           heavy_chunk=64, host_rounds=0 if device_rounds else 1)
            assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, src))
            edges = sum(int(deg[f].sum()) for f in rows if f.any())
            heavy = sum(int(deg[f & (deg > light_max)].sum()) for f in rows if f.any())
            _say(f"src {src} light_max {light_max}: edges {st['edges_total']} (oracle {edges}), "
                 f"chunk edges {st['heavy_edges']} (oracle {heavy})")
            assert st["edges_total"] == edges and st["heavy_edges"] == heavy
            assert (heavy > 0) == (light_max < deg.max())


@pytest.mark.parametrize("tol", [0.25, 0.5, 3.0, 40.0])
@pytest.mark.parametrize("scale", [10, 13])
def test_tolerance_rule_bitexact(gpu, abi, orc, rng, tol, scale):
    """Tolerance > 0: the original's order-dependent rule (candidate + tol <
    best in incoming-edge order) gives the original's distance bits,
    iteration count and convergence flag, also when truncated, on float
    weights with fractions and on an unweighted graph (w = 1)."""
    hg = rmat(abi, scale, 60 + scale)
    w = (hg.weights.astype(np.float32) * rng.choice(np.array([0.1, 0.37, 1.75, 1.0], np.float32),
                                                    size=hg.E)).astype(np.float32)
    g = abi.DeviceGraph(hg.offsets, hg.columns, w, weights_f32=True)
    gu = abi.DeviceGraph(hg.offsets, hg.columns, None, weights_f32=True)
    use_ref = orc.ref_available()
    for src in [0, 5, 901]:
        for graph, weights in ((g, w), (gu, None)):
            ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, weights, src, tol, 0, use_ref=use_ref)
            d, st = graph.sssp(src, mode=0, tolerance=tol)
            assert np.array_equal(d, ref.view(np.uint32)), (src, tol)
            assert st["rounds"] == it and bool(st["converged"]) == conv
            tr, it2, conv2 = orc.sssp_jacobi_f32(hg.offsets, hg.columns, weights, src, tol, 2, use_ref=use_ref)
            d2, st2 = graph.sssp(src, mode=0, tolerance=tol, max_rounds=2)
            assert np.array_equal(d2, tr.view(np.uint32)) and st2["rounds"] == it2
            assert bool(st2["converged"]) == conv2
    # integer weights, ASYNC mode, push direction and batches reject it
    gi = abi.DeviceGraph.from_host(hg)
    for bad in (lambda: gi.sssp(0, tolerance=tol), lambda: g.sssp(0, mode=1, tolerance=tol),
                lambda: g.sssp(0, direction=1, tolerance=tol), lambda: g.sssp([0, 5], tolerance=tol),
                lambda: g.sssp(0, tolerance=-1.0)):
        with pytest.raises(abi.InvalidArgument):
            bad()


@pytest.mark.parametrize("max_rounds", [1, 2, 3, 5])
@pytest.mark.parametrize("direction", [1, 2])
def test_truncated_rounds_match_original_state(gpu, abi, orc, max_rounds, direction):
    """SYNC mode is round-for-round Jacobi (push or pull): the state after k
    rounds equals the original's state after max_iterations = k."""
    hg = rmat(abi, 12, 9)
    g = abi.DeviceGraph.from_host(hg, weights_f32=True)
    ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 1, 0.0, max_rounds)
    d, st = g.sssp(1, mode=0, max_rounds=max_rounds, direction=direction)
    assert np.array_equal(d, ref.view(np.uint32))
    assert st["rounds"] == it and bool(st["converged"]) == conv


def test_unweighted_and_er(gpu, abi, orc):
    hg = rmat(abi, 13, 3, wmin=0, wmax=0)
    g = abi.DeviceGraph.from_host(hg)
    d, st = g.sssp(0)
    assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, None, 0))
    er = rmat(abi, 13, 4, kind=1, wmax=255)
    g = abi.DeviceGraph.from_host(er)
    for mode in (0, 1):
        d, st = g.sssp(11, mode=mode)
        assert np.array_equal(d, orc.sssp_dijkstra(er.offsets, er.columns, er.weights, 11))


@pytest.mark.parametrize("light_max,chunk", [(0, 0), (4, 64), (64, 128), (1 << 20, 1 << 16)])
def test_degree_binning_thresholds(gpu, abi, orc, light_max, chunk):
    """Every work split (light rows only, tiny chunks, huge chunks) gives the
    same distances, push or pull; hub rows span many chunks."""
    hg = rmat(abi, 14, 6, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 0)
    for mode in (0, 1):
        for direction in (1, 2):
            d, st = g.sssp(0, mode=mode, light_max_degree=light_max, heavy_chunk=chunk, direction=direction)
            assert np.array_equal(d, dj), (mode, direction)


def test_star_hub_and_edge_cases(gpu, abi, orc):
    # A hub with 5000 out-edges (chunked), a chain behind it, isolated vertices.
    V = 6000
    edges = [(0, v) for v in range(1, 5001)] + [(5000 + i, 5001 + i) for i in range(998)]
    weights = [((v * 7919) % 50) + 1 for v in range(1, 5001)] + [3] * 998
    off, col, w = csr_from_edges(V, edges, weights)
    g = abi.DeviceGraph(off, col, w.astype(np.uint32))
    dj = orc.sssp_dijkstra(off, col, w.astype(np.uint32), 0)
    ref, it, _ = orc.sssp_jacobi_f32(off, col, w.astype(np.float32), 0)
    for direction in (1, 2):
        d, st = g.sssp(0, mode=0, direction=direction)
        assert np.array_equal(d, dj) and st["rounds"] == it
    # source with no out-edges: only itself reached, one quiet round
    d, st = g.sssp(5999)
    assert d[5999] == 0 and np.all(np.delete(d, 5999) == 0xFFFFFFFF) and st["rounds"] == 1
    # single vertex, no edges
    g1 = abi.DeviceGraph(np.array([0, 0], np.uint64), np.array([], np.uint32))
    d, st = g1.sssp(0)
    assert d.tolist() == [0] and st["converged"] == 1


def test_invalid_arguments(gpu, abi):
    g = abi.DeviceGraph(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([5], np.uint32))
    with pytest.raises(abi.InvalidArgument):
        g.sssp(2)
    with pytest.raises(abi.InvalidArgument):
        abi.DeviceGraph(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([-1.0], np.float32),
                        weights_f32=True)


def test_u32_overflow_is_reported(gpu, abi):
    off, col, w = csr_from_edges(3, [(0, 1), (1, 2)], [0xFFFFFFF0, 0x20])
    g = abi.DeviceGraph(off, col, np.asarray(w, np.uint32))
    with pytest.raises(abi.HytError) as ei:
        g.sssp(0)
    assert ei.value.status == abi.HYT_ERR_OVERFLOW


def test_scale20_invariants(gpu, abi, orc):
    """Size-independent properties at a larger scale: d[src] = 0, no edge can
    improve any distance, every reached vertex has a tight in-edge; plus exact
    agreement with Dijkstra."""
    hg = rmat(abi, 20, 11, wmax=255)
    g = abi.DeviceGraph.from_host(hg)
    src = int(abi.sample_sources(hg.offsets, 12, 1)[0])
    d, st = g.sssp(src, mode=1)
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


@pytest.mark.parametrize("wait_degree", [8, 32, 100])
def test_waiting_light_rows_same_distances(gpu, abi, orc, wait_degree):
    """ASYNC rounds where light rows above wait_degree edges also wait under
    the heavy_delta bound: the same distances (RMAT, ER, float weights;
    device-driven and host-driven rounds)."""
    for kind in (0, 1):
        hg = rmat(abi, 14, 17, wmax=255, kind=kind)
        g = abi.DeviceGraph.from_host(hg)
        for src in abi.sample_sources(hg.offsets, 18, 2):
            ref = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
            for hd in (1, 2, 8):
                _say(f"kind {kind} src {src} wait_degree {wait_degree} heavy_delta {hd}")
                d, st = g.sssp(int(src), mode=1, heavy_delta=hd, wait_degree=wait_degree)
                assert np.array_equal(d, ref) and st["converged"] == 1
                dh, _ = g.sssp(int(src), mode=1, heavy_delta=hd, wait_degree=wait_degree, host_rounds=1)
                assert np.array_equal(dh, ref)
    hf = rmat(abi, 12, 19)
    gf = abi.DeviceGraph.from_host(hf, weights_f32=True)
    ref, _, _ = orc.sssp_jacobi_f32(hf.offsets, hf.columns, hf.weights.astype(np.float32), 0)
    d, _ = gf.sssp(0, mode=1, wait_degree=wait_degree)
    assert np.array_equal(d, ref.view(np.uint32))
