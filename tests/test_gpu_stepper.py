"""GPU parity: round-by-round runs (hyt_stepper_*, behind the SEP drivers).

Bar: a SYNC SSSP run stepped one round at a time passes through exactly the
states of the original's sssp_cpu with max_iterations = 1, 2, ... (distance
bits), for push and pull rounds and both round drivers, and finishes on the
original's iteration count; ASYNC steps end at the Dijkstra distances; a
PageRank / PPR run read between its steps ends with the bits of the one-call
form (reading does not disturb the state).
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)


def _say(msg):
    print(f"[gpu-stepper] {msg}", flush=True)


@pytest.mark.parametrize("direction", [1, 2])
@pytest.mark.parametrize("host_rounds", [0, 1])
def test_sync_sssp_steps_match_original_states(gpu, abi, orc, rng, direction, host_rounds):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=91, weight_min=1, weight_max=64)
    w = (hg.weights.astype(np.float32) * rng.choice(np.array([0.1, 0.37, 1.0], np.float32), size=hg.E)).astype(
        np.float32)
    g = abi.DeviceGraph(hg.offsets, hg.columns, w, weights_f32=True)
    use_ref = orc.ref_available()
    src = 7
    _, iters, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src, 0.0, 0, use_ref=use_ref)
    s = abi.Stepper(g, abi.HYT_STEP_SSSP, src, mode=abi.HYT_SSSP_SYNC, direction=direction, host_rounds=host_rounds)
    for k in range(1, iters + 2):
        st, fin = s.step(1)
        d, tot = s.read()
        ref, it, conv = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src, 0.0, k, use_ref=use_ref)
        assert np.array_equal(d, ref.view(np.uint32)), k
        assert tot["rounds"] == min(k, iters)
        assert fin == (k >= iters), (k, iters)
        if fin:
            assert conv and it == iters
    _say(f"dir {direction} host_rounds {host_rounds}: {iters} rounds stepped one at a time")
    # several rounds per step land on the same states
    s2 = abi.Stepper(g, abi.HYT_STEP_SSSP, src, mode=abi.HYT_SSSP_SYNC, direction=direction)
    s2.step(2)
    ref, _, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, w, src, 0.0, 2, use_ref=use_ref)
    assert np.array_equal(s2.read()[0], ref.view(np.uint32))
    st, fin = s2.step(10 ** 6)
    assert fin and s2.read()[1]["rounds"] == iters


def test_async_sssp_steps_reach_dijkstra(gpu, abi, orc):
    hg = abi.HostGraph(scale=14, edge_factor=16, seed=92, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    for src in abi.sample_sources(hg.offsets, 3, 3):
        s = abi.Stepper(g, abi.HYT_STEP_SSSP, int(src), mode=abi.HYT_SSSP_ASYNC)
        n = 0
        while True:
            _, fin = s.step(1)
            n += 1
            if fin:
                break
        d, tot = s.read()
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src)))
        assert tot["rounds"] == n and tot["converged"] == 1
        s.free()


@pytest.mark.parametrize("rounds", [1, 3])
def test_async_steps_with_hub_rows_in_resumed_queues(gpu, abi, orc, rounds):
    """Device-driven ASYNC rounds list the queue's rows above light_max apart
    (the compaction's hub list); a call that resumes a run lists the rows of
    the queue it inherits again. A small split puts such rows in most queues;
    stepping 1 or 3 rounds at a time, with and without waiting rows, must
    still reach Dijkstra's distances, and a whole call must equal it too."""
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=93, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    for src in abi.sample_sources(hg.offsets, 4, 2):
        want = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(src))
        for heavy_delta in (0, 1):
            opt = dict(mode=abi.HYT_SSSP_ASYNC, light_max_degree=8, heavy_chunk=64, heavy_delta=heavy_delta)
            s = abi.Stepper(g, abi.HYT_STEP_SSSP, int(src), **opt)
            n = 0
            while True:
                _, fin = s.step(rounds)
                n += 1
                assert n < 10000
                if fin:
                    break
            d, tot = s.read()
            s.free()
            _say(f"src {int(src)} heavy_delta {heavy_delta}: {tot['rounds']} rounds in {n} calls")
            assert np.array_equal(d, want)
            assert tot["converged"] == 1
            whole, _ = g.sssp(int(src), **opt)
            assert np.array_equal(whole, want)


@pytest.mark.parametrize("seed", [-1, 11])
def test_pagerank_steps_equal_one_call(gpu, abi, orc, seed):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=93, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    eps = 1e-8
    one = g.pagerank(D, eps)[0] if seed < 0 else g.ppr(seed, D, eps)[0]
    s = abi.Stepper(g, abi.HYT_STEP_PAGERANK, seed, damping=D, epsilon=eps)
    steps = 0
    while True:
        _, fin = s.step(3)
        r, tot = s.read()  # a read between steps leaves the run's state as it is
        steps += 1
        assert abs(float(r.sum()) - 1.0) < 1e-5
        if fin:
            break
    assert np.array_equal(r, one)
    same, rounds, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=seed)
    assert np.abs(r - same).sum() <= 1e-6
    _say(f"pagerank seed {seed}: {tot['rounds']} rounds in {steps} steps, residual {tot['residual']:.2e}")


def test_stepper_errors(gpu, abi):
    hg = abi.HostGraph(scale=10, edge_factor=16, seed=94, weight_min=1, weight_max=9)
    g = abi.DeviceGraph.from_host(hg)
    with pytest.raises(abi.InvalidArgument):
        abi.Stepper(g, abi.HYT_STEP_SSSP, hg.V)
    with pytest.raises(abi.InvalidArgument):
        abi.Stepper(g, abi.HYT_STEP_PAGERANK, -1, damping=1.0)
    with pytest.raises(abi.InvalidArgument):
        abi.Stepper(g, abi.HYT_STEP_SSSP, 0, delta=8)
    s = abi.Stepper(g, abi.HYT_STEP_SSSP, 0)
    with pytest.raises(abi.InvalidArgument):
        s.read()  # no round yet
