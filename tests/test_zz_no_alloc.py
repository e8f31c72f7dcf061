"""GPU parity: hyt_sssp_noalloc (no device allocation) against hyt_sssp.

The no-allocation entry point runs the same rounds as hyt_sssp with the
library's static scratch array instead of the graph's workspace. Bar: the
distances (and, in SYNC mode, the round statistics) equal hyt_sssp's bit for
bit on two calls in a row (the second on a smaller graph, so the shared scratch keeps a larger
call's contents), for every weight mode and both round modes; the call
leaves the graph's device footprint and the device's free memory unchanged;
graphs beyond the scratch and options it cannot run fail cleanly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


def _check(abi, hip, hg, g, sources, **opt):
    out = hip.malloc(hg.V * 4)
    for s in sources:
        ref, rst = g.sssp(int(s), **opt)
        st = g.sssp_noalloc(int(s), out, **opt)
        got = hip.to_host_u32(out, hg.V)
        assert np.array_equal(got, ref), (s, opt)
        assert st["converged"] == rst["converged"] == 1
        if opt.get("mode", 0) == abi.HYT_SSSP_SYNC:
            # SYNC rounds are deterministic (ASYNC counts depend on timing)
            for k in ("rounds", "frontier_total", "edges_total", "updates_total"):
                assert st[k] == rst[k], (k, s, opt)
    hip.free(out)


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_matches_hyt_sssp_two_calls(gpu, abi, orc, hip, mode):
    big = abi.HostGraph(scale=14, edge_factor=16, seed=71, weight_min=1, weight_max=255)
    small = abi.HostGraph(scale=11, edge_factor=16, seed=72, weight_min=1, weight_max=64)
    gb, gs = abi.DeviceGraph.from_host(big), abi.DeviceGraph.from_host(small)
    src_b = abi.sample_sources(big.offsets, 3, 2)
    src_s = abi.sample_sources(small.offsets, 4, 2)
    _say(f"mode {mode}: scale 14 then scale 11")
    _check(abi, hip, big, gb, src_b, mode=mode)
    _check(abi, hip, small, gs, src_s, mode=mode)
    # against the oracle as well (the second, smaller call)
    out = hip.malloc(small.V * 4)
    gs.sssp_noalloc(int(src_s[0]), out, mode=mode)
    assert np.array_equal(hip.to_host_u32(out, small.V),
                          orc.sssp_dijkstra(small.offsets, small.columns, small.weights, int(src_s[0])))
    hip.free(out)


def test_noalloc_weight_modes_and_splits(gpu, abi, orc, hip):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=73, weight_min=1, weight_max=64)
    for graph in (abi.DeviceGraph.from_host(hg), abi.DeviceGraph.from_host(hg, weights_f32=True),
                  abi.DeviceGraph(hg.offsets, hg.columns), abi.DeviceGraph.from_host(hg, off64=True)):
        _check(abi, hip, hg, graph, [0, 77], mode=1)
        _check(abi, hip, hg, graph, [5], mode=0, light_max_degree=4, heavy_chunk=64)
        _check(abi, hip, hg, graph, [5], mode=1, delta=16)


def test_noalloc_allocates_nothing_and_rejects(gpu, abi, hip):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=74, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    out = hip.malloc(hg.V * 4)
    for s in (0, 1, 2):                        # the first call on the device takes the scratch address
        g.sssp_noalloc(s, out, mode=1)
    before_graph = g.info()["device_bytes"]
    before_free = hip.free_bytes()
    for s in (0, 1, 2):
        g.sssp_noalloc(s, out, mode=1)
    assert g.info()["device_bytes"] == before_graph  # no workspace was created
    assert hip.free_bytes() == before_free
    with pytest.raises(abi.InvalidArgument):
        g.sssp_noalloc(0, out, direction=abi.HYT_DIR_PULL)
    with pytest.raises(abi.InvalidArgument):
        g.sssp_noalloc(hg.V, out)
    with pytest.raises(abi.InvalidArgument):
        g.sssp_noalloc(0, out, tolerance=0.5)
    hip.free(out)
    # 2^23 vertices need more scratch than the array holds
    V = 1 << 23
    huge = abi.DeviceGraph(np.zeros(V + 1, np.uint64), np.zeros(0, np.uint32))
    big_out = hip.malloc(V * 4)
    with pytest.raises(abi.InvalidArgument):
        huge.sssp_noalloc(0, big_out)
    hip.free(big_out)
