"""GPU: the autotuner picks among its candidates, never loses to the
default split it measured, and the settings it returns keep SSSP bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tune_split_and_lanes(gpu, abi, orc):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=12, weight_min=1, weight_max=255)
    g = abi.DeviceGraph.from_host(hg)
    src = abi.sample_sources(hg.offsets, 3, 4)
    r = abi.tune(g, src, max_lanes=16, reps=1)
    print(f"[gpu-tune] {r}", flush=True)
    assert r["light_max_degree"] in (64, 128, 256, 512, 1024, 2048)
    assert r["heavy_chunk"] in (256, 512, 1024, 2048, 4096, 8192)
    assert r["lanes"] in (2, 4, 8, 16) and r["lanes_qps"] > 0
    assert r["candidates"] == 11 + 4
    assert 0 < r["sssp_ms"] <= r["default_sssp_ms"]
    for s in src:
        d, _ = g.sssp(int(s), mode=abi.HYT_SSSP_ASYNC, light_max_degree=r["light_max_degree"],
                      heavy_chunk=r["heavy_chunk"])
        assert np.array_equal(d, orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s)))
    with pytest.raises(abi.InvalidArgument):
        abi.tune(g, [hg.V])
    only_lanes = abi.tune(g, src, tune_split=False, max_lanes=6, reps=1)
    assert only_lanes["lanes"] in (2, 4, 6) and only_lanes["candidates"] == 3  # K = 2, 4, 6


def test_tune_hybrid_partition_size(gpu, abi):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=5, weight_min=1, weight_max=255,
                       memory=abi.HYT_HOST_PINNED)
    src = abi.sample_sources(hg.offsets, 4, 2)
    cands = [1024, 2048, 4096]
    r = abi.tune_hybrid(hg, cands, src, pack_edges=8192, device_budget_bytes=8 << 20, calibrate=0)
    print(f"[gpu-tune] hybrid {r}", flush=True)
    assert r["partition_edges"] in cands and r["candidates"] == 3 and r["hybrid_ms"] > 0
