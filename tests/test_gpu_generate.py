"""GPU: the device generator (hyt_generate_device) builds the host
generator's graphs bit for bit - offsets, columns and weights - for RMAT and
ER, odd and even scales, weighted and unweighted, pageable and pinned host
memory, one bucket or many."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-generate] {msg}", flush=True)


CASES = [
    # kind, scale, edge_factor, seed, wmin, wmax, budget (bytes; 0 = default)
    (0, 12, 16, 1, 1, 64, 0),
    (0, 13, 4, 5, 1, 255, 0),
    (0, 16, 16, 3, 1, 255, 0),
    (1, 14, 8, 2, 1, 255, 0),
    (0, 10, 16, 7, 0, 0, 0),
    (0, 18, 16, 11, 1, 255, 17 * 600_000),  # many buckets
    (1, 17, 16, 4, 1, 255, 17 * 300_000),   # many ER buckets
    (0, 1, 16, 1, 1, 64, 0),                # two vertices
]


@pytest.mark.parametrize("case", CASES)
def test_device_generator_matches_host(gpu, abi, case):
    kind, scale, ef, seed, wmin, wmax, budget = case
    _say(f"kind {kind} scale {scale} ef {ef} seed {seed} budget {budget}")
    kw = dict(kind=kind, scale=scale, edge_factor=ef, seed=seed, weight_min=wmin, weight_max=wmax)
    h = abi.HostGraph(**kw)
    d = abi.HostGraph(on_device=True, device_budget=budget, **kw)
    assert d.V == h.V and d.E == h.E
    assert np.array_equal(d.offsets, h.offsets)
    assert np.array_equal(d.columns, h.columns)
    if wmax == 0:
        assert d.weights is None and h.weights is None
    else:
        assert np.array_equal(d.weights, h.weights)


def test_pinned_output_and_small_budget_error(gpu, abi):
    kw = dict(kind=0, scale=14, edge_factor=16, seed=9, weight_min=1, weight_max=255)
    h = abi.HostGraph(**kw)
    d = abi.HostGraph(on_device=True, memory=abi.HYT_HOST_PINNED, **kw)
    assert np.array_equal(d.offsets, h.offsets) and np.array_equal(d.columns, h.columns)
    assert np.array_equal(d.weights, h.weights)
    # a bucket smaller than the densest source bin is refused, not truncated
    with pytest.raises(abi.HytError) as ei:
        abi.HostGraph(on_device=True, device_budget=17 * 16, **kw)
    assert ei.value.status == abi.HYT_ERR_OUT_OF_MEMORY
