"""GPU parity: delta-push PageRank / PPR (FP32) through the C ABI.

Tolerances (normalised L1 unless stated):
  * vs the FP64 restatement of the same synchronous delta-push procedure
    (oracle.delta_push_f64, same float damping and eps): <= 1e-6 (residuals
    and ranks are 64-bit fixed-point words; the per-row float share and the
    float output are the only roundings: bound d/(1-d) * 2^-23 + 2^-24);
  * vs the exact FP64 fixed point (oracle.pagerank_f64): <= 2*d*eps/(1-d)^2 plus
    the FP32 term above;
  * vs the original pagerank_cpu on its GPU test graph: max-abs <= 1e-5
    (cuda_algorithm_tests.cu:54).
"""
import numpy as np
import pytest

from conftest import f32

pytestmark = pytest.mark.gpu
D = f32(0.85)  # the damping the C ABI carries (float)


def _say(msg):
    print(f"[gpu-pr] {msg}", flush=True)


def test_reference_cuda_test_graph(gpu, abi, orc):
    # cuda_algorithm_tests.cu:23-56: {0->1, 0->2, 1->2, 2->0}, tolerance 1e-6
    off, col = np.array([0, 2, 3, 4], np.uint64), np.array([1, 2, 2, 0], np.uint32)
    cpu, it, res, conv = orc.pagerank_jacobi_f32(off, col, D, 1e-6, 200)
    g = abi.DeviceGraph(off, col)
    r, st = g.pagerank(D, 1e-9)
    assert st["converged"] == 1
    assert np.max(np.abs(r - cpu)) <= 1e-5
    assert abs(r.sum() - 1) < 1e-6


def test_closed_forms(gpu, abi):
    # 3-cycle -> 1/3 (algorithm_tests.cpp:22-47); dangling graph sums to 1
    g = abi.DeviceGraph(np.array([0, 1, 2, 3], np.uint64), np.array([1, 2, 0], np.uint32))
    r, _ = g.pagerank(D, 1e-9)
    assert np.all(np.abs(r - 1 / 3) <= 1e-6)
    g = abi.DeviceGraph(np.array([0, 1, 2, 2], np.uint64), np.array([1, 0], np.uint32))
    r, _ = g.pagerank(D, 1e-9)
    assert abs(r.sum() - 1) <= 1e-6
    # damping 0 -> uniform
    r, st = g.pagerank(0.0, 1e-9)
    assert np.allclose(r, 1 / 3, atol=1e-7)


@pytest.mark.parametrize("scale,eps", [(12, 1e-7), (16, 1e-7)])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_rmat_pagerank_vs_fp64(gpu, abi, orc, scale, eps, direction):
    hg = abi.HostGraph(scale=scale, edge_factor=16, seed=3, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    r, st = g.pagerank(D, eps, direction=direction)
    assert st["converged"] == 1
    same, rounds, edges = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps))
    exact = orc.pagerank_f64(hg.offsets, hg.columns, D)
    l1_same = np.abs(r - same).sum()
    l1_exact = np.abs(r - exact).sum()
    _say(f"scale {scale}: L1 vs same-procedure FP64 {l1_same:.3e}, vs exact {l1_exact:.3e}, "
         f"rounds {st['rounds']} (fp64 {rounds})")
    assert l1_same <= 1e-6
    assert l1_exact <= 2 * D * eps / (1 - D) ** 2 + 2e-6
    assert abs(st["rounds"] - rounds) <= 2
    # the original's pull iteration lands within its own stopping error
    cpu, it, res, conv = orc.pagerank_jacobi_f32(hg.offsets, hg.columns, D, 1e-6, 200)
    assert np.max(np.abs(r - cpu)) <= 1e-5


@pytest.mark.parametrize("light_max,chunk", [(4, 64), (1 << 20, 1 << 16)])
@pytest.mark.parametrize("direction", [1, 2])
def test_pagerank_work_split_invariance(gpu, abi, orc, light_max, chunk, direction):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=8, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    r, _ = g.pagerank(D, 1e-8, light_max_degree=light_max, heavy_chunk=chunk, direction=direction)
    exact = orc.pagerank_f64(hg.offsets, hg.columns, D)
    assert np.abs(r - exact).sum() <= 2 * D * 1e-8 / (1 - D) ** 2 + 2e-6


@pytest.mark.parametrize("seed_vertex", [0, 100])
@pytest.mark.parametrize("direction", [0, 2])
def test_ppr_single_seed(gpu, abi, orc, seed_vertex, direction):
    hg = abi.HostGraph(scale=13, edge_factor=16, seed=2, weight_min=0, weight_max=0)
    g = abi.DeviceGraph.from_host(hg)
    eps = 1e-9
    r, st = g.ppr(seed_vertex, D, eps, direction=direction)
    same, rounds, _ = orc.delta_push_f64(hg.offsets, hg.columns, D, f32(eps), seed=seed_vertex)
    exact = orc.pagerank_f64(hg.offsets, hg.columns, D, seed=seed_vertex)
    _say(f"ppr seed {seed_vertex}: L1 same {np.abs(r - same).sum():.3e} exact {np.abs(r - exact).sum():.3e}")
    assert np.abs(r - same).sum() <= 1e-6
    assert np.abs(r - exact).sum() <= 1e-5
    assert abs(r.sum() - 1) < 1e-5


def test_pagerank_invalid(gpu, abi):
    g = abi.DeviceGraph(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32))
    with pytest.raises(abi.InvalidArgument):
        g.pagerank(1.0, 1e-6)
    with pytest.raises(abi.InvalidArgument):
        g.pagerank(0.85, 0.0)
    with pytest.raises(abi.InvalidArgument):
        g.ppr(5)
