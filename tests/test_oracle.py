"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Pins, per function:
  sssp_jacobi_f32       original known answers (tests/algorithm_tests.cpp:100-206),
                        bit-exact against the compiled original (oracle/_ref),
                        Dijkstra and scipy on random graphs
  sssp_dijkstra         scipy.sparse.csgraph.dijkstra
  pagerank_jacobi_f32   original known answers (algorithm_tests.cpp:22-98), bit-exact
                        against oracle/_ref (ranks, residual, iterations)
  pagerank_f64          dense linear solve of (I - d P~^T) y = (1-d) t
  delta_push_f64        convergence to pagerank_f64 as eps -> 0; mass balance
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from conftest import csr_from_edges

GOLDEN = Path(__file__).resolve().parent / "golden"


def rmat(abi, scale, seed, wmax=64, ef=16, kind=0):
    g = abi.HostGraph(kind=kind, scale=scale, edge_factor=ef, seed=seed, weight_min=1 if wmax else 0,
                      weight_max=wmax)
    return g.offsets.copy(), g.columns.copy(), None if g.weights is None else g.weights.copy()


# ---------------------------------------------------------------- known answers
def test_sssp_weighted_known_answer(orc):
    # algorithm_tests.cpp:100-129
    off, col, w = np.array([0, 2, 4, 5, 5], np.uint64), np.array([1, 2, 2, 3, 3], np.uint32), \
        np.array([1, 4, 2, 6, 1], np.float32)
    d, it, conv = orc.sssp_jacobi_f32(off, col, w, 0)
    assert conv and d.tolist() == [0.0, 1.0, 3.0, 4.0]


def test_sssp_unweighted_unreachable_known_answer(orc):
    # algorithm_tests.cpp:131-157
    d, it, conv = orc.sssp_jacobi_f32(np.array([0, 1, 2, 2, 2], np.uint64), np.array([1, 2], np.uint32), None, 0)
    assert conv and d[:3].tolist() == [0.0, 1.0, 2.0] and np.isinf(d[3])


def test_sssp_rejects_negative_weight_and_bad_source(orc):
    # algorithm_tests.cpp:159-206
    with pytest.raises(orc.OracleError):
        orc.sssp_jacobi_f32(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([-1.0], np.float32))
    with pytest.raises(orc.OracleError):
        orc.sssp_jacobi_f32(np.array([0, 0, 0], np.uint64), np.array([], np.uint32), None, source=2)


def test_pagerank_known_answers(orc):
    # algorithm_tests.cpp:22-75: 3-cycle -> 1/3 each; dangling graph sums to 1
    r, it, res, conv = orc.pagerank_jacobi_f32(np.array([0, 1, 2, 3], np.uint64), np.array([1, 2, 0], np.uint32))
    assert conv and np.all(np.abs(r - 1 / 3) <= 1e-5)
    r, it, res, conv = orc.pagerank_jacobi_f32(np.array([0, 1, 2, 2], np.uint64), np.array([1, 0], np.uint32),
                                               max_iter=200)
    assert conv and abs(r.sum() - 1) <= 1e-5
    with pytest.raises(orc.OracleError):  # algorithm_tests.cpp:77-98, damping = 1
        orc.pagerank_jacobi_f32(np.array([0, 0], np.uint64), np.array([], np.uint32), damping=1.0)


def test_golden_fixtures(orc):
    """Committed golden vectors (written by tests/golden/make_golden.py from the
    compiled original) reproduce bit-exactly."""
    data = json.loads((GOLDEN / "reference_cases.json").read_text())
    for case in data["sssp"]:
        w = None if case["weights"] is None else np.array(case["weights"], np.float32)
        d, it, conv = orc.sssp_jacobi_f32(np.array(case["offsets"], np.uint64), np.array(case["columns"], np.uint32),
                                          w, case["source"], case["tolerance"], case["max_iterations"])
        assert d.view(np.uint32).tolist() == case["distance_bits"], case["name"]
        assert it == case["iterations"] and conv == case["converged"], case["name"]
    for case in data["pagerank"]:
        r, it, res, conv = orc.pagerank_jacobi_f32(np.array(case["offsets"], np.uint64),
                                                   np.array(case["columns"], np.uint32), case["damping"],
                                                   case["tolerance"], case["max_iterations"])
        assert r.view(np.uint32).tolist() == case["rank_bits"], case["name"]
        assert it == case["iterations"] and conv == case["converged"], case["name"]
        assert np.float32(res).view(np.uint32) == case["residual_bits"], case["name"]


# ------------------------------------------------ against the compiled original
needs_ref = pytest.mark.skipif(not __import__("oracle").ref_available(),
                               reason="oracle/_ref not built (original source tree absent)")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sssp_matches_original_bitexact(orc, abi, seed):
    off, col, w = rmat(abi, 11, seed)
    rng = np.random.default_rng(seed)
    wf = (w.astype(np.float32) * rng.choice([1.0, 0.37, 1.5], size=len(w))).astype(np.float32)
    for weights, tol, max_it in [(w.astype(np.float32), 0.0, 0), (wf, 0.0, 0), (wf, 0.75, 0), (None, 0.0, 0),
                                 (wf, 0.0, 3)]:
        a = orc.sssp_jacobi_f32(off, col, weights, 5, tol, max_it)
        b = orc.sssp_jacobi_f32(off, col, weights, 5, tol, max_it, use_ref=True)
        assert a[0].view(np.uint32).tolist() == b[0].view(np.uint32).tolist()
        assert a[1:] == b[1:]


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_pagerank_matches_original_bitexact(orc, abi, seed):
    off, col, _ = rmat(abi, 11, seed, wmax=0)
    for d, tol, it in [(0.85, 1e-6, 100), (0.5, 1e-4, 100), (0.85, 1e-9, 7)]:
        a = orc.pagerank_jacobi_f32(off, col, d, tol, it)
        b = orc.pagerank_jacobi_f32(off, col, d, tol, it, use_ref=True)
        assert a[0].view(np.uint32).tolist() == b[0].view(np.uint32).tolist()
        assert a[1] == b[1] and a[3] == b[3] and np.float32(a[2]) == np.float32(b[2])


# ----------------------------------------------------------- mathematical pins
@pytest.mark.parametrize("seed", [1, 7])
def test_dijkstra_matches_scipy(orc, abi, seed):
    off, col, w = rmat(abi, 12, seed)
    V = len(off) - 1
    A = sp.csr_matrix((w.astype(np.float64), col, off.astype(np.int64)), shape=(V, V))
    for s in [0, 17, V - 1]:
        ref = cg.dijkstra(A, indices=s)
        mine = orc.sssp_dijkstra(off, col, w, s)
        exp = np.where(np.isinf(ref), 0xFFFFFFFF, ref).astype(np.uint64)
        assert np.array_equal(mine.astype(np.uint64), exp)


def test_sssp_jacobi_equals_dijkstra_on_integer_weights(orc, abi):
    off, col, w = rmat(abi, 12, 3)
    d, it, conv = orc.sssp_jacobi_f32(off, col, w.astype(np.float32), 9)
    assert conv
    assert np.array_equal(orc.f32_to_u32_dist(d), orc.sssp_dijkstra(off, col, w, 9))


def test_sssp_jacobi_iteration_count_is_hop_bound(orc):
    # A path of n vertices converges after n-1 improving rounds + 1 quiet round.
    n = 9
    off, col, _ = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    d, it, conv = orc.sssp_jacobi_f32(off, col, None, 0)
    assert conv and it == n and d.tolist() == list(range(n))


def _dense_pagerank(off, col, d, seed=-1):
    V = len(off) - 1
    P = np.zeros((V, V))
    for u in range(V):
        deg = int(off[u + 1] - off[u])
        for e in range(int(off[u]), int(off[u + 1])):
            P[u, col[e]] += 1.0 / deg
    t = np.full(V, 1.0 / V) if seed < 0 else np.eye(V)[seed]
    y = np.linalg.solve(np.eye(V) - d * P.T, (1 - d) * t)
    return y / y.sum()


@pytest.mark.parametrize("seed", [-1, 0, 5])
def test_pagerank_f64_matches_dense_solve(orc, abi, seed):
    off, col, _ = rmat(abi, 8, 4, wmax=0, ef=4)  # has dangling vertices
    assert np.any(np.diff(off.astype(np.int64)) == 0)
    x = orc.pagerank_f64(off, col, 0.85, seed=seed)
    y = _dense_pagerank(off, col, 0.85, seed)
    assert np.abs(x - y).sum() < 1e-11


def test_pagerank_f64_closed_forms(orc):
    off, col, _ = csr_from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0)])
    assert np.allclose(orc.pagerank_f64(off, col, 0.85), 0.25, atol=1e-14)
    # damping 0: uniform regardless of structure
    off, col, _ = csr_from_edges(3, [(0, 1), (0, 2)])
    assert np.allclose(orc.pagerank_f64(off, col, 0.0), 1 / 3, atol=1e-15)


@pytest.mark.parametrize("seed", [-1, 3])
def test_delta_push_converges_to_fixed_point(orc, abi, seed):
    off, col, _ = rmat(abi, 10, 2, wmax=0)
    exact = orc.pagerank_f64(off, col, 0.85, seed=seed)
    errs = []
    for eps in [1e-4, 1e-6, 1e-8, 1e-10]:
        y, rounds, edges = orc.delta_push_f64(off, col, 0.85, eps, seed=seed)
        errs.append(np.abs(y - exact).sum())
    assert errs[-1] < 1e-7
    assert all(errs[i + 1] < errs[i] for i in range(len(errs) - 1))


def test_delta_push_error_bound(orc, abi):
    # |y/|y| - x|_1 <= 2 eps / (1-d)^2 for the global form (y >= (1-d) 1).
    off, col, _ = rmat(abi, 10, 5, wmax=0)
    exact = orc.pagerank_f64(off, col, 0.85)
    for eps in [1e-5, 1e-7]:
        y, _, _ = orc.delta_push_f64(off, col, 0.85, eps)
        assert np.abs(y - exact).sum() <= 2 * eps / 0.15 ** 2
