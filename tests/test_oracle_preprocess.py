"""Pins the oracle's vertex-order and contribution restatements, and the
device hub score's integer emulation of long double division, on the CPU.

  hub_sort                original known answers (tests/unit_tests.cpp:2468-2602,
                          2710-2750), bit-exact against oracle/_ref on RMAT / ER
                          graphs and hub fractions 0 ... 1
  reorder                 original known answers (unit_tests.cpp:2604-2708),
                          oracle/_ref, error precedence, and SSSP distances
                          permuting with the vertex ids
  pagerank_contributions  oracle/_ref bit-exact; equal to |x1 - x0| of the first
                          pagerank_cpu iteration from x0 = 1/n
  sssp_contributions      oracle/_ref bit-exact; hand-worked inf / NaN /
                          tolerance cases
  extended_quotient       (hytgraph_amd/csrc/device/extended_quotient.hpp)
                          against g++'s x87 long double on ~10 M pairs
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
needs_ref = pytest.mark.skipif(not __import__("oracle").ref_available(),
                               reason="oracle/_ref not built (original source tree absent)")


def rmat(abi, scale, seed, wmax=64, kind=0):
    g = abi.HostGraph(kind=kind, scale=scale, edge_factor=16, seed=seed, weight_min=1 if wmax else 0,
                      weight_max=wmax)
    return g.offsets.copy(), g.columns.copy(), None if g.weights is None else g.weights.copy()


# ---------------------------------------------------------------- known answers
def test_hub_sort_known_answers(orc):
    # unit_tests.cpp:2468-2496: scores 0.5, 0.25, 0.5
    _, s, _ = orc.hub_sort([0, 2, 3, 4], [1, 2, 2, 0], 2.0 / 3.0)
    assert np.allclose(s, [0.5, 0.25, 0.5], atol=1e-12, rtol=0)
    # unit_tests.cpp:2498-2537 and 2710-2750: two hubs 0, 1 then 2, 3, 4
    order, _, hubs = orc.hub_sort([0, 3, 4, 5, 6, 7], [1, 2, 3, 0, 0, 0, 0], 0.4)
    assert hubs == 2 and order.tolist() == [0, 1, 2, 3, 4]
    off2, col2, _ = orc.reorder([0, 3, 4, 5, 6, 7], [1, 2, 3, 0, 0, 0, 0], None, order)
    assert off2.tolist() == [0, 3, 4, 5, 6, 7] and col2.tolist() == [1, 2, 3, 0, 0, 0, 0]
    # unit_tests.cpp:2539-2573: a permutation of all 6 vertices (default fraction)
    order, _, _ = orc.hub_sort([0, 2, 3, 4, 5, 6, 6], [1, 2, 2, 3, 0, 0])
    assert sorted(order.tolist()) == list(range(6))
    # unit_tests.cpp:2575-2602: edgeless graph, one hub, natural order, zero scores
    order, s, hubs = orc.hub_sort([0, 0, 0, 0, 0], [])
    assert hubs == 1 and order.tolist() == [0, 1, 2, 3] and not s.any()
    with pytest.raises(orc.OracleError):
        orc.hub_sort([0, 0], [], 1.5)


def test_reorder_known_answers(orc):
    # unit_tests.cpp:2604-2660
    w = np.array([1, 2, 3, 4], np.float32).view(np.uint32)
    off2, col2, w2 = orc.reorder([0, 2, 3, 4], [1, 2, 2, 0], w, [2, 0, 1])
    assert off2.tolist() == [0, 1, 3, 4] and col2.tolist() == [1, 2, 0, 0]
    assert w2.view(np.float32).tolist() == [4.0, 1.0, 2.0, 3.0]
    # unit_tests.cpp:2662-2708: duplicate and out-of-range ids; the first
    # offending position decides which
    for order, kind in (([0, 1, 1], 2), ([0, 1, 3], 1), ([3, 0, 0], 1), ([0, 0, 3], 2)):
        with pytest.raises(orc.OracleError) as e:
            orc.reorder([0, 1, 2, 3], [1, 2, 0], None, order)
        assert e.value.kind == kind, order


def test_sssp_contributions_hand_cases(orc):
    inf, nan = np.inf, np.nan
    cur = [5.0, 5.0, inf, inf, 3.0, nan, 2.0, 1e30]
    cand = [3.0, 4.75, 7.0, inf, inf, 1.0, 2.0, 1e29]
    out = orc.sssp_contributions(cur, cand, 0.25)
    # 2 > 0.25; 0.25 is not > 0.25; newly reachable; both infinite; candidate
    # infinite; current NaN (not finite) and candidate finite: unit; no gain;
    # large finite gain
    assert out.tolist() == [2.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, float(np.float32(1e30) - np.float32(1e29))]
    with pytest.raises(orc.OracleError):
        orc.sssp_contributions([1.0], [0.0], -1.0)


def test_pagerank_contributions_equal_first_jacobi_step(orc, abi):
    off, col, _ = rmat(abi, 10, 4, wmax=0)
    V = len(off) - 1
    x0 = np.full(V, np.float32(1.0) / np.float32(V), np.float32)
    x1, _, _, _ = orc.pagerank_jacobi_f32(off, col, 0.85, 1e-30, 1)
    c = orc.pagerank_contributions(off, col, x0, 0.85)
    assert np.array_equal(c, np.abs(x1 - x0).astype(np.float64))


# ---------------------------------------------------------- the compiled original
@needs_ref
@pytest.mark.parametrize("kind,scale,seed", [(0, 10, 1), (0, 12, 2), (1, 11, 3)])
def test_hub_sort_and_reorder_match_original(orc, abi, kind, scale, seed):
    off, col, w = rmat(abi, scale, seed, kind=kind)
    for frac in (0.0, 0.08, 1.0 / 3.0, 0.5, 1.0):
        a = orc.hub_sort(off, col, frac)
        b = orc.hub_sort(off, col, frac, use_ref=True)
        assert np.array_equal(a[0], b[0]) and a[2] == b[2], frac
        assert a[1].view(np.uint64).tolist() == b[1].view(np.uint64).tolist()
    order = orc.hub_sort(off, col, 0.08)[0]
    rng = np.random.default_rng(seed)
    wf = (w.astype(np.float32) * np.float32(0.37)).astype(np.float32)
    for perm in (order, rng.permutation(len(off) - 1).astype(np.uint32)):
        a = orc.reorder(off, col, wf.view(np.uint32), perm)
        b = orc.reorder(off, col, wf.view(np.uint32), perm, use_ref=True)  # weights travel as float bits
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for bad in ([0] * (len(off) - 1), list(range(len(off) - 2)) + [len(off) + 5]):
        with pytest.raises(orc.OracleError):
            orc.reorder(off, col, None, bad, use_ref=True)


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_contributions_match_original(orc, abi, seed):
    off, col, _ = rmat(abi, 11, seed, wmax=0)
    V = len(off) - 1
    rng = np.random.default_rng(seed)
    for x in (np.full(V, np.float32(1.0 / V), np.float32), rng.random(V, np.float32) / np.float32(V),
              (rng.random(V) * 10).astype(np.float32)):
        for d in (0.85, 0.5, 0.0):
            a = orc.pagerank_contributions(off, col, x, d)
            b = orc.pagerank_contributions(off, col, x, d, use_ref=True)
            assert a.view(np.uint64).tolist() == b.view(np.uint64).tolist()
    cur = rng.random(5000).astype(np.float32) * 100
    cand = cur - rng.random(5000).astype(np.float32) * 2
    cur[::7] = np.inf
    cand[::11] = np.inf
    for tol in (0.0, 0.5, 1.5):
        a = orc.sssp_contributions(cur, cand, tol)
        b = orc.sssp_contributions(cur, cand, tol, use_ref=True)
        assert a.view(np.uint64).tolist() == b.view(np.uint64).tolist()


# ------------------------------------------------------------- mathematical pins
def test_reorder_permutes_sssp_distances(orc, abi):
    off, col, w = rmat(abi, 11, 9)
    rng = np.random.default_rng(9)
    perm = rng.permutation(len(off) - 1).astype(np.uint32)
    off2, col2, w2 = orc.reorder(off, col, w, perm)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm), dtype=np.uint32)
    for src in (0, 17, 300):
        d = orc.sssp_dijkstra(off, col, w, src)
        d2 = orc.sssp_dijkstra(off2, col2, w2, int(inv[src]))
        assert np.array_equal(d2[inv], d)


def test_extended_quotient_matches_long_double():
    exe = ROOT / "hytgraph_amd" / "lib" / "extended_quotient_test"  # built from tests/cpp
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
