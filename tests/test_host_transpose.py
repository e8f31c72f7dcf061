"""Host-built incoming CSR (hyt_host_graph_transpose): the transpose the
original's pull kernels read (build_incoming_csr, src/algorithms/sssp.cpp:21-80),
sources of each destination in ascending order. Checked against a stable
numpy sort of the edge list by destination, for several thread counts."""
import numpy as np
import pytest

from conftest import csr_from_edges


def reference_transpose(offsets, columns, V):
    src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(offsets.astype(np.int64)))
    order = np.argsort(columns, kind="stable")  # edges are already in source order
    in_src = src[order]
    in_off = np.zeros(V + 1, np.uint64)
    np.cumsum(np.bincount(columns, minlength=V), out=in_off[1:])
    return in_off, in_src


@pytest.mark.parametrize("threads", [1, 3, 16, 0])
def test_rmat_transpose_matches_stable_sort(abi, threads):
    g = abi.HostGraph(scale=12, edge_factor=16, seed=4, weight_min=1, weight_max=9)
    in_off, in_src = g.transpose(threads)
    want_off, want_src = reference_transpose(g.offsets, g.columns, g.V)
    assert np.array_equal(in_off, want_off)
    assert np.array_equal(in_src, want_src)
    # a second call returns the kept transpose
    o2, s2 = g.transpose(2)
    assert np.array_equal(o2, in_off) and np.array_equal(s2, in_src)


def test_er_and_edge_cases(abi):
    e = abi.HostGraph(kind=abi.HYT_GEN_ER, scale=11, edge_factor=8, seed=2, weight_min=0, weight_max=0)
    in_off, in_src = e.transpose(7)
    want_off, want_src = reference_transpose(e.offsets, e.columns, e.V)
    assert np.array_equal(in_off, want_off) and np.array_equal(in_src, want_src)
    # parallel edges, self-loops, a vertex with no edges either way, a hub
    edges = [(0, 1), (0, 1), (2, 2), (3, 0), (1, 0), (5, 0), (5, 3)] + [(4, v) for v in (0, 1, 2, 3, 5, 5)]
    off, col, _ = csr_from_edges(7, edges)
    h = abi.HostGraph.from_csr(off, col)
    in_off, in_src = h.transpose(4)
    want_off, want_src = reference_transpose(off, col, 7)
    assert np.array_equal(in_off, want_off) and np.array_equal(in_src, want_src)
    assert in_off[-1] == len(edges) and in_off[6] == in_off[7]  # vertex 6: no in-edges
    one = abi.HostGraph.from_csr(np.array([0, 0], np.uint64), np.array([], np.uint32))
    o1, s1 = one.transpose()
    assert o1.tolist() == [0, 0] and len(s1) == 0
