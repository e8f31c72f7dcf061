"""Graph files (CPU): the original's text edge-list rules and the binary CSR.

Text loader pins: the original's own known answers
(tests/unit_tests.cpp:373-592 - unweighted, weighted, parallel edges,
out-of-range / missing-weight / malformed records), its messages, and, on
seeded random edge lists, numpy's stable sort by source (the grouping rule of
src/graph/graph_loader.cpp:130-179). Binary CSR: bit-exact round trips of
generated graphs, and every corruption the header and checksums must catch.
"""
import os

import numpy as np
import pytest


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_text_known_answers(abi, tmp_path):
    g = abi.HostGraph.from_edge_list(_write(tmp_path, "u.el", "# simple directed graph\n0 1\n0 2\n1 2\n2 0\n"), 4)
    assert (g.V, g.E, g.weights) == (4, 4, None)
    assert g.offsets.tolist() == [0, 2, 3, 4, 4] and g.columns.tolist() == [1, 2, 2, 0]
    g = abi.HostGraph.from_edge_list(_write(tmp_path, "w.el", "# weighted graph\n0 1 0.5\n0 2 1.25\n1 2 2.5\n"), 3,
                                     weighted=True)
    assert g.weights_f32 and g.weights.view(np.float32).tolist() == [0.5, 1.25, 2.5]
    g = abi.HostGraph.from_edge_list(_write(tmp_path, "p.el", "0 1\n0 1\n0 2\n"), 3)
    assert g.offsets.tolist() == [0, 3, 3, 3] and g.columns.tolist() == [1, 1, 2]


@pytest.mark.parametrize("text,V,weighted,msg", [
    ("0 3\n", 3, False, "vertex ID out of range at line 1"),
    ("0 1\n", 2, True, "missing edge weight at line 1"),
    ("this is not an edge\n", 2, False, "malformed edge at line 1"),
    ("# c\n\n0 1\n1 x\n", 2, False, "malformed edge at line 4"),
    ("0 1 2\n", 3, False, "unexpected data at line 1"),
    ("0 1 2.5 7\n", 3, True, "unexpected data at line 1"),
    ("0 1 2.5.3\n", 3, True, "unexpected data at line 1"),
    ("1.5 2\n", 3, False, "malformed edge at line 1"),
    ("0 1 inf\n", 3, True, "missing edge weight at line 1"),
    ("0 1 nan\n", 3, True, "missing edge weight at line 1"),
    ("0 99999999999999999999\n", 3, False, "malformed edge at line 1"),
])
def test_text_rejects(abi, tmp_path, text, V, weighted, msg):
    with pytest.raises(abi.InvalidArgument) as ei:
        abi.HostGraph.from_edge_list(_write(tmp_path, "bad.el", text), V, weighted=weighted)
    assert str(ei.value).endswith("GraphLoader::load_edge_list: " + msg), str(ei.value)


def test_text_unreadable_file(abi, tmp_path):
    with pytest.raises(abi.HytError) as ei:
        abi.HostGraph.from_edge_list(tmp_path / "missing.el", 3)
    assert ei.value.status == abi.HYT_ERR_IO and "unable to open file" in str(ei.value)


def test_text_whitespace_and_comments(abi, tmp_path):
    text = "  # indented comment\r\n\t0\t2 \t1e1\r\n\n+1 0 .5\n  2   1   3.\n0 1 -2E-1\n"
    g = abi.HostGraph.from_edge_list(_write(tmp_path, "ws.el", text), 3, weighted=True)
    assert g.offsets.tolist() == [0, 2, 3, 4]
    assert g.columns.tolist() == [2, 1, 0, 1]
    assert g.weights.view(np.float32).tolist() == pytest.approx([10.0, -0.2, 0.5, 3.0])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_text_grouping_matches_stable_sort(abi, tmp_path, seed):
    rng = np.random.default_rng(seed)
    V, E = 200, 3000
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    w = (rng.integers(1, 1000, E) / 8.0).astype(np.float32)   # exact in decimal
    lines = [f"{s} {d} {x!r}" for s, d, x in zip(src, dst, w.tolist())]
    for k in range(0, E, 97):
        lines.insert(k, "# noise")
    g = abi.HostGraph.from_edge_list(_write(tmp_path, "r.el", "\n".join(lines) + "\n"), V, weighted=True)
    order = np.argsort(src, kind="stable")
    assert np.array_equal(g.columns, dst[order].astype(np.uint32))
    assert np.array_equal(g.weights.view(np.float32), w[order])
    assert np.array_equal(g.offsets, np.concatenate([[0], np.cumsum(np.bincount(src, minlength=V))]))


@pytest.mark.parametrize("weighted", [True, False])
def test_binary_round_trip(abi, tmp_path, weighted):
    hg = abi.HostGraph(scale=13, seed=5, weight_min=1 if weighted else 0, weight_max=255 if weighted else 0)
    p = tmp_path / "g.hytcsr"
    hg.save(p)
    assert os.path.getsize(p) == 128 + (hg.V + 1) * 8 + hg.E * 4 * (2 if weighted else 1)
    g = abi.HostGraph.load(p)
    assert (g.V, g.E) == (hg.V, hg.E)
    assert np.array_equal(g.offsets, hg.offsets) and np.array_equal(g.columns, hg.columns)
    assert (g.weights is None) == (not weighted)
    if weighted:
        assert np.array_equal(g.weights, hg.weights) and not g.weights_f32
    # the float flag travels with the file
    hg.save(p, weights_f32=True)
    assert abi.HostGraph.load(p).weights_f32 == weighted


def _corrupt(path, at, byte=None):
    data = bytearray(path.read_bytes())
    data[at] ^= 0x5A if byte is None else byte
    path.write_bytes(bytes(data))


def test_binary_detects_corruption(abi, tmp_path):
    hg = abi.HostGraph(scale=11, seed=9)
    p = tmp_path / "g.hytcsr"
    col0 = 128 + (hg.V + 1) * 8
    for where, what in [(col0 + 4 * 100, "columns checksum"), (128 + 8 * 7, "row offsets checksum"),
                        (col0 + hg.E * 4 + 12, "weights checksum"), (20, "header checksum"),
                        (0, "magic")]:
        hg.save(p)
        _corrupt(p, where)
        with pytest.raises(abi.InvalidArgument) as ei:
            abi.HostGraph.load(p)
        assert what in str(ei.value), (where, str(ei.value))
    # without verification a flipped column byte loads (the caller opted out)
    hg.save(p)
    _corrupt(p, col0 + 4 * 100)
    g = abi.HostGraph.load(p, verify=False)
    assert g.E == hg.E
    # truncated file
    hg.save(p)
    p.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(abi.InvalidArgument, match="size mismatch"):
        abi.HostGraph.load(p)
    with pytest.raises(abi.HytError) as ei:
        abi.HostGraph.load(tmp_path / "none.hytcsr")
    assert ei.value.status == abi.HYT_ERR_IO


def test_from_csr_validates(abi):
    off = np.array([0, 2, 3, 3], np.uint64)
    g = abi.HostGraph.from_csr(off, np.array([1, 2, 0], np.uint32), np.array([5, 6, 7], np.uint32))
    assert g.offsets.tolist() == [0, 2, 3, 3] and g.weights.tolist() == [5, 6, 7]
    with pytest.raises(abi.InvalidArgument, match="non-decreasing"):
        abi.HostGraph.from_csr(np.array([0, 2, 1, 3], np.uint64), np.array([1, 2, 0], np.uint32))
    with pytest.raises(abi.InvalidArgument, match="out of range"):
        abi.HostGraph.from_csr(off, np.array([1, 3, 0], np.uint32))
    with pytest.raises(abi.InvalidArgument, match="edge count"):
        abi.HostGraph.from_csr(np.array([0, 2, 3, 4], np.uint64), np.array([1, 2, 0], np.uint32))
