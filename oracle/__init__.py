"""Oracle: CPU restatement of the original project's hot-path algorithms.

TEST INFRASTRUCTURE ONLY. Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker of the HIP path and
never as the thing measured or shipped. The product (hytgraph_amd) does not
import it.

Two libraries:
  lib/liborc.so       oracle.c, the restatement (built by `make` here)
  _ref/libhyt_ref.so  the original CPU sources compiled in place plus
                      ref_driver.cpp (built by `make ref` where the original
                      source tree exists; the built file travels with the tree)

Pins of each restated function are listed in oracle.c's header.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORC_LIB = HERE / "lib" / "liborc.so"
REF_LIB = HERE / "_ref" / "libhyt_ref.so"
REFERENCE_TREE = Path("/root/reference")

_P = C.c_void_p
_u64 = C.c_uint64
_orc = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compiles liborc.so, and (when the original tree is present) _ref/libhyt_ref.so
    and the original's SEP header tests built against this repository's headers."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if with_ref and REFERENCE_TREE.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REF={REFERENCE_TREE}"], check=True)
        subprocess.run(["make", "-s", "-C", str(HERE), "sep", f"REF={REFERENCE_TREE}"], check=True)
        subprocess.run(["make", "-s", "-C", str(HERE), "unit", f"REF={REFERENCE_TREE}"], check=True)


def _load_orc():
    global _orc
    if _orc is None:
        if not ORC_LIB.exists():
            raise FileNotFoundError(f"{ORC_LIB} missing: run `make -C oracle`")
        lib = C.CDLL(str(ORC_LIB))
        lib.orc_sssp_jacobi_f32.argtypes = [_u64, _P, _P, _P, C.c_uint32, C.c_float, _u64, _P, _P, _P]
        lib.orc_sssp_dijkstra.argtypes = [_u64, _P, _P, _P, C.c_uint32, _P]
        lib.orc_pagerank_jacobi_f32.argtypes = [_u64, _P, _P, C.c_float, C.c_float, _u64, _P, _P, _P, _P]
        lib.orc_pagerank_f64.argtypes = [_u64, _P, _P, C.c_double, C.c_int64, C.c_double, _u64, _P, _P]
        lib.orc_delta_push_f64.argtypes = [_u64, _P, _P, C.c_double, C.c_double, C.c_int64, _u64, _P, _P, _P]
        lib.orc_hub_sort.argtypes = [_u64, _P, _P, C.c_double, _P, _P, _P]
        lib.orc_reorder.argtypes = [_u64, _P, _P, _P, _P, _P, _P, _P, _P]
        lib.orc_pagerank_contributions.argtypes = [_u64, _P, _P, C.c_float, _P, _P]
        lib.orc_sssp_contributions.argtypes = [_u64, _P, _P, C.c_float, _P]
        _orc = lib
    return _orc


def ref_available() -> bool:
    return REF_LIB.exists()


def _load_ref():
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise FileNotFoundError(f"{REF_LIB} missing: run `make -C oracle ref`")
        lib = C.CDLL(str(REF_LIB))
        lib.ref_sssp_cpu.argtypes = [_u64, _P, _P, _P, C.c_uint32, C.c_float, _u64, _P, _P, _P]
        lib.ref_pagerank_cpu.argtypes = [_u64, _P, _P, C.c_float, C.c_float, _u64, _P, _P, _P, _P]
        lib.ref_partition.argtypes = [_u64, _P, _P, _u64, _P, _P, _P, _P, _u64, _P]
        lib.ref_cost_evaluate.argtypes = [_u64, _P, _P, _P, _u64, _P, _P, _P, _P, _P, _u64, _P]
        lib.ref_combine.argtypes = [_P, _u64, _u64, _P, _P, _P, _P]
        lib.ref_hub_sort.argtypes = [_u64, _P, _P, C.c_double, _P, _P, _P]
        lib.ref_reorder.argtypes = [_u64, _P, _P, _P, _P, _P, _P, _P]
        lib.ref_pagerank_contributions.argtypes = [_u64, _P, _P, C.c_float, _P, _P]
        lib.ref_sssp_contributions.argtypes = [_u64, _P, _P, C.c_float, _P]
        _ref = lib
    return _ref


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


def _arr(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _check(rc: int, what: str):
    if rc != 0:
        raise OracleError(rc, what)


def sssp_jacobi_f32(off, col, w=None, source=0, tol=0.0, max_iter=0, use_ref=False):
    """sssp_cpu semantics. Returns (float32 distances, iterations, converged)."""
    off, col, w = _arr(off, np.uint64), _arr(col, np.uint32), _arr(w, np.float32)
    V = len(off) - 1
    out = np.empty(V, dtype=np.float32)
    it, conv = C.c_uint64(), C.c_int()
    lib = _load_ref() if use_ref else _load_orc()
    fn = lib.ref_sssp_cpu if use_ref else lib.orc_sssp_jacobi_f32
    _check(fn(V, _p(off), _p(col), _p(w), source, tol, max_iter, _p(out), C.byref(it), C.byref(conv)),
           "sssp_jacobi_f32")
    return out, it.value, bool(conv.value)


def sssp_dijkstra(off, col, w=None, source=0):
    """Exact u32 distances (UINT32_MAX unreachable)."""
    off, col, w = _arr(off, np.uint64), _arr(col, np.uint32), _arr(w, np.uint32)
    V = len(off) - 1
    out = np.empty(V, dtype=np.uint32)
    _check(_load_orc().orc_sssp_dijkstra(V, _p(off), _p(col), _p(w), source, _p(out)), "sssp_dijkstra")
    return out


def pagerank_jacobi_f32(off, col, damping=0.85, tol=1e-6, max_iter=100, use_ref=False):
    """pagerank_cpu semantics. Returns (ranks f32, iterations, residual, converged)."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    V = len(off) - 1
    out = np.empty(V, dtype=np.float32)
    it, res, conv = C.c_uint64(), C.c_float(), C.c_int()
    lib = _load_ref() if use_ref else _load_orc()
    fn = lib.ref_pagerank_cpu if use_ref else lib.orc_pagerank_jacobi_f32
    _check(fn(V, _p(off), _p(col), damping, tol, max_iter, _p(out), C.byref(it), C.byref(res), C.byref(conv)),
           "pagerank_jacobi_f32")
    return out, it.value, res.value, bool(conv.value)


def pagerank_f64(off, col, damping=0.85, seed=-1, tol=1e-14, max_iter=100000):
    """FP64 fixed point (PageRank, or PPR for seed >= 0), normalised to sum 1."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    V = len(off) - 1
    out = np.empty(V, dtype=np.float64)
    it = C.c_uint64()
    _check(_load_orc().orc_pagerank_f64(V, _p(off), _p(col), damping, seed, tol, max_iter, _p(out), C.byref(it)),
           "pagerank_f64")
    return out


def delta_push_f64(off, col, damping=0.85, eps=1e-7, seed=-1, max_rounds=0):
    """FP64 synchronous delta-push. Returns (normalised ranks, rounds, edges pushed)."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    V = len(off) - 1
    out = np.empty(V, dtype=np.float64)
    rounds, edges = C.c_uint64(), C.c_uint64()
    _check(_load_orc().orc_delta_push_f64(V, _p(off), _p(col), damping, eps, seed, max_rounds, _p(out),
                                          C.byref(rounds), C.byref(edges)), "delta_push_f64")
    return out, rounds.value, edges.value


def f32_to_u32_dist(d: np.ndarray) -> np.ndarray:
    """Float distances (+inf unreachable) -> u32 (UINT32_MAX unreachable).
    Exact only for integral values below 2^24."""
    out = np.full(d.shape, 0xFFFFFFFF, dtype=np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


# ----------------------------------------------- original planning code (via _ref)
def ref_partition(off, col, target_bytes):
    """The original LogicalPartitioner: (n, 4) u64 [vb, ve, eb, ee]."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    V = len(off) - 1
    cap = max(V, 1)
    vb, ve = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    eb, ee = np.zeros(cap, np.uint64), np.zeros(cap, np.uint64)
    n = C.c_uint64()
    _check(_load_ref().ref_partition(V, _p(off), _p(col), target_bytes, _p(vb), _p(ve), _p(eb), _p(ee), cap,
                                     C.byref(n)), "ref_partition")
    k = n.value
    return np.stack([vb[:k].astype(np.uint64), ve[:k].astype(np.uint64), eb[:k], ee[:k]], axis=1)


def ref_cost_evaluate(off, col, active, target_bytes, alpha, beta, gamma, rtt, cpu, m, mr, d1, d2, align):
    """The original HyTMCostModel over its own partitions: (metrics (n,8), costs (n,4), engines (n))."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    act = _arr(active, np.uint8)
    V = len(off) - 1
    cap = max(V, 1)
    dp = np.array([alpha, beta, gamma, rtt, cpu], np.float64)
    up = np.array([m, mr, d1, d2, align], np.uint64)
    metrics = np.zeros((cap, 8), np.uint64)
    costs = np.zeros((cap, 4), np.float64)
    engines = np.zeros(cap, np.int32)
    n = C.c_uint64()
    _check(_load_ref().ref_cost_evaluate(V, _p(off), _p(col), _p(act), target_bytes, _p(dp), _p(up), _p(metrics),
                                         _p(costs), _p(engines), cap, C.byref(n)), "ref_cost_evaluate")
    k = n.value
    return metrics[:k], costs[:k], engines[:k]


def ref_combine(engines, k=4):
    e = _arr(engines, np.int32)
    n = len(e)
    te = np.zeros(max(n, 1), np.int32)
    to = np.zeros(n + 1, np.uint64)
    mem = np.zeros(max(n, 1), np.uint64)
    nt = C.c_uint64()
    _check(_load_ref().ref_combine(_p(e), n, k, _p(te), _p(to), _p(mem), C.byref(nt)), "ref_combine")
    return [(int(te[t]), mem[int(to[t]):int(to[t + 1])].tolist()) for t in range(nt.value)]


def hub_sort(off, col, fraction=0.08, use_ref=False):
    """HubSorter::sort: (vertex_order u32[V], scores f64[V], hub_count)."""
    off, col = _arr(off, np.uint64), _arr(col, np.uint32)
    V = len(off) - 1
    order = np.empty(V, np.uint32)
    scores = np.empty(V, np.float64)
    hubs = C.c_uint64()
    fn = _load_ref().ref_hub_sort if use_ref else _load_orc().orc_hub_sort
    _check(fn(V, _p(off), _p(col), C.c_double(fraction), _p(order), _p(scores), C.byref(hubs)), "hub_sort")
    return order, scores, hubs.value


def reorder(off, col, w, order, use_ref=False):
    """CSRGraph::reorder_vertices: (offsets u64[V+1], columns, weights or
    None). w: u32 bit patterns (float bits for the compiled original) or None. OracleError with code 1 on an invalid
    order; .kind = 1 (out of range) or 2 (duplicate) from the restatement."""
    off, col, order = _arr(off, np.uint64), _arr(col, np.uint32), _arr(order, np.uint32)
    w = _arr(w, np.uint32)
    V = len(off) - 1
    off2 = np.empty(V + 1, np.uint64)
    col2 = np.empty(len(col), np.uint32)
    w2 = None if w is None else np.empty(len(col), np.uint32)
    if use_ref:
        rc = _load_ref().ref_reorder(V, _p(off), _p(col), _p(w), _p(order), _p(off2), _p(col2), _p(w2))
        _check(rc, "reorder")
    else:
        kind = C.c_int()
        rc = _load_orc().orc_reorder(V, _p(off), _p(col), _p(w), _p(order), _p(off2), _p(col2), _p(w2),
                                     C.byref(kind))
        if rc != 0:
            err = OracleError(rc, "reorder")
            err.kind = kind.value
            raise err
    return off2, col2, w2


def pagerank_contributions(off, col, x, damping=0.85, use_ref=False):
    off, col, x = _arr(off, np.uint64), _arr(col, np.uint32), _arr(x, np.float32)
    V = len(off) - 1
    out = np.empty(V, np.float64)
    fn = _load_ref().ref_pagerank_contributions if use_ref else _load_orc().orc_pagerank_contributions
    _check(fn(V, _p(off), _p(col), C.c_float(damping), _p(x), _p(out)), "pagerank_contributions")
    return out


def sssp_contributions(cur, cand, tol=0.0, use_ref=False):
    cur, cand = _arr(cur, np.float32), _arr(cand, np.float32)
    out = np.empty(len(cur), np.float64)
    fn = _load_ref().ref_sssp_contributions if use_ref else _load_orc().orc_sssp_contributions
    _check(fn(len(cur), _p(cur), _p(cand), C.c_float(tol), _p(out)), "sssp_contributions")
    return out
