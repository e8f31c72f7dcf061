"""Writes tests/golden/reference_cases.json from the ORIGINAL project's CPU code.

Runs in the build container only (needs oracle/_ref/libhyt_ref.so, compiled
from the original sources by `make -C oracle ref`). The fixture holds inputs
and the exact outputs (float bit patterns, iteration counts) of sssp_cpu and
pagerank_cpu on the original's own test graphs
(tests/algorithm_tests.cpp, tests/cuda_algorithm_tests.cu) and on a few small
seeded random graphs.

Usage: python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402


def random_graph(rng, V, E, frac=True):
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    keep = src != dst
    pairs = sorted(set(zip(src[keep].tolist(), dst[keep].tolist())))
    off = np.zeros(V + 1, np.uint64)
    for s, _ in pairs:
        off[s + 1] += 1
    off = np.cumsum(off, dtype=np.uint64)
    col = np.array([d for _, d in pairs], np.uint32)
    w = rng.integers(1, 20, len(pairs)).astype(np.float32)
    if frac:
        w = (w * np.float32(0.37)).astype(np.float32)
    return off, col, w


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libhyt_ref.so missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20240601)
    sssp_inputs = [
        # algorithm_tests.cpp:100-129 and cuda_algorithm_tests.cu:58-87
        ("weighted4", [0, 2, 4, 5, 5], [1, 2, 2, 3, 3], [1.0, 4.0, 2.0, 6.0, 1.0], 0, 0.0, 0),
        # algorithm_tests.cpp:131-157
        ("chain_unreachable", [0, 1, 2, 2, 2], [1, 2], None, 0, 0.0, 0),
    ]
    off, col, w = random_graph(rng, 64, 400)
    sssp_inputs += [
        ("random64_frac", off.tolist(), col.tolist(), w.tolist(), 3, 0.0, 0),
        ("random64_tol", off.tolist(), col.tolist(), w.tolist(), 3, 0.5, 0),
        ("random64_trunc", off.tolist(), col.tolist(), w.tolist(), 3, 0.0, 2),
    ]
    off2, col2, _ = random_graph(rng, 200, 900, frac=False)
    sssp_inputs.append(("random200_unweighted", off2.tolist(), col2.tolist(), None, 7, 0.0, 0))

    out = {"sssp": [], "pagerank": [], "source": "original sssp_cpu / pagerank_cpu compiled from source"}
    for name, o, c, ww, s, tol, mi in sssp_inputs:
        d, it, conv = oracle.sssp_jacobi_f32(o, c, ww, s, tol, mi, use_ref=True)
        out["sssp"].append({"name": name, "offsets": o, "columns": c, "weights": ww, "source": s,
                            "tolerance": tol, "max_iterations": mi,
                            "distance_bits": d.view(np.uint32).tolist(), "iterations": it, "converged": conv})

    pr_inputs = [
        ("cycle3", [0, 1, 2, 3], [1, 2, 0], 0.85, 1e-6, 100),            # algorithm_tests.cpp:22-47
        ("dangling3", [0, 1, 2, 2], [1, 0], 0.85, 1e-6, 200),            # algorithm_tests.cpp:49-75
        ("cuda3", [0, 2, 3, 4], [1, 2, 2, 0], 0.85, 1e-6, 200),          # cuda_algorithm_tests.cu:23-56
        ("random200", off2.tolist(), col2.tolist(), 0.85, 1e-6, 100),
        ("random200_d05", off2.tolist(), col2.tolist(), 0.5, 1e-5, 100),
    ]
    for name, o, c, d, tol, mi in pr_inputs:
        r, it, res, conv = oracle.pagerank_jacobi_f32(o, c, d, tol, mi, use_ref=True)
        out["pagerank"].append({"name": name, "offsets": o, "columns": c, "damping": d, "tolerance": tol,
                                "max_iterations": mi, "rank_bits": r.view(np.uint32).tolist(),
                                "iterations": it, "converged": conv,
                                "residual_bits": int(np.float32(res).view(np.uint32))})
    path = Path(__file__).resolve().parent / "reference_cases.json"
    path.write_text(json.dumps(out))
    print(f"wrote {path} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
