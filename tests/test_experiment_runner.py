"""The experiment executable and runner (counterparts of the original's
src/main.cpp and experiments/run_experiment.py): command line, the result
record on failure (CPU), and on the GPU the measurements of real runs against
the oracle on the same generated graph."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "hytgraph_amd" / "lib" / "hyt_experiment"


def _run(*args):
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=300)


def _runner(tmp_path, yaml_text):
    cfg = tmp_path / "exp.yaml"
    cfg.write_text(yaml_text)
    out = tmp_path / "out" / "result.json"
    done = subprocess.run([sys.executable, "-m", "hytgraph_amd.run_experiment", "--config", str(cfg),
                           "--output", str(out)], capture_output=True, text=True, cwd=ROOT, timeout=600)
    return done, (json.loads(out.read_text()) if out.exists() else None)


def test_command_line():
    assert EXE.exists(), "build the package first (python -m hytgraph_amd.build)"
    r = _run("--help")
    assert r.returncode == 0 and "Usage:" in r.stdout
    assert _run().returncode == 2                                   # --output is required
    assert _run("--output", "x.json", "--config", "novalue").returncode == 2
    assert _run("--output", "x.json", "--bogus").returncode == 2


def test_failure_is_recorded(tmp_path):
    """An experiment that fails (here before any device work: the edge list
    is missing) still writes its record, with status failure and the error."""
    out = tmp_path / "r.json"
    r = _run("--output", out, "--experiment-id", "missing-file", "--algorithm", "sssp",
             "--config", f"graph={tmp_path / 'none.el'}", "--config", "vertices=4")
    assert r.returncode == 1
    rec = json.loads(out.read_text())
    assert rec["status"] == "failure" and rec["experiment_id"] == "missing-file"
    assert "unable to open file" in rec["config"]["error"]
    assert rec["schema_version"] == "0.1" and rec["config"]["algorithm"] == "sssp"
    r = _run("--output", out, "--algorithm", "nosuch")
    assert r.returncode == 1 and "unknown algorithm" in json.loads(out.read_text())["config"]["error"]


def test_runner_adds_metadata(tmp_path):
    done, rec = _runner(tmp_path, f"experiment_id: meta\nalgorithm: sssp\ngraph: {tmp_path / 'none.el'}\n"
                                  "vertices: 4\nweighted: false\n")
    assert done.returncode == 1 and rec is not None
    assert rec["status"] == "failure"
    assert rec["git_commit"] and rec["runner"]["name"] == "hytgraph_amd/run_experiment.py"
    assert rec["experiment_config"]["vertices"] == 4 and rec["config"]["weighted"] == "false"


@pytest.mark.gpu
def test_sssp_and_pagerank_runs(gpu, abi, orc, tmp_path):
    hg = abi.HostGraph(scale=12, edge_factor=16, seed=3, weight_min=1, weight_max=64)
    dj = orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, 5)
    ref, it, _ = orc.sssp_jacobi_f32(hg.offsets, hg.columns, hg.weights.astype(np.float32), 5)
    fin = dj != 0xFFFFFFFF
    done, rec = _runner(tmp_path, "experiment_id: sssp12\nalgorithm: sssp\ngraph: rmat\nscale: 12\n"
                                  "seed: 3\nweight_min: 1\nweight_max: 64\nsource: 5\nmode: sync\n")
    assert done.returncode == 0, done.stderr
    m = rec["measurements"]
    assert rec["status"] == "success" and m["vertices"] == hg.V and m["edges"] == hg.E
    assert m["reached_vertices"] == fin.sum() and m["distance_sum"] == dj[fin].astype(np.float64).sum()
    assert m["rounds"] == it and m["converged"] == 1 and m["gteps"] > 0
    # the same graph from a binary CSR file, async rounds: same distances
    path = tmp_path / "g.hytcsr"
    hg.save(path)
    done, rec = _runner(tmp_path, f"algorithm: sssp\ngraph: {path}\nsource: 5\nmode: async\n")
    assert done.returncode == 0, done.stderr
    assert rec["measurements"]["distance_sum"] == dj[fin].astype(np.float64).sum()
    # batched: 4 degree-weighted sources in one call
    done, rec = _runner(tmp_path, "algorithm: sssp_batched\nscale: 12\nseed: 3\nsources: 4\n")
    srcs = abi.sample_sources(hg.offsets, 4, 4)
    want = sum(int((orc.sssp_dijkstra(hg.offsets, hg.columns, hg.weights, int(s)) != 0xFFFFFFFF).sum())
               for s in srcs)
    assert done.returncode == 0 and rec["measurements"]["reached_vertices"] == want
    done, rec = _runner(tmp_path, "algorithm: pagerank\nscale: 12\nseed: 3\nepsilon: 1e-7\n")
    assert done.returncode == 0, done.stderr
    m = rec["measurements"]
    assert abs(m["rank_sum"] - 1.0) < 1e-5 and m["converged"] == 1 and m["residual"] < 1e-5
