"""C ABI surface on the CPU: the library loads, exports every declared symbol,
and validates host-side arguments before touching a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(hyt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported(abi):
    lib = abi.load()
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows each one's signature
    assert set(names) <= set(abi.SIGNATURES), set(names) - set(abi.SIGNATURES)


def test_version_and_device_count(abi):
    assert b"gfx950" in abi.load().hyt_version()
    assert abi.device_count() >= 0


def test_csr_validation_precedes_device(abi):
    # malformed CSR is rejected as invalid_argument with or without a device
    bad = [
        (np.array([1, 1], np.uint64), np.array([0], np.uint32)),     # offsets[0] != 0
        (np.array([0, 2, 1], np.uint64), np.array([0], np.uint32)),  # non-monotone
        (np.array([0, 1, 1], np.uint64), np.array([5], np.uint32)),  # column out of range
    ]
    for off, col in bad:
        with pytest.raises(abi.InvalidArgument):
            abi.DeviceGraph(off, col)
    with pytest.raises(abi.InvalidArgument):
        abi.DeviceGraph(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([np.nan], np.float32),
                        weights_f32=True)


def test_no_device_is_reported(abi):
    if abi.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(abi.HytError) as ei:
        abi.DeviceGraph(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32))
    assert ei.value.status == abi.HYT_ERR_NO_DEVICE


def test_generator_is_deterministic_and_clean(abi):
    a = abi.HostGraph(scale=12, edge_factor=16, seed=5, weight_min=1, weight_max=64, threads=1)
    b = abi.HostGraph(scale=12, edge_factor=16, seed=5, weight_min=1, weight_max=64, threads=5)
    assert a.E == b.E and np.array_equal(a.columns, b.columns) and np.array_equal(a.weights, b.weights)
    off = a.offsets.astype(np.int64)
    src = np.repeat(np.arange(a.V), np.diff(off))
    key = src.astype(np.uint64) << np.uint64(32) | a.columns.astype(np.uint64)
    assert np.all(np.diff(key.astype(np.int64)) > 0)  # sorted rows, no duplicates
    assert np.all(src != a.columns)                    # no self-loops
    assert a.weights.min() >= 1 and a.weights.max() <= 64
    c = abi.HostGraph(scale=12, edge_factor=16, seed=6, weight_min=1, weight_max=64)
    assert not np.array_equal(c.columns[:1000], a.columns[:1000])


def test_generator_rmat_skew_and_er(abi):
    r = abi.HostGraph(scale=14, edge_factor=16, seed=1, weight_min=0, weight_max=0)
    e = abi.HostGraph(kind=abi.HYT_GEN_ER, scale=14, edge_factor=16, seed=1, weight_min=0, weight_max=0)
    assert r.weights is None and e.weights is None
    dr, de = np.diff(r.offsets.astype(np.int64)), np.diff(e.offsets.astype(np.int64))
    # RMAT (a=.57): heavy-tailed, vertex 0 is the largest hub; ER is near-uniform
    assert dr.max() > 20 * de.max() / 2 and dr.argmax() == 0
    assert de.max() < 60 and e.E > 0.99 * 16 * (1 << 14)
    # RMAT quadrant mass: the fraction of edges with src < V/2 is ~ a + b = .76
    frac = r.offsets[1 << 13] / r.E
    assert 0.70 < frac < 0.82


def test_generator_rejects_bad_params(abi):
    with pytest.raises(abi.InvalidArgument):
        abi.HostGraph(scale=0)
    with pytest.raises(abi.InvalidArgument):
        abi.HostGraph(scale=8, a=0.6, b=0.3, c=0.3)
    with pytest.raises(abi.InvalidArgument):
        abi.HostGraph(scale=8, weight_min=5, weight_max=2)


def test_source_sampling_is_degree_weighted(abi):
    g = abi.HostGraph(scale=12, edge_factor=16, seed=3, weight_min=0, weight_max=0)
    s = abi.sample_sources(g.offsets, 9, 4000)
    deg = np.diff(g.offsets.astype(np.int64))
    assert np.all(deg[s] > 0)
    assert np.array_equal(s, abi.sample_sources(g.offsets, 9, 4000))
    # vertex 0 (largest hub) is drawn about deg(0)/E of the time
    p0 = deg[0] / g.E
    assert abs((s == 0).mean() - p0) < 4 * np.sqrt(p0 / 4000) + 1e-3


def test_hybrid_config_defaults_and_host_checks(abi):
    c = abi.hybrid_config()
    assert c.device_budget_bytes == 4 << 30 and c.partition_edges == 1 << 20 and c.force_engine == -1
    assert c.cost.alpha == 0.8 and c.cost.entry_bytes == 8
    # a pageable host graph runs without zero-copy: forcing it is refused
    # before any device work
    pageable = abi.HostGraph(scale=8, edge_factor=4, seed=1)
    with pytest.raises(abi.InvalidArgument):
        abi.HybridEngine(pageable, force_engine=abi.HYT_ENGINE_ZERO_COPY)


def test_struct_layouts_match_header(abi, tmp_path):
    # every ctypes mirror has the size and field offsets of its C struct
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {
        "hyt_gen_params": abi.GenParams, "hyt_run_options": abi.RunOptions, "hyt_run_stats": abi.RunStats,
        "hyt_cost_params": abi.CostParams, "hyt_partition": abi.Partition,
        "hyt_partition_metrics": abi.PartitionMetrics, "hyt_partition_costs": abi.PartitionCosts,
        "hyt_hybrid_config": abi.HybridConfig, "hyt_hybrid_stats": abi.HybridStats,
        "hyt_request": abi.Request, "hyt_completion": abi.Completion, "hyt_serve_config": abi.ServeConfig,
        "hyt_serve_stats": abi.ServeStats, "hyt_tune_options": abi.TuneOptions,
        "hyt_tune_result": abi.TuneResult,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hytgraph_amd.h"', "int main(void){"]
    expect = []
    for cname, py in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(py))
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
            expect.append(getattr(py, f).offset)
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect
