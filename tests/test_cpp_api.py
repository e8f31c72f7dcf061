"""Runs the C++ API test executable (tests/cpp/cpp_api_test.cpp, built by
hytgraph_amd.build): host-side API on the CPU, drop-in GPU functions on a GPU."""
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent.parent / "hytgraph_amd" / "lib" / "cpp_api_test"


def _run(arg):
    assert EXE.exists(), f"{EXE} missing: run python -m hytgraph_amd.build"
    return subprocess.run([str(EXE), arg], capture_output=True, text=True, timeout=300)


def test_cpp_host_api():
    r = _run("cpu")
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_gpu_api_skips_without_device(abi):
    if abi.device_count() > 0:
        pytest.skip("device present: covered by the gpu-marked test")
    assert _run("gpu").returncode == 77


@pytest.mark.gpu
def test_cpp_gpu_api(gpu):
    r = _run("gpu")
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_planning_api():
    exe = EXE.parent / "planning_api_test"
    assert exe.exists(), f"{exe} missing: run python -m hytgraph_amd.build"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
