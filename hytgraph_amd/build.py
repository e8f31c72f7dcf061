"""Builds libhytgraph_amd.so (HIP kernels for gfx950 + C-ABI + host code).

Every source is compiled by hipcc; device code only for gfx950. The shared
library lands next to this file (hytgraph_amd/lib/) so it travels with the
source tree. Run: python -m hytgraph_amd.build  [--jobs N] [--debug]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"
LIBNAME = "libhytgraph_amd.so"
CPPLIB = "libhytgraph_cpp.so"
CPPSRC = PKG / "cpp"
TESTS_CPP = ROOT / "tests" / "cpp"
APPS = PKG / "apps"  # executables (hyt_experiment) built beside the libraries

ARCH = "gfx950"
COMMON = [
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wextra",
    "-Wno-unused-parameter",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    f"-I{CSRC / 'host'}",
    f"-I{CSRC / 'device'}",
]
# Packed-FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build hytgraph_amd")


def sources() -> list[Path]:
    out = sorted((CSRC / "host").glob("*.cpp")) + sorted((CSRC / "device").glob("*.hip"))
    return out


def compile_one(src: Path, opt: list[str]) -> Path:
    rel = src.relative_to(CSRC)
    obj = OBJDIR / (str(rel).replace(os.sep, "__") + ".o")
    obj.parent.mkdir(parents=True, exist_ok=True)
    cmd = [hipcc(), *COMMON, *opt, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip", *DEVICE]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 8, debug: bool = False) -> Path:
    opt = ["-O0", "-g"] if debug else ["-O3"]
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, opt), srcs))
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-o", str(tmp), "-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    build_cpp(debug)
    return out


def build_cpp(debug: bool = False) -> Path:
    """The C++ API (include/graph, include/algorithms, ...) as libhytgraph_cpp.so,
    host-only C++ built with g++ and linked against libhytgraph_amd.so, plus the
    C++ test executables under tests/cpp."""
    cxx = os.environ.get("CXX", "g++")
    opt = ["-O0", "-g"] if debug else ["-O2"]
    srcs = sorted(CPPSRC.rglob("*.cpp"))
    out = LIBDIR / CPPLIB
    tmp = LIBDIR / (CPPLIB + ".tmp")
    cmd = [cxx, "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-pthread", *opt,
           f"-I{ROOT / 'include'}", f"-I{CSRC / 'host'}", *map(str, srcs), f"-L{LIBDIR}", "-lhytgraph_amd",
           "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C++ library build failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    for test_src in [*sorted(TESTS_CPP.glob("*.cpp")), *sorted(APPS.glob("*.cpp"))]:
        exe = LIBDIR / test_src.stem
        cmd = [cxx, "-std=c++17", "-Wall", "-Wextra", *opt, f"-I{ROOT / 'include'}", f"-I{CSRC / 'device'}",
               str(test_src),
               f"-L{LIBDIR}", "-lhytgraph_cpp", "-lhytgraph_amd", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"C++ executable build failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--debug", action="store_true")
    args = ap.parse_args(argv)
    out = build(args.jobs, args.debug)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
