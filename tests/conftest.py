import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_built():
    """Tests load what is built; they build only when nothing is."""
    from hytgraph_amd import abi

    if not abi.LIB_PATH.exists():
        from hytgraph_amd import build as b

        b.build()
    import oracle

    if not oracle.ORC_LIB.exists():
        oracle.build(with_ref=True)


_ensure_built()


@pytest.fixture(scope="session")
def abi():
    from hytgraph_amd import abi as a

    a.load()
    return a


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(abi):
    """The HIP path is the product: on a GPU test a missing device is a failure."""
    n = abi.device_count()
    assert n >= 1, "no HIP device visible to a gpu-marked test"
    return n


def csr_from_edges(V, edges, weights=None):
    """CSR (u64 offsets) from (src, dst) pairs, keeping the given edge order per row."""
    edges = list(edges)
    order = sorted(range(len(edges)), key=lambda i: edges[i][0])
    off = np.zeros(V + 1, dtype=np.uint64)
    for s, _ in edges:
        off[s + 1] += 1
    off = np.cumsum(off, dtype=np.uint64)
    col = np.array([edges[i][1] for i in order], dtype=np.uint32)
    w = None if weights is None else np.array([weights[i] for i in order])
    return off, col, w


class HipRuntime:
    """Device buffers for tests that pass device pointers to the C ABI, from
    the HIP runtime libhytgraph_amd.so itself links (torch bundles its own
    runtime; the two do not share one process well)."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.lib = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.lib.hipDeviceSynchronize.argtypes = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def malloc(self, nbytes):
        p = self.C.c_void_p()
        self._ok(self.lib.hipMalloc(self.C.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, ptr):
        self._ok(self.lib.hipFree(self.C.c_void_p(ptr)), "hipFree")

    def to_host_u32(self, ptr, n):
        out = np.empty(n, np.uint32)
        self._ok(self.lib.hipMemcpy(out.ctypes.data, self.C.c_void_p(ptr), n * 4, 2), "hipMemcpy")
        return out

    def to_host(self, ptr, n, dtype):
        out = np.empty(n, dtype)
        self._ok(self.lib.hipMemcpy(out.ctypes.data, self.C.c_void_p(ptr), out.nbytes, 2), "hipMemcpy")
        return out

    def from_host(self, arr):
        """A device copy of a host array (free it with free())."""
        arr = np.ascontiguousarray(arr)
        ptr = self.malloc(max(arr.nbytes, 4))
        self._ok(self.lib.hipMemcpy(self.C.c_void_p(ptr), arr.ctypes.data, arr.nbytes, 1), "hipMemcpy")
        return ptr

    def free_bytes(self):
        f, t = self.C.c_size_t(), self.C.c_size_t()
        self._ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.lib.hipMemGetInfo(self.C.byref(f), self.C.byref(t)), "hipMemGetInfo")
        return f.value


@pytest.fixture(scope="session")
def hip(gpu):
    return HipRuntime()


def f32(x):
    """The value a float parameter of the C ABI carries: the FP64 oracles
    are given the damping and epsilon the GPU computes with."""
    return float(np.float32(x))


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
