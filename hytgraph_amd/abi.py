"""ctypes binding of the C ABI declared in include/hytgraph_amd.h.

Only plain pointers and sizes cross the boundary. Host arrays are numpy
arrays; device outputs may be any object exposing a device address
(e.g. ``torch.Tensor.data_ptr()``) together with ``out_on_device=1``.
The library is loaded from hytgraph_amd/lib; a missing library is an error
(there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libhytgraph_amd.so"

HYT_OK = 0
HYT_ERR_INVALID_ARGUMENT = 1
HYT_ERR_NO_DEVICE = 2
HYT_ERR_HIP = 3
HYT_ERR_OUT_OF_MEMORY = 4
HYT_ERR_OVERFLOW = 5
HYT_ERR_UNSUPPORTED = 6
HYT_ERR_INTERNAL = 7
HYT_ERR_IO = 8

HYT_GEN_RMAT = 0
HYT_GEN_ER = 1
HYT_HOST_PAGEABLE = 0
HYT_HOST_PINNED = 1
HYT_GRAPH_WEIGHTS_U32 = 0
HYT_GRAPH_WEIGHTS_F32 = 1
HYT_GRAPH_OFF64 = 2
HYT_SSSP_SYNC = 0
HYT_SSSP_ASYNC = 1
HYT_DIR_AUTO = 0
HYT_DIR_PUSH = 1
HYT_DIR_PULL = 2

U32_INF = 0xFFFFFFFF
F32_INF_BITS = 0x7F800000


class GenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("scale", C.c_uint32),
        ("edge_factor", C.c_uint32),
        ("a", C.c_double),
        ("b", C.c_double),
        ("c", C.c_double),
        ("seed", C.c_uint64),
        ("weight_min", C.c_uint32),
        ("weight_max", C.c_uint32),
        ("threads", C.c_int32),
        ("memory", C.c_int32),
    ]


class RunOptions(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32),
        ("out_on_device", C.c_uint32),
        ("max_rounds", C.c_uint64),
        ("time_kernels", C.c_uint32),
        ("light_max_degree", C.c_uint32),
        ("heavy_chunk", C.c_uint32),
        ("direction", C.c_uint32),
        ("pull_percent", C.c_uint32),
        ("delta", C.c_uint32),
        ("lane_form", C.c_uint32),
        ("tolerance", C.c_float),
        ("host_rounds", C.c_uint32),
        ("task_order", C.c_uint32),
        ("heavy_delta", C.c_uint32),
        ("persistent", C.c_uint32),
        ("wait_degree", C.c_uint32),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("rounds", C.c_uint64),
        ("converged", C.c_int32),
        ("overflow", C.c_int32),
        ("frontier_total", C.c_uint64),
        ("edges_total", C.c_uint64),
        ("updates_total", C.c_uint64),
        ("expand_ms", C.c_double),
        ("expand_launches", C.c_uint64),
        ("wall_ms", C.c_double),
        ("residual", C.c_double),
        ("pull_rounds", C.c_uint64),
        ("pull_active_edges", C.c_uint64),
        ("pull_edges", C.c_uint64),
        ("pull_frontier", C.c_uint64),
        ("heavy_edges", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class CostParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("request_bytes", C.c_uint64), ("requests_per_tlp", C.c_uint64),
        ("rtt", C.c_double), ("cpu_bytes_per_second", C.c_double),
        ("entry_bytes", C.c_uint64), ("index_bytes", C.c_uint64), ("alignment_bytes", C.c_uint64),
    ]


class Partition(C.Structure):
    _fields_ = [("vertex_begin", C.c_uint32), ("vertex_end", C.c_uint32),
                ("edge_begin", C.c_uint64), ("edge_end", C.c_uint64)]


class PartitionMetrics(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("active_vertices", "active_edges", "total_edges", "filter_bytes",
                                         "compaction_bytes", "zc_requests", "zc_alignment", "zc_total_requests")]


class PartitionCosts(C.Structure):
    _fields_ = [("filter", C.c_double), ("compaction", C.c_double), ("zero_copy", C.c_double),
                ("zc_rtt", C.c_double), ("engine", C.c_int32), ("pad", C.c_int32)]


class HybridConfig(C.Structure):
    _fields_ = [
        ("device_budget_bytes", C.c_uint64), ("partition_edges", C.c_uint64), ("pack_edges", C.c_uint64),
        ("filter_combine_k", C.c_uint32), ("force_engine", C.c_int32), ("host_threads", C.c_uint32),
        ("calibrate", C.c_uint32), ("cache", C.c_uint32), ("flags", C.c_uint32), ("cost", CostParams),
    ]


class HybridStats(C.Structure):
    _fields_ = [
        ("run", RunStats), ("partitions", C.c_uint64), ("parts_cached", C.c_uint64),
        ("parts_filter", C.c_uint64), ("parts_compaction", C.c_uint64), ("parts_zero_copy", C.c_uint64),
        ("tasks", C.c_uint64), ("bytes_filter", C.c_uint64), ("bytes_compaction", C.c_uint64),
        ("bytes_zero_copy", C.c_uint64), ("copy_ms", C.c_double), ("pack_ms", C.c_double), ("plan_ms", C.c_double),
        ("link_gbs_peak", C.c_double), ("cpu_pack_gbs", C.c_double), ("zero_copy_gbs", C.c_double),
        ("device_edge_bytes", C.c_uint64), ("device_state_bytes", C.c_uint64), ("runs_reordered", C.c_uint64),
        ("transpose_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "run"}
        d["run"] = self.run.as_dict()
        return d


HYT_ENGINE_FILTER = 0
HYT_ENGINE_COMPACTION = 1
HYT_ENGINE_ZERO_COPY = 2

HYT_QUERY_SSSP = 0
HYT_QUERY_PPR = 1


class Request(C.Structure):
    _fields_ = [("graph", C.c_uint32), ("kind", C.c_uint32), ("source", C.c_uint32), ("reserved", C.c_uint32),
                ("arrival_s", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("done_s", C.c_double), ("session", C.c_uint32), ("lane", C.c_uint32),
                ("rounds", C.c_uint32), ("reserved", C.c_uint32), ("reached", C.c_uint64),
                ("checksum", C.c_uint64)]


class ServeConfig(C.Structure):
    _fields_ = [("device_budget_bytes", C.c_uint64), ("max_lanes", C.c_uint32), ("lane_form", C.c_uint32),
                ("damping", C.c_float), ("epsilon", C.c_float), ("reserved", C.c_uint32 * 4)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("busy_s", C.c_double), ("steps", C.c_uint64), ("rounds", C.c_uint64),
                ("edges", C.c_uint64), ("admitted", C.c_uint64), ("graph_bytes", C.c_uint64),
                ("state_bytes", C.c_uint64), ("frontier_rows", C.c_uint64), ("lane_rows", C.c_uint64),
                ("item_rounds", C.c_uint64),
                ("sessions", C.c_uint32), ("lanes", C.c_uint32 * 16),
                ("reserved", C.c_uint32), ("host_bytes", C.c_uint64), ("host_zero_copy_bytes", C.c_uint64),
                ("parts_cached", C.c_uint64), ("parts_moved", C.c_uint64)]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("lanes", "reserved")}
        d["lanes"] = list(self.lanes)[: self.sessions]
        return d


class ServeGraph(C.Structure):
    _fields_ = [("resident", C.c_void_p), ("sssp_engine", C.c_void_p), ("ppr_engine", C.c_void_p)]


class TuneOptions(C.Structure):
    _fields_ = [("tune_split", C.c_uint32), ("tune_lanes", C.c_uint32), ("max_lanes", C.c_uint32),
                ("reps", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("light_max_degree", C.c_uint32), ("heavy_chunk", C.c_uint32), ("lanes", C.c_uint32),
                ("candidates", C.c_uint32), ("sssp_ms", C.c_double), ("default_sssp_ms", C.c_double),
                ("lanes_qps", C.c_double), ("partition_edges", C.c_uint64), ("hybrid_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class HytError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class InvalidArgument(HytError, ValueError):
    pass


# Every exported symbol and its signature: (restype, argtypes).
_P = C.c_void_p
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
SIGNATURES = {
    "hyt_last_error": (C.c_char_p, []),
    "hyt_version": (C.c_char_p, []),
    "hyt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "hyt_generate": (C.c_int, [C.POINTER(GenParams), C.POINTER(_P)]),
    "hyt_generate_device": (C.c_int, [C.POINTER(GenParams), C.c_uint64, C.POINTER(_P)]),
    "hyt_host_graph_view": (
        C.c_int,
        [_P, _u64p, _u64p, C.POINTER(_u64p), C.POINTER(_u32p), C.POINTER(_u32p)],
    ),
    "hyt_host_graph_transpose": (C.c_int, [_P, C.c_uint32, C.POINTER(_u64p), C.POINTER(_u32p)]),
    "hyt_host_graph_free": (C.c_int, [_P]),
    "hyt_sample_sources": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint32, _P]),
    "hyt_graph_upload": (
        C.c_int,
        [C.c_uint64, C.c_uint64, _P, _P, _P, C.c_uint32, C.POINTER(_P)],
    ),
    "hyt_graph_free": (C.c_int, [_P]),
    "hyt_graph_trim": (C.c_int, [_P]),
    "hyt_graph_info": (C.c_int, [_P, _u64p, _u64p, _u64p]),
    "hyt_sssp": (
        C.c_int,
        [_P, _P, C.c_uint32, _P, C.POINTER(RunOptions), C.POINTER(RunStats)],
    ),
    "hyt_sssp_noalloc": (
        C.c_int,
        [_P, C.c_uint32, _P, C.POINTER(RunOptions), C.POINTER(RunStats)],
    ),
    "hyt_stream_probe": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "hyt_graph_download": (C.c_int, [_P, _P, _P, _P]),
    "hyt_hub_sort": (C.c_int, [_P, C.c_double, _P, _P, _u64p, C.c_uint32]),
    "hyt_graph_reorder": (C.c_int, [_P, _P, C.c_uint32]),
    "hyt_pagerank_contributions": (C.c_int, [_P, _P, C.c_float, _P, C.c_uint32]),
    "hyt_sssp_contributions": (C.c_int, [_P, _P, C.c_uint64, C.c_float, _P, C.c_uint32]),
    "hyt_stepper_create": (
        C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.POINTER(RunOptions), C.POINTER(_P)]),
    "hyt_stepper_step": (C.c_int, [_P, C.c_uint64, C.POINTER(RunStats), C.POINTER(C.c_int32)]),
    "hyt_stepper_read": (C.c_int, [_P, _P, C.POINTER(RunStats)]),
    "hyt_stepper_free": (C.c_int, [_P]),
    "hyt_pagerank": (
        C.c_int,
        [_P, C.c_float, C.c_float, _P, C.POINTER(RunOptions), C.POINTER(RunStats)],
    ),
    "hyt_ppr": (
        C.c_int,
        [_P, _P, C.c_uint32, C.c_float, C.c_float, _P, C.POINTER(RunOptions), C.POINTER(RunStats)],
    ),
    "hyt_cost_params_default": (C.c_int, [C.POINTER(CostParams)]),
    "hyt_partition_graph": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P, C.c_uint64, _u64p]),
    "hyt_partition_metrics_host": (
        C.c_int, [_P, C.c_uint64, _P, _P, C.c_uint64, C.POINTER(CostParams), _P]),
    "hyt_cost_evaluate": (C.c_int, [C.POINTER(CostParams), _P, C.c_uint64, _P]),
    "hyt_select_engine": (
        C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int32)]),
    "hyt_combine_tasks": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P, _P, _P, _u64p]),
    "hyt_load_edge_list": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_int32, C.POINTER(_P)]),
    "hyt_host_graph_save": (C.c_int, [_P, C.c_char_p, C.c_uint32]),
    "hyt_host_graph_load": (C.c_int, [C.c_char_p, C.c_int32, C.c_uint32, C.POINTER(_P), _u32p]),
    "hyt_host_graph_from_csr": (C.c_int, [C.c_uint64, C.c_uint64, _P, _P, _P, C.c_int32, C.POINTER(_P)]),
    "hyt_hybrid_config_default": (C.c_int, [C.POINTER(HybridConfig)]),
    "hyt_hybrid_create": (C.c_int, [_P, C.c_uint32, C.POINTER(HybridConfig), C.POINTER(_P)]),
    "hyt_hybrid_free": (C.c_int, [_P]),
    "hyt_hybrid_info": (C.c_int, [_P, _u64p, C.POINTER(CostParams), C.POINTER(HybridStats)]),
    "hyt_hybrid_sssp": (C.c_int, [_P, C.c_uint32, _P, C.POINTER(RunOptions), C.POINTER(HybridStats)]),
    "hyt_hybrid_pagerank": (
        C.c_int, [_P, C.c_float, C.c_float, _P, C.POINTER(RunOptions), C.POINTER(HybridStats)]),
    "hyt_hybrid_sssp_k": (C.c_int, [_P, _P, C.c_uint32, _P, C.POINTER(RunOptions), C.POINTER(HybridStats)]),
    "hyt_hybrid_ppr_k": (
        C.c_int, [_P, _P, C.c_uint32, C.c_float, C.c_float, _P, C.POINTER(RunOptions), C.POINTER(HybridStats)]),
    "hyt_lanes_state_bytes": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _u64p]),
    "hyt_lanes_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.POINTER(_P)]),
    "hyt_lanes_free": (C.c_int, [_P]),
    "hyt_lanes_admit": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "hyt_lanes_step": (C.c_int, [C.POINTER(_P), C.c_uint32, C.c_uint32, _u64p, _u64p]),
    "hyt_lanes_read": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _u64p, _u64p, _u32p]),
    "hyt_lanes_release": (C.c_int, [_P, C.c_uint32]),
    "hyt_lanes_set_form": (C.c_int, [_P, C.c_uint32]),
    "hyt_lanes_info": (C.c_int, [_P, _u32p, _u64p, _u64p, _u32p, _u64p]),
    "hyt_serve_config_default": (C.c_int, [C.POINTER(ServeConfig)]),
    "hyt_tune": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(TuneOptions), C.POINTER(TuneResult)]),
    "hyt_tune_hybrid": (C.c_int, [_P, C.c_uint32, C.POINTER(HybridConfig), _P, C.c_uint32, _P, C.c_uint32,
                                  C.POINTER(TuneResult)]),
    "hyt_serve_trace": (
        C.c_int, [C.POINTER(_P), C.c_uint32, C.POINTER(ServeConfig), C.POINTER(Request), C.c_uint64,
                  C.POINTER(_P), C.POINTER(Completion), C.POINTER(ServeStats)]),
    "hyt_serve_trace_mixed": (
        C.c_int, [C.POINTER(ServeGraph), C.c_uint32, C.POINTER(ServeConfig), C.POINTER(Request), C.c_uint64,
                  C.POINTER(_P), C.POINTER(Completion), C.POINTER(ServeStats)]),
    "hyt_lanes_create_hybrid": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.POINTER(_P)]),
}

_lib = None


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Loads the built library (once). Raises FileNotFoundError when absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -m hytgraph_amd.build` (no fallback exists)"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status == HYT_OK:
        return
    msg = load().hyt_last_error().decode(errors="replace")
    if status == HYT_ERR_INVALID_ARGUMENT:
        raise InvalidArgument(status, msg)
    raise HytError(status, msg)


def device_count() -> int:
    n = C.c_int(0)
    check(load().hyt_device_count(C.byref(n)))
    return n.value


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


class HostGraph:
    """Generator output owned by the library (views stay valid while alive)."""

    def __init__(self, kind=HYT_GEN_RMAT, scale=16, edge_factor=16, a=0.57, b=0.19, c=0.19,
                 seed=1, weight_min=1, weight_max=64, threads=0, memory=HYT_HOST_PAGEABLE, _handle=None,
                 on_device=False, device_budget=0):
        """on_device: the same graph built on the current HIP device
        (hyt_generate_device, key buckets within device_budget bytes of HBM)."""
        self.weights_f32 = False
        if _handle is None:
            p = GenParams(kind, scale, edge_factor, a, b, c, seed, weight_min, weight_max, threads, memory)
            _handle = C.c_void_p()
            if on_device:
                check(load().hyt_generate_device(C.byref(p), C.c_uint64(device_budget), C.byref(_handle)))
            else:
                check(load().hyt_generate(C.byref(p), C.byref(_handle)))
        self._attach(_handle)

    @classmethod
    def load(cls, path, memory=HYT_HOST_PAGEABLE, verify=True):
        """Binary CSR file written by save() (hyt_host_graph_load)."""
        h, f32 = C.c_void_p(), C.c_uint32()
        check(load().hyt_host_graph_load(str(path).encode(), memory, 1 if verify else 0, C.byref(h), C.byref(f32)))
        g = cls(_handle=h)
        g.weights_f32 = bool(f32.value)
        return g

    @classmethod
    def from_edge_list(cls, path, num_vertices, weighted=False, memory=HYT_HOST_PAGEABLE):
        """Text edge list with the original loader's rules (hyt_load_edge_list);
        weights are float bit patterns."""
        h = C.c_void_p()
        check(load().hyt_load_edge_list(str(path).encode(), num_vertices, 1 if weighted else 0, memory, C.byref(h)))
        g = cls(_handle=h)
        g.weights_f32 = bool(weighted)
        return g

    @classmethod
    def from_csr(cls, offsets, columns, weights=None, memory=HYT_HOST_PAGEABLE, weights_f32=False):
        """Copy of caller CSR arrays (hyt_host_graph_from_csr)."""
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        col = np.ascontiguousarray(columns, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        h = C.c_void_p()
        check(load().hyt_host_graph_from_csr(len(off) - 1, len(col), _ptr(off), _ptr(col),
                                             _ptr(w) if w is not None else None, memory, C.byref(h)))
        g = cls(_handle=h)
        g.weights_f32 = bool(weights_f32 and w is not None)
        return g

    def save(self, path, weights_f32=None):
        f32 = self.weights_f32 if weights_f32 is None else weights_f32
        check(load().hyt_host_graph_save(self._h, str(path).encode(), 1 if f32 else 0))

    def _attach(self, h):
        self._h = h
        V, E = C.c_uint64(), C.c_uint64()
        off, col, w = _u64p(), _u32p(), _u32p()
        check(load().hyt_host_graph_view(h, C.byref(V), C.byref(E), C.byref(off), C.byref(col), C.byref(w)))
        self.V, self.E = V.value, E.value
        self.offsets = np.ctypeslib.as_array(off, shape=(self.V + 1,))
        self.columns = np.ctypeslib.as_array(col, shape=(max(self.E, 1),))[: self.E]
        self.weights = (np.ctypeslib.as_array(w, shape=(max(self.E, 1),))[: self.E]
                        if bool(w) else None)

    def transpose(self, threads=0):
        """The incoming CSR (hyt_host_graph_transpose): (in_offsets u64[V+1],
        in_sources u32[E]) views, valid until free()."""
        off, col = _u64p(), _u32p()
        check(load().hyt_host_graph_transpose(self._h, threads, C.byref(off), C.byref(col)))
        return (np.ctypeslib.as_array(off, shape=(self.V + 1,)),
                np.ctypeslib.as_array(col, shape=(max(self.E, 1),))[: self.E])

    def free(self):
        if self._h:
            check(load().hyt_host_graph_free(self._h))
            self._h = None
            self.offsets = self.columns = self.weights = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sample_sources(offsets: np.ndarray, seed: int, count: int) -> np.ndarray:
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = np.zeros(count, dtype=np.uint32)
    check(load().hyt_sample_sources(_ptr(offsets), len(offsets) - 1, seed, count, _ptr(out)))
    return out


HYT_LANES_AUTO = 0
HYT_LANES_GROUP = 1
HYT_LANES_ITEMS = 2


def _options(mode=HYT_SSSP_SYNC, out_on_device=0, max_rounds=0, time_kernels=0,
             light_max_degree=0, heavy_chunk=0, direction=HYT_DIR_AUTO, pull_percent=0, delta=0,
             lane_form=HYT_LANES_AUTO, tolerance=0.0, host_rounds=0, task_order=0,
             heavy_delta=0, persistent=0, wait_degree=0) -> RunOptions:
    return RunOptions(mode, out_on_device, max_rounds, time_kernels, light_max_degree, heavy_chunk,
                      direction, pull_percent, delta, lane_form, tolerance, host_rounds, task_order,
                      heavy_delta, persistent, wait_degree)


class DeviceGraph:
    """A CSR graph resident in HBM (hyt_graph_t)."""

    def __init__(self, offsets, columns, weights=None, weights_f32=False, off64=False):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        columns = np.ascontiguousarray(columns, dtype=np.uint32)
        if weights is not None:
            if weights_f32:
                weights = np.ascontiguousarray(weights, dtype=np.float32).view(np.uint32)
            else:
                weights = np.ascontiguousarray(weights, dtype=np.uint32)
        self.V = len(offsets) - 1
        self.E = int(offsets[-1]) if self.V >= 0 and len(offsets) else 0
        self.weights_f32 = weights_f32
        self.weighted = weights is not None
        h = C.c_void_p()
        flags = HYT_GRAPH_WEIGHTS_F32 if weights_f32 else HYT_GRAPH_WEIGHTS_U32
        if off64:
            flags |= HYT_GRAPH_OFF64
        check(load().hyt_graph_upload(self.V, len(columns), _ptr(offsets), _ptr(columns),
                                      _ptr(weights), flags, C.byref(h)))
        self._h = h

    @classmethod
    def from_host(cls, hg: HostGraph, weights_f32=False, off64=False):
        w = hg.weights
        if w is not None and weights_f32:
            w = w.astype(np.float32)
        return cls(hg.offsets, hg.columns, w, weights_f32=weights_f32, off64=off64)

    def info(self) -> dict:
        V, E, B = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(load().hyt_graph_info(self._h, C.byref(V), C.byref(E), C.byref(B)))
        return {"V": V.value, "E": E.value, "device_bytes": B.value}

    def trim(self):
        """Frees the traversal workspaces and the transpose (rebuilt on demand)."""
        check(load().hyt_graph_trim(self._h))

    def free(self):
        if getattr(self, "_h", None):
            check(load().hyt_graph_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def download(self):
        """(offsets u64[V+1], columns u32[E], weights u32[E] or None) of the device CSR."""
        off = np.empty(self.V + 1, np.uint64)
        col = np.empty(self.E, np.uint32)
        w = np.empty(self.E, np.uint32) if self.weighted else None
        check(load().hyt_graph_download(self._h, _ptr(off), _ptr(col), _ptr(w)))
        return off, col, w

    # -- vertex order -------------------------------------------------------
    def hub_sort(self, hub_fraction=0.08, out_ptr=None, scores_ptr=None):
        """hyt_hub_sort: (vertex_order u32[V], scores f64[V], hub_count); with
        out_ptr (device u32[V], scores_ptr device f64[V] or None) the results
        stay on the device and (None, None, hub_count) is returned."""
        hubs = C.c_uint64()
        if out_ptr is not None:
            check(load().hyt_hub_sort(self._h, hub_fraction, C.c_void_p(out_ptr),
                                      C.c_void_p(scores_ptr) if scores_ptr else None, C.byref(hubs), 1))
            return None, None, hubs.value
        order = np.empty(self.V, np.uint32)
        scores = np.empty(self.V, np.float64)
        check(load().hyt_hub_sort(self._h, hub_fraction, _ptr(order), _ptr(scores), C.byref(hubs), 0))
        return order, scores, hubs.value

    def reorder(self, vertex_order=None, order_ptr=None):
        """hyt_graph_reorder (in place): vertex_order[new] = old, host array or
        device address."""
        if order_ptr is not None:
            check(load().hyt_graph_reorder(self._h, C.c_void_p(order_ptr), 1))
            return
        order = np.ascontiguousarray(vertex_order, dtype=np.uint32)
        if len(order) != self.V:
            raise InvalidArgument(HYT_ERR_INVALID_ARGUMENT,
                                  "CSRGraph::reorder_vertices: vertex_order must contain exactly one entry per vertex")
        check(load().hyt_graph_reorder(self._h, _ptr(order), 0))

    # -- contributions ------------------------------------------------------
    def pagerank_contributions(self, current_ranks, damping=0.85):
        """hyt_pagerank_contributions: f64[V] from f32[V] ranks (host arrays)."""
        x = np.ascontiguousarray(current_ranks, dtype=np.float32)
        if len(x) != self.V:
            raise InvalidArgument(HYT_ERR_INVALID_ARGUMENT,
                                  "PageRank current_ranks size must match graph vertex count")
        out = np.empty(self.V, np.float64)
        check(load().hyt_pagerank_contributions(self._h, _ptr(x), damping, _ptr(out), 0))
        return out

    # -- SSSP ---------------------------------------------------------------
    def sssp(self, sources, mode=HYT_SSSP_SYNC, out=None, out_ptr=None, **opt):
        """Returns (dist u32[V] or [V][K], stats dict). With out_ptr (device
        address) the result stays on the device and None is returned."""
        src = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
        K = len(src)
        o = _options(mode=mode, out_on_device=1 if out_ptr is not None else 0, **opt)
        st = RunStats()
        if out_ptr is not None:
            check(load().hyt_sssp(self._h, _ptr(src), K, C.c_void_p(out_ptr), C.byref(o), C.byref(st)))
            return None, st.as_dict()
        if out is None:
            out = np.empty((self.V, K) if K > 1 else self.V, dtype=np.uint32)
        check(load().hyt_sssp(self._h, _ptr(src), K, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    # -- PageRank -------------------------------------------------------------
    def sssp_noalloc(self, source, out_ptr, mode=HYT_SSSP_SYNC, **opt):
        """hyt_sssp_noalloc: one source into device memory (out_ptr, u32[V]),
        with the library's static scratch instead of the graph's workspace."""
        o = _options(mode=mode, out_on_device=1, **opt)
        st = RunStats()
        check(load().hyt_sssp_noalloc(self._h, int(source), C.c_void_p(out_ptr), C.byref(o), C.byref(st)))
        return st.as_dict()

    def pagerank(self, damping=0.85, epsilon=1e-7, out_ptr=None, **opt):
        o = _options(out_on_device=1 if out_ptr is not None else 0, **opt)
        st = RunStats()
        if out_ptr is not None:
            check(load().hyt_pagerank(self._h, damping, epsilon, C.c_void_p(out_ptr), C.byref(o), C.byref(st)))
            return None, st.as_dict()
        out = np.empty(self.V, dtype=np.float32)
        check(load().hyt_pagerank(self._h, damping, epsilon, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    def ppr(self, seeds, damping=0.85, epsilon=1e-6, out_ptr=None, **opt):
        sd = np.ascontiguousarray(np.atleast_1d(seeds), dtype=np.uint32)
        K = len(sd)
        o = _options(out_on_device=1 if out_ptr is not None else 0, **opt)
        st = RunStats()
        if out_ptr is not None:
            check(load().hyt_ppr(self._h, _ptr(sd), K, damping, epsilon, C.c_void_p(out_ptr),
                                 C.byref(o), C.byref(st)))
            return None, st.as_dict()
        out = np.empty((self.V, K) if K > 1 else self.V, dtype=np.float32)
        check(load().hyt_ppr(self._h, _ptr(sd), K, damping, epsilon, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()


# ------------------------------------------------------------ transfer planning
def cost_params(**overrides) -> CostParams:
    p = CostParams()
    check(load().hyt_cost_params_default(C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def partition_graph(offsets, target_bytes) -> np.ndarray:
    """Returns an (n, 4) u64 array of [vertex_begin, vertex_end, edge_begin, edge_end]."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    V = len(off) - 1
    cap = max(V, 1)
    out = (Partition * cap)()
    n = C.c_uint64()
    check(load().hyt_partition_graph(_ptr(off), V, target_bytes, out, cap, C.byref(n)))
    return np.array([[p.vertex_begin, p.vertex_end, p.edge_begin, p.edge_end] for p in out[: n.value]],
                    dtype=np.uint64).reshape(-1, 4)


def partition_metrics(offsets, active, parts: np.ndarray, params: CostParams) -> np.ndarray:
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    act = np.ascontiguousarray(active, dtype=np.uint8)
    n = len(parts)
    pa = (Partition * max(n, 1))(*[Partition(int(a), int(b), int(c), int(d)) for a, b, c, d in parts])
    out = (PartitionMetrics * max(n, 1))()
    check(load().hyt_partition_metrics_host(_ptr(off), len(off) - 1, _ptr(act), pa, n, C.byref(params), out))
    return np.array([[getattr(m, f) for f, _ in PartitionMetrics._fields_] for m in out[:n]],
                    dtype=np.uint64).reshape(-1, 8)


def cost_evaluate(metrics: np.ndarray, params: CostParams):
    n = len(metrics)
    ma = (PartitionMetrics * max(n, 1))(*[PartitionMetrics(*map(int, row)) for row in metrics])
    out = (PartitionCosts * max(n, 1))()
    check(load().hyt_cost_evaluate(C.byref(params), ma, n, out))
    costs = np.array([[c.filter, c.compaction, c.zero_copy, c.zc_rtt] for c in out[:n]]).reshape(-1, 4)
    engines = np.array([c.engine for c in out[:n]], dtype=np.int32)
    return costs, engines


def select_engine(f, c, z, alpha, beta) -> int:
    e = C.c_int32()
    check(load().hyt_select_engine(f, c, z, alpha, beta, C.byref(e)))
    return e.value


def combine_tasks(engines, k=4):
    """Returns [(engine, [members...]), ...] in task order."""
    e = np.ascontiguousarray(engines, dtype=np.int32)
    n = len(e)
    te = np.zeros(max(n, 1), np.int32)
    to = np.zeros(n + 1, np.uint64)
    mem = np.zeros(max(n, 1), np.uint64)
    nt = C.c_uint64()
    check(load().hyt_combine_tasks(_ptr(e), n, k, _ptr(te), _ptr(to), _ptr(mem), C.byref(nt)))
    return [(int(te[t]), mem[int(to[t]):int(to[t + 1])].tolist()) for t in range(nt.value)]


# ------------------------------------------------------------ hybrid engine
def hybrid_config(**overrides) -> HybridConfig:
    c = HybridConfig()
    check(load().hyt_hybrid_config_default(C.byref(c)))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class HybridEngine:
    """Out-of-HBM traversal over a pinned HostGraph (kept alive by this object)."""

    def __init__(self, host: HostGraph, weights_f32=False, **config):
        self.host = host
        self.V = host.V
        cfg = hybrid_config(**config)
        h = C.c_void_p()
        check(load().hyt_hybrid_create(host._h, 1 if weights_f32 else 0, C.byref(cfg), C.byref(h)))
        self._h = h

    def info(self) -> dict:
        n, p, st = C.c_uint64(), CostParams(), HybridStats()
        check(load().hyt_hybrid_info(self._h, C.byref(n), C.byref(p), C.byref(st)))
        d = st.as_dict()
        d["params"] = {f: getattr(p, f) for f, _ in CostParams._fields_}
        return d

    def sssp(self, source, out_ptr=None, **opt):
        o = _options(mode=HYT_SSSP_ASYNC, out_on_device=1 if out_ptr is not None else 0, **opt)
        st = HybridStats()
        if out_ptr is not None:
            check(load().hyt_hybrid_sssp(self._h, int(source), C.c_void_p(out_ptr), C.byref(o), C.byref(st)))
            return None, st.as_dict()
        out = np.empty(self.V, np.uint32)
        check(load().hyt_hybrid_sssp(self._h, int(source), _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    def pagerank(self, damping=0.85, epsilon=1e-7, out_ptr=None, **opt):
        o = _options(out_on_device=1 if out_ptr is not None else 0, **opt)
        st = HybridStats()
        if out_ptr is not None:
            check(load().hyt_hybrid_pagerank(self._h, damping, epsilon, C.c_void_p(out_ptr), C.byref(o),
                                             C.byref(st)))
            return None, st.as_dict()
        out = np.empty(self.V, np.float32)
        check(load().hyt_hybrid_pagerank(self._h, damping, epsilon, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    def sssp_k(self, sources, out_ptr=None, **opt):
        """K sources in one run (hyt_hybrid_sssp_k): u32[V][K] distances."""
        src = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
        K = len(src)
        o = _options(out_on_device=1 if out_ptr is not None else 0, **opt)
        st = HybridStats()
        if out_ptr is not None:
            check(load().hyt_hybrid_sssp_k(self._h, _ptr(src), K, C.c_void_p(out_ptr), C.byref(o), C.byref(st)))
            return None, st.as_dict()
        out = np.empty((self.V, K), np.uint32)
        check(load().hyt_hybrid_sssp_k(self._h, _ptr(src), K, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    def ppr_k(self, seeds, damping=0.85, epsilon=1e-6, out_ptr=None, **opt):
        """K personalized-PageRank seeds in one run (hyt_hybrid_ppr_k): f32[V][K]."""
        sd = np.ascontiguousarray(np.atleast_1d(seeds), dtype=np.uint32)
        K = len(sd)
        o = _options(out_on_device=1 if out_ptr is not None else 0, **opt)
        st = HybridStats()
        if out_ptr is not None:
            check(load().hyt_hybrid_ppr_k(self._h, _ptr(sd), K, damping, epsilon, C.c_void_p(out_ptr), C.byref(o),
                                          C.byref(st)))
            return None, st.as_dict()
        out = np.empty((self.V, K), np.float32)
        check(load().hyt_hybrid_ppr_k(self._h, _ptr(sd), K, damping, epsilon, _ptr(out), C.byref(o), C.byref(st)))
        return out, st.as_dict()

    def free(self):
        if getattr(self, "_h", None):
            check(load().hyt_hybrid_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def stream_probe(nbytes: int, reps: int = 5) -> float:
    """GB/s of a read-only streaming kernel over nbytes of device memory."""
    g = C.c_double()
    check(load().hyt_stream_probe(nbytes, reps, C.byref(g)))
    return g.value


def sssp_contributions(current, candidate, tolerance=0.0):
    """hyt_sssp_contributions on the current device: f64[n] from two f32[n]
    distance arrays (host arrays)."""
    cur = np.ascontiguousarray(current, dtype=np.float32)
    cand = np.ascontiguousarray(candidate, dtype=np.float32)
    if len(cur) != len(cand):
        raise InvalidArgument(HYT_ERR_INVALID_ARGUMENT, "SSSP contribution distance vectors must have equal size")
    out = np.empty(len(cur), np.float64)
    check(load().hyt_sssp_contributions(_ptr(cur), _ptr(cand), len(cur), tolerance, _ptr(out), 0))
    return out


# ------------------------------------------------------------ round-by-round runs
HYT_STEP_SSSP = 0
HYT_STEP_PAGERANK = 1


class Stepper:
    """A single-source SSSP or PageRank/PPR run advanced round by round
    (hyt_stepper_*); keeps its DeviceGraph alive."""

    def __init__(self, graph, kind=HYT_STEP_SSSP, vertex=0, damping=0.85, epsilon=1e-7, **opt):
        self.graph = graph
        self.kind = kind
        self.V = graph.V
        o = _options(**opt)
        h = C.c_void_p()
        check(load().hyt_stepper_create(graph._h, kind, int(vertex) & 0xFFFFFFFF, damping, epsilon, C.byref(o),
                                        C.byref(h)))
        self._h = h

    def step(self, rounds=1):
        """Runs up to `rounds` rounds; returns (this call's stats, finished)."""
        st = RunStats()
        fin = C.c_int32()
        check(load().hyt_stepper_step(self._h, rounds, C.byref(st), C.byref(fin)))
        return st.as_dict(), bool(fin.value)

    def read(self):
        """(current distances u32[V] or normalised ranks f32[V], the run's stats so far)."""
        out = np.empty(self.V, np.uint32 if self.kind == HYT_STEP_SSSP else np.float32)
        st = RunStats()
        check(load().hyt_stepper_read(self._h, _ptr(out), C.byref(st)))
        return out, st.as_dict()

    def free(self):
        if getattr(self, "_h", None):
            check(load().hyt_stepper_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------ continuous batching
class LaneSession:
    """K query lanes of one kind over a DeviceGraph (kept alive by this object)."""

    def __init__(self, graph, kind=HYT_QUERY_SSSP, lanes=8, damping=0.85, epsilon=1e-6):
        """graph: a DeviceGraph, or a HybridEngine (hyt_lanes_create_hybrid)."""
        self.graph = graph
        self.kind = kind
        self.V = graph.V
        h = C.c_void_p()
        if isinstance(graph, HybridEngine):
            check(load().hyt_lanes_create_hybrid(graph._h, kind, lanes, damping, epsilon, C.byref(h)))
        else:
            check(load().hyt_lanes_create(graph._h, kind, lanes, damping, epsilon, C.byref(h)))
        self._h = h

    def admit(self, lane: int, source: int) -> None:
        check(load().hyt_lanes_admit(self._h, lane, int(source)))

    @staticmethod
    def step(sessions, rounds=1):
        """Runs up to `rounds` rounds on every session; returns (finished, running) lane masks."""
        n = len(sessions)
        arr = (C.c_void_p * max(n, 1))(*[s._h for s in sessions])
        fin = (C.c_uint64 * max(n, 1))()
        run = (C.c_uint64 * max(n, 1))()
        check(load().hyt_lanes_step(arr, n, rounds, fin, run))
        return list(fin)[:n], list(run)[:n]

    def read(self, lane: int, out_ptr=None):
        """Returns (result u32[V] / f32[V] or None with out_ptr, reached, checksum, rounds)."""
        reached, checksum, rounds = C.c_uint64(), C.c_uint64(), C.c_uint32()
        if out_ptr is not None:
            check(load().hyt_lanes_read(self._h, lane, C.c_void_p(out_ptr), 1, C.byref(reached),
                                        C.byref(checksum), C.byref(rounds)))
            return None, reached.value, checksum.value, rounds.value
        out = np.empty(self.V, dtype=np.float32 if self.kind == HYT_QUERY_PPR else np.uint32)
        check(load().hyt_lanes_read(self._h, lane, _ptr(out), 0, C.byref(reached), C.byref(checksum),
                                    C.byref(rounds)))
        return out, reached.value, checksum.value, rounds.value

    def release(self, lane: int) -> None:
        check(load().hyt_lanes_release(self._h, lane))

    def set_form(self, form: int) -> None:
        check(load().hyt_lanes_set_form(self._h, form))

    def info(self) -> dict:
        k, occ, run, F, b = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint64()
        check(load().hyt_lanes_info(self._h, C.byref(k), C.byref(occ), C.byref(run), C.byref(F), C.byref(b)))
        return {"lanes": k.value, "occupied": occ.value, "running": run.value, "frontier": F.value,
                "state_bytes": b.value}

    def free(self):
        if getattr(self, "_h", None):
            check(load().hyt_lanes_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def lanes_state_bytes(V: int, E: int, max_degree: int, kind: int, lanes: int) -> int:
    b = C.c_uint64()
    check(load().hyt_lanes_state_bytes(V, E, max_degree, kind, lanes, C.byref(b)))
    return b.value


def serve_config(**overrides) -> ServeConfig:
    c = ServeConfig()
    check(load().hyt_serve_config_default(C.byref(c)))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def serve_trace(graphs, requests, keep_results=False, **config):
    """Replays requests = [(arrival_s, graph_index, kind, source), ...] (sorted by
    arrival). graphs: DeviceGraph, or (sssp_engine, ppr_engine) HybridEngine
    pairs (either may be None) for graphs served from host memory
    (hyt_serve_trace_mixed). Returns (completions as a structured numpy
    array, stats dict, results list or None)."""
    n = len(requests)
    req = (Request * max(n, 1))()
    for i, (t, gi, kind, src) in enumerate(requests):
        req[i] = Request(int(gi), int(kind), int(src), 0, float(t))
    mixed = any(not isinstance(g, DeviceGraph) for g in graphs)
    sizes = [g.V if isinstance(g, DeviceGraph) else next(e for e in g if e is not None).V for g in graphs]
    if mixed:
        garr = (ServeGraph * len(graphs))()
        for i, g in enumerate(graphs):
            if isinstance(g, DeviceGraph):
                garr[i] = ServeGraph(g._h, None, None)
            else:
                garr[i] = ServeGraph(None, g[0]._h if g[0] is not None else None,
                                     g[1]._h if g[1] is not None else None)
    else:
        garr = (C.c_void_p * len(graphs))(*[g._h for g in graphs])
    comp = (Completion * max(n, 1))()
    st = ServeStats()
    cfg = serve_config(**config)
    results = None
    res_arr = None
    if keep_results:
        results = [np.empty(sizes[r[1]], dtype=np.float32 if r[2] == HYT_QUERY_PPR else np.uint32)
                   for r in requests]
        res_arr = (C.c_void_p * max(n, 1))(*[a.ctypes.data for a in results])
    fn = load().hyt_serve_trace_mixed if mixed else load().hyt_serve_trace
    check(fn(garr, len(graphs), C.byref(cfg), req, n, res_arr, comp, C.byref(st)))
    dt = np.dtype([("admit_s", "f8"), ("done_s", "f8"), ("session", "u4"), ("lane", "u4"), ("rounds", "u4"),
                   ("reserved", "u4"), ("reached", "u8"), ("checksum", "u8")])
    arr = np.frombuffer(bytes(comp), dtype=dt, count=n).copy() if n else np.zeros(0, dt)
    return arr, st.as_dict(), results


# ------------------------------------------------------------------ autotuner
def tune(graph: DeviceGraph, sources, tune_split=True, tune_lanes=True, max_lanes=64, reps=2) -> dict:
    src = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
    o = TuneOptions(1 if tune_split else 0, 1 if tune_lanes else 0, max_lanes, reps)
    r = TuneResult()
    check(load().hyt_tune(graph._h, _ptr(src), len(src), C.byref(o), C.byref(r)))
    return r.as_dict()


def tune_hybrid(host: HostGraph, candidates, sources, weights_f32=False, **config) -> dict:
    cand = np.ascontiguousarray(np.atleast_1d(candidates), dtype=np.uint64)
    src = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
    cfg = hybrid_config(**config)
    r = TuneResult()
    check(load().hyt_tune_hybrid(host._h, 1 if weights_f32 else 0, C.byref(cfg), _ptr(cand), len(cand), _ptr(src),
                                 len(src), C.byref(r)))
    return r.as_dict()
