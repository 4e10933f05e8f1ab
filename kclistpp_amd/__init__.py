"""MI355X-native Sim-kClist++ (k-clique densest subgraph) - Python host side.

Thin ctypes layer over ``libkclpp.so`` (the C ABI in ``include/kclpp.h``).
The names mirror the reference programs (btsun/kclistpp):

* ``read_edge_list``  - ReadEdgeList (Graph.c:34-53)
* ``Graph``           - SortByCore + Relabel + MakeGraph (Graph.c:57-156)
* ``Graph.clique_enum`` - CDF_CliqueEnum (kCDensestCmpSync.c:298-325)
* ``Sim``             - AllocCdf + the CDF_Main loop body + CDF_FindDensestSubset
* ``cmpsync``         - CDF_Main (kCDensestCmpSync.c:385-419)

The library is required: importing works without it, but every call raises
``KclError`` if ``libkclpp.so`` is missing (no CPU fallback exists).
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# KCLPP_LIBRARY selects another build of the library (same ABI).
LIB_PATH = os.environ.get("KCLPP_LIBRARY") or os.path.join(PKG, "libkclpp.so")

FRANK_WOLFE = 2
PAVA_PREPROCESS = 3
SCHED_COUNTED = 0
SCHED_LEARNED = 1
SCHED_DEGREE = 2
SCHEDULES = {"counted": SCHED_COUNTED, "learned": SCHED_LEARNED, "degree": SCHED_DEGREE}
CM_FEATURES = 8
CM_OUTPUTS = 2
CM_PARAMS = 4866
CM_NORM = 20
NUM_CLASSES = 9  # word classes of the enumeration kernels (KCL_NUM_CLASSES)

_STATUS = {
    0: "KCL_OK", 1: "KCL_E_IO", 2: "KCL_E_PARSE", 3: "KCL_E_INVALID_ARG",
    4: "KCL_E_OOM_HOST", 5: "KCL_E_OOM_DEVICE", 6: "KCL_E_HIP", 7: "KCL_E_OVERFLOW",
    8: "KCL_E_UNSUPPORTED",
}


class KclError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status
        self.name = _STATUS.get(status, str(status))


class _GraphInfo(C.Structure):
    _fields_ = [("n_in", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint64),
                ("core", C.c_uint32)]


class _SimOpts(C.Structure):
    _fields_ = [("schedule", C.c_uint32), ("split_work", C.c_uint32), ("model", C.c_void_p),
                ("reserved", C.c_uint32 * 4)]


class _TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("fit_norm", C.c_uint32), ("seed", C.c_uint64), ("reserved", C.c_uint32 * 4)]


class _Densest(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("density", C.c_double),
                ("ub", C.c_double), ("sum", C.c_uint64)]


class _PlanInfo(C.Structure):
    _fields_ = [("roots_total", C.c_uint32), ("roots_kept", C.c_uint32), ("tasks", C.c_uint32),
                ("reserved", C.c_uint32), ("cliques", C.c_uint64), ("plan_ms", C.c_double)]


class _TrimInfo(C.Structure):
    _fields_ = [("n_in", C.c_uint32), ("n2", C.c_uint32), ("e2", C.c_uint32),
                ("core", C.c_uint32), ("ncktot", C.c_uint64), ("ckcore", C.c_uint64),
                ("ckrho", C.c_double), ("rho", C.c_double), ("size", C.c_uint32),
                ("ntrim", C.c_uint32), ("ncktrim", C.c_uint64), ("mtrim", C.c_uint32),
                ("gpu_pops", C.c_uint32), ("seconds_degrees", C.c_double),
                ("seconds_peel", C.c_double)]


class _BatchOpts(C.Structure):
    _fields_ = [("schedule", C.c_uint32), ("split_work", C.c_uint32), ("model", C.c_void_p),
                ("tune", C.c_uint32), ("row_bits", C.c_uint32), ("per_cu", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]


class _BatchInfo(C.Structure):
    _fields_ = [("graphs", C.c_uint32), ("distinct_k", C.c_uint32), ("nodes", C.c_uint64),
                ("edges", C.c_uint64), ("cliques", C.c_uint64), ("tasks", C.c_uint32),
                ("row_bits", C.c_uint32), ("per_cu", C.c_uint32), ("reserved", C.c_uint32),
                ("plan_ms", C.c_double), ("tune_ms", C.c_double)]


class _TrimOpts(C.Structure):
    _fields_ = [("gpu_peel_min_cliques", C.c_uint64), ("reserved", C.c_uint32 * 4)]


class _ExactOpts(C.Structure):
    _fields_ = [("max_iter", C.c_uint32), ("max_seconds", C.c_uint32),
                ("max_network_cliques", C.c_uint64), ("schedule", C.c_uint32),
                ("gpu_list_min", C.c_uint32), ("model", C.c_void_p), ("reserved", C.c_uint32 * 2)]


class _ExactResult(C.Structure):
    _fields_ = [("certified", C.c_uint32), ("num_iter", C.c_uint32), ("n", C.c_uint32),
                ("m", C.c_uint32), ("cliques", C.c_uint64), ("density", C.c_double),
                ("max_flow_calls", C.c_uint32), ("region", C.c_uint32), ("seconds", C.c_double),
                ("upper_bound", C.c_double), ("stop", C.c_uint32),
                ("degree_passes", C.c_uint32), ("host_seconds", C.c_double)]


class _SamplingOpts(C.Structure):
    _fields_ = [("target", C.c_uint64), ("seed", C.c_uint64), ("reserved", C.c_uint32 * 4)]


class _SamplingInfo(C.Structure):
    _fields_ = [("total", C.c_uint64), ("sampled", C.c_uint64), ("prob", C.c_double),
                ("rounds", C.c_uint32)]


class _SamplingRecord(C.Structure):
    _fields_ = [("T", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint32), ("prefix", C.c_uint32),
                ("prefix_sampled", C.c_uint64), ("cliques", C.c_uint64),
                ("edge_density", C.c_double), ("density", C.c_double), ("ub_sample", C.c_double),
                ("ub", C.c_double), ("seconds", C.c_double)]


class _Record(C.Structure):
    _fields_ = [("T", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint32),
                ("density", C.c_double), ("ub", C.c_double), ("sum", C.c_uint64),
                ("seconds", C.c_double)]


class _DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("cu_count", C.c_uint32),
                ("hbm_bytes", C.c_uint64), ("lds_bytes_per_cu", C.c_uint32),
                ("clock_khz", C.c_uint32)]


class _ClassInfo(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("words", "k", "first", "count", "waves", "per_cu",
                                          "blocks", "lds_bytes")]


_lib = None

EXPORTS = [
    "kcl_last_error", "kcl_status_name", "kcl_read_edge_list", "kcl_write_edge_list", "kcl_free",
    "kcl_graph_create", "kcl_graph_destroy", "kcl_graph_get_info", "kcl_graph_rank",
    "kcl_graph_set_edge_split",
    "kcl_graph_set_dense_threshold",
    "kcl_graph_csr", "kcl_graph_edges", "kcl_count_edges", "kcl_clique_enum",
    "kcl_clique_enum_noalloc", "kcl_sim_create",
    "kcl_sim_destroy", "kcl_sim_step", "kcl_sim_counts", "kcl_sim_rho", "kcl_sim_set_rho",
    "kcl_sim_totals", "kcl_sim_last_pass", "kcl_sim_plan_info", "kcl_sim_root_cliques", "kcl_sim_profile_pass",
    "kcl_sim_plan_classes", "kcl_sim_find_densest", "kcl_cmpsync_run",
    "kcl_device_get_info", "kcl_version", "kcl_trim_create", "kcl_trim_create_opts",
    "kcl_trim_destroy",
    "kcl_trim_get_info", "kcl_trim_nodes", "kcl_trim_edges", "kcl_trim_write", "kcl_exact_run",
    "kcl_cost_model_create", "kcl_cost_model_destroy", "kcl_cost_model_load",
    "kcl_cost_model_save", "kcl_cost_model_get", "kcl_cost_model_set", "kcl_cost_model_predict",
    "kcl_cost_model_gradient", "kcl_cost_model_train", "kcl_sim_root_features",
    "kcl_sim_cost_samples", "kcl_sim_replan", "kcl_batch_create", "kcl_batch_destroy",
    "kcl_batch_get_info", "kcl_batch_graph_cliques", "kcl_batch_step", "kcl_batch_reset",
    "kcl_batch_last_pass",
    "kcl_batch_counts", "kcl_batch_rho", "kcl_batch_totals", "kcl_batch_find_densest",
    "kcl_batch_run", "kcl_sampling_create", "kcl_sampling_destroy", "kcl_sampling_get_info",
    "kcl_sampling_cliques", "kcl_sampling_step", "kcl_sampling_loads", "kcl_sampling_report",
    "kcl_sampling_run",
]


def lib() -> C.CDLL:
    """Load libkclpp.so (raises KclError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KclError(6, f"{LIB_PATH} is not built; run `python -m kclistpp_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int64
    sigs = {
        "kcl_last_error": ([], C.c_char_p),
        "kcl_status_name": ([C.c_int], C.c_char_p),
        "kcl_version": ([], C.c_char_p),
        "kcl_read_edge_list": ([C.c_char_p, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(u64),
                                C.POINTER(u32)], C.c_int),
        "kcl_write_edge_list": ([C.c_char_p, vp, u64], C.c_int),
        "kcl_free": ([vp], None),
        "kcl_graph_create": ([vp, u64, u32, C.POINTER(vp)], C.c_int),
        "kcl_graph_destroy": ([vp], None),
        "kcl_graph_get_info": ([vp, C.POINTER(_GraphInfo)], C.c_int),
        "kcl_graph_rank": ([vp, vp], C.c_int),
        "kcl_graph_set_edge_split": ([vp, u32], C.c_int),
        "kcl_graph_set_dense_threshold": ([vp, C.c_float], C.c_int),
        "kcl_graph_csr": ([vp, vp, vp], C.c_int),
        "kcl_graph_edges": ([vp, vp], C.c_int),
        "kcl_count_edges": ([vp, u32, vp, C.POINTER(u32)], C.c_int),
        "kcl_clique_enum": ([vp, u32, C.c_int, vp, vp, vp, C.POINTER(u64)], C.c_int),
        "kcl_clique_enum_noalloc": ([vp, u32, C.c_int, vp, vp, vp, C.POINTER(u64)], C.c_int),
        "kcl_sim_create": ([vp, u32, C.POINTER(_SimOpts), C.POINTER(vp)], C.c_int),
        "kcl_sim_destroy": ([vp], None),
        "kcl_sim_step": ([vp, u32], C.c_int),
        "kcl_sim_counts": ([vp, vp], C.c_int),
        "kcl_sim_rho": ([vp, vp], C.c_int),
        "kcl_sim_set_rho": ([vp, vp], C.c_int),
        "kcl_sim_totals": ([vp, vp], C.c_int),
        "kcl_sim_last_pass": ([vp, C.POINTER(u64), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)], C.c_int),
        "kcl_sim_plan_info": ([vp, C.POINTER(_PlanInfo)], C.c_int),
        "kcl_sim_root_cliques": ([vp, vp], C.c_int),
        "kcl_sim_profile_pass": ([vp, C.POINTER(u32), vp, vp, vp, vp, vp], C.c_int),
        "kcl_sim_find_densest": ([vp, C.POINTER(_Densest), vp, vp, vp], C.c_int),
        "kcl_cmpsync_run": ([vp, u32, u32, C.POINTER(_SimOpts), C.c_char_p, i64,
                             C.POINTER(_Record), C.POINTER(u32)], C.c_int),
        "kcl_device_get_info": ([C.POINTER(_DeviceInfo)], C.c_int),
        "kcl_trim_create": ([vp, u64, u32, u32, C.POINTER(vp)], C.c_int),
        "kcl_trim_create_opts": ([vp, u64, u32, u32, C.POINTER(_TrimOpts), C.POINTER(vp)],
                                 C.c_int),
        "kcl_trim_destroy": ([vp], None),
        "kcl_trim_get_info": ([vp, C.POINTER(_TrimInfo)], C.c_int),
        "kcl_trim_nodes": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "kcl_trim_edges": ([vp, vp], C.c_int),
        "kcl_trim_write": ([vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                            i64], C.c_int),
        "kcl_exact_run": ([vp, u32, C.POINTER(_ExactOpts), C.c_char_p, i64,
                           C.POINTER(_ExactResult), vp], C.c_int),
        "kcl_cost_model_create": ([u64, C.POINTER(vp)], C.c_int),
        "kcl_cost_model_destroy": ([vp], None),
        "kcl_cost_model_load": ([C.c_char_p, C.POINTER(vp)], C.c_int),
        "kcl_cost_model_save": ([vp, C.c_char_p], C.c_int),
        "kcl_cost_model_get": ([vp, vp, vp], C.c_int),
        "kcl_cost_model_set": ([vp, vp, vp], C.c_int),
        "kcl_cost_model_predict": ([vp, vp, u64, vp], C.c_int),
        "kcl_cost_model_gradient": ([vp, vp, vp, u64, C.POINTER(C.c_float), vp], C.c_int),
        "kcl_cost_model_train": ([vp, vp, vp, u64, C.POINTER(_TrainOpts), vp], C.c_int),
        "kcl_sim_root_features": ([vp, vp], C.c_int),
        "kcl_sim_cost_samples": ([vp, C.POINTER(u32), vp, vp, vp], C.c_int),
        "kcl_sim_replan": ([vp, u32, vp], C.c_int),
        "kcl_batch_create": ([vp, vp, u32, C.POINTER(_BatchOpts), C.POINTER(vp)], C.c_int),
        "kcl_batch_destroy": ([vp], None),
        "kcl_batch_get_info": ([vp, C.POINTER(_BatchInfo)], C.c_int),
        "kcl_batch_graph_cliques": ([vp, vp], C.c_int),
        "kcl_batch_step": ([vp, u32], C.c_int),
        "kcl_batch_reset": ([vp], C.c_int),
        "kcl_batch_last_pass": ([vp, C.POINTER(u64), C.POINTER(C.c_double)], C.c_int),
        "kcl_batch_counts": ([vp, u32, vp], C.c_int),
        "kcl_batch_rho": ([vp, u32, vp], C.c_int),
        "kcl_batch_totals": ([vp, u32, vp], C.c_int),
        "kcl_batch_find_densest": ([vp, C.POINTER(_Densest)], C.c_int),
        "kcl_batch_run": ([vp, u32, C.POINTER(_Record), vp], C.c_int),
        "kcl_sim_plan_classes": ([vp, vp], C.c_int),
        "kcl_sampling_create": ([vp, u32, C.POINTER(_SamplingOpts), C.POINTER(vp)], C.c_int),
        "kcl_sampling_destroy": ([vp], None),
        "kcl_sampling_get_info": ([vp, C.POINTER(_SamplingInfo)], C.c_int),
        "kcl_sampling_cliques": ([vp, vp], C.c_int),
        "kcl_sampling_step": ([vp], C.c_int),
        "kcl_sampling_loads": ([vp, vp], C.c_int),
        "kcl_sampling_report": ([vp, C.POINTER(_SamplingRecord), vp], C.c_int),
        "kcl_sampling_run": ([vp, u32, C.POINTER(_SamplingRecord), C.POINTER(u32)], C.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().kcl_last_error().decode(errors="replace")
        raise KclError(rc, msg)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def version() -> str:
    return lib().kcl_version().decode()


def device_info() -> dict:
    d = _DeviceInfo()
    _check(lib().kcl_device_get_info(C.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(), "cu_count": d.cu_count,
            "hbm_bytes": d.hbm_bytes, "lds_bytes_per_cu": d.lds_bytes_per_cu,
            "clock_khz": d.clock_khz}


def read_edge_list(path: str) -> tuple[np.ndarray, int]:
    """ReadEdgeList: returns ((m, 2) uint32 pairs, n_in = max id + 1)."""
    L = lib()
    p = C.POINTER(C.c_uint32)()
    m = C.c_uint64(0)
    n = C.c_uint32(0)
    _check(L.kcl_read_edge_list(path.encode(), C.byref(p), C.byref(m), C.byref(n)))
    try:
        if m.value:
            arr = np.ctypeslib.as_array(p, shape=(m.value * 2,)).copy().reshape(-1, 2)
        else:
            arr = np.zeros((0, 2), np.uint32)
    finally:
        L.kcl_free(C.cast(p, C.c_void_p))
    return arr, int(n.value)


def write_edge_list(path: str, pairs: np.ndarray) -> None:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    _check(lib().kcl_write_edge_list(path.encode(), _ptr(pairs), pairs.shape[0]))


@dataclass
class Densest:
    n: int
    m: int
    density: float
    ub: float
    sum: int


class Graph:
    """Degeneracy-ordered DAG on the GPU (SortByCore + Relabel + MakeGraph)."""

    def __init__(self, pairs: np.ndarray, n_in: int = 0):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        h = C.c_void_p()
        _check(lib().kcl_graph_create(_ptr(pairs) if pairs.size else None, pairs.shape[0],
                                      n_in, C.byref(h)))
        self._h = h
        info = _GraphInfo()
        _check(lib().kcl_graph_get_info(h, C.byref(info)))
        self.n_in, self.n, self.m, self.core = info.n_in, info.n, int(info.m), info.core

    @classmethod
    def from_file(cls, path: str) -> "Graph":
        pairs, n_in = read_edge_list(path)
        return cls(pairs, n_in)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().kcl_graph_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        if not self._h:
            raise KclError(3, "graph is closed")
        return self._h

    def rank(self) -> np.ndarray:
        out = np.zeros(max(self.n_in, 1), np.uint32)
        _check(lib().kcl_graph_rank(self.handle, _ptr(out)))
        return out[:self.n_in]

    def csr(self) -> tuple[np.ndarray, np.ndarray]:
        cd = np.zeros(self.n + 1, np.uint32)
        adj = np.zeros(max(self.m, 1), np.uint32)
        _check(lib().kcl_graph_csr(self.handle, _ptr(cd), _ptr(adj)))
        return cd, adj[:self.m]

    def edges(self) -> np.ndarray:
        st = np.zeros(max(2 * self.m, 2), np.uint32)
        _check(lib().kcl_graph_edges(self.handle, _ptr(st)))
        return st[:2 * self.m].reshape(-1, 2)

    def set_edge_split(self, max_vertex_degree: int) -> None:
        """Roots with more DAG out-neighbours run as edge tasks (timing only)."""
        _check(lib().kcl_graph_set_edge_split(self.handle, max_vertex_degree))

    def set_dense_threshold(self, min_items: float) -> None:
        """k >= 6 roots whose expected common out-set per edge item is at
        least min_items are counted a wavefront per item (0: all; timing
        only)."""
        _check(lib().kcl_graph_set_dense_threshold(self.handle, float(min_items)))

    def count_edges(self, nodes: np.ndarray) -> int:
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        out = C.c_uint32(0)
        _check(lib().kcl_count_edges(self.handle, nodes.size, _ptr(nodes) if nodes.size else None,
                                     C.byref(out)))
        return int(out.value)

    def clique_enum(self, k: int, task: int, rho: np.ndarray | None = None,
                    level: np.ndarray | None = None, acc: np.ndarray | None = None):
        """CDF_CliqueEnum: returns (acc, visits); acc is accumulated into."""
        if acc is None:
            acc = np.zeros(max(self.n, 1), np.float64)
        rho_a = None if rho is None else np.ascontiguousarray(rho, np.float64)
        lev_a = None if level is None else np.ascontiguousarray(level, np.uint32)
        v = C.c_uint64(0)
        _check(lib().kcl_clique_enum(self.handle, k, task, _ptr(rho_a), _ptr(lev_a), _ptr(acc),
                                     C.byref(v)))
        return acc[:self.n], int(v.value)

    def clique_enum_noalloc(self, k: int, task: int, rho: np.ndarray | None = None,
                            level: np.ndarray | None = None, acc: np.ndarray | None = None):
        """kcl_clique_enum_noalloc: the same pass with no device allocation
        (intermediates in the library's static scratch); returns (acc,
        visits)."""
        if acc is None:
            acc = np.zeros(max(self.n, 1), np.float64)
        rho_a = None if rho is None else np.ascontiguousarray(rho, np.float64)
        lev_a = None if level is None else np.ascontiguousarray(level, np.uint32)
        v = C.c_uint64(0)
        _check(lib().kcl_clique_enum_noalloc(self.handle, k, task, _ptr(rho_a), _ptr(lev_a),
                                             _ptr(acc), C.byref(v)))
        return acc[:self.n], int(v.value)


def _opts(schedule: int = SCHED_COUNTED, split_work: int = 0, model: "CostModel | None" = None):
    o = _SimOpts()
    o.schedule = schedule
    o.split_work = split_work
    o.model = model.handle if model is not None else None
    return o


def _f32(a, cols: int) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an (rows, {cols}) array, got shape {a.shape}")
    return a


class CostModel:
    """Learned root-cost model: BF16 MLP 8 -> 64 -> 64 -> 2 with GPU training
    and inference kernels (kcl_cost_model_*)."""

    def __init__(self, seed: int = 0, path: str | None = None):
        h = C.c_void_p()
        if path is not None:
            _check(lib().kcl_cost_model_load(path.encode(), C.byref(h)))
        else:
            _check(lib().kcl_cost_model_create(seed, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        if not self._h:
            raise KclError(3, "cost model is closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().kcl_cost_model_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def save(self, path: str) -> None:
        _check(lib().kcl_cost_model_save(self.handle, path.encode()))

    def get(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.zeros(CM_PARAMS, np.float32)
        nrm = np.zeros(CM_NORM, np.float32)
        _check(lib().kcl_cost_model_get(self.handle, _ptr(p), _ptr(nrm)))
        return p, nrm

    def set(self, params: np.ndarray | None = None, norm: np.ndarray | None = None) -> None:
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        nrm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        if p is not None and p.size != CM_PARAMS:
            raise ValueError(f"params: expected {CM_PARAMS} floats")
        if nrm is not None and nrm.size != CM_NORM:
            raise ValueError(f"norm: expected {CM_NORM} floats")
        _check(lib().kcl_cost_model_set(self.handle, _ptr(p), _ptr(nrm)))

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = _f32(x, CM_FEATURES)
        y = np.zeros((max(len(x), 1), CM_OUTPUTS), np.float32)
        _check(lib().kcl_cost_model_predict(self.handle, _ptr(x), len(x), _ptr(y)))
        return y[:len(x)]

    def gradient(self, x: np.ndarray, y: np.ndarray) -> tuple[float, np.ndarray]:
        x = _f32(x, CM_FEATURES)
        y = _f32(y, CM_OUTPUTS)
        if len(x) != len(y):
            raise ValueError("x and y differ in rows")
        g = np.zeros(CM_PARAMS, np.float32)
        loss = C.c_float(0)
        _check(lib().kcl_cost_model_gradient(self.handle, _ptr(x), _ptr(y), len(x),
                                             C.byref(loss), _ptr(g)))
        return float(loss.value), g

    def train(self, x: np.ndarray, y: np.ndarray, epochs: int = 0, batch: int = 0,
              lr: float = 0.0, fit_norm: bool = True, seed: int = 0) -> np.ndarray:
        """Adam on the GPU; returns the mean loss of each epoch."""
        x = _f32(x, CM_FEATURES)
        y = _f32(y, CM_OUTPUTS)
        if len(x) != len(y):
            raise ValueError("x and y differ in rows")
        o = _TrainOpts()
        o.epochs, o.batch, o.lr, o.fit_norm, o.seed = epochs, batch, lr, int(fit_norm), seed
        losses = np.zeros(epochs or 60, np.float32)
        _check(lib().kcl_cost_model_train(self.handle, _ptr(x), _ptr(y), len(x), C.byref(o),
                                          _ptr(losses)))
        return losses


class Sim:
    """Device-resident Sim-kClist++ state for one graph and k."""

    def __init__(self, graph: Graph, k: int, schedule: int = SCHED_COUNTED, split_work: int = 0,
                 model: CostModel | None = None):
        self.graph = graph
        self.k = k
        h = C.c_void_p()
        o = _opts(schedule, split_work, model)
        _check(lib().kcl_sim_create(graph.handle, k, C.byref(o), C.byref(h)))
        self._h = h
        self.t = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().kcl_sim_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, t: int | None = None) -> None:
        """One iteration (FW pass + rho update); t defaults to the next one."""
        t = self.t + 1 if t is None else t
        _check(lib().kcl_sim_step(self._h, t))
        self.t = t

    def counts(self) -> np.ndarray:
        out = np.zeros(max(self.graph.n, 1), np.uint64)
        _check(lib().kcl_sim_counts(self._h, _ptr(out)))
        return out[:self.graph.n]

    def rho(self) -> np.ndarray:
        out = np.zeros(max(self.graph.n, 1), np.float64)
        _check(lib().kcl_sim_rho(self._h, _ptr(out)))
        return out[:self.graph.n]

    def set_rho(self, rho: np.ndarray) -> None:
        rho = np.ascontiguousarray(rho, np.float64)
        _check(lib().kcl_sim_set_rho(self._h, _ptr(rho)))

    def totals(self) -> np.ndarray:
        out = np.zeros(max(self.graph.n, 1), np.uint64)
        _check(lib().kcl_sim_totals(self._h, _ptr(out)))
        return out[:self.graph.n]

    def last_pass(self) -> tuple[int, float, float]:
        v = C.c_uint64(0)
        km = C.c_double(0)
        pm = C.c_double(0)
        _check(lib().kcl_sim_last_pass(self._h, C.byref(v), C.byref(km), C.byref(pm)))
        return int(v.value), float(km.value), float(pm.value)

    def plan_info(self) -> dict:
        p = _PlanInfo()
        _check(lib().kcl_sim_plan_info(self._h, C.byref(p)))
        return {"roots_total": p.roots_total, "roots_kept": p.roots_kept, "tasks": p.tasks,
                "cliques": int(p.cliques), "plan_ms": p.plan_ms}

    def root_cliques(self) -> np.ndarray:
        out = np.zeros(max(self.graph.n, 1), np.uint64)
        _check(lib().kcl_sim_root_cliques(self._h, _ptr(out)))
        return out[:self.graph.n]

    def profile_pass(self) -> dict:
        """One pass with per-task shader-clock timing (state unchanged)."""
        nt = C.c_uint32(0)
        _check(lib().kcl_sim_profile_pass(self._h, C.byref(nt), None, None, None, None, None))
        n = nt.value
        out = {"root": np.zeros(max(n, 1), np.uint32), "part": np.zeros(max(n, 1), np.uint32),
               "nparts": np.zeros(max(n, 1), np.uint32),
               "build_cycles": np.zeros(max(n, 1), np.uint64),
               "total_cycles": np.zeros(max(n, 1), np.uint64)}
        _check(lib().kcl_sim_profile_pass(self._h, C.byref(nt), _ptr(out["root"]),
                                          _ptr(out["part"]), _ptr(out["nparts"]),
                                          _ptr(out["build_cycles"]), _ptr(out["total_cycles"])))
        return {k: v[:n] for k, v in out.items()}

    def plan_classes(self) -> list[dict]:
        """Launch shape of the plan's word classes: vertex-task classes,
        then edge-task classes (only those with tasks)."""
        arr = (_ClassInfo * (2 * NUM_CLASSES))()
        _check(lib().kcl_sim_plan_classes(self._h, C.cast(arr, C.c_void_p)))
        out = []
        for i, c in enumerate(arr):
            if c.count:
                d = {f: int(getattr(c, f)) for f, _ in _ClassInfo._fields_}
                d["edge"] = i >= NUM_CLASSES
                out.append(d)
        return out

    def root_features(self) -> np.ndarray:
        """Cost-model features of every node as a root ((n, 8) float32)."""
        x = np.zeros((max(self.graph.n, 1), CM_FEATURES), np.float32)
        _check(lib().kcl_sim_root_features(self._h, _ptr(x)))
        return x[:self.graph.n]

    def cost_samples(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(roots, features, labels) from one timed pass of the current plan."""
        nr = C.c_uint32(0)
        _check(lib().kcl_sim_cost_samples(self._h, C.byref(nr), None, None, None))
        n = nr.value
        root = np.zeros(max(n, 1), np.uint32)
        x = np.zeros((max(n, 1), CM_FEATURES), np.float32)
        y = np.zeros((max(n, 1), CM_OUTPUTS), np.float32)
        _check(lib().kcl_sim_cost_samples(self._h, C.byref(nr), _ptr(root), _ptr(x), _ptr(y)))
        return root[:n], x[:n], y[:n]

    def replan(self, schedule: int, model: CostModel | None = None) -> None:
        _check(lib().kcl_sim_replan(self._h, schedule,
                                    model.handle if model is not None else None))

    def find_densest(self, want_vectors: bool = False):
        d = _Densest()
        n = max(self.graph.n, 1)
        reo = np.zeros(n, np.uint32) if want_vectors else None
        lev = np.zeros(n, np.uint32) if want_vectors else None
        tent = np.zeros(n, np.uint64) if want_vectors else None
        _check(lib().kcl_sim_find_densest(self._h, C.byref(d), _ptr(reo), _ptr(lev), _ptr(tent)))
        info = Densest(d.n, d.m, d.density, d.ub, int(d.sum))
        if want_vectors:
            m = self.graph.n
            return info, reo[:m], lev[:m], tent[:m]
        return info


def cmpsync(graph: Graph, k: int, num_iter: int, stat_path: str | None = None,
            schedule: int = SCHED_COUNTED, t0_unix: int = 0,
            model: CostModel | None = None) -> list[dict]:
    """CDF_Main: one record per power of two <= num_iter."""
    recs = (_Record * 32)()
    nrec = C.c_uint32(0)
    o = _opts(schedule, 0, model)
    _check(lib().kcl_cmpsync_run(graph.handle, k, num_iter, C.byref(o),
                                 stat_path.encode() if stat_path else None, t0_unix, recs,
                                 C.byref(nrec)))
    return [{"T": r.T, "n": r.n, "m": r.m, "density": r.density, "ub": r.ub, "sum": r.sum,
             "seconds": r.seconds} for r in recs[:nrec.value]]


class Batch:
    """Many graphs, each with its own k, iterated together (one captured
    hipGraph replay per iteration of every graph)."""

    def __init__(self, graphs: list[Graph], ks: list[int], schedule: int = SCHED_COUNTED,
                 model: CostModel | None = None, tune: bool = False, row_bits: int = 0,
                 per_cu: int = 0, split_work: int = 0):
        if len(graphs) != len(ks) or not graphs:
            raise ValueError("one k per graph, at least one graph")
        self.graphs = list(graphs)
        self.ks = [int(k) for k in ks]
        self.n = [g.n for g in graphs]
        o = _BatchOpts()
        o.schedule, o.split_work, o.tune, o.row_bits, o.per_cu = (
            schedule, split_work, int(tune), row_bits, per_cu)
        o.model = model.handle if model is not None else None
        hs = (C.c_void_p * len(graphs))(*[g.handle for g in graphs])
        karr = np.ascontiguousarray(self.ks, np.uint32)
        h = C.c_void_p()
        _check(lib().kcl_batch_create(C.cast(hs, C.c_void_p), _ptr(karr), len(graphs),
                                      C.byref(o), C.byref(h)))
        self._h = h
        self.t = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().kcl_batch_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        i = _BatchInfo()
        _check(lib().kcl_batch_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _BatchInfo._fields_ if f != "reserved"}

    def graph_cliques(self) -> np.ndarray:
        out = np.zeros(len(self.graphs), np.uint64)
        _check(lib().kcl_batch_graph_cliques(self._h, _ptr(out)))
        return out

    def step(self, t: int | None = None) -> None:
        t = self.t + 1 if t is None else t
        _check(lib().kcl_batch_step(self._h, t))
        self.t = t

    def reset(self) -> None:
        """rho, counts and totals back to zero (the state after create)."""
        _check(lib().kcl_batch_reset(self._h))
        self.t = 0

    def last_pass(self) -> tuple[int, float]:
        v = C.c_uint64(0)
        ms = C.c_double(0)
        _check(lib().kcl_batch_last_pass(self._h, C.byref(v), C.byref(ms)))
        return int(v.value), float(ms.value)

    def _seg(self, fn, i, dtype):
        out = np.zeros(max(self.n[i], 1), dtype)
        _check(fn(self._h, i, _ptr(out)))
        return out[:self.n[i]]

    def counts(self, i: int) -> np.ndarray:
        return self._seg(lib().kcl_batch_counts, i, np.uint64)

    def rho(self, i: int) -> np.ndarray:
        return self._seg(lib().kcl_batch_rho, i, np.float64)

    def totals(self, i: int) -> np.ndarray:
        return self._seg(lib().kcl_batch_totals, i, np.uint64)

    def find_densest(self) -> list[Densest]:
        d = (_Densest * len(self.graphs))()
        _check(lib().kcl_batch_find_densest(self._h, d))
        return [Densest(x.n, x.m, x.density, x.ub, int(x.sum)) for x in d]

    def run(self, num_iter: int) -> list[list[dict]]:
        """CDF_Main of every graph: per graph, one record per power of two."""
        recs = (_Record * (32 * len(self.graphs)))()
        nrec = np.zeros(len(self.graphs), np.uint32)
        _check(lib().kcl_batch_run(self._h, num_iter, recs, _ptr(nrec)))
        out = []
        for i in range(len(self.graphs)):
            out.append([{"T": r.T, "n": r.n, "m": r.m, "density": r.density, "ub": r.ub,
                         "sum": r.sum, "seconds": r.seconds}
                        for r in recs[32 * i:32 * i + int(nrec[i])]])
        return out


class Trim:
    """k-clique core trimming (kClistCoreTrim): k-clique degrees on the GPU,
    the reference's heaps on the host."""

    def __init__(self, pairs: np.ndarray, k: int, n_in: int = 0, gpu_peel_min_cliques: int = 0):
        """gpu_peel_min_cliques: peel pops whose node lies in at least this
        many alive k-cliques are counted on the GPU (0 = the library's
        default); results do not depend on it."""
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        h = C.c_void_p()
        o = _TrimOpts()
        o.gpu_peel_min_cliques = gpu_peel_min_cliques
        _check(lib().kcl_trim_create_opts(_ptr(pairs) if pairs.size else None, pairs.shape[0],
                                          n_in, k, C.byref(o), C.byref(h)))
        self._h = h
        self.k = k
        info = _TrimInfo()
        _check(lib().kcl_trim_get_info(h, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in _TrimInfo._fields_ if f != "reserved"}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().kcl_trim_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def nodes(self):
        """(map, nck, peel_key, peel_core, kept), n2 entries each."""
        n2 = self.info["n2"]
        m = max(n2, 1)
        mp = np.zeros(m, np.uint32)
        nck = np.zeros(m, np.uint64)
        pk = np.zeros(m, np.uint32)
        pc = np.zeros(m, np.uint64)
        kept = np.zeros(m, np.uint8)
        _check(lib().kcl_trim_nodes(self._h, _ptr(mp), _ptr(nck), _ptr(pk), _ptr(pc), _ptr(kept)))
        return mp[:n2], nck[:n2], pk[:n2], pc[:n2], kept[:n2]

    def edges(self) -> np.ndarray:
        mt = self.info["mtrim"]
        out = np.zeros(max(2 * mt, 2), np.uint32)
        _check(lib().kcl_trim_edges(self._h, _ptr(out)))
        return out[:2 * mt].reshape(-1, 2)

    def write(self, ckdeg=None, ckcore=None, ckdens=None, trimmed=None, stat=None, t0_unix=0):
        enc = lambda p: p.encode() if p else None  # noqa: E731
        _check(lib().kcl_trim_write(self._h, enc(ckdeg), enc(ckcore), enc(ckdens), enc(trimmed),
                                    enc(stat), t0_unix))


class Sampling:
    """Sampling++ (kCDensestSamplingKclistpp): k-cliques kept with
    probability min(1, target / count) by a seeded hash of their ids,
    synchronous "++" rounds over them, approximate densest subgraphs."""

    def __init__(self, graph: Graph, k: int, target: int = 0, seed: int = 0):
        self.graph = graph
        self.k = k
        o = _SamplingOpts()
        o.target = target
        o.seed = seed
        h = C.c_void_p()
        _check(lib().kcl_sampling_create(graph.handle, k, C.byref(o), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            lib().kcl_sampling_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def info(self) -> dict:
        i = _SamplingInfo()
        _check(lib().kcl_sampling_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _SamplingInfo._fields_}

    def cliques(self) -> np.ndarray:
        """Kept cliques, (sampled, k) uint32, each row ascending (rows in no
        particular order)."""
        n = self.info["sampled"]
        out = np.zeros((max(n, 1), self.k), np.uint32)
        _check(lib().kcl_sampling_cliques(self._h, _ptr(out)))
        return out[:n]

    def step(self) -> None:
        _check(lib().kcl_sampling_step(self._h))

    def loads(self) -> np.ndarray:
        out = np.zeros(max(self.graph.n, 1), np.uint64)
        _check(lib().kcl_sampling_loads(self._h, _ptr(out)))
        return out[:self.graph.n]

    def report(self, want_order: bool = False):
        r = _SamplingRecord()
        order = np.zeros(max(self.graph.n, 1), np.uint32) if want_order else None
        _check(lib().kcl_sampling_report(self._h, C.byref(r), _ptr(order)))
        rec = {f: getattr(r, f) for f, _ in _SamplingRecord._fields_}
        return (rec, order[:self.graph.n]) if want_order else rec

    def run(self, num_iter: int) -> list[dict]:
        cap = max(1, num_iter.bit_length())
        recs = (_SamplingRecord * cap)()
        nrec = C.c_uint32(cap)
        _check(lib().kcl_sampling_run(self._h, num_iter, recs, C.byref(nrec)))
        return [{f: getattr(r, f) for f, _ in _SamplingRecord._fields_}
                for r in recs[:nrec.value]]


def exact(graph: Graph, k: int, max_iter: int = 0, max_network_cliques: int = 0,
          stat_path: str | None = None, t0_unix: int = 0, max_seconds: int = 0,
          schedule: int = SCHED_COUNTED, model: CostModel | None = None,
          gpu_list_min: int = 0):
    """Exact k-clique densest subgraph (kCDensestMem's result).  Returns
    (result dict, input ids of the maximal densest subgraph)."""
    o = _ExactOpts()
    o.max_iter = max_iter
    o.max_network_cliques = max_network_cliques
    o.max_seconds = max_seconds
    o.schedule = schedule
    o.gpu_list_min = gpu_list_min
    o.model = model.handle if model is not None else None
    r = _ExactResult()
    nodes = np.zeros(max(graph.n, 1), np.uint32)
    _check(lib().kcl_exact_run(graph.handle, k, C.byref(o),
                               stat_path.encode() if stat_path else None, t0_unix, C.byref(r),
                               _ptr(nodes)))
    res = {f: getattr(r, f) for f, _ in _ExactResult._fields_ if f != "reserved"}
    return res, nodes[:r.n].copy()
