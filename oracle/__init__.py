"""ctypes bindings for the oracle (test infrastructure only).

``Oracle`` wraps ``oracle/_build/libkcloracle.so`` (the C restatement of
the reference's Sim-kClist++ path and trimming); ``RefHarness`` wraps
``oracle/_ref/libkclref.so`` (the reference's own sources compiled with
oracle/ref_harness.c).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only as the checker.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libkcloracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkclref.so")
REF_TRIM_BIN = os.path.join(HERE, "_ref", "kClistCoreTrim")
REF_SAMPLING_BIN = os.path.join(HERE, "_ref", "kCDensestSamplingKclistpp")
REF_CMPSYNC_BIN = os.path.join(HERE, "_ref", "kCDensestCmpSync")
REF_CMPSYNC_BIN = os.path.join(HERE, "_ref", "kCDensestCmpSync")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p_raw = C.POINTER(C.c_uint32)


class _Densest(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("density", C.c_double),
                ("ub", C.c_double), ("sum", C.c_uint64)]


class _Record(C.Structure):
    _fields_ = [("T", C.c_uint32), ("n", C.c_uint32), ("m", C.c_uint32),
                ("density", C.c_double), ("ub", C.c_double), ("sum", C.c_uint64)]


class _TrimInfo(C.Structure):
    _fields_ = [("n_in", C.c_uint32), ("n2", C.c_uint32), ("e2", C.c_uint32),
                ("core", C.c_uint32), ("ncktot", C.c_uint64), ("ckcore", C.c_uint64),
                ("ckrho", C.c_double), ("rho", C.c_double), ("size", C.c_uint32),
                ("ntrim", C.c_uint32), ("ncktrim", C.c_uint64), ("mtrim", C.c_uint32)]


@dataclass
class Dag:
    """Relabelled DAG: sorted root edges (es, et), CSR (cd, adj)."""
    n_in: int
    n: int
    core: int
    rank: np.ndarray
    es: np.ndarray
    et: np.ndarray
    cd: np.ndarray
    adj: np.ndarray

    @property
    def m(self) -> int:
        return int(self.es.size)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return C.CDLL(path)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        L = _load(path)
        L.orc_degeneracy_rank.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p]
        L.orc_build_dag.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _u32p, _u32p,
                                    _u32p, _u32p, C.POINTER(C.c_uint32)]
        L.orc_build_dag.restype = C.c_uint32
        L.orc_clique_enum.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, _u32p, _u32p,
                                      C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_clique_enum.restype = C.c_uint64
        L.orc_rho_update.argtypes = [C.c_uint32, C.c_uint32, _f64p, _u64p]
        L.orc_find_densest.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, _u32p, _u32p,
                                       C.c_uint32, _f64p, _u32p, _u32p, _u64p,
                                       C.POINTER(_Densest)]
        L.orc_cmpsync.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, _u32p, _u32p,
                                  C.c_uint32, C.c_uint32, C.POINTER(_Record), C.c_void_p,
                                  C.c_void_p]
        L.orc_cmpsync.restype = C.c_int
        L.orc_trim.argtypes = [C.c_uint32, C.c_uint64, _u32p, C.c_uint32, _u32p, _u64p, _u32p,
                               _u64p, _u32p, C.POINTER(_TrimInfo)]
        L.orc_trim.restype = C.c_int
        L.orc_max_flow.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _u64p, C.c_uint32,
                                   C.c_uint32]
        L.orc_max_flow.restype = C.c_uint64
        L.orc_count_in_set.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint64, _u32p, _u32p,
                                       C.c_uint32, _u8p]
        L.orc_count_in_set.restype = C.c_uint64
        L.orc_read_edge_list.argtypes = [C.c_char_p, C.POINTER(_u32p_raw), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32)]
        L.orc_free.argtypes = [C.c_void_p]
        self.L = L

    # -- input -----------------------------------------------------------
    def read_edge_list(self, path: str) -> tuple[np.ndarray, int]:
        """fscanf("%u %u") over the whole file (Graph.c:34-53): pairs, max id + 1."""
        p = _u32p_raw()
        m, n = C.c_uint64(0), C.c_uint32(0)
        if self.L.orc_read_edge_list(path.encode(), C.byref(p), C.byref(m), C.byref(n)) != 0:
            raise OSError(path)
        try:
            out = np.ctypeslib.as_array(p, shape=(2 * m.value,)).copy() if m.value \
                else np.zeros(0, np.uint32)
        finally:
            self.L.orc_free(p)
        return out.reshape(-1, 2), int(n.value)

    # -- graph ---------------------------------------------------------
    def rank(self, pairs: np.ndarray, n: int | None = None) -> np.ndarray:
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
        if n is None:
            n = int(pairs.max()) + 1 if pairs.size else 1
        rank = np.zeros(n, dtype=np.uint32)
        self.L.orc_degeneracy_rank(n, pairs.shape[0], pairs.reshape(-1), rank)
        return rank

    def dag(self, pairs: np.ndarray) -> Dag:
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n_in = int(pairs.max()) + 1 if pairs.size else 1
        rank = self.rank(pairs, n_in)
        m = pairs.shape[0]
        es = np.zeros(max(m, 1), dtype=np.uint32)
        et = np.zeros(max(m, 1), dtype=np.uint32)
        cd = np.zeros(n_in + 1, dtype=np.uint32)
        adj = np.zeros(max(m, 1), dtype=np.uint32)
        core = C.c_uint32(0)
        n2 = self.L.orc_build_dag(n_in, m, pairs.reshape(-1), rank, es, et, cd, adj,
                                  C.byref(core))
        return Dag(n_in, int(n2), int(core.value), rank, es[:m].copy(), et[:m].copy(),
                   cd[:n2 + 1].copy(), adj[:m].copy())

    # -- enumeration ---------------------------------------------------
    def _dag_args(self, g: Dag):
        es = g.es if g.m else np.zeros(1, np.uint32)
        et = g.et if g.m else np.zeros(1, np.uint32)
        adj = g.adj if g.m else np.zeros(1, np.uint32)
        return g.n, np.ascontiguousarray(g.cd), adj, g.m, es, et

    def fw_counts(self, g: Dag, k: int, rho: np.ndarray) -> tuple[np.ndarray, int]:
        acc = np.zeros(max(g.n, 1), dtype=np.uint64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        v = self.L.orc_clique_enum(*self._dag_args(g), k, 2, rho.ctypes.data, None,
                                   acc.ctypes.data)
        return acc[:g.n], int(v)

    def pava_counts(self, g: Dag, k: int, level: np.ndarray) -> np.ndarray:
        acc = np.zeros(max(g.n, 1), dtype=np.uint64)
        level = np.ascontiguousarray(level, dtype=np.uint32)
        self.L.orc_clique_enum(*self._dag_args(g), k, 3, None, level.ctypes.data,
                               acc.ctypes.data)
        return acc[:g.n]

    def count(self, g: Dag, k: int) -> int:
        return int(self.L.orc_clique_enum(*self._dag_args(g), k, 2, None, None, None))

    def count_in_set(self, g: Dag, k: int, mask: np.ndarray) -> int:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        return int(self.L.orc_count_in_set(*self._dag_args(g), k, mask))

    def rho_update(self, rho: np.ndarray, t: int, cnt: np.ndarray) -> np.ndarray:
        rho = np.array(rho, dtype=np.float64, copy=True)
        self.L.orc_rho_update(rho.size, t, rho, np.ascontiguousarray(cnt, dtype=np.uint64))
        return rho

    def find_densest(self, g: Dag, k: int, rho: np.ndarray):
        n = max(g.n, 1)
        reordered = np.zeros(n, np.uint32)
        level = np.zeros(n, np.uint32)
        tent = np.zeros(n, np.uint64)
        d = _Densest()
        self.L.orc_find_densest(*self._dag_args(g), k, np.ascontiguousarray(rho, np.float64),
                                reordered, level, tent, C.byref(d))
        return ({"n": d.n, "m": d.m, "density": d.density, "ub": d.ub, "sum": d.sum},
                reordered[:g.n], level[:g.n], tent[:g.n])

    def cmpsync(self, g: Dag, k: int, num_iter: int, keep_counts: bool = False):
        recs = (_Record * 33)()
        tmax = 1 << (num_iter.bit_length() - 1) if num_iter > 0 else 0
        counts = np.zeros((max(tmax, 1), max(g.n, 1)), np.uint64) if keep_counts else None
        rho = np.zeros(max(g.n, 1), np.float64)
        nrec = self.L.orc_cmpsync(*self._dag_args(g), k, num_iter, recs,
                                  counts.ctypes.data if keep_counts else None, rho.ctypes.data)
        out = [{"T": r.T, "n": r.n, "m": r.m, "density": r.density, "ub": r.ub, "sum": r.sum}
               for r in recs[:nrec]]
        return out, (counts[:tmax, :g.n] if keep_counts else None), rho[:g.n]

    # -- trimming --------------------------------------------------------
    def trim(self, pairs: np.ndarray, k: int):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        n_in = int(pairs.max()) + 1 if pairs.size else 1
        m = pairs.shape[0]
        mp = np.zeros(n_in, np.uint32)
        nck = np.zeros(n_in, np.uint64)
        ick = np.zeros(n_in, np.uint32)
        icv = np.zeros(n_in, np.uint64)
        tp = np.zeros(max(2 * m, 2), np.uint32)
        info = _TrimInfo()
        flat = pairs.reshape(-1) if m else np.zeros(2, np.uint32)
        rc = self.L.orc_trim(n_in, m, flat, k, mp, nck, ick, icv, tp, C.byref(info))
        if rc != 0:
            raise ValueError("k must be >= 3")
        n2 = info.n2
        d = {f: getattr(info, f) for f, _ in _TrimInfo._fields_}
        return d, mp[:n2], nck[:n2], ick[:n2], icv[:n2], tp[:2 * info.mtrim].reshape(-1, 2)

    def max_flow(self, nv, frm, to, cap, s, t) -> int:
        return int(self.L.orc_max_flow(nv, len(frm), np.ascontiguousarray(frm, np.uint32),
                                       np.ascontiguousarray(to, np.uint32),
                                       np.ascontiguousarray(cap, np.uint64), s, t))


class RefHarness:
    """The reference's own CmpSync code (built from its sources)."""

    def __init__(self, path: str = REF_SO):
        L = _load(path)
        L.kclref_load.argtypes = [C.c_char_p]
        L.kclref_load.restype = C.c_void_p
        L.kclref_load_pairs.argtypes = [_u32p, C.c_uint]
        L.kclref_load_pairs.restype = C.c_void_p
        L.kclref_load_csr.argtypes = [C.c_uint, _u32p, _u32p]
        L.kclref_load_csr.restype = C.c_void_p
        for f in ("kclref_n_in", "kclref_n", "kclref_e", "kclref_core"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_uint
        L.kclref_rank.argtypes = [C.c_void_p, _u32p]
        L.kclref_edges.argtypes = [C.c_void_p, _u32p]
        L.kclref_csr.argtypes = [C.c_void_p, _u32p, _u32p]
        L.kclref_threads.argtypes = [C.c_int]
        L.kclref_start.argtypes = [C.c_void_p, C.c_uint]
        L.kclref_step.argtypes = [C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
        L.kclref_report.argtypes = [C.c_void_p, _u32p, _f64p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
        L.kclref_set_rho.argtypes = [C.c_void_p, _f64p]
        L.kclref_pava_level.argtypes = [C.c_void_p, _u32p, _f64p]
        L.kclref_time_fw_roots.argtypes = [C.c_void_p, _u32p, C.c_uint, C.POINTER(C.c_double)]
        L.kclref_time_fw_roots.restype = C.c_double
        L.kclref_time_fw_each.argtypes = [C.c_void_p, _u32p, C.c_uint, _f64p,
                                          C.POINTER(C.c_double)]
        L.kclref_time_fw_each.restype = C.c_double
        L.kclref_free.argtypes = [C.c_void_p]
        self.L = L

    def load(self, path: str):
        h = self.L.kclref_load(path.encode())
        if not h:
            raise RuntimeError("reference load failed")
        return h

    def load_pairs(self, pairs: np.ndarray):
        """The reference's graph from edge pairs in memory (same EdgeList as
        its reader would build from their text)."""
        p = np.ascontiguousarray(pairs, np.uint32)
        h = self.L.kclref_load_pairs(p, len(p))
        if not h:
            raise RuntimeError("reference load failed")
        return h

    def load_csr(self, cd: np.ndarray, adj: np.ndarray):
        """The reference's Graph from a DAG CSR (no degeneracy ordering)."""
        cd = np.ascontiguousarray(cd, np.uint32)
        adj = np.ascontiguousarray(adj if adj.size else np.zeros(1, np.uint32), np.uint32)
        h = self.L.kclref_load_csr(cd.size - 1, cd, adj)
        if not h:
            raise RuntimeError("reference load failed")
        return h

    def info(self, h):
        L = self.L
        return {"n_in": L.kclref_n_in(h), "n": L.kclref_n(h), "e": L.kclref_e(h),
                "core": L.kclref_core(h)}

    def rank(self, h):
        out = np.zeros(self.L.kclref_n_in(h), np.uint32)
        self.L.kclref_rank(h, out)
        return out

    def edges(self, h):
        out = np.zeros(2 * max(self.L.kclref_e(h), 1), np.uint32)
        self.L.kclref_edges(h, out)
        return out[:2 * self.L.kclref_e(h)].reshape(-1, 2)

    def csr(self, h):
        n, e = self.L.kclref_n(h), self.L.kclref_e(h)
        cd = np.zeros(n + 1, np.uint32)
        adj = np.zeros(max(e, 1), np.uint32)
        self.L.kclref_csr(h, cd, adj)
        return cd, adj[:e]

    def threads(self, p: int):
        self.L.kclref_threads(p)

    def start(self, h, k: int):
        self.L.kclref_start(h, k)

    def step(self, h, t: int):
        n = self.L.kclref_n(h)
        add = np.zeros(max(n, 1), np.float64)
        rho = np.zeros(max(n, 1), np.float64)
        self.L.kclref_step(h, t, add.ctypes.data, rho.ctypes.data)
        return add[:n], rho[:n]

    def report(self, h):
        n = self.L.kclref_n(h)
        nm = np.zeros(2, np.uint32)
        du = np.zeros(2, np.float64)
        level = np.zeros(max(n, 1), np.uint32)
        reo = np.zeros(max(n, 1), np.uint32)
        tent = np.zeros(max(n, 1), np.float64)
        self.L.kclref_report(h, nm, du, level.ctypes.data, reo.ctypes.data, tent.ctypes.data)
        return ({"n": int(nm[0]), "m": int(nm[1]), "density": float(du[0]), "ub": float(du[1])},
                level[:n], reo[:n], tent[:n])

    def set_rho(self, h, rho):
        self.L.kclref_set_rho(h, np.ascontiguousarray(rho, np.float64))

    def pava_level(self, h, level: np.ndarray) -> np.ndarray:
        """One PAVA-preprocess pass with the given level vector; returns
        rho_tentative (indexed by level)."""
        n = self.L.kclref_n(h)
        out = np.zeros(max(n, 1), np.float64)
        self.L.kclref_pava_level(h, np.ascontiguousarray(level, np.uint32), out)
        return out[:n]

    def time_fw_roots(self, h, roots: np.ndarray):
        v = C.c_double(0)
        roots = np.ascontiguousarray(roots, np.uint32)
        s = self.L.kclref_time_fw_roots(h, roots, roots.size, C.byref(v))
        return float(s), float(v.value)

    def time_fw_each(self, h, roots: np.ndarray):
        """Per-edge thread-seconds of CDF_CliqueEnum(FRANK_WOLFE), edges run
        one per thread: (wall seconds, per-edge seconds, visits)."""
        roots = np.ascontiguousarray(roots, np.uint32)
        sec = np.zeros(max(roots.size, 1), np.float64)
        v = C.c_double(0)
        wall = self.L.kclref_time_fw_each(h, roots, roots.size, sec, C.byref(v))
        return float(wall), sec[:roots.size], float(v.value)

    def free(self, h):
        self.L.kclref_free(h)
