"""Seeded synthetic graph generators for the Sim-kClist++ workloads.

Every generator returns an ``(m, 2)`` ``uint32`` array of undirected edges:
no self-loops, no duplicate pairs (in either orientation), each pair in a
random orientation, rows in a seeded random order.  The reference reads
exactly this as a text edge list (``"%u %u"`` per line, Graph.c:41), and its
relabelling depends on the row order (the heap in Graph.c:57-91 consumes
adjacency in file order), so the row order is part of the workload.

Generators use numpy's PCG64 (``numpy.random.default_rng(seed)``); the same
seed gives the same bytes on every machine with this numpy.
"""

from __future__ import annotations

import math
import os
from typing import Iterable

import numpy as np

__all__ = [
    "erdos_renyi",
    "chung_lu",
    "rmat",
    "bipartite_plus",
    "plant_clique",
    "plant_community",
    "finalize",
    "write_text",
    "write_binary",
    "read_binary",
    "workload",
    "batch_specs",
]


def _canon_keys(u: np.ndarray, v: np.ndarray, n: int) -> np.ndarray:
    lo = np.minimum(u, v).astype(np.uint64)
    hi = np.maximum(u, v).astype(np.uint64)
    return lo * np.uint64(n) + hi


def _unique_in_draw_order(keys: np.ndarray) -> np.ndarray:
    _, first = np.unique(keys, return_index=True)
    first.sort()
    return keys[first]


def _sample_unique(rng: np.random.Generator, n: int, m: int, draw) -> np.ndarray:
    """Draw endpoint pairs with ``draw(rng, size)`` until ``m`` distinct
    loop-free undirected pairs exist; keep the first ``m`` in draw order.
    Returns canonical keys ``lo * n + hi``."""
    keys = np.empty(0, dtype=np.uint64)
    need = m
    batch = max(1024, int(m * 1.15) + 1024)
    for _ in range(64):
        u, v = draw(rng, batch)
        ok = u != v
        k = _canon_keys(u[ok], v[ok], n)
        keys = _unique_in_draw_order(np.concatenate([keys, k]))
        if keys.size >= m:
            return keys[:m]
        need = m - keys.size
        batch = max(1024, int(need * 2.0) + 1024)
    raise RuntimeError(f"could not draw {m} distinct edges on {n} vertices")


def _keys_to_pairs(keys: np.ndarray, n: int) -> np.ndarray:
    lo = (keys // np.uint64(n)).astype(np.uint32)
    hi = (keys % np.uint64(n)).astype(np.uint32)
    return np.stack([lo, hi], axis=1)


def erdos_renyi(n: int, m: int, seed: int) -> np.ndarray:
    """G(n, m): ``m`` distinct uniform pairs, as canonical keys."""
    rng = np.random.default_rng(seed)

    def draw(r, size):
        return r.integers(0, n, size, dtype=np.int64), r.integers(0, n, size, dtype=np.int64)

    return _keys_to_pairs(_sample_unique(rng, n, m, draw), n), rng


def _chung_lu_weights(n: int, avg_degree: float, gamma: float, w_max: float) -> np.ndarray:
    """w_i = c (i + i0)^(-1/(gamma-1)), with i0 and c fixed so that
    w_0 = w_max and sum(w) = n * avg_degree."""
    alpha = 1.0 / (gamma - 1.0)
    idx = np.arange(n, dtype=np.float64)
    target = n * avg_degree

    def total(i0):
        return w_max * np.sum(((idx + i0) / i0) ** (-alpha))

    lo, hi = 1e-9, 1e12
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if total(mid) < target:
            lo = mid
        else:
            hi = mid
    i0 = math.sqrt(lo * hi)
    return w_max * ((idx + i0) / i0) ** (-alpha)


def chung_lu(n: int, m: int, gamma: float, w_max: float, seed: int):
    """Chung-Lu power law by endpoint sampling: both endpoints drawn
    independently with probability proportional to the weights
    (average degree 2m/n, largest expected degree ``w_max``); loops and
    repeats are redrawn until ``m`` distinct pairs exist."""
    rng = np.random.default_rng(seed)
    w = _chung_lu_weights(n, 2.0 * m / n, gamma, w_max)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]

    def draw(r, size):
        a = np.searchsorted(cdf, r.random(size), side="right")
        b = np.searchsorted(cdf, r.random(size), side="right")
        np.minimum(a, n - 1, out=a)
        np.minimum(b, n - 1, out=b)
        return a.astype(np.int64), b.astype(np.int64)

    return _keys_to_pairs(_sample_unique(rng, n, m, draw), n), rng


def rmat(scale: int, m: int, a: float, b: float, c: float, seed: int):
    """R-MAT with quadrant probabilities (a, b, c, 1-a-b-c), ``2**scale``
    vertices, ``m`` distinct undirected pairs after dropping loops and
    repeats; vertex ids are then permuted by a seeded permutation."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    ab, abc = a + b, a + b + c

    def draw(r, size):
        u = np.zeros(size, dtype=np.int64)
        v = np.zeros(size, dtype=np.int64)
        for _ in range(scale):
            x = r.random(size)
            ubit = x >= ab
            vbit = ((x >= a) & (x < ab)) | (x >= abc)
            u = (u << 1) | ubit
            v = (v << 1) | vbit
        return u, v

    pairs = _keys_to_pairs(_sample_unique(rng, n, m, draw), n)
    perm = rng.permutation(n).astype(np.uint32)
    return perm[pairs], rng


def bipartite_plus(part: int, p_in: float, seed: int) -> np.ndarray:
    """Complete bipartite K_{part,part} plus each pair inside a side with
    probability ``p_in``, finalized: DAG out-degrees reach about ``part``
    (wide bitmap rows) while the cliques stay few (a clique takes at most a
    clique of each side)."""
    rng = np.random.default_rng(seed)
    a = np.arange(part)
    b = np.arange(part, 2 * part)
    cross = np.stack(np.meshgrid(a, b, indexing="ij"), -1).reshape(-1, 2)
    intra = []
    for side in (a, b):
        iu, ju = np.triu_indices(part, 1)
        keep = rng.random(iu.size) < p_in
        intra.append(np.stack([side[iu[keep]], side[ju[keep]]], 1))
    pairs = np.concatenate([cross, *intra]).astype(np.uint32)
    return finalize(pairs, rng)


def plant_clique(pairs: np.ndarray, n: int, size: int, rng: np.random.Generator) -> np.ndarray:
    """Add every pair among ``size`` distinct random vertices."""
    return plant_community(pairs, n, size, 1.0, rng)


def plant_community(pairs: np.ndarray, n: int, size: int, p: float,
                    rng: np.random.Generator) -> np.ndarray:
    """Add each pair among ``size`` distinct random vertices with
    probability ``p`` (p = 1: a clique), then drop repeats."""
    ids = rng.choice(n, size=size, replace=False).astype(np.int64)
    iu, ju = np.triu_indices(size, k=1)
    keep = rng.random(iu.size) < p if p < 1.0 else np.ones(iu.size, dtype=bool)
    u, v = ids[iu[keep]], ids[ju[keep]]
    base = _canon_keys(pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64), n)
    extra = _canon_keys(u, v, n)
    keys = _unique_in_draw_order(np.concatenate([base, extra]))
    return _keys_to_pairs(keys, n)


def finalize(pairs: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Random row order and random orientation of each pair."""
    perm = rng.permutation(pairs.shape[0])
    out = pairs[perm]
    flip = rng.random(out.shape[0]) < 0.5
    out[flip] = out[flip][:, ::-1]
    return np.ascontiguousarray(out, dtype=np.uint32)


def write_text(pairs: np.ndarray, path: str, progress=None) -> None:
    """The reference's input format: one ``"%u %u\\n"`` line per edge.
    progress(lines_written) is called after every 2^24 lines."""
    with open(path, "w") as f:
        if pairs.shape[0] == 0:
            return
        step = 1 << 20
        for i in range(0, pairs.shape[0], step):
            blk = pairs[i:i + step]
            f.write("\n".join(f"{a} {b}" for a, b in blk.tolist()))
            f.write("\n")
            if progress and ((i + step) >> 24) != (i >> 24):
                progress(min(i + step, pairs.shape[0]))


def write_binary(pairs: np.ndarray, path: str) -> None:
    """Little-endian uint32 pairs, no header."""
    np.ascontiguousarray(pairs, dtype="<u4").tofile(path)


def read_binary(path: str) -> np.ndarray:
    return np.fromfile(path, dtype="<u4").reshape(-1, 2)


# ---------------------------------------------------------------------------
# Named workloads.


def _er_planted(n, m, clique, seed):
    pairs, rng = erdos_renyi(n, m, seed)
    pairs = plant_clique(pairs, n, clique, rng)
    return finalize(pairs, rng)


def _cl_planted(n, m, gamma, w_max, size, p, seed):
    pairs, rng = chung_lu(n, m, gamma, w_max, seed)
    pairs = plant_community(pairs, n, size, p, rng)
    return finalize(pairs, rng)


def workload(name: str, seed: int = 1, scale_down: int = 1) -> np.ndarray:
    """Edge arrays of the named workloads.

    ``er20k``: G(20000, 100000) plus a planted 30-clique (k=4 case).
    ``cl1m``:  Chung-Lu n=1e6, m=1.6e7, gamma 2.3, w_max 2e4, plus a
               150-vertex community with p=0.8 (k=5 case).
    ``cl4m``:  Chung-Lu n=4e6, m=6e7, gamma 2.1, plus a 200-vertex
               near-clique with p=0.9 (k=6 exact case).
    ``rmat23``: R-MAT(0.57, 0.19, 0.19, 0.05), scale 23, m=2e8 (k=7/8).
    ``scale_down`` divides n and m (for quick tests of the same shape).
    """
    s = max(1, int(scale_down))
    if name == "er20k":
        return _er_planted(20000 // s, 100000 // s, 30, seed)
    if name == "cl1m":
        return _cl_planted(1_000_000 // s, 16_000_000 // s, 2.3, 20000.0 / math.sqrt(s),
                           150, 0.8, seed)
    if name == "cl4m":
        return _cl_planted(4_000_000 // s, 60_000_000 // s, 2.1, 4e4 / math.sqrt(s),
                           200, 0.9, seed)
    if name == "rmat23":
        scale = 23 - int(round(math.log2(s)))
        pairs, rng = rmat(scale, 200_000_000 // s, 0.57, 0.19, 0.19, seed)
        return finalize(pairs, rng)
    raise KeyError(name)


def batch_specs(count: int = 512, first_seed: int = 1000) -> Iterable[dict]:
    """The mixed batch: n log-uniform in [1e3, 2e5], average degree
    uniform in [4, 40], gamma uniform in [2.1, 2.8], k = 3 + (i mod 5)."""
    meta = np.random.default_rng(first_seed)
    for i in range(count):
        n = int(round(math.exp(meta.uniform(math.log(1e3), math.log(2e5)))))
        deg = float(meta.uniform(4.0, 40.0))
        gamma = float(meta.uniform(2.1, 2.8))
        m = int(n * deg / 2)
        yield {"index": i, "seed": first_seed + i, "n": n, "m": m, "gamma": gamma,
               "w_max": float(math.sqrt(n * deg)), "k": 3 + (i % 5)}


def batch_graph(spec: dict) -> np.ndarray:
    pairs, rng = chung_lu(spec["n"], spec["m"], spec["gamma"], spec["w_max"], spec["seed"])
    return finalize(pairs, rng)


if __name__ == "__main__":  # pragma: no cover - manual generation helper
    import argparse

    ap = argparse.ArgumentParser(description="write a named workload")
    ap.add_argument("name")
    ap.add_argument("out")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--scale-down", type=int, default=1)
    ap.add_argument("--binary", action="store_true")
    a = ap.parse_args()
    e = workload(a.name, a.seed, a.scale_down)
    (write_binary if a.binary else write_text)(e, a.out)
    print(f"{a.name}: {e.shape[0]} edges -> {a.out}", flush=True)
    os._exit(0)
