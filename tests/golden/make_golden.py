"""Write the golden vectors of the Sim-kClist++ path from the reference itself.

Runs in the build container only (it needs /root/reference): builds the
reference's CmpSync sources and kClistCoreTrim.c with oracle/Makefile into
oracle/_ref/, then for each seeded case writes tests/golden/<case>.npz with
what the reference computed:

  pairs            input edges (small cases; larger ones store a SHA-256
                   of the regenerated input instead)
  rank             SortByCore ranks (Graph.c:82)
  es, et           sorted DAG edges (Graph.c:125)
  counts           per-iteration rho_add as uint64, rows t = 1..Tmax
                   (or count_sha: SHA-256 of each row for large cases)
  rho_bits         final rho as IEEE bits
  records          one row per report: T, n, m, density bits, ub bits
  level, tent      level and rho_tentative of the last report
and for trimming cases the outputs of the reference kClistCoreTrim program
(ckdeg, ckcore, ckdens, trimmed edges, stat line).

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kclistpp_amd import graphgen as gg  # noqa: E402
from oracle import REF_TRIM_BIN, RefHarness  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def small_graphs():
    """Named small inputs: (name, pairs)."""
    out = []

    def cl(name, n, m, gamma, wmax, size, p, seed):
        pairs, rng = gg.chung_lu(n, m, gamma, wmax, seed)
        if size:
            pairs = gg.plant_community(pairs, n, size, p, rng)
        out.append((name, gg.finalize(pairs, rng)))

    def er(name, n, m, clique, seed):
        pairs, rng = gg.erdos_renyi(n, m, seed)
        if clique:
            pairs = gg.plant_clique(pairs, n, clique, rng)
        out.append((name, gg.finalize(pairs, rng)))

    er("er2k_k12", 2000, 10000, 12, 11)
    cl("cl3k", 3000, 30000, 2.3, 300.0, 40, 0.7, 12)
    cl("cl1k_dense", 1000, 20000, 2.1, 400.0, 30, 0.9, 13)
    er("er300", 300, 3000, 0, 14)
    # Hand-made shapes.
    k7 = np.array([(i, j) for i in range(7) for j in range(i + 1, 7)], np.uint32)
    out.append(("k7", k7))
    two = np.concatenate([k7, np.array([(i + 10, j + 10) for i in range(5)
                                        for j in range(i + 1, 5)], np.uint32),
                          np.array([(6, 10)], np.uint32)])
    out.append(("k7_k5_bridge", two))
    path = np.array([(i, i + 1) for i in range(20)], np.uint32)
    out.append(("path21", path))
    star = np.array([(0, i) for i in range(1, 30)], np.uint32)
    out.append(("star30", star))
    return out


SIM_CASES = [
    # (graph name, k, num_iter)
    ("er2k_k12", 4, 8), ("er2k_k12", 3, 4),
    ("cl3k", 2, 4), ("cl3k", 3, 8), ("cl3k", 4, 8), ("cl3k", 5, 4), ("cl3k", 6, 4),
    ("cl1k_dense", 4, 4), ("cl1k_dense", 5, 4), ("cl1k_dense", 7, 2),
    ("er300", 3, 8), ("k7", 3, 4), ("k7", 5, 4), ("k7", 7, 2), ("k7", 8, 2),
    ("k7_k5_bridge", 4, 8), ("path21", 3, 4), ("path21", 2, 4), ("star30", 2, 4),
]

TRIM_CASES = [("cl3k", 4), ("cl3k", 5), ("cl1k_dense", 5), ("er2k_k12", 4), ("k7_k5_bridge", 4)]

# Full-size workload cases: vectors stored as digests.
BIG_CASES = [
    # (workload name, seed, scale_down, k, num_iter)
    ("er20k", 1, 1, 4, 10),
    ("cl1m", 3, 64, 5, 4),
]


def run_sim_case(ref: RefHarness, path: str, k: int, num_iter: int, store_vectors: bool):
    h = ref.load(path)
    info = ref.info(h)
    rank = ref.rank(h)
    e = ref.edges(h)
    ref.start(h, k)
    counts, count_sha, records = [], [], []
    rho = np.zeros(info["n"])
    T, t = 1, 1
    level = tent = None
    while T <= num_iter:
        while t <= T:
            add, rho = ref.step(h, t)
            c = add.astype(np.uint64)
            assert (c.astype(np.float64) == add).all()
            if store_vectors:
                counts.append(c)
            count_sha.append(sha(c))
            t += 1
        rec, level, _, tent = ref.report(h)
        records.append([T, rec["n"], rec["m"],
                        int(np.float64(rec["density"]).view(np.uint64)),
                        int(np.float64(rec["ub"]).view(np.uint64))])
        T <<= 1
    ref.free(h)
    out = {
        "n_in": np.uint32(info["n_in"]), "n": np.uint32(info["n"]), "core": np.uint32(info["core"]),
        "k": np.uint32(k), "num_iter": np.uint32(num_iter),
        "records": np.array(records, dtype=np.uint64),
        "rho_bits": rho.view(np.uint64).copy() if store_vectors else np.array([], np.uint64),
        "rho_sha": np.array([sha(rho.view(np.uint64))]),
        "count_sha": np.array(count_sha),
    }
    if store_vectors:
        out["rank"] = rank
        out["es"] = e[:, 0].copy()
        out["et"] = e[:, 1].copy()
        out["counts"] = np.array(counts, dtype=np.uint64).reshape(len(counts), -1)
        out["level"] = level
        out["tent"] = tent.astype(np.uint64)
    else:
        out["rank_sha"] = np.array([sha(rank)])
        out["edges_sha"] = np.array([sha(e)])
    return out


def run_trim_case(workdir: str, path: str, k: int, tag: str):
    ckdeg = os.path.join(workdir, "ckdeg.txt")
    ckcore = os.path.join(workdir, "ckcore.txt")
    ckdens = os.path.join(workdir, "ckdens.txt")
    subprocess.run([REF_TRIM_BIN, str(k), path, ckdeg, ckcore, ckdens, tag], cwd=workdir,
                   check=True, capture_output=True)
    deg = np.loadtxt(ckdeg, dtype=np.uint64, ndmin=2)
    core = np.loadtxt(ckcore, dtype=np.uint64, ndmin=2)
    with open(ckdens) as f:
        dens_line = f.read().strip()
    trimmed_path = os.path.join(workdir, f"trimmed_{tag}_{k}.txt")
    trimmed = np.loadtxt(trimmed_path, dtype=np.uint32, ndmin=2) if os.path.getsize(trimmed_path) \
        else np.zeros((0, 2), np.uint32)
    with open(os.path.join(workdir, f"stat_trim_{tag}_{k}.txt")) as f:
        stat = f.read().strip().split("\n")[-1].split("\t")
    return {
        "ckdeg": deg.reshape(-1, 2), "ckcore": core.reshape(-1, 2),
        "ckdens": np.array([dens_line]), "trimmed": trimmed.reshape(-1, 2),
        "stat": np.array([int(stat[0]), int(stat[1]), int(stat[2])], dtype=np.uint64),
    }


def one_sim_case(path: str, k: int, T: int, store_vectors: bool, out: str) -> None:
    """Child process: the reference keeps per-thread scratch sized by the
    first graph it sees (kCDensestCmpSync.c:74-80), so each graph gets a
    fresh process."""
    ref = RefHarness()
    ref.threads(8)
    data = run_sim_case(ref, path, k, T, store_vectors)
    np.savez(out, **data)


def sim_case_in_child(path: str, k: int, T: int, store_vectors: bool) -> dict:
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "case.npz")
        subprocess.run([sys.executable, os.path.abspath(__file__), "--one", path, str(k), str(T),
                        "1" if store_vectors else "0", out], check=True, capture_output=True)
        with np.load(out, allow_pickle=False) as z:
            return {key: z[key] for key in z.files}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True,
                   capture_output=True)
    graphs = dict(small_graphs())
    index = {"sim": [], "trim": [], "big": []}
    with tempfile.TemporaryDirectory() as td:
        paths = {}
        for name, pairs in graphs.items():
            p = os.path.join(td, name + ".txt")
            gg.write_text(pairs, p)
            paths[name] = p
        for name, k, T in SIM_CASES:
            case = f"sim_{name}_k{k}_T{T}"
            data = sim_case_in_child(paths[name], k, T, store_vectors=True)
            data["pairs"] = graphs[name]
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            index["sim"].append(case)
            print(case, "records", data["records"][:, :3].tolist(), flush=True)
        for name, k in TRIM_CASES:
            case = f"trim_{name}_k{k}"
            data = run_trim_case(td, paths[name], k, name)
            data["pairs"] = graphs[name]
            data["k"] = np.uint32(k)
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            index["trim"].append(case)
            print(case, "stat", data["stat"].tolist(), flush=True)
        for wname, seed, sd, k, T in BIG_CASES:
            case = f"big_{wname}_s{seed}_d{sd}_k{k}_T{T}"
            pairs = gg.workload(wname, seed=seed, scale_down=sd)
            p = os.path.join(td, case + ".txt")
            gg.write_text(pairs, p)
            data = sim_case_in_child(p, k, T, store_vectors=False)
            data["pairs_sha"] = np.array([sha(pairs)])
            data["workload"] = np.array([wname])
            data["seed"] = np.uint32(seed)
            data["scale_down"] = np.uint32(sd)
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            index["big"].append(case)
            print(case, "records", data["records"][:, :3].tolist(), flush=True)
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--one":
        a = sys.argv[2:]
        one_sim_case(a[0], int(a[1]), int(a[2]), a[3] == "1", a[4])
    else:
        main()
