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

    python tests/golden/make_golden.py [small] [wide] [big] [batch]
    python tests/golden/make_golden.py --big-case <case> ...   # one workload case

Groups: small (hand-made and small seeded graphs, full vectors), wide
(wide bitmap rows at k = 6-8 and a wide trimming case), big (the benchmark
workloads and their scaled families, as digests), batch (the 512-graph mixed
batch, per-graph digests).  Existing index entries are kept.
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

# Workload cases: vectors stored as digests.  One per benchmark config at
# the largest size the reference finishes here in minutes: the k = 4 case
# and the k = 5 case at full size, the k = 6 exact case's Chung-Lu family
# (gamma 2.1 plus the 200-vertex near-clique with p = 0.9, 1.7e10 6-cliques)
# at 1/256 of its nodes and edges, and the R-MAT case at scales 10 and 11
# (k = 7 and 8; 2e9-1.2e10 cliques per pass).
BIG_CASES = [
    # (workload name, seed, scale_down, k, num_iter)
    ("er20k", 1, 1, 4, 10),
    ("cl1m", 3, 64, 5, 4),
    ("cl1m", 1, 1, 5, 4),
    ("cl4m", 1, 256, 6, 2),
    ("rmat23", 1, 8192, 7, 2),
    ("rmat23", 1, 8192, 8, 1),
    ("rmat23", 1, 4096, 7, 1),
]

# Wide out-neighbourhoods: K_{part,part} plus sparse edges inside each side,
# so that roots fill the 6-, 8-, 12- and 16-word bitmap classes; one
# Frank-Wolfe pass at a seeded rho with many ties and one PAVA pass at a
# seeded level permutation per k.
WIDE_CASES = [
    # (part, p_in, seed, ks)
    (340, 0.06, 340, (6, 7, 8)),
    (480, 0.05, 480, (6, 7, 8)),
    (700, 0.04, 700, (6, 7, 8)),
    (940, 0.035, 940, (6, 7, 8)),
]

# Trimming of a wide graph (k-clique degrees over rows of 12 words).
WIDE_TRIM_CASES = [(700, 0.04, 700, 6)]

# The mixed batch: every graph of graphgen.batch_specs(512, 1000) with its
# own k, num_iter = 4 (iterations 1-4, reports at T = 1, 2, 4).
BATCH_T = 4


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


def run_wide_case(ref: RefHarness, pairs: np.ndarray, k: int, seed: int) -> dict:
    h = ref.load_pairs(pairs)
    info = ref.info(h)
    ref.start(h, k)
    rng = np.random.default_rng(seed * 10 + k)
    rho = rng.integers(0, 3, info["n"]).astype(np.float64)
    level = rng.permutation(info["n"]).astype(np.uint32)
    ref.set_rho(h, rho)
    add, _ = ref.step(h, 1)  # the pass runs at the rho just set
    tent = ref.pava_level(h, level)
    ref.free(h)
    fw = add.astype(np.uint64)
    assert (fw.astype(np.float64) == add).all() and (tent.astype(np.uint64) == tent).all()
    return {"n": np.uint32(info["n"]), "core": np.uint32(info["core"]), "k": np.uint32(k),
            "rho": rho, "level": level, "fw": fw, "visits": np.uint64(fw.sum()),
            "tent": tent.astype(np.uint64)}


def one_wide_case(part: int, p_in: float, seed: int, k: int, out: str) -> None:
    ref = RefHarness()
    ref.threads(8)
    pairs = gg.bipartite_plus(part, p_in, seed)
    data = run_wide_case(ref, pairs, k, seed)
    data["pairs_sha"] = np.array([sha(pairs)])
    data["gen"] = np.array([part, seed], np.uint32)
    data["p_in"] = np.float64(p_in)
    np.savez(out, **data)


def one_batch_graph(i: int, T: int, out: str) -> None:
    """Child process: graph i of the mixed batch through the reference's own
    load (SortByCore/Relabel/MakeGraph), T iterations and the reports."""
    spec = list(gg.batch_specs(512, first_seed=1000))[i]
    pairs = gg.batch_graph(spec)
    ref = RefHarness()
    ref.threads(2)
    h = ref.load_pairs(pairs)
    info = ref.info(h)
    rank, e = ref.rank(h), ref.edges(h)
    ref.start(h, spec["k"])
    shas, recs = [], []
    t, Tp = 1, 1
    rho = np.zeros(info["n"])
    while Tp <= T:
        while t <= Tp:
            add, rho = ref.step(h, t)
            c = add.astype(np.uint64)
            assert (c.astype(np.float64) == add).all()
            shas.append(sha(c))
            t += 1
        rec, _, _, _ = ref.report(h)
        recs.append([Tp, rec["n"], rec["m"], int(np.float64(rec["density"]).view(np.uint64)),
                     int(np.float64(rec["ub"]).view(np.uint64))])
        Tp <<= 1
    ref.free(h)
    np.savez(out, n=np.uint32(info["n"]), core=np.uint32(info["core"]),
             pairs_sha=np.array([sha(pairs)]), rank_sha=np.array([sha(rank)]),
             edges_sha=np.array([sha(e)]), count_sha=np.array(shas),
             rho_sha=np.array([sha(rho.view(np.uint64))]), records=np.array(recs, np.uint64))


def _child(args: list[str]) -> dict:
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "case.npz")
        subprocess.run([sys.executable, os.path.abspath(__file__), *args, out], check=True,
                       capture_output=True)
        with np.load(out, allow_pickle=False) as z:
            return {key: z[key] for key in z.files}


def make_batch(workers: int = 4) -> str:
    import concurrent.futures as cf
    count = 512
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(lambda i: _child(["--one-batch", str(i), str(BATCH_T)]),
                            range(count)))
    data = {key: np.stack([p[key] for p in parts]) for key in parts[0]}
    data["num_iter"] = np.uint32(BATCH_T)
    case = f"batch512_T{BATCH_T}"
    np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
    print(case, "graphs", count, flush=True)
    return case


def make_wide() -> list[str]:
    cases = []
    for part, p_in, seed, ks in WIDE_CASES:
        for k in ks:
            case = f"wide_bip{part}_k{k}"
            data = _child(["--one-wide", str(part), repr(p_in), str(seed), str(k)])
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            cases.append(case)
            print(case, "core", int(data["core"]), "visits", int(data["visits"]), flush=True)
    return cases


def make_wide_trim() -> list[str]:
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for part, p_in, seed, k in WIDE_TRIM_CASES:
            name = f"bip{part}"
            pairs = gg.bipartite_plus(part, p_in, seed)
            path = os.path.join(td, name + ".txt")
            gg.write_text(pairs, path)
            data = run_trim_case(td, path, k, name)
            data["pairs_sha"] = np.array([sha(pairs)])
            data["gen"] = np.array([part, seed], np.uint32)
            data["p_in"] = np.float64(p_in)
            data["k"] = np.uint32(k)
            case = f"trim_{name}_k{k}"
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            cases.append(case)
            print(case, "stat", data["stat"].tolist(), flush=True)
    return cases


def make_big(only=None) -> list[str]:
    cases = []
    with tempfile.TemporaryDirectory() as td:
        for wname, seed, sd, k, T in BIG_CASES:
            case = f"big_{wname}_s{seed}_d{sd}_k{k}_T{T}"
            if only and case not in only:
                continue
            pairs = gg.workload(wname, seed=seed, scale_down=sd)
            p = os.path.join(td, case + ".txt")
            gg.write_text(pairs, p)
            data = sim_case_in_child(p, k, T, store_vectors=False)
            os.remove(p)
            data["pairs_sha"] = np.array([sha(pairs)])
            data["workload"] = np.array([wname])
            data["seed"] = np.uint32(seed)
            data["scale_down"] = np.uint32(sd)
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **data)
            cases.append(case)
            print(case, "records", data["records"][:, :3].tolist(), flush=True)
    return cases


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


def _update_index(group: str, cases: list[str]) -> None:
    path = os.path.join(HERE, "index.json")
    index = json.load(open(path)) if os.path.exists(path) else {}
    have = index.get(group, [])
    index[group] = have + [c for c in cases if c not in have]
    with open(path, "w") as f:
        json.dump(index, f, indent=1)


def main(groups: list[str]):
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True,
                   capture_output=True)
    if "small" in groups:
        graphs = dict(small_graphs())
        index = {"sim": [], "trim": []}
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
        _update_index("sim", index["sim"])
        _update_index("trim", index["trim"])
    if "wide" in groups:
        _update_index("wide", make_wide())
        _update_index("trim", make_wide_trim())
    if "big" in groups:
        _update_index("big", make_big())
    if "batch" in groups:
        _update_index("batch", [make_batch()])


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--one":
        a = sys.argv[2:]
        one_sim_case(a[0], int(a[1]), int(a[2]), a[3] == "1", a[4])
    elif len(sys.argv) > 1 and sys.argv[1] == "--one-wide":
        a = sys.argv[2:]
        one_wide_case(int(a[0]), float(a[1]), int(a[2]), int(a[3]), a[4])
    elif len(sys.argv) > 1 and sys.argv[1] == "--one-batch":
        a = sys.argv[2:]
        one_batch_graph(int(a[0]), int(a[1]), a[2])
    elif len(sys.argv) > 1 and sys.argv[1] == "--big-case":
        _update_index("big", make_big(set(sys.argv[2:])))
    else:
        main(sys.argv[1:] or ["small", "wide", "big", "batch"])
