"""Benchmark: k-clique visits/s of one Sim-kClist++ iteration on one MI355X.

Workload (the reference's k=5 case): seeded Chung-Lu graph, n = 1,000,000,
m = 16,000,000, degree exponent 2.3, largest expected degree 20,000, plus a
planted 150-vertex community with internal edge probability 0.8; k = 5.
A step is one iteration t of kCDensestCmpSync.c:396-404: the enumeration
pass over all k-cliques with the per-clique min-rho credit, then the fp64
rho update.  The graph, rho and the counts are resident in HBM when the
timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cl1m]

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "cl1m": ("cl1m", 1, 1, 5,
             "Chung-Lu n=1e6 m=1.6e7 gamma=2.3 w_max=2e4 + 150-vertex community p=0.8, k=5"),
    "er20k": ("er20k", 1, 1, 4, "Erdos-Renyi n=2e4 m=1e5 + planted K30, k=4"),
    "cl1m_d16": ("cl1m", 1, 16, 5, "Chung-Lu cl1m scaled down 16x, k=5 (quick check)"),
    "cl1m_k6": ("cl1m", 1, 1, 6,
                "Chung-Lu n=1e6 m=1.6e7 gamma=2.3 w_max=2e4 + 150-vertex community p=0.8, k=6"),
    "cl4m": ("cl4m", 1, 1, 6,
             "Chung-Lu n=4e6 m=6e7 gamma=2.1 w_max=4e4 + 200-vertex near-clique p=0.9, k=6"),
}


def load_model(args, kcl):
    """The cost model for --schedule learned (the shipped one by default)."""
    if args.schedule != "learned":
        return None
    return kcl.CostModel(path=args.model or os.path.join(kcl.PKG, "cost_model.kclcm"))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(cd: np.ndarray, adj: np.ndarray, roots: np.ndarray) -> int:
    """Bytes one pass must move, per the vertex-rooted formulation: for each
    scheduled root s: its extent and key (16 B), and for each out-neighbour
    x: the id (4 B), its extent and key (16 B) and its adjacency slice
    (4 d+(x)); plus one 8-byte count write per node."""
    dout = np.diff(cd).astype(np.int64)
    n = dout.size
    src = np.repeat(np.arange(n), dout)
    per_edge = 20 + 4 * dout[adj.astype(np.int64)]
    return 16 * int(roots.sum()) + int(per_edge[roots[src]].sum()) + 8 * n


def profiled(workload: str, key: str):
    """A figure from the committed rocprofv3 summary of this workload
    (traffic_bytes_per_pass: FETCH_SIZE + WRITE_SIZE scaled by the
    copy-kernel factors; lds_frac_upper_per_pass), or None."""
    path = os.path.join(ROOT, "profiles", f"pass_kernel_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1].split()[0])
    return None


def cpu_baseline(pairs: np.ndarray, k: int, target_s: float = 12.0):
    """Times the reference's own CDF_CliqueEnum(FRANK_WOLFE) (built from its
    sources into oracle/_ref) on a random sample of root edges at rho = 0,
    with all host threads.  Falls back to the oracle's single-threaded
    restatement when the reference build is absent."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        from oracle import RefHarness
        ref = RefHarness()
    except Exception as e:  # pragma: no cover - reference build missing
        log("reference build unavailable:", e)
        return cpu_baseline_port(pairs, k, target_s)
    ref.threads(threads)
    t0 = time.time()
    h = ref.load_pairs(pairs)  # the reference's SortByCore/Relabel/MakeGraph
    log(f"cpu baseline: reference graph built in {time.time() - t0:.1f}s")
    info = ref.info(h)
    ref.start(h, k)
    e = info["e"]
    rng = np.random.default_rng(12345)
    chunk = 256
    spent, visits, tried = 0.0, 0.0, 0
    perm = rng.permutation(e).astype(np.uint32)
    # Grow the sample (chunks doubling from 256 root edges) until about
    # target_s seconds of reference work.
    while spent < target_s and tried < e:
        take = int(min(e - tried, chunk))
        roots = np.sort(perm[tried:tried + take])
        s, v = ref.time_fw_roots(h, roots)
        spent += s
        visits += v
        tried += take
        log(f"cpu baseline: {tried} roots, {visits:.3e} visits, {spent:.1f}s")
        chunk *= 2
    ref.free(h)
    return {"value": visits / spent if spent > 0 else 0.0, "unit": "k-clique visits/s",
            "cores": threads, "kind": "reference",
            "sample": f"CDF_CliqueEnum(FRANK_WOLFE) of the reference over {tried} of {e} root "
                      f"edges drawn uniformly (seed 12345), rho = 0, {threads} OpenMP threads, "
                      f"{spent:.1f}s"}


def cpu_baseline_port(pairs, k, target_s):  # pragma: no cover - fallback only
    from oracle import Oracle
    o = Oracle()
    g = o.dag(pairs)
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(g.m, size=min(g.m, 20000), replace=False))
    from oracle import Dag
    sub = Dag(g.n_in, g.n, g.core, g.rank, g.es[idx].copy(), g.et[idx].copy(), g.cd, g.adj)
    t0 = time.time()
    _, v = o.fw_counts(sub, k, np.zeros(g.n))
    dt = time.time() - t0
    return {"value": v / dt, "unit": "k-clique visits/s", "cores": 1, "kind": "port",
            "sample": f"oracle restatement over {idx.size} random root edges, 1 thread, {dt:.1f}s"}


_REF_GRAPH_CHILD = """
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import RefHarness
ref = RefHarness()
ref.threads(int(sys.argv[4]))
h = ref.load_pairs(np.load(sys.argv[2]))
ref.start(h, int(sys.argv[3]))
e = ref.info(h)["e"]
s, v = ref.time_fw_roots(h, np.arange(e, dtype=np.uint32))
print(json.dumps({"seconds": s, "visits": v, "edges": e}))
"""


def cpu_baseline_batch(pairs_list, ks, target_s: float = 12.0, wall_s: float = 90.0):
    """The reference's CDF_CliqueEnum(FRANK_WOLFE) over whole graphs of the
    batch (seeded order, all root edges, rho = 0), one child process per
    graph (the reference sizes its per-thread scratch by the first graph it
    sees), until about target_s seconds of reference work or wall_s seconds
    of wall clock (loading each graph dominates for these sparse graphs)."""
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    order = np.random.default_rng(12345).permutation(len(pairs_list))
    spent, visits, graphs = 0.0, 0.0, 0
    w0 = time.time()
    with tempfile.TemporaryDirectory() as td:
        for i in order:
            if spent >= target_s or time.time() - w0 > wall_s:
                break
            path = os.path.join(td, "g.npy")
            np.save(path, np.ascontiguousarray(pairs_list[i], np.uint32))
            r = subprocess.run([sys.executable, "-c", _REF_GRAPH_CHILD, ROOT, path, str(ks[i]),
                                str(threads)], capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                log(f"cpu baseline: reference failed on graph {i}: {r.stderr[-300:]}")
                return None
            res = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])
            spent += res["seconds"]
            visits += res["visits"]
            graphs += 1
            log(f"cpu baseline: graph {i} (k={ks[i]}, {res['edges']} edges): "
                f"{res['visits']:.3e} visits in {res['seconds']:.2f}s")
    return {"value": visits / spent if spent > 0 else 0.0, "unit": "k-clique visits/s",
            "cores": threads, "kind": "reference",
            "sample": f"CDF_CliqueEnum(FRANK_WOLFE) of the reference over every root edge of "
                      f"{graphs} of the {len(pairs_list)} graphs (seeded order), rho = 0, "
                      f"{threads} OpenMP threads, {spent:.1f}s"}


def run_batch(args, kcl, gg):
    """Mixed batch: 512 Chung-Lu graphs (n log-uniform in [1e3, 2e5], average
    degree 4..40, gamma 2.1..2.8, k = 3 + (i mod 5)); a step is one
    iteration of every graph (one hipGraph replay)."""
    specs = list(gg.batch_specs(512, first_seed=1000))
    t0 = time.time()
    pairs_list, graphs = [], []
    for i, sp in enumerate(specs):
        p = gg.batch_graph(sp)
        pairs_list.append(p)
        graphs.append(kcl.Graph(p))
        if i % 64 == 63:
            log(f"bench batch: {i + 1} graphs generated and built ({time.time() - t0:.0f}s)")
    ks = [sp["k"] for sp in specs]
    model = load_model(args, kcl)
    t1 = time.time()
    b = kcl.Batch(graphs, ks, schedule=kcl.SCHEDULES[args.schedule], model=model,
                  tune=not args.no_tune)
    info = b.info()
    log(f"bench batch: set-up {time.time() - t1:.1f}s: {info}")
    for g in graphs:
        g.close()
    t = 0
    for _ in range(args.warmup):
        t += 1
        b.step(t)
        log(f"bench batch: warmup t={t} {b.last_pass()}")
    visits = 0
    kern_ms = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t += 1
        b.step(t)
        v, ms = b.last_pass()
        if v != info["cliques"]:
            raise RuntimeError(f"step {t} visited {v} cliques, the batch has {info['cliques']}")
        visits += v
        kern_ms += ms
    elapsed = time.perf_counter() - t_start
    log(f"bench batch: {args.steps} steps in {elapsed:.3f}s")
    # The C4 run itself: CDF_Main with T = 50 for every graph (32 iterations,
    # reports at T = 1, 2, 4, ..., 32), from rho = 0.
    b.reset()
    r0 = time.perf_counter()
    recs = b.run(50)
    run_s = time.perf_counter() - r0
    log(f"bench batch: run(50) of all graphs in {run_s:.2f}s, "
        f"{sum(len(r) for r in recs)} records")
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(pairs_list, ks)
    b.close()
    if model is not None:
        model.close()
    out = {
        "metric": "k-clique visits/sec per kClist++ iteration",
        "value": visits / elapsed if elapsed > 0 else 0.0,
        "unit": "k-clique visits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
        "config": {"workload": "batch512: 512 Chung-Lu graphs, n log-uniform [1e3, 2e5], avg "
                               "degree U[4, 40], gamma U[2.1, 2.8], k = 3 + (i mod 5)",
                   "graphs": 512, "nodes": info["nodes"], "edges": info["edges"],
                   "cliques_per_step": info["cliques"], "schedule": args.schedule,
                   "row_bits": info["row_bits"], "per_cu": info["per_cu"],
                   "tasks": info["tasks"], "plan_ms": info["plan_ms"],
                   "tune_ms": info["tune_ms"], "gpu_ms_per_step": kern_ms / max(args.steps, 1),
                   "run50_s": run_s,
                   "parallelism": "1 GPU"},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_exact(args, kcl, gg):
    """Time-to-exact: binary edge file -> graph -> certified maximal densest
    subgraph (kcl_exact_run), wall clock."""
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    log(f"bench exact: generating {args.workload} ({desc})")
    pairs = gg.workload(wname, seed=seed, scale_down=sd)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.u32")
        kcl.write_edge_list(path, pairs)
        del pairs
        t0 = time.perf_counter()
        pr, n_in = kcl.read_edge_list(path)
        t1 = time.perf_counter()
        g = kcl.Graph(pr, n_in)
        t2 = time.perf_counter()
        log(f"bench exact: read {t1 - t0:.2f}s, graph n={g.n} m={g.m} core={g.core} "
            f"{t2 - t1:.2f}s; running the exact pipeline")
        model = load_model(args, kcl)
        res, nodes = kcl.exact(g, k, stat_path=os.path.join(td, "stat_exact.txt"),
                               max_seconds=args.exact_budget,
                               schedule=kcl.SCHEDULES[args.schedule], model=model)
        if model is not None:
            model.close()
        t3 = time.perf_counter()
        stat = open(os.path.join(td, "stat_exact.txt")).read()
    log(stat)
    out = {
        "metric": f"time-to-exact k-CDS (k={k})",
        "value": t3 - t0,
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
        "config": {"workload": f"{args.workload}: {desc}", "k": k, "n": g.n, "m": g.m,
                   "schedule": args.schedule,
                   "core": g.core},
        "phases_s": {"read": t1 - t0, "graph": t2 - t1, "exact": t3 - t2},
        "result": res,
        "cpu_baseline": None,
        "cpu_baseline_note": "the reference's exact program (kCDensestMem.c) stores every "
                             "clique and needs Boost.Graph, which is absent; not timed",
    }
    g.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cl1m", choices=sorted(WORKLOADS) + ["batch512"])
    ap.add_argument("--no-tune", action="store_true",
                    help="batch512: keep the default row width and occupancy")
    ap.add_argument("--schedule", default="counted", choices=["counted", "learned", "degree"],
                    help="root order/splitting: counting-pass costs, the learned cost "
                         "model, or plain d+^2 (counts are identical under all three)")
    ap.add_argument("--model", default=None,
                    help="learned schedule: cost model file (default: the shipped one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-budget", type=int, default=240,
                    help="exact mode: seconds before giving up on a certificate")
    ap.add_argument("--mode", default="sim", choices=["sim", "exact"],
                    help="sim: visits/s of one iteration (default); exact: time-to-exact")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)

    import kclistpp_amd as kcl
    from kclistpp_amd import graphgen as gg

    if args.mode == "exact":
        run_exact(args, kcl, gg)
        return
    if args.workload == "batch512":
        run_batch(args, kcl, gg)
        return
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    log(f"bench: generating {args.workload} ({desc})")
    t0 = time.time()
    pairs = gg.workload(wname, seed=seed, scale_down=sd)
    log(f"bench: {pairs.shape[0]} edges in {time.time() - t0:.1f}s; building the graph")
    t0 = time.time()
    g = kcl.Graph(pairs)
    log(f"bench: graph n={g.n} m={g.m} core={g.core} built in {time.time() - t0:.1f}s")
    cd, adj = g.csr()
    sched = kcl.SCHEDULES[args.schedule]
    model = load_model(args, kcl)
    t0 = time.time()
    sim = kcl.Sim(g, k, schedule=sched, model=model)
    plan = sim.plan_info()
    alg_bytes = algorithmic_bytes(cd, adj, sim.root_cliques() > 0)
    log(f"bench: plan ready in {time.time() - t0:.1f}s: {plan}; warmup {args.warmup}")
    t = 0
    for _ in range(args.warmup):
        t += 1
        sim.step(t)
        v, kms, pms = sim.last_pass()
        log(f"bench: warmup t={t} visits={v} pass={pms:.2f}ms kernels={kms:.2f}ms")
    visits, kern_ms = 0, 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t += 1
        sim.step(t)  # synchronous: returns after the pass and update completed
        v, kms, pms = sim.last_pass()
        if v != plan["cliques"]:
            raise RuntimeError(f"step {t} visited {v} cliques, the graph has {plan['cliques']}")
        visits += v
        kern_ms += kms
    elapsed = time.perf_counter() - t_start
    log(f"bench: {args.steps} steps in {elapsed:.3f}s, {visits} visits")
    ms_per_step = 1000.0 * elapsed / max(args.steps, 1)
    value = visits / elapsed if elapsed > 0 else 0.0
    kern_avg_ms = kern_ms / max(args.steps, 1)
    achieved = alg_bytes / (kern_avg_ms * 1e-3) / 1e9 if kern_avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": (lambda v: int(v) if v is not None else None)(
                    profiled(args.workload, "traffic_bytes_per_pass")),
                "lds_frac_upper": profiled(args.workload, "lds_frac_upper_per_pass"),
                "kernel": "pass_kernel (all word classes of one enumeration pass)",
                "kernel_ms": round(kern_avg_ms, 4), "algorithmic_bytes": alg_bytes}
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the reference on the host cores")
        cpu = cpu_baseline(pairs, k)
    sim.close()
    if model is not None:
        model.close()
    g.close()
    out = {
        "metric": "k-clique visits/sec per kClist++ iteration",
        "value": value,
        "unit": "k-clique visits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
        "config": {"workload": f"{args.workload}: {desc}", "k": k, "n": g.n, "m": g.m,
                   "core": g.core, "roots": plan["roots_kept"], "visits_per_step": visits // max(args.steps, 1),
                   "schedule": args.schedule, "parallelism": "1 GPU",
                   "plan": plan},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
