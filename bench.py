"""Benchmark: k-clique visits/s of one Sim-kClist++ iteration on one MI355X.

Default workload (the k = 6 case of the exact pipeline): seeded Chung-Lu
graph, n = 4,000,000, m = 60,000,000, degree exponent 2.1, largest expected
degree 40,000, plus a planted 200-vertex near-clique with internal edge
probability 0.9; k = 6, roots ordered and split by the learned cost model.
A step is one iteration t of kCDensestCmpSync.c:396-404: the enumeration
pass over all k-cliques with the per-clique min-rho credit, then the fp64
rho update.  The graph, rho and the counts are resident in HBM when the
timed region starts.  The same run times the plain (d+^2) schedule on the
same states, and a short leg of the k = 5 workload (Chung-Lu 1M/16M).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cl4m]

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
    "rmat15": ("rmat23", 1, 256, 7,
               "R-MAT (0.57, 0.19, 0.19, 0.05) scale 15, m=2e8/256 before dedup, k=7"),
    "rmat16": ("rmat23", 1, 128, 7,
               "R-MAT (0.57, 0.19, 0.19, 0.05) scale 16, m=2e8/128 before dedup, k=7"),
    "rmat21": ("rmat23", 1, 4, 7,
               "R-MAT (0.57, 0.19, 0.19, 0.05) scale 21, m=2e8/4 before dedup, k=7"),
    "rmat23": ("rmat23", 1, 1, 7,
               "R-MAT (0.57, 0.19, 0.19, 0.05) scale 23, m=2e8 before dedup, k=7"),
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


def profiled(workload: str, key: str, text: bool = False):
    """A figure from the committed rocprofv3 summary of this workload
    (profiles/pass_kernel_<workload>.md, lines "key: value ..."), or None:
    traffic_bytes_per_pass (FETCH_SIZE + WRITE_SIZE scaled by the copy
    kernel's factors), valu_busy_frac, lds_busy_frac, lds_bank_conflict_frac,
    hbm_traffic_frac, bound."""
    path = os.path.join(ROOT, "profiles", f"pass_kernel_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            v = line.split(":", 1)[1].split()[0]
            return v if text else float(v)
    return None


def run_reference(cmd: list, cwd: str, timeout_s: float):
    """Runs a reference program with its output on a pseudo-terminal (so its
    stdio is line-buffered) and passes each line on as it is printed.
    Returns (finished, seconds, lines)."""
    import pty
    import select
    import subprocess
    master, slave = pty.openpty()
    c0 = time.perf_counter()
    proc = subprocess.Popen(cmd, cwd=cwd, stdout=slave, stderr=slave, close_fds=True)
    os.close(slave)
    buf, lines, killed = b"", [], False
    while True:
        ready, _, _ = select.select([master], [], [], 1.0)
        if ready:
            try:
                data = os.read(master, 65536)
            except OSError:  # the child closed its side
                data = b""
            if not data:
                break
            buf += data
            while b"\n" in buf:
                line, buf = buf.split(b"\n", 1)
                text = line.decode(errors="replace").rstrip()
                lines.append(text)
                log("reference: " + text)
        elif proc.poll() is not None:
            break
        if time.perf_counter() - c0 > timeout_s:
            proc.kill()
            killed = True
            break
    rc = proc.wait()
    os.close(master)
    return (not killed and rc == 0), time.perf_counter() - c0, lines


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(cd: np.ndarray, adj: np.ndarray, root_cliques: np.ndarray, k: int,
                 budget_s: float = 30.0):
    """Whole-pass time of the reference's own CDF_CliqueEnum(FRANK_WOLFE)
    (built from its sources into oracle/_ref) on all host threads, estimated
    from a bounded stratified sample of its root edges at rho = 0.

    The reference runs one task per DAG edge (s, t) under a dynamic OpenMP
    schedule (kCDensestCmpSync.c:300-305).  Edges are grouped by their
    source's k-clique count (exact, from the GPU plan): stratum 0 holds the
    edges of roots without a k-clique, stratum b >= 1 those of roots with
    [2^(b-1), 2^b) cliques.  Each stratum is sampled uniformly, in batches
    sized for about half a second of all threads and doubling within the
    stratum's share of the budget (a small stratum is run whole).  The
    sample's cost is taken in thread-seconds: a large batch runs on all
    threads (wall x threads), a small one edge by edge, one edge per thread
    under its own clock.  A stratum's time is its thread-seconds scaled by
    its edge count and divided by the thread count (ideal scaling, which
    favours the CPU); value = the pass's k-cliques / the summed estimate."""
    threads = host_threads()
    from oracle import RefHarness
    ref = RefHarness()
    ref.threads(threads)
    t0 = time.time()
    h = ref.load_csr(cd, adj)  # the same DAG as the reference's MakeGraph (tests check it)
    ref.start(h, k)
    log(f"cpu baseline: reference graph from the CSR in {time.time() - t0:.1f}s")
    n = cd.size - 1
    dout = np.diff(cd).astype(np.int64)
    rc = root_cliques[:n].astype(np.float64)
    stratum = np.zeros(n, np.int64)
    nz = rc > 0
    stratum[nz] = 1 + np.floor(np.log2(rc[nz])).astype(np.int64)
    levels = [b for b in np.unique(stratum) if dout[stratum == b].sum() > 0]
    share = budget_s / max(len(levels), 1)
    rng = np.random.default_rng(12345)
    total_s, sampled_all, wall_all = 0.0, 0, 0.0
    per_thread_rate = 4e7  # visits/s, only to size the first batch
    big_batch = 64 * threads  # from here on, batches run on all threads at once
    for b in levels:
        roots = np.nonzero(stratum == b)[0]
        nb = int(dout[roots].sum())
        vb = float(rc[roots].sum())
        csum = np.cumsum(dout[roots])
        batch = int(min(max(per_thread_rate * threads * 0.5 / max(vb / nb, 1.0), threads),
                        1 << 22))
        thread_s, wall, taken = 0.0, 0.0, 0
        while wall < share and taken < nb:
            if nb <= batch and taken == 0:  # the whole stratum
                idx = np.nonzero(np.repeat(stratum == b, dout))[0] if roots.size > 50000 else \
                    np.concatenate([np.arange(cd[r], cd[r + 1]) for r in roots])
            else:
                pos = rng.integers(0, nb, min(batch, nb))
                ri = np.searchsorted(csum, pos, side="right")
                idx = cd[roots[ri]] + (pos - (csum[ri] - dout[roots[ri]]))
            idx = np.sort(idx).astype(np.uint32)
            if idx.size >= big_batch:
                sec, _ = ref.time_fw_roots(h, idx)
                thread_s += sec * threads
            else:
                sec, per, _ = ref.time_fw_each(h, idx)
                thread_s += float(per.sum())
            wall += sec
            taken += idx.size
            batch *= 2
            if idx.size == nb and taken == nb:
                break
        est = thread_s * nb / taken / threads if taken else 0.0
        total_s += est
        sampled_all += taken
        wall_all += wall
        log(f"cpu baseline: stratum {b}: {roots.size} roots, {nb} edges, {vb:.3e} cliques; "
            f"sampled {taken} edges, {thread_s:.2f} thread-s in {wall:.2f}s -> {est:.3e}s")
    ref.free(h)
    total_v = float(rc.sum())
    model = cpu_model()
    return {"value": total_v / total_s if total_s > 0 else 0.0, "unit": "k-clique visits/s",
            "cores": threads, "kind": "reference", "cpu_model": model,
            "pass_seconds_estimate": total_s,
            "sample": f"CDF_CliqueEnum(FRANK_WOLFE) of the reference (built from its sources) "
                      f"at rho = 0, {threads} OpenMP threads of a {model}: {sampled_all} of "
                      f"{int(dout.sum())} root edges in {len(levels)} strata by their root's "
                      f"k-clique count (log2), {wall_all:.1f}s of sampling; per stratum, "
                      f"sampled thread-seconds x edges / sampled edges / threads; whole-pass "
                      f"estimate {total_s:.4g}s for {total_v:.4g} cliques"}


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


def cpu_baseline_exact(pairs_file, res: dict, k: int, workload: str, gg) -> dict:
    """Time-to-exact of the same storage-free pipeline with the reference's
    own CPU passes: the reference's reader and graph build (ReadEdgeList,
    SortByCore, Relabel, MakeGraph on the same text file), then as many
    enumeration passes as the GPU run made (one Frank-Wolfe pass per
    iteration, one PAVA pass per round, the region tests' k-clique degree
    passes), each priced at a measured CDF_CliqueEnum(FRANK_WOLFE) pass of
    the reference over every root edge on all host threads, plus the same
    host certificate time (clique listing and max-flow run on the host in
    both).  The reference's own exact program cannot run here: it stores
    every clique and needs Boost.Graph, which is absent."""
    from oracle import RefHarness
    threads = host_threads()
    wname, seed, sd, _, _ = WORKLOADS[workload]
    pairs = gg.workload(wname, seed=seed, scale_down=sd)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.txt")
        gg.write_text(pairs, path)
        del pairs
        ref = RefHarness()
        ref.threads(threads)
        t0 = time.perf_counter()
        h = ref.load(path)
        load_s = time.perf_counter() - t0
    ref.start(h, k)
    e = ref.info(h)["e"]
    pass_s, visits = ref.time_fw_roots(h, np.arange(e, dtype=np.uint32))
    ref.free(h)
    rounds = int(res["num_iter"]).bit_length()  # T = 1, 2, 4, ..., num_iter
    passes = int(res["num_iter"]) + rounds + int(res["degree_passes"])
    total = load_s + passes * pass_s + float(res["host_seconds"])
    log(f"cpu baseline: reference load {load_s:.2f}s, pass {pass_s:.3f}s ({visits:.4g} "
        f"cliques), {passes} passes, host certificate {res['host_seconds']:.2f}s")
    return {"value": total, "unit": "s", "cores": threads, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": f"reference ReadEdgeList+SortByCore+MakeGraph {load_s:.2f}s + {passes} "
                      f"passes x {pass_s:.3f}s (its CDF_CliqueEnum(FRANK_WOLFE) over all "
                      f"{e} root edges, {threads} OpenMP threads of a {cpu_model()}) + "
                      f"{float(res['host_seconds']):.2f}s host certificate"}


def run_trim(args, kcl, gg):
    """k-clique core trimming (kClistCoreTrim) of a workload: (k-1)-core,
    k-clique degrees on the GPU, the peel with heavy pops on the GPU; wall
    clock from the edges in memory to the trimmed set.  cpu_baseline: the
    reference's own kClistCoreTrim program (built from its sources) on the
    same text file, single-threaded as it is written, with a time limit."""
    import subprocess
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    pairs = workload_pairs(args, gg, wname, seed, sd)
    log(f"bench trim: {args.workload} ({desc}), {pairs.shape[0]} edges")
    t0 = time.perf_counter()
    with kcl.Trim(pairs, k, gpu_peel_min_cliques=args.gpu_peel_min) as t:
        elapsed = time.perf_counter() - t0
        info = dict(t.info)
    log(f"bench trim: {elapsed:.2f}s {info}")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import REF_TRIM_BIN
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "edges.txt")
            gg.write_text(pairs, path)
            c0 = time.perf_counter()
            try:
                r = subprocess.run([REF_TRIM_BIN, str(k), path, "d.txt", "c.txt", "e.txt", "b"],
                                   cwd=td, capture_output=True, text=True,
                                   timeout=args.cpu_timeout)
                ok = r.returncode == 0
            except subprocess.TimeoutExpired:
                ok = False
            cs = time.perf_counter() - c0
        cpu = {"value": cs, "unit": "s", "cores": 1, "kind": "reference", "cpu_model": cpu_model(),
               "sample": f"the reference kClistCoreTrim (one thread, as written) on the same edge "
                         f"file: {'finished' if ok else 'stopped at the time limit'} after "
                         f"{cs:.1f}s" + ("" if ok else " (a lower bound)")}
        log(f"bench trim: reference {cs:.1f}s ({'finished' if ok else 'time limit'})")
    out = {"metric": f"k-clique core trimming time (k={k})", "value": elapsed, "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "u64",
           "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
           "config": {"workload": f"{args.workload}: {desc}", "k": k}, "result": info,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def makespan_of(sim, clock_hz: float, cus: int) -> dict:
    """One profiled pass of sim's plan at its current rho: per-task
    shader-clock cycles priced against the GPU's resident workgroups.

    A class-c workgroup holds 1/per_cu_c of a CU for its task's cycles, so
    the pass needs at least sum_t cycles_t / per_cu_c CU-cycles; spread over
    every CU at the peak shader clock that is the bound below which no order
    or packing of these tasks can finish (the timers count shader cycles,
    whose rate is at most the peak, so the bound errs low)."""
    prof = sim.profile_pass()
    cyc = prof["total_cycles"].astype(np.float64)
    cu_cycles, classes = 0.0, []
    for c in sim.plan_classes():
        seg = cyc[c["first"]:c["first"] + c["count"]]
        work = float(seg.sum())
        cu_cycles += work / c["per_cu"]
        classes.append({"words": c["words"], "k": c["k"], "edge": c["edge"], "tasks": c["count"],
                        "waves": c["waves"], "per_cu": c["per_cu"], "blocks": c["blocks"],
                        "work_cu_ms": round(work / c["per_cu"] / clock_hz * 1e3, 3),
                        "alone_lb_ms": round(work / max(c["blocks"], 1) / clock_hz * 1e3, 3),
                        "max_task_ms": round(float(seg.max()) / clock_hz * 1e3, 3)})
    return {"tasks": int(cyc.size), "lb_ms": cu_cycles / cus / clock_hz * 1e3,
            "max_task_ms": float(cyc.max()) / clock_hz * 1e3 if cyc.size else 0.0,
            "build_share": float(prof["build_cycles"].sum()) / max(float(cyc.sum()), 1.0),
            "classes": classes}


def run_makespan(args, kcl, gg):
    """Pass time of each schedule against a makespan lower bound, on the
    same rho state (after --warmup iterations of the learned schedule).
    Each schedule's pass is timed --steps times from that state, then one
    profiled pass gives its per-task cycles (makespan_of).  The smallest
    bound over the schedules caps what any order or split of the same work
    could gain over the plain schedule: plain pass / that bound."""
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    pairs = workload_pairs(args, gg, wname, seed, sd)
    dev = kcl.device_info()
    clock_hz, cus = dev["clock_khz"] * 1e3, dev["cu_count"]
    g = kcl.Graph(pairs)
    del pairs
    log(f"bench makespan: {args.workload} n={g.n} m={g.m} core={g.core}; clock "
        f"{clock_hz / 1e6:.0f} MHz, {cus} CUs")
    model = kcl.CostModel(path=args.model or os.path.join(kcl.PKG, "cost_model.kclcm"))
    main = kcl.Sim(g, k, schedule=kcl.SCHED_LEARNED, model=model)
    t = 0
    for _ in range(args.warmup):
        t += 1
        main.step(t)
    rho = main.rho()
    t += 1
    out = {}
    retrained = None
    for name in ("degree", "counted", "learned", "online"):
        if name == "online":
            # The shipped model fine-tuned on this graph's own pass timings
            # (the learned plan's profiled pass at this state), then replanned.
            from kclistpp_amd import train_cost_model as tcm
            main.set_rho(rho)
            retrained = tcm.online(main, model, epochs=10)
        s = main if name in ("learned", "online") else kcl.Sim(g, k, schedule=kcl.SCHEDULES[name])
        ms, counts = [], None
        for _ in range(max(args.steps, 1)):
            s.set_rho(rho)
            s.step(t)
            ms.append(s.last_pass()[2])
            c = s.counts()
            if counts is not None and not np.array_equal(c, counts):
                raise RuntimeError(f"{name}: counts differ between repeated passes")
            counts = c
        if name == "degree":
            ref_counts = counts
        elif not np.array_equal(counts, ref_counts):
            raise RuntimeError(f"{name}: counts differ from the plain schedule's")
        s.set_rho(rho)
        m = makespan_of(s, clock_hz, cus)
        m["pass_ms"] = float(np.median(ms))
        m["pass_over_lb"] = m["pass_ms"] / m["lb_ms"] if m["lb_ms"] > 0 else None
        out[name] = m
        log(f"bench makespan: {name}: pass {m['pass_ms']:.1f} ms (runs {ms}), bound "
            f"{m['lb_ms']:.1f} ms, longest task {m['max_task_ms']:.1f} ms, {m['tasks']} tasks")
        if s is not main:
            s.close()
    main.close()
    model.close()
    if retrained is not None:
        retrained.close()
    lb = min(v["lb_ms"] for v in out.values())
    res = {"metric": "pass makespan vs lower bound, per schedule", "unit": "ms",
           "workload": f"{args.workload}: {desc}", "k": k, "t": t, "clock_mhz": clock_hz / 1e6,
           "cu_count": cus, "schedules": out, "lb_min_ms": lb,
           "plain_over_lb_min": out["degree"]["pass_ms"] / lb if lb > 0 else None,
           "learned_vs_plain": out["degree"]["pass_ms"] / out["learned"]["pass_ms"],
           "online_vs_plain": out["degree"]["pass_ms"] / out["online"]["pass_ms"]}
    g.close()
    print(json.dumps(res), flush=True)


def run_sampling(args, kcl, gg):
    """Sampling++ (kCDensestSamplingKclistpp): count, keep ~1e7 cliques,
    T = --steps rounds with reports after 1, 2, 4, ...; wall clock from the
    edges in memory to the last report.  cpu_baseline: the reference program
    (built from its sources) on the same text file, with a time limit, on
    one thread: its clique counter (kCDensestSamplingKclistpp.c:156, 194) is
    a shared global incremented inside the OpenMP region without a
    reduction, so with several threads it loses counts (16 threads on the
    1M/16M graph printed 1,546,586,300 of 5,935,593,639 5-cliques) and the
    sampling probability it derives from that count is wrong."""
    import subprocess
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    pairs = workload_pairs(args, gg, wname, seed, sd)
    T = args.steps
    log(f"bench sampling: {args.workload} ({desc}), {pairs.shape[0]} edges, T = {T}")
    t0 = time.perf_counter()
    with kcl.Graph(pairs) as g:
        t1 = time.perf_counter()
        with kcl.Sampling(g, k, seed=1) as smp:
            t2 = time.perf_counter()
            recs = smp.run(T)
            info = smp.info
        t3 = time.perf_counter()
    log(f"bench sampling: graph {t1 - t0:.2f}s, count+sample {t2 - t1:.2f}s, rounds+reports "
        f"{t3 - t2:.2f}s; {info}; last {recs[-1]}")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import REF_SAMPLING_BIN
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "edges.txt")
            gg.write_text(pairs, path)
            ok, cs, out = run_reference([REF_SAMPLING_BIN, "1", str(T), str(k), path], td,
                                        args.cpu_timeout)
            lines = [x for x in out if x.startswith("Approximate")]
            counted = [x for x in out if x.startswith(f"Number of {k}-cliques")]
        cpu = {"value": cs, "unit": "s", "cores": 1, "kind": "reference",
               "cpu_model": cpu_model(),
               "sample": f"the reference kCDensestSamplingKclistpp on one thread (its clique "
                         f"counter races under OpenMP) on the same edge file, T = {T}: "
                         f"{'finished' if ok else 'stopped at the time limit'} after {cs:.1f}s"
                         + ("" if ok else " (a lower bound)"),
               "count_line": counted[-1] if counted else None,
               "last_line": lines[-1] if lines else None}
        log(f"bench sampling: reference {cs:.1f}s ({'finished' if ok else 'time limit'})")
    out = {"metric": f"Sampling++ time (k={k}, T={T})", "value": t3 - t0, "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "u64",
           "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
           "config": {"workload": f"{args.workload}: {desc}", "k": k, "T": T},
           "phases_s": {"graph": t1 - t0, "count_sample": t2 - t1, "rounds_reports": t3 - t2},
           "result": {"info": info, "records": recs}, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def run_cmpsync(args, kcl, gg):
    """kCDensestCmpSync end to end: the text edge file -> reader -> graph ->
    plan -> T = --steps iterations (rounded down to a power of two) with a
    report after 1, 2, 4, ... -> the stat file; wall clock over all of it.
    cpu_baseline: the reference program on the same file and host threads
    when it finishes within --cpu-timeout, else (kind "reference") its
    measured load time plus the run's passes priced at one measured
    CDF_CliqueEnum(FRANK_WOLFE) pass of the reference on all threads."""
    import subprocess
    from oracle import REF_CMPSYNC_BIN, RefHarness
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    T = args.steps
    pairs = workload_pairs(args, gg, wname, seed, sd)
    model = load_model(args, kcl)
    kcl.device_info()  # the HIP runtime starts here, outside the timed region
    # One untimed run on a 6-clique: the first use of the library's kernels in
    # a process loads their code objects (~0.15 s), a once-per-process cost.
    six = np.array([(a, b) for a in range(6) for b in range(a + 1, 6)], np.uint32)
    with tempfile.TemporaryDirectory() as td, kcl.Graph(six) as g0:
        kcl.cmpsync(g0, min(k, 6), 2, stat_path=os.path.join(td, "stat.txt"))
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.txt")
        gg.write_text(pairs, path)
        del pairs
        t0 = time.perf_counter()
        pr, n_in = kcl.read_edge_list(path)
        t1 = time.perf_counter()
        with kcl.Graph(pr, n_in) as g:
            del pr
            t2 = time.perf_counter()
            recs = kcl.cmpsync(g, k, T, stat_path=os.path.join(td, "stat.txt"),
                               schedule=kcl.SCHEDULES[args.schedule], model=model)
            t3 = time.perf_counter()
        log(f"bench cmpsync: read {t1 - t0:.2f}s, graph {t2 - t1:.2f}s, run {t3 - t2:.2f}s; "
            f"last record {recs[-1]}")
        cpu = None
        if not args.no_cpu_baseline:
            threads = host_threads()
            ok, cs, _ = run_reference([REF_CMPSYNC_BIN, str(threads), str(T), str(k), path, "b"],
                                      td, args.cpu_timeout)
            if ok:
                cpu = {"value": cs, "unit": "s", "cores": threads, "kind": "reference",
                       "cpu_model": cpu_model(),
                       "sample": f"the reference kCDensestCmpSync ({threads} OpenMP threads) on "
                                 f"the same edge file, T = {T}: {cs:.2f}s"}
            else:
                ref = RefHarness()
                ref.threads(threads)
                l0 = time.perf_counter()
                h = ref.load(path)
                load_s = time.perf_counter() - l0
                ref.start(h, k)
                e = ref.info(h)["e"]
                pass_s, _ = ref.time_fw_roots(h, np.arange(e, dtype=np.uint32))
                ref.free(h)
                iters = 1 << (T.bit_length() - 1)
                passes = iters + iters.bit_length()  # FW passes, one PAVA pass per report
                est = load_s + passes * pass_s
                cpu = {"value": est, "unit": "s", "cores": threads, "kind": "reference",
                       "cpu_model": cpu_model(),
                       "sample": f"the reference kCDensestCmpSync did not finish within "
                                 f"{args.cpu_timeout:.0f}s; estimate = its ReadEdgeList+SortByCore"
                                 f"+MakeGraph {load_s:.2f}s + {passes} passes x {pass_s:.3f}s "
                                 f"(one CDF_CliqueEnum(FRANK_WOLFE) pass over all {e} root "
                                 f"edges, {threads} threads)"}
            log(f"bench cmpsync: reference {cpu['value']:.2f}s ({cpu['sample']})")
    if model is not None:
        model.close()
    out = {"metric": f"kCDensestCmpSync end to end (k={k}, T={T})", "value": t3 - t0,
           "unit": "s", "n_gpus": 1, "higher_is_better": False, "vs_baseline": None,
           "dtype": "u64", "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
           "config": {"workload": f"{args.workload}: {desc}", "k": k, "T": T,
                      "schedule": args.schedule},
           "phases_s": {"read": t1 - t0, "graph": t2 - t1, "plan_iterations_reports": t3 - t2},
           "note": "the library's kernels were loaded by an untimed run on a 6-clique",
           "records": recs, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def run_input(args, kcl, gg):
    """The input path at scale: text edge file -> reader -> degeneracy order
    -> relabel -> CSR (kcl_read_edge_list + kcl_graph_create), wall clock.
    cpu_baseline: the reference's own ReadEdgeList + SortByCore + Relabel +
    MakeGraph (Graph.c:34-156, built from its sources, single-threaded as
    written) on the same file; its ranks and DAG edges are compared with
    this build's bit for bit."""
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    log(f"bench input: generating {args.workload} ({desc})")
    t = time.time()
    pairs = workload_pairs(args, gg, wname, seed, sd)
    log(f"bench input: {pairs.shape[0]} edges in {time.time() - t:.1f}s")
    six = np.array([(a, b) for a in range(6) for b in range(a + 1, 6)], np.uint32)
    with kcl.Graph(six):  # loads the library's kernels outside the timed region
        pass
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.txt")
        gg.write_text(pairs, path, progress=lambda i: log(f"bench input: {i} lines written"))
        m_in = int(pairs.shape[0])
        del pairs
        log(f"bench input: text file {os.path.getsize(path) / 1e6:.0f} MB")
        t0 = time.perf_counter()
        pr, n_in = kcl.read_edge_list(path)
        t1 = time.perf_counter()
        g = kcl.Graph(pr, n_in)
        t2 = time.perf_counter()
        del pr
        info = {"n_in": n_in, "n": g.n, "m": g.m, "core": g.core}
        log(f"bench input: read {t1 - t0:.2f}s, graph {t2 - t1:.2f}s, {info}")
        cpu, parity = None, None
        if not args.no_cpu_baseline:
            from oracle import RefHarness
            ref = RefHarness()
            l0 = time.perf_counter()
            h = ref.load(path)
            load_s = time.perf_counter() - l0
            log(f"bench input: reference load {load_s:.2f}s")
            parity = {"rank_equal": bool(np.array_equal(g.rank(), ref.rank(h))),
                      "edges_equal": bool(np.array_equal(g.edges(), ref.edges(h)))}
            ref.free(h)
            cpu = {"value": load_s, "unit": "s", "cores": 1, "kind": "reference",
                   "cpu_model": cpu_model(),
                   "sample": "the reference's ReadEdgeList + SortByCore + Relabel + MakeGraph "
                             "(built from its sources, single-threaded as written) on the same "
                             f"text file: {load_s:.2f}s"}
        g.close()
    out = {"metric": "input path time (read + degeneracy order + relabel + CSR)",
           "value": t2 - t0, "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (seeded generator, kclistpp_amd/graphgen.py)",
           "config": {"workload": f"{args.workload}: {desc}", "edges_in_file": m_in, **info},
           "phases_s": {"read": t1 - t0, "graph": t2 - t1}, "parity": parity,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def run_exact(args, kcl, gg):
    """Time-to-exact: text edge file -> graph -> certified maximal densest
    subgraph (kcl_exact_run), wall clock."""
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    log(f"bench exact: generating {args.workload} ({desc})")
    pairs = gg.workload(wname, seed=seed, scale_down=sd)
    model = load_model(args, kcl)
    # One untimed run on a 6-clique: the first use of the library's kernels in
    # a process loads their code objects, a once-per-process cost.
    six = np.array([(a, b) for a in range(6) for b in range(a + 1, 6)], np.uint32)
    with tempfile.TemporaryDirectory() as td, kcl.Graph(six) as g0:
        kcl.exact(g0, min(k, 6), stat_path=os.path.join(td, "stat.txt"))
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.txt")  # the reference's text format
        gg.write_text(pairs, path)
        del pairs
        t0 = time.perf_counter()
        pr, n_in = kcl.read_edge_list(path)
        t1 = time.perf_counter()
        g = kcl.Graph(pr, n_in)
        t2 = time.perf_counter()
        log(f"bench exact: read {t1 - t0:.2f}s, graph n={g.n} m={g.m} core={g.core} "
            f"{t2 - t1:.2f}s; running the exact pipeline")
        res, nodes = kcl.exact(g, k, stat_path=os.path.join(td, "stat_exact.txt"),
                               max_seconds=args.exact_budget,
                               schedule=kcl.SCHEDULES[args.schedule], model=model)
        if model is not None:
            model.close()
        t3 = time.perf_counter()
        stat = open(os.path.join(td, "stat_exact.txt")).read()
    log(stat)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_exact(pairs_file=None, res=res, k=k, workload=args.workload, gg=gg)
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
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=15)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cl4m", choices=sorted(WORKLOADS) + ["batch512"])
    ap.add_argument("--no-tune", action="store_true",
                    help="batch512: keep the default row width and occupancy")
    ap.add_argument("--schedule", default="learned", choices=["counted", "learned", "degree"],
                    help="root order/splitting: counting-pass costs, the learned cost "
                         "model, or plain d+^2 (counts are identical under all three)")
    ap.add_argument("--model", default=None,
                    help="learned schedule: cost model file (default: the shipped one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact-budget", type=int, default=240,
                    help="exact mode: seconds before giving up on a certificate")
    ap.add_argument("--mode", default="sim",
                    choices=["sim", "exact", "trim", "makespan", "sampling", "cmpsync",
                             "input"],
                    help="sim: visits/s of one iteration (default); exact: time-to-exact; "
                         "trim: k-clique core trimming time; makespan: each schedule's "
                         "pass against a makespan lower bound; sampling: Sampling++ time "
                         "(--steps rounds); cmpsync: kCDensestCmpSync end to end "
                         "(T = --steps); input: reader + graph build against the "
                         "reference's loader")
    ap.add_argument("--gpu-peel-min", type=int, default=0,
                    help="trim mode: pops through at least this many cliques go to the GPU "
                         "(0: the library's default)")
    ap.add_argument("--cpu-timeout", type=float, default=900.0,
                    help="trim mode: time limit of the reference program")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="iterations of the plain (d+^2) schedule timed on the main run's "
                         "states (0: skip)")
    ap.add_argument("--secondary", default="cl1m",
                    help="a second workload run for a few steps ('' to skip)")
    ap.add_argument("--cpu-budget", type=float, default=24.0,
                    help="seconds of reference work in the CPU baseline's sample")
    ap.add_argument("--cache", default="",
                    help="directory for the generated edges (reused by later runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB on the device (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)

    import kclistpp_amd as kcl
    from kclistpp_amd import graphgen as gg

    if args.mode == "exact":
        run_exact(args, kcl, gg)
        return
    if args.mode == "trim":
        run_trim(args, kcl, gg)
        return
    if args.mode == "makespan":
        run_makespan(args, kcl, gg)
        return
    if args.mode == "sampling":
        run_sampling(args, kcl, gg)
        return
    if args.mode == "cmpsync":
        run_cmpsync(args, kcl, gg)
        return
    if args.mode == "input":
        run_input(args, kcl, gg)
        return
    if args.workload == "batch512":
        run_batch(args, kcl, gg)
        return
    run_sim(args, kcl, gg)


def timed_steps(sim, t: int, steps: int, cliques: int, label: str):
    """steps synchronous iterations from t+1; returns (t, seconds, visits,
    enumeration-kernel ms summed, pass ms list)."""
    visits, kern_ms, passes = 0, 0.0, []
    t_start = time.perf_counter()
    for _ in range(steps):
        t += 1
        sim.step(t)  # synchronous: returns after the pass and update completed
        v, kms, pms = sim.last_pass()
        if v != cliques:
            raise RuntimeError(f"{label}: step {t} visited {v} cliques, the graph has {cliques}")
        visits += v
        kern_ms += kms
        passes.append(pms)
    return t, time.perf_counter() - t_start, visits, kern_ms, passes


def plain_leg(kcl, g, k, sim, t: int, steps: int):
    """The plain schedule (d+^2 order and splits) on the same states as the
    main schedule: per step, the plain sim takes the main sim's rho, both run
    iteration t, their counts must agree, and their pass times are kept."""
    t0 = time.time()
    plain = kcl.Sim(g, k, schedule=kcl.SCHED_DEGREE)
    plan_s = time.time() - t0
    main_ms, plain_ms = [], []
    for _ in range(steps):
        t += 1
        plain.set_rho(sim.rho())
        sim.step(t)
        plain.step(t)
        if not np.array_equal(sim.counts(), plain.counts()):
            raise RuntimeError(f"plain schedule counts differ at t={t}")
        main_ms.append(sim.last_pass()[2])
        plain_ms.append(plain.last_pass()[2])
        log(f"bench: t={t} main pass {main_ms[-1]:.2f}ms, plain pass {plain_ms[-1]:.2f}ms")
    plain.close()
    m, p = float(np.mean(main_ms)), float(np.mean(plain_ms))
    return t, {"steps": steps, "t_from": t - steps + 1, "plan_s": round(plan_s, 2),
               "main_pass_ms": round(m, 3), "plain_pass_ms": round(p, 3),
               "main_vs_plain": round(p / m, 4) if m > 0 else None}


def workload_pairs(args, gg, wname: str, seed: int, sd: int) -> np.ndarray:
    """The workload's edges: generated, or loaded from --cache DIR where an
    earlier run stored them (the generator is seeded, so both are the same
    array)."""
    path = os.path.join(args.cache, f"{wname}_s{seed}_d{sd}.npy") if args.cache else None
    if path and os.path.exists(path):
        return np.load(path, allow_pickle=False)
    pairs = gg.workload(wname, seed=seed, scale_down=sd)
    if path:
        os.makedirs(args.cache, exist_ok=True)
        np.save(path + ".tmp.npy", pairs)
        os.replace(path + ".tmp.npy", path)
    return pairs


def secondary_leg(args, kcl, gg, name: str, steps: int = 10, warmup: int = 2):
    """A short run of another workload: under the library's default schedule
    (COUNTED) and, when it differs, under the main run's schedule."""
    wname, seed, sd, k, desc = WORKLOADS[name]
    pairs = workload_pairs(args, gg, wname, seed, sd)
    model = load_model(args, kcl)
    out = {"workload": f"{name}: {desc}", "steps": steps, "unit": "k-clique visits/s",
           "schedules": {}}
    with kcl.Graph(pairs) as g:
        del pairs
        for sched in dict.fromkeys(["counted", args.schedule]):
            sim = kcl.Sim(g, k, schedule=kcl.SCHEDULES[sched],
                          model=model if sched == "learned" else None)
            plan = sim.plan_info()
            t, _, _, _, _ = timed_steps(sim, 0, warmup, plan["cliques"], name)
            t, sec, visits, kern_ms, _ = timed_steps(sim, t, steps, plan["cliques"], name)
            sim.close()
            out["schedules"][sched] = {"value": visits / sec if sec > 0 else 0.0,
                                       "ms_per_step": 1000.0 * sec / steps,
                                       "kernel_ms": kern_ms / steps}
            out["visits_per_step"] = plan["cliques"]
    if model is not None:
        model.close()
    log(f"bench: secondary {name}: {out}")
    return out


def roofline_of(workload: str, alg_bytes: int, kern_avg_ms: float, dev: dict | None = None) -> dict:
    """The dominant kernel's roofline: algorithmic HBM bytes per pass over
    the enumeration kernels' time per pass (HIP events on the pass's
    stream), plus what the committed counter profile of this workload
    measured (HBM traffic, VALU and LDS busy fractions, the bound).

    valu_issue: the pass's VALU instructions (SQ_INSTS_VALU of one profiled
    pass, wavefront instructions) at 4 cycles each on 4 SIMDs per CU at the
    device's peak clock give the shortest pass the VALU can issue; frac is
    that time over the measured one."""
    achieved = alg_bytes / (kern_avg_ms * 1e-3) / 1e9 if kern_avg_ms > 0 else 0.0
    traffic = profiled(workload, "traffic_bytes_per_pass")
    # The roof: HBM, the only one the integer pass has (it issues no MFMA).
    # What actually binds is read from the committed counter profile of the
    # workload ("limiter": valu / lds / latency), with the busy fractions.
    out = {"bound": "hbm",
           "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
           "traffic": int(traffic) if traffic is not None else None,
           "kernel": "pass_kernel (all word classes of one enumeration pass)",
           "kernel_ms": round(kern_avg_ms, 4), "algorithmic_bytes": alg_bytes,
           "limiter": profiled(workload, "limiter", text=True),
           "profile": f"profiles/pass_kernel_{workload}.md"}
    for key in ("valu_busy_frac", "lds_busy_frac", "salu_busy_frac", "lds_bank_conflict_frac",
                "hbm_traffic_frac"):
        v = profiled(workload, key)
        if v is not None:
            out[key] = v
    insts = profiled(workload, "valu_insts_per_pass")
    if insts and dev and dev.get("cu_count") and dev.get("clock_khz") and kern_avg_ms > 0:
        floor_ms = insts * 4.0 / (4.0 * dev["cu_count"]) / (dev["clock_khz"] * 1e3) * 1e3
        out["valu_issue"] = {"insts_per_pass": insts, "cycles_per_inst": 4,
                             "simds": 4 * dev["cu_count"], "clock_mhz": dev["clock_khz"] / 1e3,
                             "floor_ms": round(floor_ms, 3),
                             "frac": round(floor_ms / kern_avg_ms, 4)}
    return out


def run_sim(args, kcl, gg):
    wname, seed, sd, k, desc = WORKLOADS[args.workload]
    log(f"bench: generating {args.workload} ({desc})")
    t0 = time.time()
    pairs = workload_pairs(args, gg, wname, seed, sd)
    log(f"bench: {pairs.shape[0]} edges in {time.time() - t0:.1f}s; building the graph")
    t0 = time.time()
    g = kcl.Graph(pairs)
    del pairs
    log(f"bench: graph n={g.n} m={g.m} core={g.core} built in {time.time() - t0:.1f}s")
    cd, adj = g.csr()
    sched = kcl.SCHEDULES[args.schedule]
    model = load_model(args, kcl)
    t0 = time.time()
    sim = kcl.Sim(g, k, schedule=sched, model=model)
    plan = sim.plan_info()
    root_cliques = sim.root_cliques()
    alg_bytes = algorithmic_bytes(cd, adj, root_cliques > 0)
    log(f"bench: plan ready in {time.time() - t0:.1f}s: {plan}; warmup {args.warmup}")
    t = 0
    for _ in range(args.warmup):
        t, _, _, _, ps = timed_steps(sim, t, 1, plan["cliques"], "warmup")
        log(f"bench: warmup t={t} pass={ps[0]:.2f}ms")
    t, elapsed, visits, kern_ms, passes = timed_steps(sim, t, args.steps, plan["cliques"], "bench")
    log(f"bench: {args.steps} steps in {elapsed:.3f}s, {visits} visits; passes "
        f"{min(passes or [0]):.2f}-{max(passes or [0]):.2f}ms")
    ms_per_step = 1000.0 * elapsed / max(args.steps, 1)
    value = visits / elapsed if elapsed > 0 else 0.0
    kern_avg_ms = kern_ms / max(args.steps, 1)
    roofline = roofline_of(args.workload, alg_bytes, kern_avg_ms, kcl.device_info())
    plain = None
    if args.plain_steps > 0 and args.schedule != "degree":
        t, plain = plain_leg(kcl, g, k, sim, t, args.plain_steps)
    sim.close()
    if model is not None:
        model.close()
    g.close()
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the reference on the host cores")
        cpu = cpu_baseline(cd, adj, root_cliques, k, budget_s=args.cpu_budget)
    del cd, adj
    secondary = None
    if args.secondary and args.secondary != args.workload:
        secondary = secondary_leg(args, kcl, gg, args.secondary)
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
                   "core": g.core, "roots": plan["roots_kept"],
                   "visits_per_step": visits // max(args.steps, 1),
                   "schedule": args.schedule, "parallelism": "1 GPU",
                   "t_timed": [args.warmup + 1, args.warmup + args.steps],
                   "plan": plan},
        "plain_schedule": plain,
        "secondary": secondary,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def copy_probe(mib: int) -> None:
    """A plain device copy of `mib` MiB (torch), for scaling the FETCH_SIZE
    and WRITE_SIZE counters: rocprofv3 --pmc over this run reads what the
    counters report for a known number of bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: {mib} MiB copied")


if __name__ == "__main__":
    main()
