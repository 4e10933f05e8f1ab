"""Exact k-clique densest subgraph (the result of kCDensestMem).  The
reference program needs Boost.Graph, which is absent, so its trajectory is
unpinned; its final result - the maximal k-clique densest subgraph, unique
by supermodularity - is pinned by exhaustive search on small graphs, by
planted structures, and by an independent max-flow (the oracle's Dinic over
every clique of the graph) on larger ones."""

import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "kclistpp_amd", "bin", "kCDensestMem")


def _say(*a):
    print(*a, flush=True)


def _brute_maximal_densest(pairs, k):
    nodes = sorted(set(pairs.reshape(-1).tolist()))
    idx = {v: i for i, v in enumerate(nodes)}
    n = len(nodes)
    adj = np.zeros((n, n), bool)
    for a, b in pairs.tolist():
        adj[idx[a], idx[b]] = adj[idx[b], idx[a]] = True
    cl = [sum(1 << i for i in sub) for sub in itertools.combinations(range(n), k)
          if all(adj[a, b] for a, b in itertools.combinations(sub, 2))]
    cl = np.array(cl, dtype=np.int64)
    best_p, best_q, union = -1, 1, 0
    for mask in range(1, 1 << n):
        p = int(((cl & mask) == cl).sum()) if cl.size else 0
        q = bin(mask).count("1")
        if p * best_q > best_p * q:
            best_p, best_q, union = p, q, mask
        elif p * best_q == best_p * q:
            union |= mask
    members = [nodes[i] for i in range(n) if union >> i & 1]
    return best_p * len(members) // best_q, members, (best_p, best_q)


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("k", [2, 3, 4])
def test_exact_brute_force(kcl, seed, k):
    rng = np.random.default_rng(100 * k + seed)
    n = 12
    p = 0.35 + 0.05 * (seed % 4)
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p],
                     np.uint32)
    pairs = gg.finalize(pairs, rng)
    cliques, members, (bp, bq) = _brute_maximal_densest(pairs, k)
    _say("exact brute", seed, k, members)
    with kcl.Graph(pairs) as g:
        r, nodes = kcl.exact(g, k)
    assert r["certified"] == 1 and r["stop"] in (1, 2)
    if bp == 0:
        return  # no k-clique: every set has density 0 (reference: the whole graph)
    assert nodes.tolist() == members
    assert r["cliques"] * bq == bp * r["n"] and r["cliques"] == cliques


def test_exact_disjoint_cliques(kcl):
    k9 = [(a, b) for a in range(9) for b in range(a + 1, 9)]
    k6 = [(20 + a, 20 + b) for a in range(6) for b in range(a + 1, 6)]
    pairs = np.array(k9 + k6 + [(8, 20)], np.uint32)
    for k in (3, 4, 5):
        _say("exact K9+K6 k", k)
        with kcl.Graph(pairs) as g:
            r, nodes = kcl.exact(g, k)
        assert r["certified"] and nodes.tolist() == list(range(9))
        assert r["cliques"] == math.comb(9, k) and r["m"] == 36


def test_exact_returns_the_maximal_densest_subgraph(kcl):
    # Two disjoint K7 joined by one edge: both halves and their union tie in
    # density for k >= 3; the maximal densest subgraph is the union.
    a = [(i, j) for i in range(7) for j in range(i + 1, 7)]
    b = [(10 + i, 10 + j) for i in range(7) for j in range(i + 1, 7)]
    pairs = np.array(a + b + [(0, 10)], np.uint32)
    with kcl.Graph(pairs) as g:
        r, nodes = kcl.exact(g, 4)
    assert r["certified"] and r["n"] == 14 and r["cliques"] == 2 * math.comb(7, 4)


# gpu_list_min: 0 lists a certificate region's cliques on the host when it
# holds fewer than 2^20, 1 always on the GPU (the walk kernels' listing).
@pytest.mark.parametrize("gpu_list_min", [0, 1])
@pytest.mark.parametrize("case,k", [("sim_er2k_k12_k4_T8", 4), ("sim_cl3k_k4_T8", 4),
                                    ("sim_cl3k_k5_T4", 5), ("sim_cl1k_dense_k4_T4", 4)])
def test_exact_vs_oracle_max_flow(kcl, oracle, case, k, gpu_list_min):
    gd = load_golden(case)
    pairs = gd["pairs"]
    _say("exact", case, k, "gpu_list_min", gpu_list_min)
    with kcl.Graph(pairs) as g:
        r, nodes = kcl.exact(g, k, gpu_list_min=gpu_list_min)
    assert r["certified"] == 1
    og = oracle.dag(pairs)
    rank = og.rank
    mask = np.zeros(og.n, np.uint8)
    mask[rank[nodes]] = 1
    p, q = oracle.count_in_set(og, k, mask), len(nodes)
    assert p == r["cliques"] and q == r["n"]
    # Independent certificate: Goldberg's network over every k-clique of the
    # graph with sink capacity p, source capacity q: saturated iff nothing
    # is denser than p/q.
    full = np.ones(og.n, np.uint8)
    cl = _list_cliques(og, k)
    nc = len(cl)
    frm, to, cap = [], [], []
    base = 2 + og.n
    for c, members in enumerate(cl):
        frm.append(0); to.append(base + c); cap.append(q)
        for v in members:
            frm.append(base + c); to.append(2 + v); cap.append(q)
    for v in range(og.n):
        frm.append(2 + v); to.append(1); cap.append(p)
    assert oracle.max_flow(base + nc, frm, to, cap, 0, 1) == q * nc
    del full


def _list_cliques(og, k):
    nbr = [set() for _ in range(og.n)]
    for a, b in zip(og.es.tolist(), og.et.tolist()):
        nbr[a].add(b)
    out = []

    def rec(chosen, cand):
        if len(chosen) == k:
            out.append(tuple(chosen))
            return
        for x in sorted(cand):
            rec(chosen + [x], cand & nbr[x])

    for s in range(og.n):
        rec([s], set(nbr[s]))
    return out


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_exact_cli_stat_file(kcl, tmp_path):
    gd = load_golden("sim_er2k_k12_k4_T8")
    path = str(tmp_path / "g.txt")
    gg.write_text(gd["pairs"], path)
    _say("exact cli")
    r = subprocess.run([CLI, "8", "4", path, "er"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = open(tmp_path / "stat_exact_er_8_4.txt").read().splitlines()
    assert lines[0].startswith("[Number of Iterations]\t[Number of Nodes]")
    assert lines[-2].startswith("[Number of Iterations]\t[Stopping Condition]")
    f = lines[-1].split("\t")
    # The planted K12 is stable (its region is itself) and a complete graph,
    # so the reference's closed-form criterion A certifies it.
    assert f[1] == "Goldberg" and int(f[2]) == 12 and int(f[3]) == 66
    assert f[4] == "%.12f" % (math.comb(12, 4) / 12)


def test_exact_over_budget_reports_bounds(kcl, tmp_path):
    """With a network budget too small for any certificate the result is
    the densest subgraph found, uncertified, with a valid upper bound."""
    rng = np.random.default_rng(9)
    pairs, _ = gg.erdos_renyi(400, 2400, 9)
    pairs = gg.plant_clique(pairs, 400, 14, rng)
    stat = str(tmp_path / "stat.txt")
    with kcl.Graph(pairs) as g:
        r_small, _ = kcl.exact(g, 4, max_network_cliques=10, max_iter=8, stat_path=stat)
        r_full, nodes = kcl.exact(g, 4)
    assert r_small["certified"] == 0 and r_small["max_flow_calls"] == 0 and r_small["stop"] == 0
    labels = [r.split("\t")[4] for r in open(stat).read().splitlines()[1:]]
    assert len(labels) == 4 and "OVER BUDGET" in labels
    assert set(labels) <= {"OVER BUDGET", "NOT STABLE", "STABLE BUT NOT DENSEST"}
    assert r_small["upper_bound"] >= r_small["density"] > 0
    assert r_full["certified"] == 1
    assert r_small["density"] <= r_full["density"] <= r_small["upper_bound"] + 1e-9
    assert r_full["n"] == len(nodes)


def test_exact_planted_clique_by_criterion_a(kcl):
    """The k = 4 workload (ER 20k/100k + planted K30): the K30 is its own
    region and complete, so criterion A certifies it without a max-flow."""
    pairs = gg.workload("er20k", seed=1)
    with kcl.Graph(pairs) as g:
        r, nodes = kcl.exact(g, 4)
    assert r["certified"] == 1 and r["stop"] == 1 and r["max_flow_calls"] == 0
    assert r["n"] == 30 and r["cliques"] == math.comb(30, 4) and r["m"] == 435


@pytest.mark.parametrize("gpu_list_min", [0, 1])
def test_exact_community_by_max_flow(kcl, oracle, gpu_list_min):
    """A 60-node community with p = 0.8 planted in a Chung-Lu graph, k = 5:
    not a clique, so the certificate is the max-flow; the result's clique
    count is checked against the oracle's, with the region's cliques listed
    on the host and on the GPU."""
    pairs, rng = gg.chung_lu(20000, 120000, 2.3, 300.0, 31)
    pairs = gg.finalize(gg.plant_community(pairs, 20000, 60, 0.8, rng), rng)
    with kcl.Graph(pairs) as g:
        r, nodes = kcl.exact(g, 5, gpu_list_min=gpu_list_min)
    assert r["certified"] == 1 and r["stop"] == 2 and r["max_flow_calls"] >= 1
    og = oracle.dag(pairs)
    mask = np.zeros(og.n, np.uint8)
    mask[og.rank[nodes]] = 1
    assert oracle.count_in_set(og, 5, mask) == r["cliques"] and len(nodes) == r["n"]
    assert r["n"] <= 60 and r["cliques"] > 10 ** 5
