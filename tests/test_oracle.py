"""Pins the oracle (oracle/kcl_oracle.c) before it is trusted as the checker.

* golden vectors written by the reference's own code (tests/golden/*.npz):
  ranks, DAG, per-iteration counts, rho bits, records, level/tent, and the
  kClistCoreTrim outputs;
* brute force over all k-subsets of small graphs, with ties in rho;
* closed forms (K_n, disjoint cliques, clique-free graphs, the planted K30);
* planted mistakes in a Python copy of the winner rule must fail the
  golden vectors (so the vectors do pin the tie rule and the update).
"""

import itertools
import math

import numpy as np
import pytest

from conftest import golden_index, golden_pairs, load_golden
from kclistpp_amd import graphgen as gg

SIM = golden_index()["sim"]
TRIM = golden_index()["trim"]
BIG = golden_index()["big"]
WIDE = golden_index()["wide"]


def _bits(x: float) -> int:
    return int(np.float64(x).view(np.uint64))


@pytest.mark.parametrize("case", SIM)
def test_oracle_matches_reference_sim(oracle, case):
    gd = load_golden(case)
    g = oracle.dag(gd["pairs"])
    assert g.n == int(gd["n"]) and g.core == int(gd["core"])
    assert np.array_equal(g.rank, gd["rank"])
    assert np.array_equal(g.es, gd["es"]) and np.array_equal(g.et, gd["et"])
    k, T = int(gd["k"]), int(gd["num_iter"])
    recs, counts, rho = oracle.cmpsync(g, k, T, keep_counts=True)
    assert np.array_equal(counts, gd["counts"])
    assert np.array_equal(rho.view(np.uint64), gd["rho_bits"])
    got = np.array([[r["T"], r["n"], r["m"], _bits(r["density"]), _bits(r["ub"])] for r in recs],
                   dtype=np.uint64)
    assert np.array_equal(got, gd["records"])
    # Level order and rho_tentative of the last report.
    _, _, level, tent = oracle.find_densest(g, k, rho)
    assert np.array_equal(level, gd["level"])
    assert np.array_equal(tent, gd["tent"])


@pytest.mark.parametrize("case", TRIM)
def test_oracle_matches_reference_trim(oracle, case):
    gd = load_golden(case)
    k = int(gd["k"])
    info, mp, nck, ick, icv, trimmed = oracle.trim(golden_pairs(gd), k)
    # ckdeg: "old id, k-clique degree" per new id (kClistCoreTrim.c:419-429).
    assert np.array_equal(np.stack([mp, nck]).T.astype(np.uint64), gd["ckdeg"])
    # ckcore: peel order with running core (kClistCoreTrim.c:672-676).
    assert np.array_equal(np.stack([mp[ick], icv]).T.astype(np.uint64), gd["ckcore"])
    # ckdens: "%lf %lf %u" then ids of the densest prefix (kClistCoreTrim.c:678-684).
    n2, size = info["n2"], info["size"]
    ids = [int(mp[ick[i]]) for i in range(n2 - 1, n2 - size - 1, -1)] if size != n2 else []
    line = "%f %f %u" % (info["ckrho"], info["rho"], size) + "".join(f" {v}" for v in ids)
    assert line == str(gd["ckdens"][0])
    assert np.array_equal(trimmed, gd["trimmed"])
    assert [info["ntrim"], info["mtrim"], info["ncktrim"]] == gd["stat"].tolist()


@pytest.mark.parametrize("case", [c for c in BIG if "er20k" in c])
def test_oracle_matches_reference_er20k(oracle, case):
    import hashlib
    gd = load_golden(case)
    pairs = gg.workload(str(gd["workload"][0]), seed=int(gd["seed"]),
                        scale_down=int(gd["scale_down"]))
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == str(gd["pairs_sha"][0])
    g = oracle.dag(pairs)
    recs, counts, rho = oracle.cmpsync(g, int(gd["k"]), int(gd["num_iter"]), keep_counts=True)
    shas = [hashlib.sha256(c.tobytes()).hexdigest() for c in counts]
    assert shas == [str(s) for s in gd["count_sha"]]
    got = np.array([[r["T"], r["n"], r["m"], _bits(r["density"]), _bits(r["ub"])] for r in recs],
                   dtype=np.uint64)
    assert np.array_equal(got, gd["records"])


@pytest.mark.parametrize("case", [c for c in WIDE if "bip340" in c])
def test_oracle_matches_reference_wide_rows(oracle, case):
    """Out-neighbourhoods of up to 350 members (the reference's FW pass at a
    rho with ties, and its PAVA pass at a level permutation)."""
    gd = load_golden(case)
    g = oracle.dag(golden_pairs(gd))
    k = int(gd["k"])
    assert g.core == int(gd["core"])
    fw, visits = oracle.fw_counts(g, k, gd["rho"])
    assert visits == int(gd["visits"]) and np.array_equal(fw, gd["fw"])
    assert np.array_equal(oracle.pava_counts(g, k, gd["level"]), gd["tent"])


# ---------------------------------------------------------------------------
# Brute force: every k-subset, winner = argmin (rho, id), with many ties.

def _brute_counts(n, edges, k, rho=None, level=None):
    adj = np.zeros((n, n), bool)
    for a, b in edges:
        adj[a, b] = adj[b, a] = True
    cnt = np.zeros(n, np.uint64)
    total = 0
    for sub in itertools.combinations(range(n), k):
        if all(adj[a, b] for a, b in itertools.combinations(sub, 2)):
            total += 1
            if rho is not None:
                w = min(sub, key=lambda x: (rho[x], x))
                cnt[w] += 1
            else:
                w = max(sub, key=lambda x: level[x])
                cnt[level[w]] += 1
    return cnt, total


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
def test_oracle_brute_force(oracle, seed, k):
    rng = np.random.default_rng(seed * 31 + k)
    n_in = 13
    p = 0.55 + 0.05 * (seed % 3)
    pairs = np.array([(a, b) for a in range(n_in) for b in range(a + 1, n_in)
                      if rng.random() < p], np.uint32)
    pairs = gg.finalize(pairs, rng)
    g = oracle.dag(pairs)
    edges = list(zip(g.es.tolist(), g.et.tolist()))
    # rho with heavy ties (few distinct values)
    rho = rng.integers(0, 3, g.n).astype(np.float64) * 0.5
    got, visits = oracle.fw_counts(g, k, rho)
    want, total = _brute_counts(g.n, edges, k, rho=rho)
    assert visits == total and np.array_equal(got, want)
    level = rng.permutation(g.n).astype(np.uint32)
    got = oracle.pava_counts(g, k, level)
    want, _ = _brute_counts(g.n, edges, k, level=level)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# Closed forms.

@pytest.mark.parametrize("n,k", [(6, 2), (8, 3), (9, 4), (10, 5), (7, 7), (7, 8)])
def test_oracle_complete_graph(oracle, n, k):
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n)], np.uint32)
    g = oracle.dag(pairs)
    assert oracle.count(g, k) == math.comb(n, k)
    # At rho = 0 every clique goes to its smallest member: node i gets C(n-1-i, k-1).
    cnt, _ = oracle.fw_counts(g, k, np.zeros(g.n))
    assert cnt.tolist() == [math.comb(n - 1 - i, k - 1) for i in range(n)]
    recs, _, _ = oracle.cmpsync(g, k, 8)
    if k <= n:
        assert all(r["n"] == n and r["sum"] == math.comb(n, k) for r in recs)
        assert all(r["density"] == math.comb(n, k) / n for r in recs)


def test_oracle_disjoint_cliques(oracle):
    blocks = [(0, 9), (20, 6), (40, 4)]
    pairs = np.array([(o + a, o + b) for o, s in blocks for a in range(s) for b in range(a + 1, s)],
                     np.uint32)
    g = oracle.dag(pairs)
    for k in (3, 4):
        assert oracle.count(g, k) == sum(math.comb(s, k) for _, s in blocks)
        recs, _, _ = oracle.cmpsync(g, k, 16)
        assert recs[-1]["n"] == 9 and recs[-1]["sum"] == math.comb(9, k)


def test_oracle_clique_free(oracle):
    # Bipartite graphs have no triangles.
    pairs = np.array([(a, 10 + b) for a in range(10) for b in range(10) if (a + b) % 3], np.uint32)
    g = oracle.dag(pairs)
    for k in (3, 4, 5):
        assert oracle.count(g, k) == 0
    assert oracle.count(g, 2) == pairs.shape[0]


def test_oracle_planted_k30(oracle):
    pairs = gg.workload("er20k", seed=1)
    g = oracle.dag(pairs)
    assert oracle.count(g, 4) == 27405
    recs, _, _ = oracle.cmpsync(g, 4, 2)
    assert recs[0]["n"] == 30 and recs[0]["density"] == 27405 / 30


# ---------------------------------------------------------------------------
# Planted mistakes: a wrong tie rule or update order must miss the vectors.

def test_golden_vectors_pin_tie_rule_and_update(oracle):
    gd = load_golden("sim_cl3k_k4_T8")
    g = oracle.dag(gd["pairs"])
    edges = list(zip(g.es.tolist(), g.et.tolist()))
    nbr = [set() for _ in range(g.n)]
    for a, b in edges:
        nbr[a].add(b)
        nbr[b].add(a)
    cliques = []
    for a, b in edges:  # 4-cliques by brute extension (small graph)
        common = sorted(nbr[a] & nbr[b])
        for i, x in enumerate(common):
            for y in common[i + 1:]:
                if y in nbr[x] and x < b and y < b:
                    cliques.append((a, b, x, y))
    cliques = [c for c in set(tuple(sorted(c)) for c in cliques)]

    def counts_for(rho, pick):
        c = np.zeros(g.n, np.uint64)
        for q in cliques:
            c[pick(q, rho)] += 1
        return c

    right = lambda q, rho: min(q, key=lambda x: (rho[x], x))   # noqa: E731
    wrong_tie = lambda q, rho: min(q, key=lambda x: (rho[x], -x))  # noqa: E731
    rho = np.zeros(g.n)
    assert np.array_equal(counts_for(rho, right), gd["counts"][0])
    assert not np.array_equal(counts_for(rho, wrong_tie), gd["counts"][0])
    assert len(cliques) == int(gd["counts"][0].sum())
    # The rho bits pin the update's operation order: the algebraically
    # equal rho + 2(cnt - rho)/(t + 2) misses them, the reference's order hits.
    C = gd["counts"].astype(np.float64)
    ref_order = np.zeros(g.n)
    alt_order = np.zeros(g.n)
    for t in range(1, C.shape[0] + 1):
        ref_order = oracle.rho_update(ref_order, t, gd["counts"][t - 1])
        alt_order = alt_order + 2.0 * (C[t - 1] - alt_order) / (t + 2.0)
    assert np.array_equal(ref_order.view(np.uint64), gd["rho_bits"])
    assert not np.array_equal(alt_order.view(np.uint64), gd["rho_bits"])


# ---------------------------------------------------------------------------
# Max-flow (Dinic) against networkx.

@pytest.mark.parametrize("seed", range(5))
def test_oracle_max_flow(oracle, seed):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(seed)
    nv = 30
    G = nx.DiGraph()
    frm, to, cap = [], [], []
    for _ in range(150):
        a, b = rng.integers(0, nv, 2)
        if a == b:
            continue
        c = int(rng.integers(1, 50))
        frm.append(a)
        to.append(b)
        cap.append(c)
        if G.has_edge(a, b):
            G[a][b]["capacity"] += c
        else:
            G.add_edge(a, b, capacity=c)
    want = nx.maximum_flow_value(G, 0, nv - 1) if G.has_node(0) and G.has_node(nv - 1) else 0
    assert oracle.max_flow(nv, frm, to, cap, 0, nv - 1) == want


def test_reference_pairs_loader_matches_its_reader(tmp_path):
    """oracle/ref_harness.c's in-memory loader builds the same graph as the
    reference's own text reader (used by bench.py's CPU baseline)."""
    import kclistpp_amd as kcl
    from kclistpp_amd import graphgen as gg
    from oracle import RefHarness
    try:
        ref = RefHarness()
    except OSError:
        pytest.skip("reference build (oracle/_ref) absent")
    pairs = gg.workload("er20k", seed=4)
    path = str(tmp_path / "g.txt")
    kcl.write_edge_list(path, pairs)
    a, b = ref.load(path), ref.load_pairs(pairs)
    assert ref.info(a) == ref.info(b)
    assert np.array_equal(ref.rank(a), ref.rank(b))
    assert np.array_equal(ref.edges(a), ref.edges(b))


def test_reference_csr_loader_matches_its_own_build():
    """The reference Graph filled from a CSR (the benchmark's CPU baseline
    path) enumerates like the one its own SortByCore/MakeGraph builds."""
    from oracle import RefHarness
    pairs, rng = gg.chung_lu(2000, 12000, 2.3, 200.0, 7)
    pairs = gg.finalize(gg.plant_community(pairs, 2000, 30, 0.8, rng), rng)
    ref = RefHarness()
    a = ref.load_pairs(pairs)
    cd, adj = ref.csr(a)
    b = ref.load_csr(cd, adj)
    assert ref.info(a)["core"] == ref.info(b)["core"] and ref.info(b)["e"] == adj.size
    assert np.array_equal(ref.edges(a), ref.edges(b))
    ref.start(a, 4)  # the reference's rho/rho_add are globals: one graph at a time
    ca, _ = ref.step(a, 1)
    ref.start(b, 4)
    cb, _ = ref.step(b, 1)
    # Per-edge timing (the benchmark's CPU baseline) visits what a batch does.
    idx = np.arange(0, adj.size, 7, dtype=np.uint32)
    _, va = ref.time_fw_roots(b, idx)
    wall, per, vb = ref.time_fw_each(b, idx)
    assert va == vb and per.size == idx.size and np.all(per >= 0) and wall > 0
    assert np.array_equal(ca, cb) and ca.sum() > 0
    ref.free(a)
    ref.free(b)


@pytest.mark.parametrize("seed", range(5))
def test_oracle_count_in_set_brute_force(oracle, seed):
    """k-cliques inside a node subset (the exact pipeline's checker) against
    every k-subset of the subset, on random 14-node graphs and masks."""
    rng = np.random.default_rng(900 + seed)
    n = 14
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n)
                      if rng.random() < 0.55], np.uint32)
    pairs = gg.finalize(pairs, rng)
    g = oracle.dag(pairs)
    adj = np.zeros((g.n_in, g.n_in), bool)
    for a, b in pairs.tolist():
        adj[a, b] = adj[b, a] = True
    for k in (2, 3, 4, 5):
        for _ in range(4):
            keep = rng.random(g.n_in) < 0.7
            want = sum(1 for sub in itertools.combinations(np.nonzero(keep)[0].tolist(), k)
                       if all(adj[a, b] for a, b in itertools.combinations(sub, 2)))
            mask = np.zeros(g.n, np.uint8)
            mask[g.rank[np.nonzero(keep)[0]]] = 1
            assert oracle.count_in_set(g, k, mask) == want
