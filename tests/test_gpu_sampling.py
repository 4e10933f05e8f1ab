"""GPU checks of Sampling++ (kCDensestSamplingKclistpp.c): the clique listing
against brute force, the sampler's kept set against its hash rule, the
synchronous rounds against the oracle's Frank-Wolfe pass, and the report
(extraction, densest prefix, exact count of the induced subgraph) against
the oracle's PAVA pass and clique count.  The reference itself draws its
sample with rand() seeded by the clock, so no run of it is reproducible;
these checks pin every step that is deterministic."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "kclistpp_amd", "bin", "kCDensestSamplingKclistpp")
M64 = (1 << 64) - 1


def _say(*a):
    print(*a, flush=True)


def _community(seed, n=400, m=2000, size=24, p=0.85):
    pairs, rng = gg.chung_lu(n, m, 2.3, 60.0, seed)
    pairs = gg.plant_community(pairs, n, size, p, rng)
    return gg.finalize(pairs, rng)


def _all_cliques(og, k):
    """Every k-clique of the relabelled DAG by brute force over bitsets,
    rows ascending, lexicographically sorted."""
    n = og.n
    nb = [0] * n
    for s in range(n):
        for j in range(og.cd[s], og.cd[s + 1]):
            t = int(og.adj[j])
            nb[s] |= 1 << t
            nb[t] |= 1 << s
    out = []

    def grow(members, cand, need):
        if need == 0:
            out.append(sorted(members))
            return
        while cand:
            v = cand.bit_length() - 1
            cand &= ~(1 << v)
            grow(members + [v], cand & nb[v], need - 1)

    for s in range(n):
        grow([s], nb[s] & ((1 << s) - 1), k - 1)
    arr = np.array(out, np.uint32).reshape(-1, k)
    return arr[np.lexsort(arr.T[::-1])] if arr.size else arr


def _sorted_rows(a):
    return a[np.lexsort(a.T[::-1])] if a.size else a


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _kept_by_hash(cliques, seed, thr53):
    """The sampler's rule: splitmix64 steps over the ascending ids from the
    seed; keep when the top 53 bits are below thr53."""
    keep = []
    for row in cliques:
        h = seed
        for x in row:
            h = _mix64(h ^ int(x))
        keep.append((h >> 11) < thr53)
    return cliques[np.array(keep, bool)] if len(keep) else cliques


@pytest.mark.parametrize("k", [3, 4, 5, 6, 8, 11])
def test_listing_matches_brute_force(kcl, oracle, k):
    pairs = _community(70 + k)
    og = oracle.dag(pairs)
    want = _all_cliques(og, k)
    _say("listing k", k, "cliques", want.shape[0])
    with kcl.Graph(pairs) as g, kcl.Sampling(g, k, target=10 ** 9) as s:
        assert s.info["prob"] == 1.0 and s.info["total"] == want.shape[0]
        got = _sorted_rows(s.cliques())
        assert np.array_equal(got, want)


def test_listing_wide_rows_and_edge_tasks(kcl, oracle):
    """Rows of several words and edge tasks (roots above the split run as
    (s, u) pairs; their cliques end with u, s)."""
    pairs = gg.bipartite_plus(150, 0.05, 9)
    og = oracle.dag(pairs)
    k = 5
    want = _all_cliques(og, k)
    _say("listing wide, cliques", want.shape[0], "core", og.core)
    with kcl.Graph(pairs) as g, kcl.Graph(pairs) as ge:
        ge.set_edge_split(64)
        with kcl.Sampling(g, k, target=10 ** 9) as a, kcl.Sampling(ge, k, target=10 ** 9) as b:
            assert np.array_equal(_sorted_rows(a.cliques()), want)
            assert np.array_equal(_sorted_rows(b.cliques()), want)


def test_sampler_keeps_by_its_hash(kcl, oracle):
    pairs = _community(5, size=26, p=0.9)
    og = oracle.dag(pairs)
    k = 5
    with kcl.Graph(pairs) as g:
        with kcl.Sampling(g, k, target=10 ** 9) as full:
            allc = _sorted_rows(full.cliques())
        total = allc.shape[0]
        target = total // 10
        _say("sampler: cliques", total, "target", target)
        p = target / total
        thr53 = int(np.ldexp(p, 53))
        for seed in (0, 7):
            with kcl.Sampling(g, k, target=target, seed=seed) as s:
                info = s.info
                assert info["total"] == total and abs(info["prob"] - p) < 1e-15
                got = _sorted_rows(s.cliques())
                assert np.array_equal(got, _kept_by_hash(allc, seed, thr53))
                sd = np.sqrt(total * p * (1 - p))
                assert abs(got.shape[0] - total * p) < 6 * sd
                if seed == 0:
                    first = got
        assert not np.array_equal(first, got)


def test_rounds_match_the_oracle(kcl, oracle):
    """With every clique kept, a round gives each clique's unit to its
    member of smallest load, ties to the smaller id: the oracle's
    Frank-Wolfe pass at rho = the loads."""
    pairs = _community(11)
    og = oracle.dag(pairs)
    k = 4
    with kcl.Graph(pairs) as g, kcl.Sampling(g, k, target=10 ** 9) as s:
        prev = np.zeros(og.n, np.uint64)
        for t in range(1, 6):
            s.step()
            cur = s.loads()
            want, _ = oracle.fw_counts(og, k, prev.astype(np.float64))
            assert np.array_equal(cur - prev, want), t
            prev = cur
        assert int(prev.sum()) == 5 * s.info["total"]


@pytest.mark.parametrize("k,target", [(4, 10 ** 9), (5, 2000)])
def test_report_matches_the_oracle(kcl, oracle, k, target):
    """Extraction: nodes by decreasing load (ties by id), each kept clique
    to its member of largest rank (the oracle's PAVA pass with level =
    rank when every clique is kept), the first strict maximum of the
    prefix density, and the exact clique count of the prefix's induced
    subgraph."""
    pairs = _community(20 + k, size=28, p=0.85)
    og = oracle.dag(pairs)
    with kcl.Graph(pairs) as g, kcl.Sampling(g, k, target=target, seed=3) as s:
        for _ in range(4):
            s.step()
        rec, order = s.report(want_order=True)
        loads = s.loads()
        ids = np.arange(og.n)
        assert np.array_equal(order, np.lexsort((ids, -loads.astype(np.int64))))
        rank = np.empty(og.n, np.int64)
        rank[order] = ids
        kept = s.cliques()
        if s.info["prob"] == 1.0:
            pushed = oracle.pava_counts(og, k, rank.astype(np.uint32))
        else:
            pushed = np.bincount(rank[kept].max(axis=1), minlength=og.n)
        cum = np.cumsum(pushed.astype(np.float64))
        dens = cum / np.arange(1, og.n + 1)
        size = int(np.argmax(dens)) + 1  # first maximum
        assert rec["prefix"] == size and rec["prefix_sampled"] == int(cum[size - 1])
        inside = rank < size
        es, et = og.es.astype(np.int64), og.et.astype(np.int64)
        sel = inside[es] & inside[et]
        sub = np.stack([rank[es[sel]], rank[et[sel]]], 1).astype(np.uint32)
        sg = oracle.dag(sub)
        assert rec["n"] == sg.n and rec["m"] == sub.shape[0]
        assert rec["cliques"] == oracle.count(sg, k)
        assert rec["density"] == rec["cliques"] / sg.n


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_prints_the_reference_lines(kcl, tmp_path):
    pairs = _community(31, size=26, p=0.9)
    path = str(tmp_path / "g.txt")
    gg.write_text(pairs, path)
    _say("sampling cli")
    r = subprocess.run([CLI, "4", "8", "5", path, "11"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [x for x in r.stdout.splitlines() if x.startswith("Approximate densest subgraph:")]
    assert len(lines) == 4  # T = 1, 2, 4, 8
    with kcl.Graph(pairs) as g, kcl.Sampling(g, 5, seed=11) as s:
        recs = s.run(8)
        sampled = s.info["sampled"]
    assert "Number of sampled 5-cliques: %d" % sampled in r.stdout
    for line, rec in zip(lines, recs):
        head = line.split(". ")[0] if ". " in line else line
        assert head.startswith("Approximate densest subgraph: %u nodes, %u edges, edge density = "
                               "%f, k-clique density = %f" % (rec["n"], rec["m"],
                                                              rec["edge_density"],
                                                              rec["density"]))
    with pytest.raises(kcl.KclError):
        with kcl.Graph(pairs) as g:
            kcl.Sampling(g, 2)
