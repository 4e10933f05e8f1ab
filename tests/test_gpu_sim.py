"""GPU parity of the Sim-kClist++ path against the reference's golden vectors
and the oracle.  Bit-exact: ranks, DAG, per-iteration uint64 counts, rho
bits, level/tent, and the report records (n, m, density bits, ub bits)."""

import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_index, golden_pairs, load_golden
from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu

SIM = golden_index()["sim"]
BIG = golden_index()["big"]
WIDE = golden_index()["wide"]


def _bits(x):
    return int(np.float64(x).view(np.uint64))


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", SIM)
def test_graph_build_matches_reference(kcl, case):
    gd = load_golden(case)
    _say("graph", case)
    with kcl.Graph(gd["pairs"]) as g:
        assert (g.n, g.core) == (int(gd["n"]), int(gd["core"]))
        assert np.array_equal(g.rank(), gd["rank"])
        e = g.edges()
        assert np.array_equal(e[:, 0], gd["es"]) and np.array_equal(e[:, 1], gd["et"])
        cd, adj = g.csr()
        assert cd[-1] == g.m and np.array_equal(adj, gd["et"])


@pytest.mark.parametrize("case", SIM)
def test_sim_iterations_match_reference(kcl, case):
    gd = load_golden(case)
    k, T = int(gd["k"]), int(gd["num_iter"])
    _say("sim", case)
    with kcl.Graph(gd["pairs"]) as g, kcl.Sim(g, k) as s:
        recs = []
        t = 1
        Tp = 1
        while Tp <= T:
            while t <= Tp:
                s.step(t)
                c = s.counts()
                assert np.array_equal(c, gd["counts"][t - 1]), f"counts differ at t={t}"
                visits, _, _ = s.last_pass()
                assert visits == int(gd["counts"][t - 1].sum())
                t += 1
            info, reo, level, tent = s.find_densest(want_vectors=True)
            recs.append([Tp, info.n, info.m, _bits(info.density), _bits(info.ub)])
            Tp <<= 1
        assert np.array_equal(s.rho().view(np.uint64), gd["rho_bits"])
        assert np.array_equal(np.array(recs, np.uint64), gd["records"])
        assert np.array_equal(level, gd["level"])
        assert np.array_equal(tent, gd["tent"])
        # Exact totals: rho_t = 2 V_t / ((t+1)(t+2)).
        V = s.totals()
        tt = t - 1
        assert np.array_equal(V, sum((i + 1) * gd["counts"][i - 1].astype(np.uint64)
                                     for i in range(1, tt + 1)))


@pytest.mark.parametrize("case", [c for c in SIM if "cl3k" in c or "k7" in c])
def test_cmpsync_driver_matches_reference(kcl, case, tmp_path):
    gd = load_golden(case)
    k, T = int(gd["k"]), int(gd["num_iter"])
    _say("cmpsync", case)
    stat = str(tmp_path / "stat.txt")
    with kcl.Graph(gd["pairs"]) as g:
        recs = kcl.cmpsync(g, k, T, stat_path=stat)
    got = np.array([[r["T"], r["n"], r["m"], _bits(r["density"]), _bits(r["ub"])] for r in recs],
                   np.uint64)
    assert np.array_equal(got, gd["records"])
    lines = open(stat).read().splitlines()
    assert lines[0].startswith("[Number of Iterations]")
    for line, r in zip(lines[1:], gd["records"]):
        f = line.split("\t")
        want = "%u\t%u\t%u\t%.12f\t%.12f" % (r[0], r[1], r[2], np.uint64(r[3]).view(np.float64),
                                            np.uint64(r[4]).view(np.float64))
        assert "\t".join(f[:5]) == want


CLI = os.path.join(ROOT, "kclistpp_amd", "bin", "kCDensestCmpSync")


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_drop_in_stat_file(kcl, tmp_path):
    gd = load_golden("sim_cl3k_k4_T8")
    path = str(tmp_path / "g.txt")
    gg.write_text(gd["pairs"], path)
    _say("cli cl3k k=4 T=10")
    r = subprocess.run([CLI, "8", "10", "4", path, "cl3k"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "stat_cmpsync_cl3k_8_10_4.txt").read().splitlines()
    assert len(lines) == 1 + len(gd["records"])
    for line, rec in zip(lines[1:], gd["records"]):
        f = line.split("\t")
        assert [int(f[0]), int(f[1]), int(f[2])] == rec[:3].tolist()
        assert f[3] == "%.12f" % np.uint64(rec[3]).view(np.float64)
        assert f[4] == "%.12f" % np.uint64(rec[4]).view(np.float64)


_BIG_PAIRS = {}


def _big_pairs(gd):
    key = (str(gd["workload"][0]), int(gd["seed"]), int(gd["scale_down"]))
    if key not in _BIG_PAIRS:
        pairs = gg.workload(key[0], seed=key[1], scale_down=key[2])
        assert hashlib.sha256(pairs.tobytes()).hexdigest() == str(gd["pairs_sha"][0])
        _BIG_PAIRS[key] = pairs
    return _BIG_PAIRS[key]


def _run_digests(kcl, s, T, after_first=None):
    shas, recs = [], []
    t, Tp = 1, 1
    while Tp <= T:
        while t <= Tp:
            s.step(t)
            shas.append(hashlib.sha256(s.counts().tobytes()).hexdigest())
            if t == 1 and after_first is not None:
                after_first(s)
            t += 1
        d = s.find_densest()
        recs.append([Tp, d.n, d.m, _bits(d.density), _bits(d.ub)])
        Tp <<= 1
    return shas, np.array(recs, np.uint64)


# Schedules of the workload cases: counts must not depend on the plan.
#   degree   the plain schedule (d+^2 order and splits)
#   counted  costs from a counting pass (the default)
#   learned  the shipped cost model
#   online   the shipped model retrained on the graph's own first-pass
#            timings, the plan rebuilt with it after iteration 1
_SCHEDULES = ["degree", "counted", "learned", "online"]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("schedule", _SCHEDULES)
@pytest.mark.parametrize("case", BIG)
def test_full_size_workloads_match_reference(kcl, case, schedule):
    """The benchmark workloads (the k = 4 and k = 5 cases at full size, the
    k = 6 exact case's Chung-Lu family with its 200-vertex near-clique, the
    R-MAT case at scales 10 and 11 with k = 7, 8) against digests of the
    reference's own vectors: ranks, DAG edges, every iteration's count
    vector, the final rho bits and each report record."""
    from kclistpp_amd import train_cost_model as tcm
    gd = load_golden(case)
    pairs = _big_pairs(gd)
    k, T = int(gd["k"]), int(gd["num_iter"])
    _say("big", case, schedule)
    model = kcl.CostModel(path=os.path.join(kcl.PKG, "cost_model.kclcm")) \
        if schedule in ("learned", "online") else None
    sched = kcl.SCHEDULES["learned" if schedule == "online" else schedule]
    retrained = []

    def retrain(s):
        retrained.append(tcm.online(s, model, epochs=10))

    with kcl.Graph(pairs) as g, kcl.Sim(g, k, schedule=sched, model=model) as s:
        assert hashlib.sha256(g.rank().tobytes()).hexdigest() == str(gd["rank_sha"][0])
        assert hashlib.sha256(g.edges().tobytes()).hexdigest() == str(gd["edges_sha"][0])
        shas, recs = _run_digests(kcl, s, T, retrain if schedule == "online" else None)
        assert shas == [str(x) for x in gd["count_sha"]]
        assert np.array_equal(recs, gd["records"])
        assert hashlib.sha256(s.rho().view(np.uint64).tobytes()).hexdigest() == \
            str(gd["rho_sha"][0])
    for m in retrained + ([model] if model is not None else []):
        m.close()


# ---------------------------------------------------------------------------
# Against the oracle: the in-process boundary with arbitrary rho / level.

def _community_graph(seed, n=2000, m=12000, size=40, p=0.8):
    pairs, rng = gg.chung_lu(n, m, 2.3, 200.0, seed)
    pairs = gg.plant_community(pairs, n, size, p, rng)
    return gg.finalize(pairs, rng)


@pytest.mark.parametrize("k", list(range(2, 11)))
def test_clique_enum_boundary_vs_oracle(kcl, oracle, k):
    pairs = _community_graph(100 + k)
    og = oracle.dag(pairs)
    rng = np.random.default_rng(k)
    rho = rng.integers(0, 4, og.n).astype(np.float64) * 0.25  # many ties
    _say("clique_enum FW k", k)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert visits == wv
        assert np.array_equal(acc, want.astype(np.float64))
        level = rng.permutation(og.n).astype(np.uint32)
        _say("clique_enum PAVA k", k)
        acc, _ = g.clique_enum(k, kcl.PAVA_PREPROCESS, level=level)
        assert np.array_equal(acc, oracle.pava_counts(og, k, level).astype(np.float64))


def _bipartite_plus(part, p_in, seed):
    """Complete bipartite K_{part,part} plus sparse edges inside each side:
    out-degrees reach ~part (the widest word classes) with few cliques."""
    return gg.bipartite_plus(part, p_in, seed)


@pytest.mark.parametrize("part,k", [(300, 3), (300, 4), (300, 5), (560, 3), (560, 4), (560, 5)])
def test_wide_neighbourhoods_vs_oracle(kcl, oracle, part, k):
    pairs = _bipartite_plus(part, 0.02, part + k)
    og = oracle.dag(pairs)
    assert og.core > (256 if part == 300 else 512)
    rng = np.random.default_rng(7)
    rho = rng.integers(0, 3, og.n).astype(np.float64)
    _say("wide", part, k, "core", og.core)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert visits == wv and np.array_equal(acc, want.astype(np.float64))


@pytest.mark.parametrize("size,p,k,above", [
    (40, 0.8, 6, 1024), (40, 0.8, 8, 1024), (40, 0.8, 10, 1024),  # 64-member rows
    (200, 0.6, 6, 1024), (200, 0.6, 7, 1024),                      # 128 (core ~104)
    (500, 0.3, 8, 1024),                                           # 192 (core ~130)
    (800, 0.33, 6, 1024), (800, 0.33, 9, 1024),                    # 256 (core ~236)
])
def test_dense_counting_vs_oracle(kcl, oracle, size, p, k, above):
    """Threshold 0: every root of rows of 4+ words counted a wavefront per
    item (members ranked by (key, id), the winner as a rank prefix, the
    lanes sharing the last level); narrower rows keep the per-lane path.
    Many key ties."""
    pairs = _community_graph(size + k, size=size, p=p)
    og = oracle.dag(pairs)
    rng = np.random.default_rng(k)
    rho = rng.integers(0, 4, og.n).astype(np.float64) * 0.25
    level = rng.permutation(og.n).astype(np.uint32)
    _say("dense size", size, "k", k, "core", og.core, "edge split", above)
    with kcl.Graph(pairs) as g:
        g.set_dense_threshold(0)
        g.set_edge_split(above)
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert visits == wv and np.array_equal(acc, want.astype(np.float64))
        acc, _ = g.clique_enum(k, kcl.PAVA_PREPROCESS, level=level)
        assert np.array_equal(acc, oracle.pava_counts(og, k, level).astype(np.float64))


def test_dense_edge_tasks_match_vertex_tasks(kcl):
    """A 350-member community with p = 0.9 at k = 7 (1.5e13 cliques, past
    the oracle's reach): edge tasks of ~250 common out-neighbours, counted
    densely by the ranked k = 6 kernel, against the vertex tasks."""
    pairs = _community_graph(21, size=350, p=0.9)
    rng = np.random.default_rng(5)
    _say("dense edge tasks k 7")
    with kcl.Graph(pairs) as g, kcl.Graph(pairs) as ge:
        rho = rng.integers(0, 4, g.n).astype(np.float64) * 0.25
        ge.set_edge_split(200)
        a, va = g.clique_enum(7, kcl.FRANK_WOLFE, rho=rho)
        b, vb = ge.clique_enum(7, kcl.FRANK_WOLFE, rho=rho)
        assert va == vb and va > 10 ** 13 and np.array_equal(a, b)


@pytest.mark.parametrize("k", [6, 7])
def test_wide_dense_rows_vs_oracle(kcl, oracle, k):
    """Rows of 10 words at k >= 6 are always counted densely, here on
    sparse out-neighbourhoods (K_{560,560} plus sparse edges per side)."""
    pairs = _bipartite_plus(560, 0.02, 600 + k)
    og = oracle.dag(pairs)
    assert og.core > 512
    rng = np.random.default_rng(11)
    rho = rng.integers(0, 3, og.n).astype(np.float64)
    _say("wide dense", k, "core", og.core)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert visits == wv and np.array_equal(acc, want.astype(np.float64))


@pytest.mark.parametrize("threshold", [64.0, 0.0])
@pytest.mark.parametrize("case", WIDE)
def test_wide_word_classes_match_reference(kcl, case, threshold):
    """Roots in the 6-, 8-, 12- and 16-word classes at k = 6, 7, 8 against
    the reference's own passes (FW at a rho with ties, PAVA at a level
    permutation), with the dense counting threshold at its default and at
    0 (every ranked root counted a wavefront per item)."""
    gd = load_golden(case)
    pairs = golden_pairs(gd)
    k = int(gd["k"])
    _say("wide", case, "threshold", threshold)
    with kcl.Graph(pairs) as g:
        assert g.core == int(gd["core"])
        g.set_dense_threshold(threshold)
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=gd["rho"])
        assert visits == int(gd["visits"])
        assert np.array_equal(acc, gd["fw"].astype(np.float64))
        acc, _ = g.clique_enum(k, kcl.PAVA_PREPROCESS, level=gd["level"])
        assert np.array_equal(acc, gd["tent"].astype(np.float64))


def test_counts_do_not_depend_on_dense_threshold(kcl):
    pairs = _community_graph(11, size=200, p=0.6)
    _say("dense threshold in sim")
    with kcl.Graph(pairs) as g, kcl.Graph(pairs) as h:
        h.set_dense_threshold(0)
        with pytest.raises(kcl.KclError):
            h.set_dense_threshold(-1)
        a, b = kcl.Sim(g, 6), kcl.Sim(h, 6, split_work=7)
        for t in (1, 2, 3):
            a.step(t)
            b.step(t)
            assert np.array_equal(a.counts(), b.counts())
            assert np.array_equal(a.rho().view(np.uint64), b.rho().view(np.uint64))
        assert a.find_densest() == b.find_densest()
        a.close()
        b.close()


@pytest.mark.parametrize("k,split,size", [(5, 1, 60), (5, 7, 60), (5, 100, 60), (3, 1, 60),
                                          (4, 1, 60), (6, 1, 60), (11, 1, 30)])
def test_counts_do_not_depend_on_splitting(kcl, k, split, size):
    """Split roots deal their items (members at k <= 4 and k > 10, edges
    otherwise) out to their parts in blocks; with split_work = 1 small roots
    get more parts than blocks of the full size, and every item must still
    be counted once."""
    pairs = _community_graph(5, size=size, p=0.9)
    _say("split", split, "k", k)
    with kcl.Graph(pairs) as g:
        base = kcl.Sim(g, k)
        other = kcl.Sim(g, k, split_work=split)
        _say("tasks", base.plan_info()["tasks"], "split into", other.plan_info()["tasks"])
        for t in (1, 2, 3):
            base.step(t)
            other.step(t)
            assert np.array_equal(base.counts(), other.counts())
            assert np.array_equal(base.rho().view(np.uint64), other.rho().view(np.uint64))
        base.close()
        other.close()


def test_trivial_and_empty_graphs(kcl):
    k5 = np.array([(i, j) for i in range(5) for j in range(i + 1, 5)], np.uint32)
    with kcl.Graph(k5) as g, kcl.Sim(g, 7) as s:  # k > core + 1: no cliques
        s.step(1)
        assert s.counts().sum() == 0
        d = s.find_densest()
        assert d.n == 1 and d.density == 0.0
    with kcl.Graph(np.zeros((0, 2), np.uint32)) as g:
        assert g.n == 0 and g.m == 0
        assert kcl.cmpsync(g, 3, 4) and True
    with pytest.raises(kcl.KclError) as e:
        kcl.Graph(np.array([[3, 3]], np.uint32))
    assert e.value.name == "KCL_E_INVALID_ARG"


@pytest.mark.parametrize("above,k", [(8, 3), (8, 4), (8, 5), (12, 6), (12, 7), (30, 5), (30, 8)])
def test_edge_tasks_vs_oracle(kcl, oracle, above, k):
    """Roots above the threshold run as edge tasks (s, u) over
    N+(s) & N+(u) with the (k-1) kernels; counts must not change."""
    pairs = _community_graph(300 + k, size=50, p=0.8)
    og = oracle.dag(pairs)
    rng = np.random.default_rng(k)
    rho = rng.integers(0, 4, og.n).astype(np.float64) * 0.25
    level = rng.permutation(og.n).astype(np.uint32)
    _say("edge tasks above", above, "k", k, "core", og.core)
    assert og.core > above
    with kcl.Graph(pairs) as g:
        g.set_edge_split(above)
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert visits == wv and np.array_equal(acc, want.astype(np.float64))
        acc, _ = g.clique_enum(k, kcl.PAVA_PREPROCESS, level=level)
        assert np.array_equal(acc, oracle.pava_counts(og, k, level).astype(np.float64))


def test_edge_tasks_in_sim_trim_and_exact(kcl):
    pairs = _community_graph(41, size=45, p=0.85)
    with kcl.Graph(pairs) as g, kcl.Graph(pairs) as ge:
        ge.set_edge_split(10)
        for k in (4, 6):
            a, b = kcl.Sim(g, k), kcl.Sim(ge, k)
            assert b.plan_info()["tasks"] > a.plan_info()["tasks"]
            for t in (1, 2, 3, 4):
                a.step(t)
                b.step(t)
                assert np.array_equal(a.counts(), b.counts())
                assert np.array_equal(a.rho().view(np.uint64), b.rho().view(np.uint64))
            assert a.find_densest() == b.find_densest()
            # cost samples aggregate edge tasks per root
            roots, x, y = b.cost_samples()
            assert len(roots) == b.plan_info()["roots_kept"] and np.all(np.isfinite(y))
            a.close()
            b.close()
        ra, na = kcl.exact(g, 4)
        rb, nb = kcl.exact(ge, 4)
        assert ra["certified"] == rb["certified"] == 1
        assert (ra["n"], ra["cliques"], ra["m"]) == (rb["n"], rb["cliques"], rb["m"])
        assert np.array_equal(na, nb)


def test_degeneracy_above_1024(kcl, oracle):
    """Complete tripartite K_{520,520,520}: every node has DAG out-degree
    up to 1040 > 1024, common out-neighbourhoods stay within one part."""
    a = 520
    parts = [np.arange(i * a, (i + 1) * a) for i in range(3)]
    pairs = np.concatenate([
        np.stack(np.meshgrid(parts[i], parts[j], indexing="ij"), -1).reshape(-1, 2)
        for i, j in ((0, 1), (0, 2), (1, 2))]).astype(np.uint32)
    rng = np.random.default_rng(3)
    pairs = gg.finalize(pairs, rng)
    og = oracle.dag(pairs)
    assert og.core > 1024
    rho = rng.integers(0, 3, og.n).astype(np.float64)
    _say("tripartite core", og.core)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(3, kcl.FRANK_WOLFE, rho=rho)
        assert visits == a ** 3
        want, wv = oracle.fw_counts(og, 3, rho)
        assert wv == a ** 3 and np.array_equal(acc, want.astype(np.float64))
        acc4, v4 = g.clique_enum(4, kcl.FRANK_WOLFE, rho=rho)
        assert v4 == 0 and not acc4.any()
        with kcl.Sim(g, 3) as s:
            s.step()
            assert s.counts().sum() == a ** 3
