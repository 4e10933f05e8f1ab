"""GPU parity of the batched path (many graphs of mixed k in one captured
hipGraph per iteration): every graph's per-iteration counts, rho bits,
totals and reports equal those of the single-graph path and of the
oracle, with each replay checked against eager work done between
replays."""

import math

import numpy as np
import pytest

from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


def _small_specs(count, seed, n_max):
    """The mixed batch's distribution (graphgen.batch_specs) with n scaled
    into [n_max/200, n_max]: Chung-Lu, average degree 4..40, gamma 2.1..2.8,
    k = 3 + (i mod 5)."""
    out = []
    for sp in gg.batch_specs(count, first_seed=seed):
        n = max(60, int(sp["n"] * n_max / 2e5))
        deg = sp["m"] * 2 / sp["n"]
        m = min(int(n * deg / 2), n * (n - 1) // 4)
        out.append(dict(sp, n=n, m=m, w_max=float(math.sqrt(n * deg))))
    return out


@pytest.fixture(scope="module")
def small(kcl):
    specs = _small_specs(20, 3000, 4000)
    pairs = [gg.batch_graph(sp) for sp in specs]
    # extra cases: a clique-free k (k > core + 1) and k = 2
    pairs.append(gg.erdos_renyi(300, 280, 5)[0])
    specs.append({"k": 9})
    pairs.append(gg.batch_graph(_small_specs(1, 77, 500)[0]))
    specs.append({"k": 2})
    graphs = [kcl.Graph(p) for p in pairs]
    ks = [sp["k"] for sp in specs]
    yield pairs, graphs, ks
    for g in graphs:
        g.close()


def test_batch_iterations_match_single_graph_path(kcl, small):
    pairs, graphs, ks = small
    _say("batch of", len(graphs), "graphs, k", ks)
    with kcl.Batch(graphs, ks) as b:
        cl = b.graph_cliques()
        sims = [kcl.Sim(g, k) for g, k in zip(graphs, ks)]
        try:
            for t in range(1, 9):
                b.step(t)
                visits, _ = b.last_pass()
                assert visits == int(cl.sum())
                for i, s in enumerate(sims):
                    s.step(t)  # eager work between the replays
                    assert np.array_equal(b.counts(i), s.counts()), (t, i)
                    assert np.array_equal(_bits(b.rho(i)), _bits(s.rho())), (t, i)
                    assert np.array_equal(b.totals(i), s.totals()), (t, i)
                if t & (t - 1) == 0:
                    reps = b.find_densest()
                    for i, s in enumerate(sims):
                        want = s.find_densest()
                        got = reps[i]
                        assert (got.n, got.m, got.sum) == (want.n, want.m, want.sum), (t, i)
                        assert _bits(got.density) == _bits(want.density)
                        assert _bits(got.ub) == _bits(want.ub)
            for i, s in enumerate(sims):
                assert cl[i] == s.plan_info()["cliques"]
        finally:
            for s in sims:
                s.close()


def test_batch_run_matches_oracle(kcl, oracle, small):
    pairs, graphs, ks = small
    with kcl.Batch(graphs, ks) as b:
        cl = b.graph_cliques()
        for _ in range(3):  # state that reset() must clear
            b.step()
        b.reset()
        recs = b.run(10)
    for i, (p, k) in enumerate(zip(pairs, ks)):
        if cl[i] > 300_000:  # keep the single-threaded oracle quick
            continue
        dag = oracle.dag(p)
        want, _, _ = oracle.cmpsync(dag, k, 10)
        got = recs[i]
        assert [r["T"] for r in got] == [r["T"] for r in want] == [1, 2, 4, 8]
        for a, w in zip(got, want):
            assert (a["n"], a["m"], a["sum"]) == (w["n"], w["m"], w["sum"]), (i, a["T"])
            assert _bits(a["density"]) == _bits(w["density"])
            assert _bits(a["ub"]) == _bits(w["ub"])


def test_batch_counts_match_oracle_per_iteration(kcl, oracle, small):
    pairs, graphs, ks = small
    with kcl.Batch(graphs, ks) as b:
        cl = b.graph_cliques()
    pick = [i for i in range(20) if cl[i] <= 200_000][:5] + [20, 21]
    with kcl.Batch([graphs[i] for i in pick], [ks[i] for i in pick]) as b:
        got = []
        for t in range(1, 5):
            b.step(t)
            got.append([b.counts(j) for j in range(len(pick))])
    for j, i in enumerate(pick):
        _, counts, _ = oracle.cmpsync(oracle.dag(pairs[i]), ks[i], 4, keep_counts=True)
        for t in range(4):
            assert np.array_equal(got[t][j], counts[t]), (i, t)


def test_configurations_do_not_change_results(kcl, small):
    _, graphs, ks = small
    res = []
    import os
    model = kcl.CostModel(path=os.path.join(kcl.PKG, "cost_model.kclcm"))
    for opts in ({}, {"tune": True}, {"row_bits": 256, "per_cu": 8},
                 {"schedule": kcl.SCHED_DEGREE, "row_bits": 128},
                 {"schedule": kcl.SCHED_LEARNED, "model": model}):
        with kcl.Batch(graphs, ks, **opts) as b:
            for _ in range(3):
                b.step()
            info = b.info()
            _say(opts, info)
            res.append([b.counts(i) for i in range(len(graphs))])
            if opts.get("row_bits"):
                assert info["row_bits"] == opts["row_bits"]
    model.close()
    for r in res[1:]:
        for a, b_ in zip(res[0], r):
            assert np.array_equal(a, b_)


def test_batch_rejects_bad_input(kcl, small):
    _, graphs, ks = small
    with pytest.raises(ValueError):
        kcl.Batch(graphs[:2], ks[:1])
    with pytest.raises(kcl.KclError):
        kcl.Batch(graphs[:1], [1])


def _batch_graph_of(i):
    spec = list(gg.batch_specs(512, first_seed=1000))[i]
    return gg.batch_graph(spec)


@pytest.mark.timeout(900)
def test_batch512_matches_reference(kcl):
    """The mixed batch as benchmarked: all 512 Chung-Lu graphs (n up to
    2e5, k = 3..7) in one batch, against digests of the reference's own run
    of every graph: ranks, DAG edges, the count vector of iterations 1-4,
    the reports at T = 1, 2, 4 and the final rho bits."""
    import concurrent.futures as cf
    import hashlib
    import multiprocessing as mp
    from conftest import golden_index, load_golden

    if "batch" not in golden_index():
        pytest.skip("no batch golden case")
    gd = load_golden(golden_index()["batch"][0])
    T = int(gd["num_iter"])
    specs = list(gg.batch_specs(512, first_seed=1000))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    _say("batch512: generating 512 graphs")
    with cf.ProcessPoolExecutor(8, mp_context=mp.get_context("spawn")) as ex:
        pairs = list(ex.map(_batch_graph_of, range(512), chunksize=8))
    graphs = []
    try:
        for i, p in enumerate(pairs):
            assert sha(p) == str(gd["pairs_sha"][i][0]), f"graph {i} input differs"
            g = kcl.Graph(p)
            graphs.append(g)
            assert sha(g.rank()) == str(gd["rank_sha"][i][0]), f"graph {i} ranks differ"
            assert sha(g.edges()) == str(gd["edges_sha"][i][0]), f"graph {i} DAG differs"
        del pairs
        ks = [sp["k"] for sp in specs]
        _say("batch512: building the batch")
        with kcl.Batch(graphs, ks, tune=True) as b:
            t, Tp, rep = 1, 1, 0
            while Tp <= T:
                while t <= Tp:
                    b.step(t)
                    for i in range(512):
                        assert sha(b.counts(i)) == str(gd["count_sha"][i][t - 1]), \
                            f"graph {i} (k={ks[i]}) counts differ at t={t}"
                    t += 1
                dens = b.find_densest()
                for i, d in enumerate(dens):
                    got = [Tp, d.n, d.m, int(_bits(d.density)), int(_bits(d.ub))]
                    assert got == gd["records"][i][rep].tolist(), f"graph {i} report T={Tp}"
                rep += 1
                Tp <<= 1
            for i in range(512):
                assert sha(b.rho(i).view(np.uint64)) == str(gd["rho_sha"][i][0]), \
                    f"graph {i} rho differs"
    finally:
        for g in graphs:
            g.close()
