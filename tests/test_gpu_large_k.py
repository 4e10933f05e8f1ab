"""GPU parity of the large-k enumeration (k above the compiled kernels, up
to the reference's unsigned char k: kCDensestCmpSync.c:425 reads k with
atoi into an unsigned char) against the oracle and closed forms: Frank-Wolfe
and PAVA credits, Sim iterations, edge tasks, k-clique degrees and the
trimming peel, for roots of every row width."""

from math import comb

import numpy as np
import pytest

from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _community(seed, n=1500, m=9000, size=30, p=0.95):
    pairs, rng = gg.chung_lu(n, m, 2.3, 150.0, seed)
    pairs = gg.plant_community(pairs, n, size, p, rng)
    return gg.finalize(pairs, rng)


def _complete(n):
    return np.array([(i, j) for i in range(n) for j in range(i + 1, n)], np.uint32)


def _wide_with_clique(part, size, seed):
    """K_{part,part} plus sparse edges inside each side and a planted
    clique: roots of ~part out-neighbours (the 10-word rows) whose cliques
    reach k > 10 only through the planted members."""
    pairs = gg.bipartite_plus(part, 0.02, seed)
    rng = np.random.default_rng(seed)
    pairs = gg.plant_clique(pairs, 2 * part, size, rng)
    return gg.finalize(pairs, rng)


@pytest.mark.parametrize("k", [11, 12, 14, 16])
def test_large_k_counts_vs_oracle(kcl, oracle, k):
    pairs = _community(300 + k)
    og = oracle.dag(pairs)
    rng = np.random.default_rng(k)
    rho = rng.integers(0, 4, og.n).astype(np.float64) * 0.25  # many ties
    level = rng.permutation(og.n).astype(np.uint32)
    _say("large k", k, "core", og.core)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert wv > 0 and visits == wv
        assert np.array_equal(acc, want.astype(np.float64))
        acc, _ = g.clique_enum(k, kcl.PAVA_PREPROCESS, level=level)
        assert np.array_equal(acc, oracle.pava_counts(og, k, level).astype(np.float64))


@pytest.mark.parametrize("n,k", [(20, 11), (20, 15), (20, 20), (20, 21), (60, 58), (66, 64)])
def test_large_k_complete_graph_closed_form(kcl, n, k):
    """K_n at rho = 0: C(n, k) cliques, relabelled node i wins the cliques
    whose smallest member it is, C(n-1-i, k-1) of them."""
    _say("K", n, "k", k)
    with kcl.Graph(_complete(n)) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert visits == comb(n, k)
        want = np.array([comb(n - 1 - i, k - 1) for i in range(n)], np.float64)
        assert np.array_equal(acc, want)


@pytest.mark.parametrize("k", [11, 13])
def test_large_k_wide_rows_vs_oracle(kcl, oracle, k):
    pairs = _wide_with_clique(560, 16, 900 + k)
    og = oracle.dag(pairs)
    assert og.core > 512
    rng = np.random.default_rng(3)
    rho = rng.integers(0, 3, og.n).astype(np.float64)
    _say("large k wide", k, "core", og.core)
    with kcl.Graph(pairs) as g:
        acc, visits = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert wv > 0 and visits == wv
        assert np.array_equal(acc, want.astype(np.float64))


def test_large_k_edge_tasks_match_vertex_tasks(kcl, oracle):
    """Roots above the edge-split threshold run as edge tasks of the
    (k-1)-clique walk over N+(s) & N+(u)."""
    pairs = _wide_with_clique(300, 15, 77)
    og = oracle.dag(pairs)
    rng = np.random.default_rng(8)
    k = 12
    _say("large k edge tasks")
    with kcl.Graph(pairs) as g, kcl.Graph(pairs) as ge:
        rho = rng.integers(0, 4, g.n).astype(np.float64) * 0.5
        ge.set_edge_split(64)
        a, va = g.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        b, vb = ge.clique_enum(k, kcl.FRANK_WOLFE, rho=rho)
        want, wv = oracle.fw_counts(og, k, rho)
        assert va == vb == wv > 0
        assert np.array_equal(a, b) and np.array_equal(a, want.astype(np.float64))


@pytest.mark.parametrize("schedule", ["degree", "counted"])
def test_large_k_sim_iterations_vs_oracle(kcl, oracle, schedule):
    k = 12
    pairs = _community(41, size=28, p=0.95)
    og = oracle.dag(pairs)
    rho = np.zeros(og.n)
    _say("large k sim", schedule)
    with kcl.Graph(pairs) as g, kcl.Sim(g, k, schedule=kcl.SCHEDULES[schedule]) as s:
        for t in (1, 2, 3):
            s.step(t)
            want, _ = oracle.fw_counts(og, k, rho)
            assert np.array_equal(s.counts(), want)
            rho = oracle.rho_update(rho, t, want)
            assert np.array_equal(s.rho().view(np.uint64), rho.view(np.uint64))
        d = s.find_densest()
        od, _, _, _ = oracle.find_densest(og, k, rho)
        assert (d.n, d.m, d.density, d.ub) == (od["n"], od["m"], od["density"], od["ub"])


@pytest.mark.parametrize("k", [11, 12])
def test_large_k_trim_vs_oracle(kcl, oracle, k):
    """k-clique degrees (every member credited) and the peel at large k."""
    pairs = _community(500 + k, n=1200, m=7000, size=26, p=0.95)
    info_o, mp_o, nck_o, pk_o, pc_o, trimmed_o = oracle.trim(pairs, k)
    assert info_o["ncktot"] > 0
    _say("large k trim", k)
    for gpu_min in (0, 1):
        with kcl.Trim(pairs, k, gpu_peel_min_cliques=gpu_min) as t:
            mp, nck, pk, pc, kept = t.nodes()
            assert np.array_equal(mp, mp_o) and np.array_equal(nck, nck_o)
            assert np.array_equal(pk, pk_o) and np.array_equal(pc, pc_o)
            assert np.array_equal(t.edges(), trimmed_o)
            for f in ("n2", "e2", "ncktot", "ckcore", "ckrho", "rho", "size", "ntrim",
                      "ncktrim", "mtrim"):
                assert t.info[f] == info_o[f], f


def test_large_k_degrees_on_narrow_rows(kcl, oracle):
    """k = 58 on K_60 plus a pendant path: the k-clique-degree walk of
    58 levels does not fit a 64-member row's share of the keys, so those
    roots run on the next wider class's kernel."""
    pairs = np.concatenate([_complete(60), np.array([[59, 60], [60, 61]], np.uint32)])
    info_o, mp_o, nck_o, pk_o, pc_o, _ = oracle.trim(pairs, 58)
    assert info_o["ncktot"] == comb(60, 58)
    _say("K60 k=58 trim")
    with kcl.Trim(pairs, 58) as t:
        mp, nck, pk, pc, _ = t.nodes()
        assert np.array_equal(mp, mp_o) and np.array_equal(nck, nck_o)
        assert np.array_equal(pk, pk_o) and np.array_equal(pc, pc_o)


def test_k_limits(kcl):
    with kcl.Graph(_complete(12)) as g:
        acc, v = g.clique_enum(255, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert v == 0 and not acc.any()
        with pytest.raises(kcl.KclError) as e:
            g.clique_enum(256, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert e.value.name == "KCL_E_INVALID_ARG"
