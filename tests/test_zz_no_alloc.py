"""kcl_clique_enum_noalloc against kcl_clique_enum, through the C ABI: the
allocation-free pass (intermediates in the library's static device scratch)
must give the same accumulated vectors and visit counts, on two calls in a
row, the second on a smaller graph (the scratch keeps the first call's
contents), and must reject sizes beyond its scratch."""

import numpy as np
import pytest

from kclistpp_amd import graphgen as gg

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _community_graph(seed, n, m, size, p):
    pairs, rng = gg.chung_lu(n, m, 2.3, 200.0, seed)
    pairs = gg.plant_community(pairs, n, size, p, rng)
    return gg.finalize(pairs, rng)


def _both(kcl, g, k, task, rho=None, level=None, acc0=None):
    a0 = None if acc0 is None else acc0.copy()
    b0 = None if acc0 is None else acc0.copy()
    want, wv = g.clique_enum(k, task, rho=rho, level=level, acc=a0)
    got, gv = g.clique_enum_noalloc(k, task, rho=rho, level=level, acc=b0)
    return want, wv, got, gv


@pytest.mark.parametrize("k", [3, 5, 6])
def test_noalloc_matches_clique_enum(kcl, k):
    big = _community_graph(70 + k, 6000, 40000, 60, 0.8)
    small = _community_graph(80 + k, 900, 5000, 25, 0.8)
    rng = np.random.default_rng(k)
    for name, pairs in (("large", big), ("smaller", small)):
        with kcl.Graph(pairs) as g:
            rho = rng.integers(0, 4, g.n).astype(np.float64) * 0.25  # many ties
            acc0 = rng.integers(0, 1000, g.n).astype(np.float64)      # accumulated into
            _say("noalloc FW", name, "k", k, "n", g.n, "core", g.core)
            want, wv, got, gv = _both(kcl, g, k, kcl.FRANK_WOLFE, rho=rho, acc0=acc0)
            assert gv == wv and wv > 0
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
            level = rng.permutation(g.n).astype(np.uint32)
            _say("noalloc PAVA", name, "k", k)
            want, wv, got, gv = _both(kcl, g, k, kcl.PAVA_PREPROCESS, level=level, acc0=acc0)
            assert gv == wv
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_noalloc_trivial_and_limits(kcl):
    k5 = np.array([(i, j) for i in range(5) for j in range(i + 1, 5)], np.uint32)
    with kcl.Graph(k5) as g:
        _say("noalloc trivial")
        acc, v = g.clique_enum_noalloc(7, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert v == 0 and not acc.any()
        acc, v = g.clique_enum_noalloc(5, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        want, wv = g.clique_enum(5, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert v == wv == 1 and np.array_equal(acc, want) and acc.sum() == 1.0
        with pytest.raises(kcl.KclError) as e:
            g.clique_enum_noalloc(3, kcl.PAVA_PREPROCESS, level=np.zeros(g.n, np.uint32))
        assert e.value.name == "KCL_E_INVALID_ARG"
        with pytest.raises(kcl.KclError) as e:
            g.clique_enum_noalloc(3, kcl.FRANK_WOLFE, rho=np.full(g.n, -1.0))
        assert e.value.name == "KCL_E_INVALID_ARG"
    # A path of 3 million nodes needs 72 MB of node vectors: over the scratch.
    n = 3_000_000
    path = np.stack([np.arange(n - 1), np.arange(1, n)], 1).astype(np.uint32)
    with kcl.Graph(path) as g:
        _say("noalloc over the scratch, n", g.n)
        with pytest.raises(kcl.KclError) as e:
            g.clique_enum_noalloc(3, kcl.FRANK_WOLFE, rho=np.zeros(g.n))
        assert e.value.name == "KCL_E_UNSUPPORTED"
