"""CPU model of the k >= 6 dense counting against the oracle.

The kernel (kclistpp_amd/csrc/pass_kernel.h, ranked kernels) relies on
three facts that this model uses the same way, so a mismatch with the
oracle (the restated kCDensestCmpSync.c:138-325 enumeration) would show a
flaw in the method rather than in the GPU code:
* numbering a root's members by (key, id) rank, every oriented row holds
  only later ranks, so the count of a member x's row inside a set S needs
  only the words of S at or above x's own word;
* the winner of a clique is its member of smallest rank when that rank is
  below nwin = #{members with key <= key_s}, else the root;
* the k-cliques with root s and source edge (u, v) are the (k-3)-cliques
  inside A'[u] & A'[v], counted as nested sets whose last two levels are
  the edges inside the final set.
"""

import numpy as np
import pytest

from kclistpp_amd import graphgen as gg


def _bits(x):
    while x:
        low = x & -x
        yield low.bit_length() - 1
        x ^= low


def _from_word(S, j):
    """S without its words below j."""
    return (S >> (64 * j)) << (64 * j)


def _edges_in(S, rows, skip):
    # Each member x of S counts its row over the words of S from x's word on
    # (skip > 0 plants a mistake: it starts that many words later).
    return sum(bin(_from_word(S, x // 64 + skip) & rows[x]).count("1") for x in _bits(S))


def _cliques(S, J, rows, skip):
    if J == 2:
        return _edges_in(S, rows, skip)
    total = 0
    for w in _bits(S):
        Q = _from_word(S, w // 64) & rows[w]
        if bin(Q).count("1") >= J - 1:
            total += _cliques(Q, J - 1, rows, skip)
    return total


def dense_model_counts(g, k, rho, skip=0, win_extra=0):
    """Per-node credits of one Frank-Wolfe pass (skip, win_extra: planted
    mistakes for the tests that show this model can fail)."""
    cnt = np.zeros(g.n, np.uint64)
    for s in range(g.n):
        L = g.adj[g.cd[s]:g.cd[s + 1]]
        d = len(L)
        if d < k - 1:
            continue
        keys = rho[L]
        order = sorted(range(d), key=lambda i: (keys[i], i))  # index order is id order
        rank = [0] * d
        for r, i in enumerate(order):
            rank[i] = r
        pos = {int(x): i for i, x in enumerate(L)}
        rows = [0] * d
        for i, x in enumerate(L):
            for y in g.adj[g.cd[x]:g.cd[x + 1]]:
                j = pos.get(int(y))
                if j is not None:
                    a, b = sorted((rank[i], rank[j]))
                    rows[a] |= 1 << b
        nwin = int(np.count_nonzero(keys <= rho[s])) + win_extra
        node_of = [int(L[i]) for i in order]
        for u in range(d):
            for v in _bits(rows[u]):
                P = rows[u] & rows[v]
                if bin(P).count("1") < k - 3:
                    continue
                c = _cliques(P, k - 3, rows, skip)
                cnt[node_of[u] if u < nwin else s] += np.uint64(c)
    return cnt


def _bipartite_plus(part, p_in, seed):
    return gg.bipartite_plus(part, p_in, seed)


def _community(seed):
    pairs, rng = gg.chung_lu(300, 1500, 2.3, 60.0, seed)
    pairs = gg.plant_community(pairs, 300, 45, 0.7, rng)
    return gg.finalize(pairs, rng)


@pytest.mark.parametrize("graph,k", [("community", 6), ("community", 7),
                                     ("bipartite", 6), ("bipartite", 7)])
def test_dense_model_matches_oracle(oracle, graph, k):
    pairs = _community(31 + k) if graph == "community" else _bipartite_plus(70, 0.12, 40 + k)
    og = oracle.dag(pairs)
    if graph == "bipartite":
        assert og.core > 64  # rows of two words
    rng = np.random.default_rng(k)
    rho = rng.integers(0, 3, og.n).astype(np.float64) * 0.5  # many key ties
    want, visits = oracle.fw_counts(og, k, rho)
    assert visits > 0
    assert np.array_equal(dense_model_counts(og, k, rho), want)


@pytest.mark.parametrize("mistake", [{"skip": 1}, {"win_extra": 1}])
def test_dense_model_fails_with_a_planted_mistake(oracle, mistake):
    """Starting the last level one word late, or taking one rank too many
    as winners, must disagree with the oracle."""
    pairs = _bipartite_plus(70, 0.12, 46)
    og = oracle.dag(pairs)
    rho = np.random.default_rng(6).integers(0, 3, og.n).astype(np.float64)
    want, _ = oracle.fw_counts(og, 6, rho)
    assert not np.array_equal(dense_model_counts(og, 6, rho, **mistake), want)
