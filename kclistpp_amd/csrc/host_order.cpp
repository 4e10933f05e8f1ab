// host_order.cpp - degeneracy ranks (Graph.c:57-91 with BinaryHeap.c).
//
// The relabelling is defined by the pop order of the reference's binary
// min-heap, including how it settles ties, so this is a sequential host
// replay of the same heap discipline: keys inserted 0..n-1 with their
// degrees; pop-min moves the last slot to the root and sifts it down
// (right child only if strictly smaller than the left, move while strictly
// smaller); each neighbour of the popped node, in edge-list order, is
// decreased by one and sifted up (move while the parent is strictly
// larger); rank = n - (pop index + 1).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "kcl_internal.h"

namespace kcl {

namespace {

struct Slot {
  uint32_t node;
  uint32_t deg;
};

struct Heap {
  std::vector<Slot> a;          // heap slots
  std::vector<uint32_t> where;  // node -> slot, UINT32_MAX once popped
  uint32_t size = 0;

  void up(uint32_t i) {
    Slot x = a[i];
    while (i > 0) {
      uint32_t p = (i - 1) >> 1;
      if (!(a[p].deg > x.deg)) break;
      a[i] = a[p];
      where[a[i].node] = i;
      i = p;
    }
    a[i] = x;
    where[x.node] = i;
  }

  void down() {
    uint32_t i = 0;
    Slot x = a[0];
    for (;;) {
      uint32_t l = 2 * i + 1;
      if (l >= size) break;
      uint32_t r = l + 1;
      uint32_t c = (r < size && a[r].deg < a[l].deg) ? r : l;
      if (!(a[c].deg < x.deg)) break;
      a[i] = a[c];
      where[a[i].node] = i;
      i = c;
    }
    a[i] = x;
    where[x.node] = i;
  }
};

}  // namespace

namespace {

// Undirected adjacency in edge order (Graph.c:62-74): for each edge (s, t) in
// turn, t is appended to s's list, then s to t's.  Threads take contiguous
// chunks of the edge list: each counts its chunk's entries per node, a
// per-node prefix over the threads places chunk c's entries after those of
// chunks before it, and each thread then writes its own entries in order, so
// every list is the sequential one.
void undirected_adjacency(uint32_t n, uint64_t m, const uint32_t *pairs,
                          std::vector<uint64_t> *off_out, std::vector<uint32_t> *nbr_out) {
  std::vector<uint64_t> &off = *off_out;
  std::vector<uint32_t> &nbr = *nbr_out;
  off.assign((size_t)n + 1, 0);
  nbr.assign(2 * m + 1, 0);
  unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // Small graphs, and sparse ids (the per-chunk counts take T n words, so
  // they must stay small next to the 2 m adjacency words), stay sequential.
  if (m < (uint64_t(1) << 22) || (uint64_t)T * n > 2 * m) T = 1;
  if (T == 1) {
    for (uint64_t i = 0; i < m; ++i) {
      off[pairs[2 * i] + 1]++;
      off[pairs[2 * i + 1] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<uint64_t> fill(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < m; ++i) {
      uint32_t s = pairs[2 * i], t = pairs[2 * i + 1];
      nbr[fill[s]++] = t;
      nbr[fill[t]++] = s;
    }
    return;
  }
  // cnt[c][v]: entries of node v in chunk c (then: entries of v in chunks < c).
  std::vector<std::vector<uint32_t>> cnt(T, std::vector<uint32_t>(n, 0));
  auto run = [&](auto &&body) {
    std::vector<std::thread> th;
    for (unsigned c = 0; c < T; ++c) th.emplace_back(body, c);
    for (auto &x : th) x.join();
  };
  run([&](unsigned c) {
    uint32_t *k = cnt[c].data();
    for (uint64_t i = m * c / T, e = m * (c + 1) / T; i < e; ++i) {
      k[pairs[2 * i]]++;
      k[pairs[2 * i + 1]]++;
    }
  });
  run([&](unsigned c) {  // per node: exclusive prefix over the chunks, total degree
    for (uint64_t v = (uint64_t)n * c / T, e = (uint64_t)n * (c + 1) / T; v < e; ++v) {
      uint64_t run_sum = 0;
      for (unsigned q = 0; q < T; ++q) {
        const uint32_t k = cnt[q][v];
        cnt[q][v] = (uint32_t)run_sum;
        run_sum += k;
      }
      off[v + 1] = run_sum;
    }
  });
  for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
  run([&](unsigned c) {
    uint32_t *k = cnt[c].data();
    for (uint64_t i = m * c / T, e = m * (c + 1) / T; i < e; ++i) {
      uint32_t s = pairs[2 * i], t = pairs[2 * i + 1];
      nbr[off[s] + k[s]++] = t;
      nbr[off[t] + k[t]++] = s;
    }
  });
}

}  // namespace

void degeneracy_rank(uint32_t n, uint64_t m, const uint32_t *pairs, uint32_t *rank) {
  if (n == 0) return;
  std::vector<uint64_t> off;
  std::vector<uint32_t> nbr;
  undirected_adjacency(n, m, pairs, &off, &nbr);
  Heap h;
  h.a.resize(n);
  h.where.assign(n, UINT32_MAX);
  for (uint32_t v = 0; v < n; ++v) {  // BH_MakeHeap inserts keys in order
    h.a[h.size] = Slot{v, (uint32_t)(off[v + 1] - off[v])};
    h.where[v] = h.size;
    ++h.size;
    h.up(h.size - 1);
  }
  for (uint32_t r = 0; r < n; ++r) {
    uint32_t u = h.a[0].node;
    h.where[u] = UINT32_MAX;
    --h.size;
    h.a[0] = h.a[h.size];
    h.where[h.a[0].node] = 0;  // also when the heap just emptied (then u itself)
    if (h.size > 0) h.down();
    rank[u] = n - (r + 1);
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
      uint32_t v = nbr[j];
      uint32_t at = h.where[v];
      if (at == UINT32_MAX) continue;
      h.a[at].deg -= 1;
      h.up(at);
    }
  }
}

}  // namespace kcl
