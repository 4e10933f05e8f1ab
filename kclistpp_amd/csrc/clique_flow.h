// clique_flow.h - maximum flow of the k-clique density network without
// materialising its arcs (host C++, no HIP).
//
// The certificate of kCDensestMem.c:434-467 is Goldberg's network for a
// candidate density p/q over a node set R: source -> clique (capacity q),
// clique -> each of its k members (capacity q), member -> sink (capacity
// p).  The flow saturates the source arcs (value q * #cliques) iff every
// subset T of R has cliques(T) * q <= p * |T|.  The reference builds it as
// a Boost graph with 2 + |R| + #cliques vertices; at tens of millions of
// cliques that is tens of gigabytes of arcs.  Here the clique side keeps
// only its k member ids and the flow on each member slot (u32, since a slot
// never carries more than q), the member side its load and the list of
// slots it appears in.  Clique -> member arcs never bind (a clique's supply
// is q), so the residual graph is: source -> clique while its supply is not
// used up, clique -> member always, member -> clique while that slot
// carries flow, member -> sink while its load is below p.
//
// Solver: a greedy initial flow (each clique's q units to its least loaded
// members) and Dinic phases (breadth-first levels, then blocking flow by an
// iterative depth-first search with current-arc pointers) on that
// implicit graph.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace kcl {

// Host threads for the certificate's scans (at most 16).
inline uint32_t host_workers() {
  const uint32_t h = std::thread::hardware_concurrency();
  return std::max<uint32_t>(1, std::min<uint32_t>(16, h ? h : 1));
}

// fn(t, lo, hi) on `workers` contiguous chunks of [0, n) in parallel.
template <class F>
void parallel_chunks(uint64_t n, uint32_t workers, F fn) {
  if (workers <= 1 || n < 4096) {
    fn(0u, (uint64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(workers);
  for (uint32_t t = 0; t < workers; ++t)
    th.emplace_back([&, t]() { fn(t, n * t / workers, n * (t + 1) / workers); });
  for (auto &x : th) x.join();
}

class CliqueFlow {
 public:
  static constexpr uint32_t kMaxK = 64;

  // mem: nc * k local member ids in [0, r).  q < 2^32, k <= kMaxK,
  // nc * k < 2^32.
  CliqueFlow(uint32_t k, uint32_t r, uint64_t nc, const uint32_t *mem, uint64_t q, uint64_t p)
      : k_(k), r_(r), nc_(nc), mem_(mem), q_(q), p_(p) {
    flow_.assign(nc * k, 0);
    sup_.assign(nc, (uint32_t)q);
    load_.assign(r, 0);
    // Member -> slot incidence (CSR, slots ascending per member), built by a
    // counting sort over contiguous chunks of the slots, one per thread.
    const uint64_t ns = nc * k;
    const uint32_t W = workers_;
    std::vector<uint64_t> cnt((size_t)W * r, 0);
    parallel_chunks(ns, W, [&](uint32_t t, uint64_t lo, uint64_t hi) {
      uint64_t *c = cnt.data() + (size_t)t * r;
      for (uint64_t i = lo; i < hi; ++i) c[mem[i]]++;
    });
    inc_off_.assign((size_t)r + 1, 0);
    // Chunk t of member x starts after every earlier chunk's entries of x.
    std::vector<uint64_t> at((size_t)W * r);
    uint64_t run = 0;
    for (uint32_t x = 0; x < r; ++x) {
      inc_off_[x] = run;
      for (uint32_t t = 0; t < W; ++t) {
        at[(size_t)t * r + x] = run;
        run += cnt[(size_t)t * r + x];
      }
    }
    inc_off_[r] = run;
    inc_.resize(ns);
    parallel_chunks(ns, W, [&](uint32_t t, uint64_t lo, uint64_t hi) {
      uint64_t *a = at.data() + (size_t)t * r;
      for (uint64_t i = lo; i < hi; ++i) inc_[a[mem[i]]++] = (uint32_t)i;
    });
  }

  // Value of a maximum flow.
  uint64_t solve() {
    greedy();
    for (;;) {
      if (!levels()) break;
      blocking_flow();
    }
#ifdef KCL_CLIQUE_FLOW_TRACE
    fprintf(stderr, "clique flow: %u phases\n", phases_);
#endif
    uint64_t f = 0;
    for (uint32_t x = 0; x < r_; ++x) f += load_[x];
    return f;
  }

  // Members with a residual path to the sink (after solve()).
  std::vector<uint8_t> reaches_sink() const {
    std::vector<uint8_t> seen(r_, 0);
    std::vector<uint32_t> queue;
    for (uint32_t x = 0; x < r_; ++x)
      if (load_[x] < p_) {
        seen[x] = 1;
        queue.push_back(x);
      }
    std::vector<uint8_t> cseen(nc_, 0);
    for (size_t h = 0; h < queue.size(); ++h) {
      const uint32_t x = queue[h];
      // c -> x has residual q - flow(c, x) > 0; then y -> c needs flow(c, y) > 0.
      for (uint64_t a = inc_off_[x]; a < inc_off_[x + 1]; ++a) {
        const uint32_t slot = inc_[a];
        const uint64_t c = slot / k_;
        if (cseen[c] || flow_[slot] >= q_) continue;
        cseen[c] = 1;
        for (uint32_t j = 0; j < k_; ++j) {
          const uint32_t y = mem_[c * k_ + j];
          if (!seen[y] && flow_[c * k_ + j] > 0) {
            seen[y] = 1;
            queue.push_back(y);
          }
        }
      }
    }
    return seen;
  }

  // Members reachable from the source in the residual graph (the source
  // side of a minimum cut, after solve()).
  std::vector<uint8_t> source_side() const {
    std::vector<uint8_t> seen(r_, 0), cseen(nc_, 0);
    std::vector<uint32_t> queue;
    auto visit_clique = [&](uint64_t c) {
      cseen[c] = 1;
      for (uint32_t j = 0; j < k_; ++j) {
        const uint32_t y = mem_[c * k_ + j];
        if (!seen[y] && flow_[c * k_ + j] < q_) {
          seen[y] = 1;
          queue.push_back(y);
        }
      }
    };
    for (uint64_t c = 0; c < nc_; ++c)
      if (sup_[c] > 0) visit_clique(c);
    for (size_t h = 0; h < queue.size(); ++h) {
      const uint32_t x = queue[h];
      for (uint64_t a = inc_off_[x]; a < inc_off_[x + 1]; ++a) {
        const uint32_t slot = inc_[a];
        const uint64_t c = slot / k_;
        if (!cseen[c] && flow_[slot] > 0) visit_clique(c);
      }
    }
    return seen;
  }

  const std::vector<uint64_t> &loads() const { return load_; }
  uint32_t phases() const { return phases_; }

 private:
  // Each clique's q units to its members in order of load (least loaded
  // first), within the sink capacities.  Cliques are taken in blocks of 64
  // consecutive ones, the blocks in a scattered order (block b at position
  // b * stride mod blocks): cliques listed together share members, and in
  // listing order the last ones would find their members already full.
  void greedy() {
    constexpr uint64_t kBlock = 64;
    const uint64_t nb = (nc_ + kBlock - 1) / kBlock;
    uint64_t stride = nb > 1 ? 2654435761ull % nb : 1;
    if (stride == 0) stride = 1;
    while (gcd(stride, nb) != 1) ++stride;
    for (uint64_t i = 0; i < nb; ++i) {
      const uint64_t b = (i * stride) % nb;
      const uint64_t end = std::min(nc_, (b + 1) * kBlock);
      for (uint64_t c = b * kBlock; c < end; ++c) greedy_one(c);
    }
  }

  static uint64_t gcd(uint64_t a, uint64_t b) {
    while (b) {
      const uint64_t t = a % b;
      a = b;
      b = t;
    }
    return a;
  }

  void greedy_one(uint64_t c) {
    uint32_t order[kMaxK];
    {
      const uint32_t *m = mem_ + c * k_;
      // Usually the least loaded member takes the whole supply.
      uint32_t jmin = 0;
      for (uint32_t j = 1; j < k_; ++j)
        if (load_[m[j]] < load_[m[jmin]]) jmin = j;
      if (p_ - load_[m[jmin]] >= sup_[c]) {
        flow_[c * k_ + jmin] += sup_[c];
        load_[m[jmin]] += sup_[c];
        sup_[c] = 0;
        return;
      }
      for (uint32_t j = 0; j < k_; ++j) order[j] = j;
      std::sort(order, order + k_,
                [&](uint32_t a, uint32_t b) { return load_[m[a]] < load_[m[b]]; });
      uint64_t left = sup_[c];
      for (uint32_t i = 0; i < k_ && left; ++i) {
        const uint32_t j = order[i];
        const uint64_t room = p_ - load_[m[j]];
        const uint64_t give = std::min(left, room);
        if (!give) continue;
        flow_[c * k_ + j] += (uint32_t)give;
        load_[m[j]] += give;
        left -= give;
      }
      sup_[c] = (uint32_t)left;
    }
  }

  // Levels: cliques with supply at level 1, members at even levels, cliques
  // reached back through carried flow at odd levels; true if some member
  // with room to the sink is reachable.
  bool levels() {
    ++phases_;
    clevel_.assign(nc_, -1);
    mlevel_.assign(r_, -1);
    const uint32_t W = workers_;
    std::vector<std::vector<uint64_t>> part(W);
    parallel_chunks(nc_, W, [&](uint32_t t, uint64_t lo, uint64_t hi) {
      for (uint64_t c = lo; c < hi; ++c)
        if (sup_[c] > 0) {
          clevel_[c] = 1;
          part[t].push_back(c);
        }
    });
    std::vector<uint64_t> cur;
    for (auto &v : part) cur.insert(cur.end(), v.begin(), v.end());
    sink_level_ = -1;
    int32_t lv = 1;
    while (!cur.empty()) {
      std::vector<uint32_t> members;
      for (uint64_t c : cur)
        for (uint32_t j = 0; j < k_; ++j) {
          const uint32_t x = mem_[c * k_ + j];
          if (mlevel_[x] < 0 && flow_[c * k_ + j] < q_) {
            mlevel_[x] = lv + 1;
            members.push_back(x);
            if (load_[x] < p_) sink_level_ = lv + 2;
          }
        }
      if (sink_level_ >= 0 || members.empty()) break;
      // Cliques reached back from these members (threads take contiguous
      // ranges of their concatenated incidence lists; a clique is claimed
      // once by compare-and-swap on its level).
      std::vector<uint64_t> starts(members.size() + 1, 0);
      for (size_t i = 0; i < members.size(); ++i)
        starts[i + 1] = starts[i] + (inc_off_[members[i] + 1] - inc_off_[members[i]]);
      for (auto &v : part) v.clear();
      const int32_t next = lv + 2;
      parallel_chunks(starts.back(), W, [&](uint32_t t, uint64_t lo, uint64_t hi) {
        size_t mi = std::upper_bound(starts.begin(), starts.end(), lo) - starts.begin() - 1;
        for (uint64_t e = lo; e < hi; ++e) {
          while (e >= starts[mi + 1]) ++mi;
          const uint32_t slot = inc_[inc_off_[members[mi]] + (e - starts[mi])];
          const uint64_t c = slot / k_;
          int32_t expect = -1;
          if (flow_[slot] > 0 && clevel_[c] < 0 &&
              __atomic_compare_exchange_n(&clevel_[c], &expect, next, false, __ATOMIC_RELAXED,
                                          __ATOMIC_RELAXED))
            part[t].push_back(c);
        }
      });
      cur.clear();
      for (auto &v : part) cur.insert(cur.end(), v.begin(), v.end());
      lv += 2;
    }
    return sink_level_ >= 0;
  }

  // Blocking flow on the level graph: paths source -> c0 -> x0 -> c1 -> x1
  // ... -> xm -> sink, alternating clique -> member (push on the slot) and
  // member -> clique (take back flow carried on the slot).
  void blocking_flow() {
    std::vector<uint32_t> cptr(nc_, 0);           // next member slot per clique
    std::vector<uint64_t> mptr(inc_off_.begin(), inc_off_.end() - 1);  // next incidence
    struct Step {
      uint64_t c;     // clique on the path
      uint32_t slot;  // its slot towards the next member
      uint32_t back;  // slot of the previous member in this clique (or UINT32_MAX)
    };
    std::vector<Step> path;
    for (uint64_t c0 = 0; c0 < nc_; ++c0) {
      if (clevel_[c0] != 1) continue;
      while (sup_[c0] > 0) {
        // Depth-first search for one augmenting path from c0.
        path.clear();
        path.push_back(Step{c0, 0, UINT32_MAX});
        bool found = false;
        while (!path.empty()) {
          Step &st = path.back();
          const uint64_t c = st.c;
          const int32_t lc = clevel_[c];
          bool advanced = false;
          while (cptr[c] < k_) {
            const uint32_t j = cptr[c];
            const uint32_t slot = (uint32_t)(c * k_ + j);
            const uint32_t x = mem_[slot];
            if (mlevel_[x] == lc + 1 && flow_[slot] < q_) {
              st.slot = slot;
              if (lc + 2 == sink_level_) {
                if (load_[x] < p_) {
                  found = true;
                  break;
                }
                ++cptr[c];
                continue;
              }
              // x -> c' through flow carried on (c', x).
              while (mptr[x] < inc_off_[x + 1]) {
                const uint32_t s2 = inc_[mptr[x]];
                const uint64_t c2 = s2 / k_;
                if (clevel_[c2] == lc + 2 && flow_[s2] > 0) {
                  path.push_back(Step{c2, 0, s2});
                  advanced = true;
                  break;
                }
                ++mptr[x];
              }
              if (advanced) break;
              mlevel_[x] = -1;  // dead end: no longer in the level graph
            }
            ++cptr[c];
          }
          if (found) break;
          if (advanced) continue;
          // Dead end at c: retreat.
          clevel_[c] = -1;
          path.pop_back();
          if (!path.empty()) {
            const uint32_t x = mem_[path.back().slot];
            ++mptr[x];
          }
        }
        if (!found) break;
        // Bottleneck: the source arc, each taken-back slot, the sink arc.
        uint64_t push = sup_[path.front().c];
        for (size_t i = 1; i < path.size(); ++i) push = std::min<uint64_t>(push, flow_[path[i].back]);
        const uint32_t xl = mem_[path.back().slot];
        push = std::min<uint64_t>(push, p_ - load_[xl]);
        for (const Step &st : path) push = std::min<uint64_t>(push, q_ - flow_[st.slot]);
        sup_[path.front().c] -= (uint32_t)push;
        for (size_t i = 0; i < path.size(); ++i) {
          flow_[path[i].slot] += (uint32_t)push;
          if (i > 0) flow_[path[i].back] -= (uint32_t)push;
        }
        load_[xl] += push;
      }
    }
  }

  uint32_t workers_ = host_workers();
  uint32_t k_, r_;
  uint64_t nc_;
  const uint32_t *mem_;
  uint64_t q_, p_;
  std::vector<uint32_t> flow_;  // per slot
  std::vector<uint32_t> sup_;   // unused supply per clique
  std::vector<uint64_t> load_;  // per member
  std::vector<uint64_t> inc_off_;
  std::vector<uint32_t> inc_;   // slots per member
  std::vector<int32_t> clevel_, mlevel_;
  int32_t sink_level_ = -1;
  uint32_t phases_ = 0;
};

}  // namespace kcl
