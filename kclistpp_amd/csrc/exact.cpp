// exact.cpp - exact k-clique densest subgraph (the result of kCDensestMem.c).
//
// The reference stores every k-clique, shuffles them, runs sequential "++"
// passes with per-clique weights, and per round either records the top
// PAVA block as NOT STABLE / STABLE BUT NOT DENSEST or certifies it by
// Goldberg's closed-form criterion (A) or a Boost max-flow (B)
// (kCDensestMem.c:193-545).  Storing every clique does not scale (a
// 200-node near-clique alone holds ~1.7e10 6-cliques), so this pipeline is
// storage-free until the final network:
//
//   1. Sim-kClist++ iterations on the GPU, doubling rounds (as :500-507);
//   2. after each round the top block S of the PAVA fit (:329-364) over the
//      level order by rho is the candidate; its density D = p/q is exact
//      (p = cliques inside S, q = |S|);
//   3. region R = the D-core by k-clique degree: repeatedly drop nodes whose
//      k-clique degree inside the remaining graph is below D (GPU passes).
//      Every node of a densest subgraph S* keeps degree >= D* >= D inside
//      S*, so S* stays inside R.  S is *stable* when R = S: no node outside
//      S can belong to a subgraph at least as dense (the storage-free
//      counterpart of the reference's stability test, :366-403, which
//      moves each clique's weight to its lowest block and compares rho
//      inside and outside S);
//   4. criterion A (:405-432) for a stable S: with b an allocation of every
//      clique (b = V_T / (T(T+3)/2): each clique's unit split over its
//      winners of rounds 1..T), no j-subset of S (j < |S|) can beat p/q
//      when C(j, k)/j <= p/q or when the top-j average of b is below the
//      next density a j-subset could have;
//   5. criterion B, for a stable S (as the reference) or, from round 8 on,
//      for any S: if the cliques of R fit the network budget, the
//      max-flow of Goldberg's network (clique_flow.h: source -> clique (q),
//      clique -> member (q), member -> sink (p), arcs implicit) decides
//      whether some T in R has density > D.  Saturating flow certifies D;
//      the maximal densest subgraph is the set of R-nodes that cannot reach
//      the sink in the residual graph.  Otherwise the min cut's source side
//      is denser than D: it becomes the new candidate (Dinkelbach-style)
//      and the region and flow repeat;
//   6. the stat file gets the reference's rows (NOT STABLE: R != S; STABLE
//      BUT NOT DENSEST: R = S and neither criterion certified; OVER BUDGET,
//      which the reference cannot reach: the candidate or its region holds
//      more cliques than the network may) and its final record (Goldberg or
//      Max Flow).
// The final subgraph is order-invariant (the maximal k-clique densest
// subgraph is unique), so it is the part compared with the reference.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "clique_flow.h"
#include "kcl_internal.h"

namespace kcl {

namespace {

using u128 = unsigned __int128;

// Clique budget of the certificate network: about (16 k + 8) bytes of host
// memory per clique (members, flows, incidence, supply) plus the clique list
// (4 k); slot indices are 32-bit, so cliques x k must stay below 2^32.
constexpr uint64_t kDefaultNetworkCliques = 100000000ull;
constexpr uint64_t kMaxNetworkCliques = 0xffffffffull;
// From this round (iterations) on, candidates that are not stable go to the
// max-flow as well.
constexpr uint32_t kFlowUnstableFrom = 8;

// Pool Adjacent Violators with the reference's tolerance (kCDensestMem.c:329-350).
uint32_t pava_top_block(const std::vector<unsigned long long> &tent, double *val0) {
  const uint32_t n = (uint32_t)tent.size();
  if (n == 0) {
    *val0 = 0;
    return 0;
  }
  std::vector<uint32_t> nag(n);
  std::vector<double> val(n);
  nag[0] = 1;
  val[0] = (double)tent[0];
  uint32_t j = 0;
  for (uint32_t i = 1; i < n; ++i) {
    j += 1;
    val[j] = (double)tent[i];
    nag[j] = 1;
    while (j > 0 && val[j] >= val[j - 1] * 0.999999) {
      val[j - 1] = (nag[j] * val[j] + nag[j - 1] * val[j - 1]) / (nag[j] + nag[j - 1]);
      nag[j - 1] += nag[j];
      --j;
    }
  }
  *val0 = val[0];
  return nag[0];
}

// k-clique degrees of the subgraph induced by `alive` (GPU: DAG of the
// relabelled graph restricted to alive nodes, credit-every-member pass).
kcl_status alive_degrees(const kcl_graph *g, uint32_t k, const std::vector<uint8_t> &alive,
                         std::vector<uint64_t> *deg, uint64_t *cliques) {
  const uint32_t n = g->n;
  deg->assign(n, 0);
  *cliques = 0;
  std::vector<uint32_t> pairs;
  for (uint32_t s = 0; s < n; ++s) {
    if (!alive[s]) continue;
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j)
      if (alive[g->adj[j]]) {
        pairs.push_back(s);
        pairs.push_back(g->adj[j]);
      }
  }
  if (pairs.empty()) return KCL_OK;
  std::vector<uint32_t> ident(n);
  for (uint32_t v = 0; v < n; ++v) ident[v] = v;
  kcl_graph *sub = nullptr;
  KCL_TRY(graph_create(pairs.data(), pairs.size() / 2, n, ident.data(), &sub));
  kcl_status st = KCL_OK;
  if (sub->n > 0 && k <= sub->core + 1) {
    PassPlan p;
    DevicePlan dp;
    DevBuf<unsigned long long> cnt, key;
    do {
      if ((st = make_plan(sub, k, 1u << 17, sub->stream, &p, nullptr, nullptr)) != KCL_OK) break;
      if ((st = upload_plan(p, &dp)) != KCL_OK) break;
      if ((st = cnt.alloc(sub->n)) != KCL_OK || (st = key.alloc(sub->n)) != KCL_OK) break;
      hipError_t e = hipMemsetAsync(cnt.p, 0, (size_t)sub->n * 8, sub->stream);
      if (e == hipSuccess) e = hipMemsetAsync(key.p, 0, (size_t)sub->n * 8, sub->stream);
      if (e != hipSuccess) {
        st = fail(KCL_E_HIP, "hipMemset: %s", hipGetErrorString(e));
        break;
      }
      if ((st = launch_pass(sub, (const uint64_t *)key.p, k, kCreditAll, dp,
                            cnt.p, sub->stream)) != KCL_OK)
        break;
      e = hipMemcpyAsync(deg->data(), cnt.p, (size_t)sub->n * 8, hipMemcpyDeviceToHost,
                         sub->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(sub->stream);
      if (e != hipSuccess) st = fail(KCL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e));
      else st = pass_status(dp, sub->stream);
    } while (0);
  }
  kcl_graph_destroy(sub);
  KCL_TRY(st);
  u128 total = 0;
  for (uint32_t v = 0; v < n; ++v) total += (*deg)[v];
  *cliques = (uint64_t)(total / k);
  return KCL_OK;
}

// The D-core (D = p/q): drop nodes with k-clique degree * q < p until none.
kcl_status density_core(const kcl_graph *g, uint32_t k, uint64_t p, uint64_t q,
                        std::vector<uint8_t> *alive, uint64_t *cliques, uint32_t *rounds) {
  std::vector<uint64_t> deg;
  *rounds = 0;
  for (;;) {
    KCL_TRY(alive_degrees(g, k, *alive, &deg, cliques));
    ++*rounds;
    bool changed = false;
    for (uint32_t v = 0; v < g->n; ++v)
      if ((*alive)[v] && (u128)deg[v] * q < (u128)p) {
        (*alive)[v] = 0;
        changed = true;
      }
    if (!changed) return KCL_OK;
  }
}

// Every k-clique of the DAG restricted to `in` (host; the region is small):
// members[0] is the clique's largest id, later members come from the
// shrinking candidate lists (ascending ids, all adjacent to the members).
struct CliqueLister {
  const kcl_graph *g;
  uint32_t k;
  std::vector<std::vector<uint32_t>> cand;
  std::vector<uint32_t> members;
  std::vector<uint32_t> *out;

  void go(uint32_t depth) {
    if (depth == k) {
      out->insert(out->end(), members.begin(), members.end());
      return;
    }
    const std::vector<uint32_t> &c = cand[depth];
    if (c.size() < k - depth) return;
    for (size_t i = 0; i < c.size(); ++i) {
      const uint32_t x = c[i];
      members[depth] = x;
      std::vector<uint32_t> &nx = cand[depth + 1];
      nx.clear();
      const uint32_t *row = g->adj.data() + g->cd[x];
      const uint32_t len = g->cd[x + 1] - g->cd[x];
      size_t a = 0;
      uint32_t b = 0;
      while (a < i && b < len) {  // out-neighbours of x among the earlier candidates
        if (c[a] < row[b]) ++a;
        else if (row[b] < c[a]) ++b;
        else {
          nx.push_back(c[a]);
          ++a;
          ++b;
        }
      }
      go(depth + 1);
    }
  }
};

// Roots in chunks of kListChunk handed to the host threads on demand; each
// chunk's cliques land in its own vector, concatenated in chunk order (the
// order a single thread would list them in).
constexpr uint32_t kListChunk = 64;

void list_cliques(const kcl_graph *g, uint32_t k, const std::vector<uint8_t> &in,
                  std::vector<uint32_t> *out) {
  out->clear();
  std::vector<uint32_t> roots;
  for (uint32_t s = 0; s < g->n; ++s)
    if (in[s]) roots.push_back(s);
  const uint32_t nchunks = (uint32_t)((roots.size() + kListChunk - 1) / kListChunk);
  std::vector<std::vector<uint32_t>> chunks(nchunks);
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    std::vector<uint32_t> *dst = nullptr;
    CliqueLister L{g, k, std::vector<std::vector<uint32_t>>(k + 1), std::vector<uint32_t>(k),
                   dst};
    for (;;) {
      const uint32_t c = next.fetch_add(1);
      if (c >= nchunks) break;
      L.out = &chunks[c];
      const size_t hi = std::min(roots.size(), (size_t)(c + 1) * kListChunk);
      for (size_t i = (size_t)c * kListChunk; i < hi; ++i) {
        const uint32_t s = roots[i];
        L.members[0] = s;
        L.cand[1].clear();
        for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j)
          if (in[g->adj[j]]) L.cand[1].push_back(g->adj[j]);
        L.go(1);
      }
    }
  };
  const uint32_t W = std::min<uint32_t>(host_workers(), std::max<uint32_t>(1, nchunks));
  std::vector<std::thread> th;
  for (uint32_t t = 1; t < W; ++t) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  size_t total = 0;
  for (auto &v : chunks) total += v.size();
  out->reserve(total);
  for (auto &v : chunks) {
    out->insert(out->end(), v.begin(), v.end());
    std::vector<uint32_t>().swap(v);
  }
}

uint64_t edges_inside(const kcl_graph *g, const std::vector<uint8_t> &in) {
  uint64_t m = 0;
  for (uint32_t s = 0; s < g->n; ++s) {
    if (!in[s]) continue;
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j) m += in[g->adj[j]] ? 1 : 0;
  }
  return m;
}

// Criterion A of kCDensestMem.c:405-432 for a candidate of n nodes holding m
// cliques, with b the candidate members' shares of an allocation of every
// clique (the same arithmetic as the reference, on b instead of its
// post-stability rho): skip the sizes j whose complete subgraph could not
// beat m/n, then require the top-j average of b to stay below the next
// density a j-subset could reach.
bool goldberg_criterion(std::vector<double> b, uint64_t m, uint32_t n, uint32_t k) {
  std::sort(b.begin(), b.end(), std::greater<double>());
  double sum_rho = 0, jck = 0;
  bool skip = true;
  const double mn = (double)m / (double)n;
  for (uint32_t j = 1; j < n; ++j) {
    sum_rho += b[j - 1];
    if (skip) {
      if (j == k)
        jck = 1;
      else if (j > k)
        jck = (jck * j) / (j - k);
      if (jck / j > mn)
        skip = false;
      else
        continue;
    }
    const double ub = sum_rho / j;
    if (ub - mn >= 1.0 / n / j && ub - mn >= (ceil((double)m * j / n) - (double)m * j / n) / j)
      return false;
  }
  return true;
}

enum Stop : uint32_t { kNotCertified = 0, kGoldberg = 1, kMaxFlow = 2 };

struct Verdict {
  uint32_t degree_passes = 0;  // k-clique degree passes run
  double host_seconds = 0;     // listing + max-flow + criterion A
  bool certified = false;
  bool tested = false;       // the candidate's region was computed
  bool stable = false;       // ... and equals the candidate
  bool over_budget = false;  // a region held more cliques than the network may
  uint32_t stop = kNotCertified;
  std::vector<uint8_t> best;  // densest found (maximal when certified by max-flow)
  uint64_t p = 0, q = 0;      // its cliques and nodes
};

// Steps 3-5 for the candidate `cand` of density p/q; b (n entries) is the
// allocation criterion A uses.
// flow_unstable: run the max-flow for a candidate that is not stable too
// (the reference tries criterion B only on stable candidates; after a few
// rounds without one, the flow settles the question instead).
kcl_status certify(const kcl_graph *g, uint32_t k, uint64_t p, uint64_t q,
                   const std::vector<uint8_t> &cand, const std::vector<double> &b,
                   uint64_t budget, uint64_t gpu_list_min, bool flow_unstable, uint32_t *flows,
                   uint32_t *region_nodes, Verdict *out) {
  const uint32_t n = g->n;
  bool first = true;
  std::vector<uint32_t> cliques;  // the last region's cliques (global ids)
  bool listed = false;
  for (;;) {
    std::vector<uint8_t> alive(n, 1);
    uint64_t cl = 0;
    uint32_t rounds = 0;
    KCL_TRY(density_core(g, k, p, q, &alive, &cl, &rounds));
    out->degree_passes += rounds;
    const auto h0 = std::chrono::steady_clock::now();
    auto host_done = [&]() {
      out->host_seconds +=
          std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
    };
    uint32_t rn = 0;
    for (uint32_t v = 0; v < n; ++v) rn += alive[v];
    *region_nodes = rn;
    if (first) {
      out->tested = true;
      out->stable = (alive == cand);
      if (out->stable) {
        std::vector<double> bs;
        bs.reserve(q);
        for (uint32_t v = 0; v < n; ++v)
          if (cand[v]) bs.push_back(b[v]);
        if (goldberg_criterion(bs, p, (uint32_t)q, k)) {
          out->certified = true;
          out->stop = kGoldberg;
          out->best = cand;
          out->p = p;
          out->q = q;
          host_done();
          return KCL_OK;
        }
      }
      first = false;
      if (!out->stable && !flow_unstable) {
        host_done();
        return KCL_OK;
      }
    }
    if (cl > budget) {
      out->over_budget = true;
      host_done();
      return KCL_OK;
    }
    if (!listed) {
      if (cl >= gpu_list_min && k >= 3) KCL_TRY(list_cliques_gpu(g, k, alive, &cliques));
      else list_cliques(g, k, alive, &cliques);
      listed = true;
    } else {
      // The region shrank (a denser candidate): keep the cliques inside it.
      size_t w = 0;
      for (size_t c = 0; c + k <= cliques.size(); c += k) {
        bool inside = true;
        for (uint32_t j = 0; j < k && inside; ++j) inside = alive[cliques[c + j]];
        if (inside)
          for (uint32_t j = 0; j < k; ++j) cliques[w++] = cliques[c + j];
      }
      cliques.resize(w);
    }
    const uint64_t nc = cliques.size() / k;
    // Local member ids 0..rn-1.
    std::vector<uint32_t> local(n, UINT32_MAX), global;
    global.reserve(rn);
    for (uint32_t v = 0; v < n; ++v)
      if (alive[v]) {
        local[v] = (uint32_t)global.size();
        global.push_back(v);
      }
    std::vector<uint32_t> lmem(cliques.size());
    for (size_t i = 0; i < cliques.size(); ++i) lmem[i] = local[cliques[i]];
    CliqueFlow net(k, rn, nc, lmem.data(), q, p);
    const uint64_t flow = net.solve();
    ++*flows;
    if ((u128)flow == (u128)q * nc) {
      // No subset of the region beats p/q; the maximal densest subgraph is
      // the set of region nodes that cannot reach the sink.
      const std::vector<uint8_t> r = net.reaches_sink();
      out->best.assign(n, 0);
      uint64_t bq = 0;
      for (uint32_t i = 0; i < rn; ++i)
        if (!r[i]) {
          out->best[global[i]] = 1;
          ++bq;
        }
      out->q = bq;
      out->p = (uint64_t)((u128)p * bq / q);
      out->certified = true;
      out->stop = kMaxFlow;
      host_done();
      return KCL_OK;
    }
    // The source side of the min cut is denser than p/q: take it.
    const std::vector<uint8_t> side = net.source_side();
    std::vector<uint8_t> t(n, 0);
    uint64_t tq = 0;
    for (uint32_t i = 0; i < rn; ++i)
      if (side[i]) {
        t[global[i]] = 1;
        ++tq;
      }
    uint64_t tp = 0;
    for (uint64_t c = 0; c < nc; ++c) {
      bool inside = true;
      for (uint32_t j = 0; j < k && inside; ++j) inside = side[lmem[c * k + j]];
      tp += inside ? 1 : 0;
    }
    if (tq == 0 || (u128)tp * q <= (u128)p * tq) {
      return fail(KCL_E_HIP, "max-flow found no denser set although the flow is not saturated");
    }
    p = tp;
    q = tq;
    out->best = t;
    out->p = p;
    out->q = q;
    host_done();
  }
}

}  // namespace

}  // namespace kcl

using kcl::fail;

extern "C" kcl_status kcl_exact_run(kcl_graph *g, uint32_t k, const kcl_exact_opts *opts,
                                    const char *stat_path, int64_t t0_unix,
                                    kcl_exact_result *res, uint32_t *nodes_out) {
  if (!g || !res) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (k < 2 || k > 255) return fail(KCL_E_INVALID_ARG, "k = %u: 2 <= k <= 255", k);
  kcl_exact_opts o{};
  o.max_iter = 1u << 16;
  o.max_network_cliques = kcl::kDefaultNetworkCliques;
  if (opts) {
    if (opts->max_iter) o.max_iter = opts->max_iter;
    if (opts->max_network_cliques) o.max_network_cliques = opts->max_network_cliques;
    if (o.max_network_cliques > kcl::kMaxNetworkCliques / k)
      o.max_network_cliques = kcl::kMaxNetworkCliques / k;
    o.max_seconds = opts->max_seconds;
    o.schedule = opts->schedule;
    o.gpu_list_min = opts->gpu_list_min;
    o.model = opts->model;
  }
  const uint64_t gpu_list_min = o.gpu_list_min ? o.gpu_list_min : (1ull << 20);
  memset(res, 0, sizeof(*res));
  const auto c0 = std::chrono::steady_clock::now();
  if (t0_unix <= 0) t0_unix = (int64_t)time(nullptr);
  FILE *f = nullptr;
  if (stat_path) {
    f = fopen(stat_path, "w");
    if (!f) return fail(KCL_E_IO, "cannot create %s", stat_path);
    fprintf(f, "[Number of Iterations]\t[Number of Nodes]\t[k-Clique Density]\t[Time (seconds)]\t[Info]\n");
    fflush(f);
  }
  const uint32_t n = g->n;
  kcl_sim *sim = nullptr;
  kcl_sim_opts so{};
  so.schedule = o.schedule;
  so.model = o.model;
  kcl_status st = kcl_sim_create(g, k, &so, &sim);
  kcl::Verdict v;
  uint32_t flows = 0, region = 0, iters = 0, degree_passes = 0;
  double host_seconds = 0;
  uint64_t best_p = 0, best_q = 0;
  std::vector<uint8_t> best;
  for (uint32_t T = 1; st == KCL_OK && n > 0 && T <= o.max_iter; T <<= 1) {
    for (uint32_t t = T / 2 + 1; st == KCL_OK && t <= T; ++t) st = kcl::sim_step(sim, t);
    if (st != KCL_OK) break;
    iters = T;
    std::vector<unsigned long long> tent;
    std::vector<uint32_t> reo;
    std::vector<double> rho;
    if ((st = kcl::sim_level_pass(sim, true, tent, reo, rho)) != KCL_OK) break;
    // The allocation b = V_T / (T(T+3)/2): every clique's unit split over
    // its winners of rounds 1..T (criterion A reads it); the best top-i
    // average of b bounds every density.
    std::vector<double> b(n);
    {
      std::vector<uint64_t> V(n);
      if ((st = kcl_sim_totals(sim, V.data())) != KCL_OK) break;
      const double w = (double)T * (double)(T + 3) / 2.0;
      for (uint32_t i = 0; i < n; ++i) b[i] = (double)V[i] / w;
      std::sort(V.begin(), V.end(), std::greater<uint64_t>());
      kcl::u128 acc = 0;
      double ub = 0;
      for (uint32_t i = 0; i < n; ++i) {
        acc += V[i];
        ub = std::max(ub, (double)acc / w / (double)(i + 1));
      }
      res->upper_bound = ub;
    }
    double val0 = 0;
    const uint32_t q = kcl::pava_top_block(tent, &val0);
    uint64_t p = 0;
    for (uint32_t i = 0; i < q; ++i) p += tent[i];
    if (p == 0) {  // no k-clique at all: every node has density 0 (the
                   // reference's criterion A holds trivially)
      best.assign(n, 1);
      best_p = 0;
      best_q = n;
      v = kcl::Verdict{};
      v.certified = true;
      v.stop = kcl::kGoldberg;
      break;
    }
    if ((kcl::u128)p * best_q > (kcl::u128)best_p * q || best_q == 0) {
      best.assign(n, 0);
      for (uint32_t i = 0; i < q; ++i) best[reo[i]] = 1;
      best_p = p;
      best_q = q;
    }
    v = kcl::Verdict{};
    // A candidate that alone holds more cliques than the network may is not
    // tested this round (its region is a superset, and computing the region
    // costs k-clique degree passes over the graph).
    if (best_p <= o.max_network_cliques) {
      st = kcl::certify(g, k, best_p, best_q, best, b, o.max_network_cliques, gpu_list_min,
                        T >= kcl::kFlowUnstableFrom, &flows, &region, &v);
      degree_passes += v.degree_passes;
      host_seconds += v.host_seconds;
      if (st != KCL_OK) break;
    }
    const int64_t now = (int64_t)time(nullptr);
    if (v.certified) break;
    if (!v.best.empty() && (kcl::u128)v.p * best_q > (kcl::u128)best_p * v.q) {
      best = v.best;
      best_p = v.p;
      best_q = v.q;
    }
    if (f) {
      const char *info = !v.tested ? "OVER BUDGET"
                         : v.stable ? "STABLE BUT NOT DENSEST"
                                    : "NOT STABLE";
      fprintf(f, "%u\t%u\t%.12f\t%ld\t%s\n", T, (uint32_t)best_q, (double)best_p / (double)best_q,
              (long)(now - t0_unix), info);
      fflush(f);
    }
    if (o.max_seconds &&
        std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count() >
            (double)o.max_seconds)
      break;
  }
  if (st == KCL_OK && v.certified && !v.best.empty() && best_p != 0) {
    best = v.best;
    best_p = v.p;
    best_q = v.q;
  }
  kcl_sim_destroy(sim);
  if (st != KCL_OK) {
    if (f) fclose(f);
    return st;
  }
  res->certified = v.certified ? 1u : 0u;
  res->stop = v.certified ? v.stop : kcl::kNotCertified;
  if (v.certified) res->upper_bound = best_q ? (double)best_p / (double)best_q : 0.0;
  res->num_iter = iters;
  res->n = (uint32_t)best_q;
  res->cliques = best_p;
  res->density = best_q ? (double)best_p / (double)best_q : 0.0;
  res->m = best.empty() ? 0 : (uint32_t)kcl::edges_inside(g, best);
  res->max_flow_calls = flows;
  res->degree_passes = degree_passes;
  res->host_seconds = host_seconds;
  res->region = region;
  res->seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
  if (nodes_out && !best.empty()) {
    // input ids of the subgraph, ascending
    uint32_t w = 0;
    for (uint32_t u = 0; u < g->n_in; ++u)
      if (g->rank[u] < n && best[g->rank[u]]) nodes_out[w++] = u;
  }
  if (f) {
    if (res->certified) {
      fprintf(f, "[Number of Iterations]\t[Stopping Condition]\t[Number of Nodes]\t[Number of "
                 "Edges]\t[k-Clique Density]\t[Number of Max-Flow Calls]\t[Time (seconds)]\n");
      fprintf(f, "%u\t%s\t%u\t%u\t%.12f\t%u\t%ld\n", res->num_iter,
              res->stop == kcl::kGoldberg ? "Goldberg" : "Max Flow", res->n, res->m, res->density,
              res->max_flow_calls, (long)((int64_t)time(nullptr) - t0_unix));
    }
    if (fclose(f) != 0) return fail(KCL_E_IO, "cannot write %s", stat_path);
  }
  return KCL_OK;
}
