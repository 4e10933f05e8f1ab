// exact.cpp - exact k-clique densest subgraph (the result of kCDensestMem.c).
//
// The reference stores every k-clique, shuffles them, runs sequential "++"
// passes with per-clique weights, and certifies the top PAVA block by a
// stability test, Goldberg's condition or a Boost max-flow
// (kCDensestMem.c:193-545).  Storing every clique does not scale (a
// 200-node near-clique alone holds ~1.7e10 6-cliques), so this pipeline is
// storage-free until the final, small network:
//
//   1. Sim-kClist++ iterations on the GPU, doubling rounds (as :500-507);
//   2. after each round the top block of the PAVA fit (:329-364) over the
//      level order by rho is the candidate S; its density D = p/q is exact
//      (p = cliques inside S, q = |S|);
//   3. region R = the D-core by k-clique degree: repeatedly drop nodes whose
//      k-clique degree inside the remaining graph is below D (GPU passes).
//      Every node of a densest subgraph S* keeps degree >= D* >= D inside
//      S*, so S* stays inside R;
//   4. if the cliques of R fit the network budget, a host max-flow decides
//      whether some T in R has density > D (Goldberg's network, as
//      :434-467: source -> clique (q), clique -> member (q), member -> sink
//      (p)).  Saturating flow certifies D; the maximal densest subgraph is
//      the set of R-nodes that cannot reach the sink in the residual graph.
//      Otherwise the min cut's source side is denser than D: it becomes the
//      new candidate (Dinkelbach-style) and steps 3-4 repeat;
//   5. the stat file gets the reference's rows and final record.
// The final subgraph is order-invariant (the maximal k-clique densest
// subgraph is unique), so it is the part compared with the reference.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "kcl_internal.h"

namespace kcl {

namespace {

using u128 = unsigned __int128;

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

// Dinic max-flow over a CSR arc array (residual capacities in place).
struct FlowNet {
  uint32_t nv = 0;
  std::vector<uint64_t> head;  // nv + 1
  std::vector<uint32_t> to;
  std::vector<uint64_t> cap;
  std::vector<uint64_t> rev;

  void build(uint32_t n, const std::vector<uint32_t> &from, const std::vector<uint32_t> &dst,
             const std::vector<uint64_t> &c) {
    nv = n;
    const size_t ne = from.size();
    head.assign((size_t)n + 1, 0);
    for (size_t i = 0; i < ne; ++i) {
      head[from[i] + 1]++;
      head[dst[i] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) head[v + 1] += head[v];
    to.resize(2 * ne);
    cap.resize(2 * ne);
    rev.resize(2 * ne);
    std::vector<uint64_t> fill(head.begin(), head.end() - 1);
    for (size_t i = 0; i < ne; ++i) {
      const uint64_t a = fill[from[i]]++, b = fill[dst[i]]++;
      to[a] = dst[i];
      cap[a] = c[i];
      to[b] = from[i];
      cap[b] = 0;
      rev[a] = b;
      rev[b] = a;
    }
  }

  uint64_t max_flow(uint32_t s, uint32_t t) {
    std::vector<int32_t> level(nv);
    std::vector<uint64_t> it(nv);
    std::vector<uint32_t> queue(nv), stack_v(nv);
    std::vector<uint64_t> stack_a(nv);
    uint64_t flow = 0;
    for (;;) {
      std::fill(level.begin(), level.end(), -1);
      uint32_t qh = 0, qt = 0;
      level[s] = 0;
      queue[qt++] = s;
      while (qh < qt) {
        const uint32_t v = queue[qh++];
        for (uint64_t a = head[v]; a < head[v + 1]; ++a)
          if (cap[a] && level[to[a]] < 0) {
            level[to[a]] = level[v] + 1;
            queue[qt++] = to[a];
          }
      }
      if (level[t] < 0) break;
      for (uint32_t v = 0; v < nv; ++v) it[v] = head[v];
      for (;;) {  // blocking flow by iterative DFS
        uint32_t depth = 0, v = s;
        bool found = false;
        for (;;) {
          if (v == t) {
            found = true;
            break;
          }
          uint64_t a = it[v];
          while (a < head[v + 1] && !(cap[a] && level[to[a]] == level[v] + 1)) ++a;
          it[v] = a;
          if (a == head[v + 1]) {
            if (depth == 0) break;
            level[v] = -1;
            v = stack_v[--depth];
            it[v]++;
            continue;
          }
          stack_v[depth] = v;
          stack_a[depth] = a;
          ++depth;
          v = to[a];
        }
        if (!found) break;
        uint64_t push = UINT64_MAX;
        for (uint32_t d = 0; d < depth; ++d) push = std::min(push, cap[stack_a[d]]);
        for (uint32_t d = 0; d < depth; ++d) {
          cap[stack_a[d]] -= push;
          cap[rev[stack_a[d]]] += push;
        }
        flow += push;
      }
    }
    return flow;
  }

  // Nodes with a residual path to t (reverse search).
  std::vector<uint8_t> reaches(uint32_t t) const {
    std::vector<uint8_t> seen(nv, 0);
    std::vector<uint32_t> queue;
    queue.push_back(t);
    seen[t] = 1;
    for (size_t qh = 0; qh < queue.size(); ++qh) {
      const uint32_t v = queue[qh];
      for (uint64_t a = head[v]; a < head[v + 1]; ++a) {
        // arc a goes v -> to[a]; its reverse to[a] -> v has capacity cap[rev[a]]
        const uint32_t u = to[a];
        if (!seen[u] && cap[rev[a]]) {
          seen[u] = 1;
          queue.push_back(u);
        }
      }
    }
    return seen;
  }

  // Nodes reachable from s in the residual graph.
  std::vector<uint8_t> reachable(uint32_t s) const {
    std::vector<uint8_t> seen(nv, 0);
    std::vector<uint32_t> queue;
    queue.push_back(s);
    seen[s] = 1;
    for (size_t qh = 0; qh < queue.size(); ++qh) {
      const uint32_t v = queue[qh];
      for (uint64_t a = head[v]; a < head[v + 1]; ++a)
        if (cap[a] && !seen[to[a]]) {
          seen[to[a]] = 1;
          queue.push_back(to[a]);
        }
    }
    return seen;
  }
};

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

void list_cliques(const kcl_graph *g, uint32_t k, const std::vector<uint8_t> &in,
                  std::vector<uint32_t> *out) {
  out->clear();
  CliqueLister L{g, k, std::vector<std::vector<uint32_t>>(k + 1), std::vector<uint32_t>(k), out};
  for (uint32_t s = 0; s < g->n; ++s) {
    if (!in[s]) continue;
    L.members[0] = s;
    L.cand[1].clear();
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j)
      if (in[g->adj[j]]) L.cand[1].push_back(g->adj[j]);
    L.go(1);
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

struct Verdict {
  bool certified = false;
  bool fits = false;
  std::vector<uint8_t> best;  // densest found (maximal when certified)
  uint64_t p = 0, q = 0;      // its cliques and nodes
};

// Steps 3-4 for candidate density p/q: region, then max-flow rounds.
kcl_status certify(const kcl_graph *g, uint32_t k, uint64_t p, uint64_t q, uint64_t budget,
                   uint32_t *flows, uint32_t *region_nodes, Verdict *out) {
  const uint32_t n = g->n;
  for (;;) {
    std::vector<uint8_t> alive(n, 1);
    uint64_t cl = 0;
    uint32_t rounds = 0;
    KCL_TRY(density_core(g, k, p, q, &alive, &cl, &rounds));
    uint32_t rn = 0;
    for (uint32_t v = 0; v < n; ++v) rn += alive[v];
    *region_nodes = rn;
    if (cl > budget) {
      out->fits = false;
      return KCL_OK;
    }
    out->fits = true;
    std::vector<uint32_t> cliques;
    list_cliques(g, k, alive, &cliques);
    const uint64_t nc = cliques.size() / k;
    // Network ids: 0 = source, 1 = sink, 2..2+n = nodes, then cliques.
    std::vector<uint32_t> from, to;
    std::vector<uint64_t> cap;
    const uint32_t base = 2 + n;
    from.reserve(nc * (k + 1) + rn);
    to.reserve(nc * (k + 1) + rn);
    cap.reserve(nc * (k + 1) + rn);
    for (uint64_t c = 0; c < nc; ++c) {
      from.push_back(0);
      to.push_back(base + (uint32_t)c);
      cap.push_back(q);
      for (uint32_t j = 0; j < k; ++j) {
        from.push_back(base + (uint32_t)c);
        to.push_back(2 + cliques[c * k + j]);
        cap.push_back(q);
      }
    }
    for (uint32_t v = 0; v < n; ++v)
      if (alive[v]) {
        from.push_back(2 + v);
        to.push_back(1);
        cap.push_back(p);
      }
    FlowNet net;
    net.build(base + (uint32_t)nc, from, to, cap);
    const uint64_t flow = net.max_flow(0, 1);
    ++*flows;
    if ((u128)flow == (u128)q * nc) {
      // No subset of the region beats p/q; the maximal densest subgraph is
      // the set of region nodes that cannot reach the sink.
      const std::vector<uint8_t> r = net.reaches(1);
      out->best.assign(n, 0);
      uint64_t bq = 0;
      for (uint32_t v = 0; v < n; ++v)
        if (alive[v] && !r[2 + v]) {
          out->best[v] = 1;
          ++bq;
        }
      out->q = bq;
      out->p = (uint64_t)((u128)p * bq / q);
      out->certified = true;
      return KCL_OK;
    }
    // The source side of the min cut is denser than p/q: take it.
    const std::vector<uint8_t> s = net.reachable(0);
    std::vector<uint8_t> t(n, 0);
    uint64_t tq = 0;
    for (uint32_t v = 0; v < n; ++v)
      if (alive[v] && s[2 + v]) {
        t[v] = 1;
        ++tq;
      }
    uint64_t tp = 0;
    for (uint64_t c = 0; c < nc; ++c) {
      bool inside = true;
      for (uint32_t j = 0; j < k && inside; ++j) inside = t[cliques[c * k + j]];
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
  o.max_network_cliques = 20000000ull;
  if (opts) {
    if (opts->max_iter) o.max_iter = opts->max_iter;
    if (opts->max_network_cliques) o.max_network_cliques = opts->max_network_cliques;
    o.max_seconds = opts->max_seconds;
    o.schedule = opts->schedule;
    o.model = opts->model;
  }
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
  uint32_t flows = 0, region = 0, iters = 0;
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
    {
      // Upper bound: V_T / (T(T+3)/2) is a convex combination of
      // allocations; the best top-i average of it bounds every density.
      std::vector<uint64_t> V(n);
      if ((st = kcl_sim_totals(sim, V.data())) != KCL_OK) break;
      std::sort(V.begin(), V.end(), std::greater<uint64_t>());
      const double w = (double)T * (double)(T + 3) / 2.0;
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
    if (p == 0) {  // no k-clique at all: every node has density 0
      best.assign(n, 1);
      best_p = 0;
      best_q = n;
      v = kcl::Verdict{};
      v.certified = true;
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
    // certified this round (its region is a superset in practice, and
    // computing the region costs k-clique degree passes over the graph).
    if (best_p <= o.max_network_cliques) {
      st = kcl::certify(g, k, best_p, best_q, o.max_network_cliques, &flows, &region, &v);
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
      fprintf(f, "%u\t%u\t%.12f\t%ld\t%s\n", T, (uint32_t)best_q, (double)best_p / (double)best_q,
              (long)(now - t0_unix), v.fits ? "STABLE BUT NOT DENSEST" : "NOT STABLE");
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
  if (v.certified) res->upper_bound = best_q ? (double)best_p / (double)best_q : 0.0;
  res->num_iter = iters;
  res->n = (uint32_t)best_q;
  res->cliques = best_p;
  res->density = best_q ? (double)best_p / (double)best_q : 0.0;
  res->m = best.empty() ? 0 : (uint32_t)kcl::edges_inside(g, best);
  res->max_flow_calls = flows;
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
              res->cliques ? "Max Flow" : "Goldberg", res->n, res->m, res->density,
              res->max_flow_calls, (long)((int64_t)time(nullptr) - t0_unix));
    }
    if (fclose(f) != 0) return fail(KCL_E_IO, "cannot write %s", stat_path);
  }
  return KCL_OK;
}
