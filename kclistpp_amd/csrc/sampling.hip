// sampling.hip - Sampling++ on the GPU (kCDensestSamplingKclistpp.c).
//
// The reference ("Seq-Sampling++", kCDensestSamplingKclistpp.c:262-489):
// count the k-cliques, keep each with probability p = min(1, 1e7 / count)
// (rand() seeded with time(NULL)), then run rounds of "++" over the kept
// cliques (each clique, in a fresh random order per round, gives one unit
// to its member of smallest load, loads updated at once), and after rounds
// 1, 2, 4, ... extract an approximate densest subgraph: nodes by decreasing
// load, each kept clique credited to its member of largest rank, the
// densest prefix, its induced subgraph rebuilt and its k-cliques counted
// exactly (:362-489).
//
// Here:
//   * the count is the plan's root-credit pass;
//   * the kept set is chosen by a seeded hash of each clique's ids, not by
//     rand(): the same graph, k, p and seed give the same set on any
//     schedule (the walk kernels list the cliques, mode kEmit);
//   * the rounds are synchronous (the Sim form of the ++ rounds): in round
//     t every kept clique gives its unit to its member of smallest load
//     after round t-1, ties to the smaller id (the cliques are stored with
//     ascending ids, and the first of equal loads wins, as in the
//     reference's update loop, :352-357); the reference's sequential,
//     shuffled rounds cannot be reproduced in parallel and depend on rand();
//   * extraction, the densest prefix, the upper bound and the exact count
//     of the prefix's induced subgraph follow :362-489 (loads summed in 64
//     bits where the reference's 32-bit sum would wrap).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "clique_flow.h"
#include "kcl_internal.h"

struct kcl_sampling {
  kcl_graph *g = nullptr;
  uint32_t k = 0;
  kcl_sampling_opts opts{};
  uint64_t total = 0;     // k-cliques of the graph
  uint64_t sampled = 0;   // cliques kept
  double prob = 1.0;      // keep probability
  uint32_t rounds = 0;    // synchronous rounds run
  kcl::DevBuf<uint32_t> ck;                  // kept cliques, k ids each (ascending)
  kcl::DevBuf<unsigned long long> load, add;  // per node: load, credits of a round
};

namespace kcl {

namespace {

constexpr uint64_t kDefaultTarget = 10000000ull;  // CLIQUES_TO_SAMPLE (:138)

// One synchronous round: every kept clique credits its member of smallest
// load (the first such member: ids ascending, so the smallest id).
__global__ void sampling_round_kernel(const uint32_t *__restrict__ ck, uint64_t nck, uint32_t k,
                                      const unsigned long long *__restrict__ load,
                                      unsigned long long *__restrict__ add) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nck;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *m = ck + c * k;
    uint32_t w = m[0];
    unsigned long long lw = load[w];
    for (uint32_t j = 1; j < k; ++j) {
      const uint32_t x = m[j];
      const unsigned long long lx = load[x];
      if (lx < lw) {
        w = x;
        lw = lx;
      }
    }
    atomicAdd(&add[w], 1ull);
  }
}

__global__ void sampling_apply_kernel(unsigned long long *__restrict__ load,
                                      unsigned long long *__restrict__ add, uint32_t n) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    load[v] += add[v];
    add[v] = 0;
  }
}

uint32_t grid_of(uint64_t work) {
  uint64_t b = (work + 255) / 256;
  if (b > 8192) b = 8192;
  return (uint32_t)(b ? b : 1);
}

// k-cliques of the graph on the given edge pairs (SortByCore, Relabel,
// MakeGraph, then a counting pass), with its node and edge counts.
kcl_status count_cliques_of(const std::vector<uint32_t> &pairs, uint32_t n_in, uint32_t k,
                            uint32_t *n, uint64_t *m, uint64_t *cliques) {
  *n = 0;
  *m = 0;
  *cliques = 0;
  if (pairs.empty()) return KCL_OK;
  kcl_graph *sub = nullptr;
  KCL_TRY(kcl_graph_create(pairs.data(), pairs.size() / 2, n_in, &sub));
  *n = sub->n;
  *m = sub->m;
  kcl_status st = KCL_OK;
  if (sub->n > 0 && k <= sub->core + 1) {
    PassPlan p;
    kcl_plan_info info{};
    st = make_plan(sub, k, 1u << 17, sub->stream, &p, nullptr, &info, KCL_SCHED_DEGREE);
    if (st == KCL_OK) *cliques = info.cliques;
  }
  kcl_graph_destroy(sub);
  return st;
}

// Lists the kept cliques of g (keep rate thr53 / 2^53, seed) into a
// device buffer; *kept = the cliques kept (the buffer holds cap of them; a
// second listing into an exact-size buffer follows when more were kept).
kcl_status emit_cliques(kcl_graph *g, uint32_t k, const PassPlan &plan, uint64_t total,
                        unsigned long long thr53, uint64_t seed, DevBuf<uint32_t> *buf,
                        uint64_t *kept) {
  *kept = 0;
  if (!total) return KCL_OK;
  DevicePlan dp;
  KCL_TRY(upload_plan(plan, &dp));
  const uint32_t n = g->n;
  DevBuf<unsigned long long> zkey;
  DevBuf<EmitCtl> ctl;
  KCL_TRY(zkey.alloc(n));
  KCL_TRY(zero_u64(zkey.p, n, g->stream));
  KCL_TRY(ctl.alloc(1));
  // Room for the expected sample plus seven standard deviations.
  const double p = thr53 >= (1ull << 53) ? 1.0 : std::ldexp((double)thr53, -53);
  const double mean = (double)total * p;
  uint64_t cap = p >= 1.0 ? total : (uint64_t)(mean + 7.0 * std::sqrt(mean) + 1024.0);
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (cap > (1ull << 40) / k)
      return fail(KCL_E_UNSUPPORTED, "%llu kept %u-cliques: over 2^40 ids",
                  (unsigned long long)cap, k);
    KCL_TRY(buf->alloc(cap ? cap * k : 1));
    EmitCtl h{0ull, (unsigned long long)cap, thr53, (unsigned long long)seed};
    KCL_HIP(hipMemcpyAsync(ctl.p, &h, sizeof(h), hipMemcpyHostToDevice, g->stream));
    KCL_TRY(launch_pass(g, (const uint64_t *)zkey.p, k, kEmit, dp,
                        reinterpret_cast<unsigned long long *>(buf->p), g->stream,
                        reinterpret_cast<unsigned long long *>(ctl.p)));
    KCL_HIP(hipMemcpyAsync(&h, ctl.p, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    KCL_TRY(pass_status(dp, g->stream));  // synchronises the stream
    *kept = h.count;
    if (h.count <= cap) {
      if (p >= 1.0 && h.count != total)
        return fail(KCL_E_HIP, "listed %llu of %llu cliques", (unsigned long long)h.count,
                    (unsigned long long)total);
      return KCL_OK;
    }
    cap = h.count;
  }
  return fail(KCL_E_HIP, "the kept cliques changed between two listings");
}

}  // namespace

// Every k-clique (k >= 3) inside the node set `in` of g, k global ids each
// (ascending within a clique, cliques in no particular order): the induced
// subgraph is built and its cliques listed by the walk kernels, then mapped
// back to g's ids.
kcl_status list_cliques_gpu(const kcl_graph *g, uint32_t k, const std::vector<uint8_t> &in,
                            std::vector<uint32_t> *out) {
  out->clear();
  const uint32_t n = g->n;
  std::vector<uint32_t> local(n, UINT32_MAX), global;
  for (uint32_t v = 0; v < n; ++v)
    if (in[v]) {
      local[v] = (uint32_t)global.size();
      global.push_back(v);
    }
  std::vector<uint32_t> pairs;
  for (uint32_t v : global)
    for (uint32_t j = g->cd[v]; j < g->cd[v + 1]; ++j) {
      const uint32_t x = g->adj[j];
      if (local[x] == UINT32_MAX) continue;
      pairs.push_back(local[v]);
      pairs.push_back(local[x]);
    }
  if (pairs.empty() || k < 3) return k < 3 ? fail(KCL_E_INVALID_ARG, "k = %u", k) : KCL_OK;
  kcl_graph *sub = nullptr;
  KCL_TRY(kcl_graph_create(pairs.data(), pairs.size() / 2, (uint32_t)global.size(), &sub));
  kcl_status st = KCL_OK;
  do {
    if (k > sub->core + 1) break;
    PassPlan plan;
    kcl_plan_info info{};
    if ((st = make_plan(sub, k, 1u << 17, sub->stream, &plan, nullptr, &info,
                        KCL_SCHED_DEGREE)) != KCL_OK)
      break;
    DevBuf<uint32_t> buf;
    uint64_t kept = 0;
    if ((st = emit_cliques(sub, k, plan, info.cliques, 1ull << 53, 0, &buf, &kept)) != KCL_OK)
      break;
    out->resize((size_t)kept * k);
    if (kept) {
      hipError_t e = hipMemcpyAsync(out->data(), buf.p, out->size() * 4, hipMemcpyDeviceToHost,
                                    sub->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(sub->stream);
      if (e != hipSuccess) {
        st = fail(KCL_E_HIP, "hipMemcpy of the listed cliques: %s", hipGetErrorString(e));
        break;
      }
    }
    // The subgraph's ids (its own degeneracy order) back to g's.
    std::vector<uint32_t> back(sub->n, UINT32_MAX);
    for (uint32_t i = 0; i < (uint32_t)global.size(); ++i)
      if (sub->rank[i] < sub->n) back[sub->rank[i]] = global[i];
    uint32_t *ids = out->data();
    parallel_chunks(out->size(), host_workers(), [&](uint32_t, uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i) ids[i] = back[ids[i]];
    });
  } while (0);
  kcl_graph_destroy(sub);
  return st;
}

}  // namespace kcl

using namespace kcl;

extern "C" kcl_status kcl_sampling_create(kcl_graph *g, uint32_t k, const kcl_sampling_opts *opts,
                                          kcl_sampling **out) {
  if (!g || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (k < 3 || k > 255)
    return fail(KCL_E_INVALID_ARG, "k = %u: Sampling++ takes 3 <= k <= 255 "
                                   "(kCDensestSamplingKclistpp.c:8)", k);
  kcl_sampling *s = new kcl_sampling();
  s->g = g;
  s->k = k;
  if (opts) s->opts = *opts;
  if (!s->opts.target) s->opts.target = kDefaultTarget;
  auto bail = [&](kcl_status st) {
    kcl_sampling_destroy(s);
    return st;
  };
  const uint32_t n = g->n;
  kcl_status st;
  if ((st = s->load.alloc(n ? n : 1)) != KCL_OK) return bail(st);
  if ((st = s->add.alloc(n ? n : 1)) != KCL_OK) return bail(st);
  if ((st = zero_u64(s->load.p, n, g->stream)) != KCL_OK) return bail(st);
  if ((st = zero_u64(s->add.p, n, g->stream)) != KCL_OK) return bail(st);
  if (n == 0 || k > g->core + 1) {  // no k-clique
    hipStreamSynchronize(g->stream);
    *out = s;
    return KCL_OK;
  }
  // Count (the plan's root-credit pass), then list the kept cliques.
  PassPlan plan;
  kcl_plan_info info{};
  if ((st = make_plan(g, k, 1u << 17, g->stream, &plan, nullptr, &info, KCL_SCHED_DEGREE)) !=
      KCL_OK)
    return bail(st);
  s->total = info.cliques;
  s->prob = s->total > s->opts.target ? (double)s->opts.target / (double)s->total : 1.0;
  const unsigned long long thr53 =
      s->prob >= 1.0 ? (1ull << 53) : (unsigned long long)std::ldexp(s->prob, 53);
  uint64_t kept = 0;
  if ((st = emit_cliques(g, k, plan, s->total, thr53, s->opts.seed, &s->ck, &kept)) != KCL_OK)
    return bail(st);
  s->sampled = kept;
  *out = s;
  return KCL_OK;
}

void kcl_sampling_destroy(kcl_sampling *s) {
  if (!s) return;
  delete s;
}

extern "C" kcl_status kcl_sampling_get_info(kcl_sampling *s, kcl_sampling_info *info) {
  if (!s || !info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  info->total = s->total;
  info->sampled = s->sampled;
  info->prob = s->prob;
  info->rounds = s->rounds;
  return KCL_OK;
}

extern "C" kcl_status kcl_sampling_cliques(kcl_sampling *s, uint32_t *out) {
  if (!s || (!out && s->sampled)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!s->sampled) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(out, s->ck.p, s->sampled * s->k * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s->g->stream));
  KCL_HIP(hipStreamSynchronize(s->g->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sampling_step(kcl_sampling *s) {
  if (!s) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const uint32_t n = s->g->n;
  if (s->sampled) {
    hipLaunchKernelGGL(sampling_round_kernel, dim3(grid_of(s->sampled)), dim3(256), 0,
                       s->g->stream, s->ck.p, (uint64_t)s->sampled, s->k, s->load.p, s->add.p);
    KCL_HIP(hipGetLastError());
    hipLaunchKernelGGL(sampling_apply_kernel, dim3(grid_of(n)), dim3(256), 0, s->g->stream,
                       s->load.p, s->add.p, n);
    KCL_HIP(hipGetLastError());
  }
  KCL_HIP(hipStreamSynchronize(s->g->stream));
  ++s->rounds;
  return KCL_OK;
}

extern "C" kcl_status kcl_sampling_loads(kcl_sampling *s, uint64_t *out) {
  if (!s || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!s->g->n) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(out, s->load.p, (size_t)s->g->n * 8, hipMemcpyDeviceToHost,
                         s->g->stream));
  KCL_HIP(hipStreamSynchronize(s->g->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sampling_report(kcl_sampling *s, kcl_sampling_record *rec,
                                          uint32_t *order) {
  if (!s || !rec) return fail(KCL_E_INVALID_ARG, "NULL argument");
  std::memset(rec, 0, sizeof(*rec));
  const uint32_t n = s->g->n, k = s->k, T = s->rounds;
  rec->T = T;
  if (n == 0) return KCL_OK;
  std::vector<uint64_t> load(n);
  KCL_TRY(kcl_sampling_loads(s, load.data()));
  // Nodes by decreasing load, ties by increasing id (glibc's qsort is a
  // stable merge sort here, :362-371).
  std::vector<uint32_t> ord(n), rank(n);
  for (uint32_t v = 0; v < n; ++v) ord[v] = v;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](uint32_t a, uint32_t b) { return load[a] > load[b]; });
  for (uint32_t i = 0; i < n; ++i) rank[ord[i]] = i;
  if (order) std::memcpy(order, ord.data(), (size_t)n * 4);
  // Each kept clique to its member of largest rank (:375-382).
  std::vector<uint32_t> ck((size_t)s->sampled * k);
  KCL_TRY(kcl_sampling_cliques(s, ck.data()));
  std::vector<uint64_t> pushed(n, 0);
  for (uint64_t c = 0; c < s->sampled; ++c) {
    uint32_t best = rank[ck[c * k]];
    for (uint32_t j = 1; j < k; ++j) best = std::max(best, rank[ck[c * k + j]]);
    ++pushed[best];
  }
  // The densest prefix: the first strict maximum of kept cliques / size
  // (:384-393).
  double density = -1.0;
  uint32_t size = 0;
  uint64_t run = 0, m_at = 0;
  for (uint32_t i = 0; i < n; ++i) {
    run += pushed[i];
    const double dd = (double)run / (double)(i + 1);
    if (density < dd) {
      density = dd;
      size = i + 1;
      m_at = run;
    }
  }
  rec->prefix = size;
  rec->prefix_sampled = m_at;
  // Upper bound (:399-413), then scaled as the output line does (:477).
  double ub = 0.0, ip1ck = 0.0;
  uint64_t sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    sum += load[ord[i]];
    if (i + 1 == k) ip1ck = 1.0;
    else if (i + 1 > k) ip1ck = (ip1ck * (double)(i + 1)) / (double)(i + 1 - k);
    if (T > 0 && ip1ck < (double)sum / T) {
      ub = ip1ck / (double)(i + 1);
    } else {
      const double cand = T > 0 ? (double)sum / T / (double)(i + 1) : 0.0;
      if (ub < cand) ub = cand;
      break;
    }
  }
  rec->ub_sample = ub;
  rec->ub = ub / s->prob / (1.0 - std::sqrt(6.0 * std::log((double)n) / ub));
  // The prefix's induced subgraph with the nodes renamed by rank (:415-435),
  // rebuilt and counted exactly.
  std::vector<uint32_t> pairs;
  const std::vector<uint32_t> &cd = s->g->cd, &adj = s->g->adj;
  for (uint32_t v = 0; v < n; ++v) {
    if (rank[v] >= size) continue;
    for (uint32_t j = cd[v]; j < cd[v + 1]; ++j) {
      const uint32_t x = adj[j];
      if (rank[x] >= size) continue;
      pairs.push_back(rank[v]);
      pairs.push_back(rank[x]);
    }
  }
  uint32_t sn = 0;
  uint64_t sm = 0, cl = 0;
  KCL_TRY(count_cliques_of(pairs, size, k, &sn, &sm, &cl));
  rec->n = sn;
  rec->m = (uint32_t)sm;
  rec->cliques = cl;
  rec->edge_density = sn > 1 ? (double)sm * 2.0 / sn / (sn - 1) : 0.0;
  rec->density = sn ? (double)cl / sn : 0.0;
  return KCL_OK;
}

extern "C" kcl_status kcl_sampling_run(kcl_sampling *s, uint32_t num_iter,
                                       kcl_sampling_record *recs, uint32_t *nrec) {
  if (!s || !nrec) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const uint32_t cap = *nrec;
  *nrec = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t T = 1; T <= num_iter; T <<= 1) {
    while (s->rounds < T) KCL_TRY(kcl_sampling_step(s));
    kcl_sampling_record r;
    KCL_TRY(kcl_sampling_report(s, &r, nullptr));
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (recs && *nrec < cap) recs[*nrec] = r;
    ++*nrec;
  }
  return KCL_OK;
}
