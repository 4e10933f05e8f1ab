// batch.hip - many graphs, each with its own k, iterated together.
//
// The graphs are laid out in HBM as one disjoint union: graph i's
// relabelled node v is node off[i] + v, its DAG rows are its own rows with
// ids shifted, so every root's out-neighbourhood stays inside its graph and
// the enumeration kernels run on the union unchanged.  Roots are grouped by
// k (one task plan per distinct k over the union).  One iteration t for
// every graph - the passes of all plans, then one rho update over all
// nodes - is captured once in a hipGraph and replayed per t (t is read from
// device memory, so the captured graph is the same for every t).
//
// A report (CDF_FindDensestSubset, kCDensestCmpSync.c:342-383, for every
// graph at once): two stable radix sorts over all nodes give each graph's
// level order (by floor(rho) descending, then id, then grouped by graph),
// one PAVA-preprocess pass over the union credits every clique to its
// graph's last-level member, and the densest prefix and upper bound are
// scanned per graph on the host in the reference's order.
//
// Results per graph are those of kcl_sim on that graph alone: cliques
// never cross graphs, the keys compare only members of one graph, and the
// rho update is elementwise.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "kcl_internal.h"

namespace kcl {
namespace {

constexpr unsigned long long kLimit53 = 1ull << 53;

uint32_t grid_of(uint64_t work) {
  uint64_t g = (work + 255) / 256;
  if (g > 16384) g = 16384;
  return (uint32_t)(g ? g : 1);
}

__global__ void set_u32_kernel(uint32_t *p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

__global__ void zero_small_kernel(unsigned long long *p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__device__ __forceinline__ void block_stats(unsigned long long sum, unsigned long long big,
                                            unsigned long long *stats) {
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off);
    big |= __shfl_xor(big, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (sum) atomicAdd(&stats[0], sum);
    if (big) atomicOr(&stats[1], big);
  }
}

// rho = (t*rho + 2*cnt)/(t + 2) (kCDensestCmpSync.c:402-403) with t from
// device memory; V += (t+1)*cnt; last = cnt; cnt = 0; stats as rho_update.
__global__ void batch_update_kernel(uint64_t n, const uint32_t *__restrict__ tp,
                                    double *__restrict__ rho, unsigned long long *__restrict__ cnt,
                                    unsigned long long *__restrict__ last,
                                    unsigned long long *__restrict__ V,
                                    unsigned long long *__restrict__ stats) {
  const uint32_t t = *tp;
  const double td = (double)t;
  unsigned long long sum = 0, big = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = cnt[i];
    const double add = (double)c;
    rho[i] = (td * rho[i] + 2.0 * add) / (td + 2.0);
    V[i] += (unsigned long long)(t + 1) * c;
    last[i] = c;
    cnt[i] = 0;
    sum += c;
    big |= (c >= kLimit53) ? 1ull : 0ull;
  }
  block_stats(sum, big, stats);
}

// CDF_NodeCmp's key (kCDensestCmpSync.c:334-340): (unsigned long long)rho,
// descending, as an ascending key.
__global__ void level_key_kernel(uint64_t n, const double *__restrict__ rho,
                                 unsigned long long *__restrict__ keys,
                                 uint32_t *__restrict__ ids) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    keys[i] = ~(unsigned long long)rho[i];
    ids[i] = (uint32_t)i;
  }
}

__global__ void graph_key_kernel(uint64_t n, const uint32_t *__restrict__ vals,
                                 const uint32_t *__restrict__ gid, uint32_t *__restrict__ gkeys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    gkeys[i] = gid[vals[i]];
}

// Position p of the grouped order holds node v of graph g at level
// i = p - off[g].
__global__ void level_scatter_kernel(uint64_t n, const uint32_t *__restrict__ order,
                                     const uint32_t *__restrict__ gid,
                                     const uint32_t *__restrict__ off,
                                     const uint32_t *__restrict__ gn, uint32_t *__restrict__ level,
                                     unsigned long long *__restrict__ pava_key,
                                     uint32_t *__restrict__ reo_local) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = order[p];
    const uint32_t g = gid[v];
    const uint32_t i = (uint32_t)p - off[g];
    level[v] = i;
    pava_key[v] = (unsigned long long)(gn[g] - 1 - i);
    reo_local[p] = v - off[g];
  }
}

__global__ void tent_scatter_kernel(uint64_t n, const uint32_t *__restrict__ level,
                                    const uint32_t *__restrict__ gid,
                                    const uint32_t *__restrict__ off,
                                    unsigned long long *__restrict__ cnt,
                                    unsigned long long *__restrict__ tent,
                                    unsigned long long *__restrict__ stats) {
  unsigned long long big = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = cnt[v];
    tent[off[gid[v]] + level[v]] = c;
    cnt[v] = 0;
    big |= (c >= kLimit53) ? 1ull : 0ull;
  }
  block_stats(0, big, stats);
}

// CountEdges (Graph.c:158-166) per graph: edges with both ends among the
// first nb[g] nodes of graph g's level order.
__global__ void count_top_edges_kernel(uint64_t n, const uint32_t *__restrict__ cd,
                                       const uint32_t *__restrict__ adj,
                                       const uint32_t *__restrict__ level,
                                       const uint32_t *__restrict__ gid,
                                       const uint32_t *__restrict__ nb,
                                       unsigned long long *__restrict__ m) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = gid[v];
    const uint32_t lim = nb[g];
    if (level[v] >= lim) continue;
    unsigned long long c = 0;
    for (uint32_t j = cd[v]; j < cd[v + 1]; ++j) c += (level[adj[j]] < lim) ? 1u : 0u;
    if (c) atomicAdd(&m[g], c);
  }
}

}  // namespace
}  // namespace kcl

struct kcl_batch {
  struct KPlan {
    uint32_t k = 0;
    kcl::PassPlan host;
    kcl::DevicePlan dev;
  };
  uint32_t count = 0;
  std::vector<uint32_t> k, n, off;  // off: count + 1 entries
  std::vector<uint64_t> cliques;
  kcl_graph *u = nullptr;
  std::vector<std::unique_ptr<KPlan>> plans;
  kcl::DevBuf<double> rho;
  kcl::DevBuf<unsigned long long> cnt, last, V, stats, tent, pava_key, keys_in, keys_out, m;
  kcl::DevBuf<uint32_t> gid, tdev, d_off, d_gn, d_nb, level, ids, vals1, gkeys_in, gkeys_out,
      order, reo_local;
  kcl::DevBuf<uint8_t> sort_tmp;
  size_t sort_tmp_bytes = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  kcl_batch_info info{};
  uint64_t last_visits = 0;
  double last_ms = 0;
  uint64_t nodes() const { return off.empty() ? 0 : off.back(); }
  ~kcl_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    plans.clear();
    kcl_graph_destroy(u);
  }
};

namespace kcl {
namespace {

// Enqueues copies of every plan's pass status word into out (read after
// the stream's next synchronisation, then checked with check_status).
kcl_status enqueue_status(kcl_batch *b, std::vector<uint32_t> &out, hipStream_t st) {
  out.assign(b->plans.size(), 0);
  for (size_t i = 0; i < b->plans.size(); ++i)
    KCL_HIP(hipMemcpyAsync(&out[i], pass_status_word(b->plans[i]->dev), 4, hipMemcpyDeviceToHost,
                           st));
  return KCL_OK;
}

kcl_status check_status(const std::vector<uint32_t> &w) {
  for (uint32_t x : w) KCL_TRY(pass_status_error(x));
  return KCL_OK;
}

kcl_status launch_all(kcl_batch *b, const uint64_t *key) {
  for (auto &kp : b->plans)
    KCL_TRY(launch_pass(b->u, key, kp->k, kCreditWinner, kp->dev, b->cnt.p, b->u->stream));
  return KCL_OK;
}

// One plan with a row-width floor and a per-CU cap applied.
kcl_status apply_plan_config(kcl_batch::KPlan *kp, uint32_t min_class, uint32_t per_cu) {
  PassPlan p = kp->host;
  promote_plan(&p, min_class);
  KCL_TRY(upload_plan(p, &kp->dev));
  kp->dev.max_per_cu = per_cu;
  kp->dev.light_waves = 32;  // residency is what the autotuner tunes here
  return KCL_OK;
}

// Every plan with the same floor and cap.
kcl_status apply_config(kcl_batch *b, uint32_t min_class, uint32_t per_cu) {
  for (auto &kp : b->plans) KCL_TRY(apply_plan_config(kp.get(), min_class, per_cu));
  return KCL_OK;
}

// Time of one pass of every plan, or of plan `only` alone (ms, best of
// two), state unchanged.
kcl_status time_pass(kcl_batch *b, float *ms, const kcl_batch::KPlan *only = nullptr) {
  hipStream_t st = b->u->stream;
  float best = 1e30f;
  for (int rep = 0; rep < 2; ++rep) {
    KCL_HIP(hipEventRecord(b->ev[0], st));
    if (only)
      KCL_TRY(launch_pass(b->u, (const uint64_t *)b->rho.p, only->k, kCreditWinner,
                          const_cast<kcl_batch::KPlan *>(only)->dev, b->cnt.p, st));
    else
      KCL_TRY(launch_all(b, (const uint64_t *)b->rho.p));
    KCL_HIP(hipEventRecord(b->ev[1], st));
    KCL_TRY(zero_u64(b->cnt.p, b->nodes(), st));
    std::vector<uint32_t> status;
    if (only) {
      status.assign(1, 0);
      KCL_HIP(hipMemcpyAsync(status.data(),
                             pass_status_word(const_cast<kcl_batch::KPlan *>(only)->dev), 4,
                             hipMemcpyDeviceToHost, st));
    } else {
      KCL_TRY(enqueue_status(b, status, st));
    }
    KCL_HIP(hipStreamSynchronize(st));
    KCL_TRY(check_status(status));
    float t = 0;
    KCL_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
    best = std::min(best, t);
  }
  *ms = best;
  return KCL_OK;
}

// Per plan (one per k: each has its own kernels and task mix), row-width
// floors 64/128/256 bits (classes 0, 1, 3) and resident workgroups per CU
// (LDS-limited, 8, 16), each plan timed alone; the fastest configuration of
// each plan wins.  The plans' passes run nearly one after another (each
// fills the device), so their best configurations add up.  Reports the
// configuration of the plan whose best pass is longest.
kcl_status autotune(kcl_batch *b, uint32_t *min_class, uint32_t *per_cu) {
  static const uint32_t kFloors[] = {0, 1, 3};
  static const uint32_t kCaps[] = {0, 8, 16};
  *min_class = 0;
  *per_cu = 0;
  KCL_TRY(apply_config(b, 0, 0));  // every plan on the device before any is timed
  float longest = -1.0f;
  for (auto &kp : b->plans) {
    float best = 1e30f;
    uint32_t bf = 0, bc = 0;
    for (uint32_t f : kFloors)
      for (uint32_t c : kCaps) {
        KCL_TRY(apply_plan_config(kp.get(), f, c));
        float ms = 0;
        KCL_TRY(time_pass(b, &ms, kp.get()));
        if (ms < best) {
          best = ms;
          bf = f;
          bc = c;
        }
      }
    KCL_TRY(apply_plan_config(kp.get(), bf, bc));
    if (best > longest) {
      longest = best;
      *min_class = bf;
      *per_cu = bc;
    }
  }
  return KCL_OK;
}

kcl_status capture(kcl_batch *b) {
  hipStream_t st = b->u->stream;
  KCL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  kcl_status s = KCL_OK;
  zero_small_kernel<<<1, 64, 0, st>>>(b->stats.p, 4);
  if (hipGetLastError() != hipSuccess) s = fail(KCL_E_HIP, "launch failed during capture");
  if (s == KCL_OK) s = launch_all(b, (const uint64_t *)b->rho.p);
  if (s == KCL_OK) {
    const uint64_t n = b->nodes();
    if (n)
      batch_update_kernel<<<grid_of(n), 256, 0, st>>>(n, b->tdev.p, b->rho.p, b->cnt.p, b->last.p,
                                                      b->V.p, b->stats.p);
    if (hipGetLastError() != hipSuccess) s = fail(KCL_E_HIP, "launch failed during capture");
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (s != KCL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess) return fail(KCL_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = g;
  KCL_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return KCL_OK;
}

kcl_status batch_create(kcl_graph *const *graphs, const uint32_t *ks, uint32_t count,
                        const kcl_batch_opts &o, kcl_batch *b) {
  const auto t0 = std::chrono::steady_clock::now();
  b->count = count;
  b->k.assign(ks, ks + count);
  b->n.resize(count);
  b->off.assign(count + 1, 0);
  uint64_t N = 0, M = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const kcl_graph *g = graphs[i];
    if (!g) return fail(KCL_E_INVALID_ARG, "graph %u is NULL", i);
    if (ks[i] < 2 || ks[i] > 255) return fail(KCL_E_INVALID_ARG, "graph %u: k = %u", i, ks[i]);
    if (ks[i] > max_pass_k() && ks[i] <= g->core + 1)
      return fail(KCL_E_UNSUPPORTED, "graph %u: k = %u beyond the compiled k <= %u", i, ks[i],
                  max_pass_k());
    b->n[i] = g->n;
    b->off[i] = (uint32_t)N;
    N += g->n;
    M += g->m;
    if (N >= 0xffffffffull || M >= 0xffffffffull)
      return fail(KCL_E_UNSUPPORTED, "the batch exceeds 2^32 nodes or edges");
  }
  b->off[count] = (uint32_t)N;
  std::vector<uint32_t> cd(N + 1), adj(M), gid(N);
  {
    uint64_t e = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const kcl_graph *g = graphs[i];
      const uint32_t base = b->off[i];
      for (uint32_t v = 0; v < g->n; ++v) {
        cd[base + v] = (uint32_t)(e + g->cd[v]);
        gid[base + v] = i;
      }
      for (uint64_t j = 0; j < g->m; ++j) adj[e + j] = base + g->adj[j];
      e += g->m;
    }
    cd[N] = (uint32_t)M;
  }
  KCL_TRY(graph_from_csr(std::move(cd), std::move(adj), &b->u));
  for (uint32_t i = 0; i < count; ++i) {
    b->u->edge_above = std::min(b->u->edge_above, graphs[i]->edge_above);
    b->u->dense_items = std::min(b->u->dense_items, graphs[i]->dense_items);
  }
  hipStream_t st = b->u->stream;
  for (int i = 0; i < 2; ++i) KCL_HIP(hipEventCreate(&b->ev[i]));
  KCL_TRY(b->rho.alloc(N));
  KCL_TRY(b->cnt.alloc(N));
  KCL_TRY(b->last.alloc(N));
  KCL_TRY(b->V.alloc(N));
  KCL_TRY(b->stats.alloc(4));
  KCL_TRY(b->tdev.alloc(1));
  KCL_TRY(b->gid.alloc(N));
  KCL_TRY(b->d_off.alloc(count + 1));
  KCL_TRY(b->d_gn.alloc(count));
  KCL_TRY(b->d_nb.alloc(count));
  KCL_TRY(b->m.alloc(count));
  KCL_HIP(hipMemsetAsync(b->rho.p, 0, (size_t)(N ? N : 1) * sizeof(double), st));
  KCL_TRY(zero_u64(b->cnt.p, N, st));
  KCL_TRY(zero_u64(b->last.p, N, st));
  KCL_TRY(zero_u64(b->V.p, N, st));
  KCL_HIP(hipMemcpyAsync(b->gid.p, gid.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  KCL_HIP(hipMemcpyAsync(b->d_off.p, b->off.data(), (count + 1) * sizeof(uint32_t),
                         hipMemcpyHostToDevice, st));
  KCL_HIP(hipMemcpyAsync(b->d_gn.p, b->n.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice,
                         st));
  KCL_HIP(hipStreamSynchronize(st));

  // One plan per distinct k, over the roots of the graphs with that k.
  std::vector<uint32_t> kv(b->k);
  std::sort(kv.begin(), kv.end());
  kv.erase(std::unique(kv.begin(), kv.end()), kv.end());
  b->cliques.assign(count, 0);
  uint32_t tasks = 0;
  const uint32_t split = o.split_work ? o.split_work : (1u << 17);
  for (uint32_t k : kv) {
    std::vector<uint64_t> cand(N, 0);
    bool any = false;
    for (uint32_t i = 0; i < count; ++i)
      if (b->k[i] == k && graphs[i]->core + 1 >= k && b->n[i]) {
        std::fill(cand.begin() + b->off[i], cand.begin() + b->off[i + 1], 1ull);
        any = true;
      }
    if (!any) continue;
    auto kp = std::make_unique<kcl_batch::KPlan>();
    kp->k = k;
    std::vector<uint64_t> ck;
    KCL_TRY(make_plan(b->u, k, split, st, &kp->host, &ck, nullptr, o.schedule, o.model, &cand));
    for (uint32_t v = 0; v < N; ++v)
      if (ck[v]) b->cliques[gid[v]] += ck[v];
    tasks += (uint32_t)(kp->host.tasks.size() + kp->host.edge_tasks.size());
    b->plans.push_back(std::move(kp));
  }
  const auto t1 = std::chrono::steady_clock::now();
  uint32_t min_class = 0, per_cu = o.per_cu;
  if (o.row_bits > 128) min_class = 3;
  else if (o.row_bits > 64) min_class = 1;
  if (o.tune) {
    KCL_TRY(autotune(b, &min_class, &per_cu));
  } else {
    KCL_TRY(apply_config(b, min_class, per_cu));
  }
  const auto t2 = std::chrono::steady_clock::now();
  KCL_TRY(capture(b));
  kcl_batch_info &in = b->info;
  in.graphs = count;
  in.distinct_k = (uint32_t)b->plans.size();
  in.nodes = N;
  in.edges = M;
  in.cliques = 0;
  for (uint64_t c : b->cliques) in.cliques += c;
  in.tasks = tasks;
  in.row_bits = 64 * kClassWords[min_class];
  in.per_cu = per_cu;
  in.plan_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  in.tune_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  return KCL_OK;
}

kcl_status batch_step(kcl_batch *b, uint32_t t) {
  hipStream_t st = b->u->stream;
  set_u32_kernel<<<1, 64, 0, st>>>(b->tdev.p, t);
  KCL_HIP(hipGetLastError());
  KCL_HIP(hipEventRecord(b->ev[0], st));
  KCL_HIP(hipGraphLaunch(b->exec, st));
  KCL_HIP(hipEventRecord(b->ev[1], st));
  unsigned long long sv[2] = {0, 0};
  std::vector<uint32_t> status;
  KCL_HIP(hipMemcpyAsync(sv, b->stats.p, 16, hipMemcpyDeviceToHost, st));
  KCL_TRY(enqueue_status(b, status, st));
  KCL_HIP(hipStreamSynchronize(st));
  KCL_TRY(check_status(status));
  float ms = 0;
  KCL_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->last_ms = ms;
  b->last_visits = sv[0];
  if (sv[1])
    return fail(KCL_E_OVERFLOW,
                "iteration %u: a node received >= 2^53 cliques; the reference's double "
                "counters are no longer exact",
                t);
  return KCL_OK;
}

kcl_status batch_densest(kcl_batch *b, kcl_densest_info *out) {
  const uint64_t N = b->nodes();
  const uint32_t count = b->count;
  for (uint32_t i = 0; i < count; ++i) {
    out[i].n = 0;
    out[i].m = 0;
    out[i].density = -1;
    out[i].ub = 0;
    out[i].sum = 0;
  }
  if (N == 0) return KCL_OK;
  hipStream_t st = b->u->stream;
  if (!b->level.p) {
    KCL_TRY(b->level.alloc(N));
    KCL_TRY(b->ids.alloc(N));
    KCL_TRY(b->vals1.alloc(N));
    KCL_TRY(b->gkeys_in.alloc(N));
    KCL_TRY(b->gkeys_out.alloc(N));
    KCL_TRY(b->order.alloc(N));
    KCL_TRY(b->reo_local.alloc(N));
    KCL_TRY(b->keys_in.alloc(N));
    KCL_TRY(b->keys_out.alloc(N));
    KCL_TRY(b->pava_key.alloc(N));
    KCL_TRY(b->tent.alloc(N));
    size_t a = 0, c = 0;
    KCL_HIP(rocprim::radix_sort_pairs(nullptr, a, b->keys_in.p, b->keys_out.p, b->ids.p,
                                      b->vals1.p, (size_t)N, 0, 64, st));
    KCL_HIP(rocprim::radix_sort_pairs(nullptr, c, b->gkeys_in.p, b->gkeys_out.p, b->vals1.p,
                                      b->order.p, (size_t)N, 0, 32, st));
    b->sort_tmp_bytes = std::max(a, c);
    KCL_TRY(b->sort_tmp.alloc(b->sort_tmp_bytes));
  }
  uint32_t gbits = 1;
  while (gbits < 32 && (1ull << gbits) < count) ++gbits;
  level_key_kernel<<<grid_of(N), 256, 0, st>>>(N, b->rho.p, b->keys_in.p, b->ids.p);
  KCL_HIP(hipGetLastError());
  size_t tb = b->sort_tmp_bytes;
  KCL_HIP(rocprim::radix_sort_pairs(b->sort_tmp.p, tb, b->keys_in.p, b->keys_out.p, b->ids.p,
                                    b->vals1.p, (size_t)N, 0, 64, st));
  graph_key_kernel<<<grid_of(N), 256, 0, st>>>(N, b->vals1.p, b->gid.p, b->gkeys_in.p);
  KCL_HIP(hipGetLastError());
  tb = b->sort_tmp_bytes;
  KCL_HIP(rocprim::radix_sort_pairs(b->sort_tmp.p, tb, b->gkeys_in.p, b->gkeys_out.p, b->vals1.p,
                                    b->order.p, (size_t)N, 0, gbits, st));
  level_scatter_kernel<<<grid_of(N), 256, 0, st>>>(N, b->order.p, b->gid.p, b->d_off.p, b->d_gn.p,
                                                   b->level.p, b->pava_key.p, b->reo_local.p);
  KCL_HIP(hipGetLastError());
  zero_small_kernel<<<1, 64, 0, st>>>(b->stats.p, 4);
  KCL_TRY(launch_all(b, (const uint64_t *)b->pava_key.p));
  tent_scatter_kernel<<<grid_of(N), 256, 0, st>>>(N, b->level.p, b->gid.p, b->d_off.p, b->cnt.p,
                                                  b->tent.p, b->stats.p);
  KCL_HIP(hipGetLastError());
  std::vector<unsigned long long> tent(N);
  std::vector<uint32_t> reo(N);
  std::vector<double> rho(N);
  unsigned long long sv[2] = {0, 0};
  KCL_HIP(hipMemcpyAsync(tent.data(), b->tent.p, N * 8, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(reo.data(), b->reo_local.p, N * 4, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(rho.data(), b->rho.p, N * 8, hipMemcpyDeviceToHost, st));
  std::vector<uint32_t> status;
  KCL_HIP(hipMemcpyAsync(sv, b->stats.p, 16, hipMemcpyDeviceToHost, st));
  KCL_TRY(enqueue_status(b, status, st));
  KCL_HIP(hipStreamSynchronize(st));
  KCL_TRY(check_status(status));
  if (sv[1]) return fail(KCL_E_OVERFLOW, "a level received >= 2^53 cliques");
  std::vector<uint32_t> nb(count, 0);
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t ni = b->n[i], o = b->off[i];
    if (ni == 0) continue;
    KCL_TRY(densest_host(tent.data() + o, reo.data() + o, rho.data() + o, ni, b->k[i], &out[i],
                         &nb[i]));
  }
  KCL_HIP(hipMemcpyAsync(b->d_nb.p, nb.data(), count * 4, hipMemcpyHostToDevice, st));
  KCL_TRY(zero_u64(b->m.p, count, st));
  count_top_edges_kernel<<<grid_of(N), 256, 0, st>>>(N, b->u->d_cd.p, b->u->d_adj.p, b->level.p,
                                                     b->gid.p, b->d_nb.p, b->m.p);
  KCL_HIP(hipGetLastError());
  std::vector<unsigned long long> m(count);
  KCL_HIP(hipMemcpyAsync(m.data(), b->m.p, count * 8, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  for (uint32_t i = 0; i < count; ++i) out[i].m = (uint32_t)m[i];
  return KCL_OK;
}

kcl_status copy_segment(kcl_batch *b, uint32_t i, const void *dev, size_t elem, void *out) {
  if (i >= b->count) return fail(KCL_E_INVALID_ARG, "graph %u of %u", i, b->count);
  if (!out && b->n[i]) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (b->n[i] == 0) return KCL_OK;
  KCL_HIP(hipMemcpy(out, (const char *)dev + (size_t)b->off[i] * elem, (size_t)b->n[i] * elem,
                    hipMemcpyDeviceToHost));
  return KCL_OK;
}

}  // namespace
}  // namespace kcl

using kcl::fail;

extern "C" kcl_status kcl_batch_create(kcl_graph *const *graphs, const uint32_t *ks,
                                       uint32_t count, const kcl_batch_opts *opts,
                                       kcl_batch **out) {
  if (!out || !graphs || !ks || count == 0) return fail(KCL_E_INVALID_ARG, "NULL argument or no graphs");
  *out = nullptr;
  kcl_batch_opts o{};
  if (opts) o = *opts;
  if (o.schedule > KCL_SCHED_DEGREE) return fail(KCL_E_INVALID_ARG, "unknown schedule");
  kcl_batch *b = new (std::nothrow) kcl_batch();
  if (!b) return fail(KCL_E_OOM_HOST, "out of host memory");
  const kcl_status st = kcl::batch_create(graphs, ks, count, o, b);
  if (st != KCL_OK) {
    delete b;
    return st;
  }
  *out = b;
  return KCL_OK;
}

extern "C" void kcl_batch_destroy(kcl_batch *b) { delete b; }

extern "C" kcl_status kcl_batch_get_info(const kcl_batch *b, kcl_batch_info *info) {
  if (!b || !info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *info = b->info;
  return KCL_OK;
}

extern "C" kcl_status kcl_batch_graph_cliques(const kcl_batch *b, uint64_t *out) {
  if (!b || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  for (uint32_t i = 0; i < b->count; ++i) out[i] = b->cliques[i];
  return KCL_OK;
}

extern "C" kcl_status kcl_batch_reset(kcl_batch *b) {
  if (!b) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const uint64_t n = b->nodes();
  hipStream_t st = b->u->stream;
  KCL_HIP(hipMemsetAsync(b->rho.p, 0, (size_t)(n ? n : 1) * sizeof(double), st));
  KCL_TRY(kcl::zero_u64(b->cnt.p, n, st));
  KCL_TRY(kcl::zero_u64(b->last.p, n, st));
  KCL_TRY(kcl::zero_u64(b->V.p, n, st));
  KCL_HIP(hipStreamSynchronize(st));
  b->last_visits = 0;
  b->last_ms = 0;
  return KCL_OK;
}

extern "C" kcl_status kcl_batch_step(kcl_batch *b, uint32_t t) {
  if (!b || t == 0) return fail(KCL_E_INVALID_ARG, "NULL batch or t = 0");
  return kcl::batch_step(b, t);
}

extern "C" kcl_status kcl_batch_last_pass(const kcl_batch *b, uint64_t *visits, double *ms) {
  if (!b) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (visits) *visits = b->last_visits;
  if (ms) *ms = b->last_ms;
  return KCL_OK;
}

extern "C" kcl_status kcl_batch_counts(kcl_batch *b, uint32_t i, uint64_t *out) {
  if (!b) return fail(KCL_E_INVALID_ARG, "NULL argument");
  return kcl::copy_segment(b, i, b->last.p, 8, out);
}

extern "C" kcl_status kcl_batch_rho(kcl_batch *b, uint32_t i, double *out) {
  if (!b) return fail(KCL_E_INVALID_ARG, "NULL argument");
  return kcl::copy_segment(b, i, b->rho.p, 8, out);
}

extern "C" kcl_status kcl_batch_totals(kcl_batch *b, uint32_t i, uint64_t *out) {
  if (!b) return fail(KCL_E_INVALID_ARG, "NULL argument");
  return kcl::copy_segment(b, i, b->V.p, 8, out);
}

extern "C" kcl_status kcl_batch_find_densest(kcl_batch *b, kcl_densest_info *out) {
  if (!b || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  return kcl::batch_densest(b, out);
}

extern "C" kcl_status kcl_batch_run(kcl_batch *b, uint32_t num_iter, kcl_record *records,
                                    uint32_t *nrec) {
  if (!b || !records || !nrec) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const auto c0 = std::chrono::steady_clock::now();
  uint32_t T = 1;
  while (T * 2 <= num_iter && T < (1u << 31)) T *= 2;
  if (num_iter == 0) T = 0;
  std::vector<kcl_densest_info> d(b->count);
  for (uint32_t i = 0; i < b->count; ++i) nrec[i] = 0;
  for (uint32_t t = 1; t <= T; ++t) {
    KCL_TRY(kcl::batch_step(b, t));
    if ((t & (t - 1)) != 0) continue;
    KCL_TRY(kcl::batch_densest(b, d.data()));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
    for (uint32_t i = 0; i < b->count; ++i) {
      if (nrec[i] >= 32) continue;
      kcl_record &r = records[(size_t)i * 32 + nrec[i]++];
      r.T = t;
      r.n = d[i].n;
      r.m = d[i].m;
      r.density = d[i].density;
      r.ub = d[i].ub;
      r.sum = d[i].sum;
      r.seconds = secs;
    }
  }
  return KCL_OK;
}
