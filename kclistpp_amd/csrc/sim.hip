// sim.hip - Sim-kClist++ state on the device and the CDF_Main driver.
//
// Replaces AllocCdf, the iteration loop of CDF_Main and
// CDF_FindDensestSubset (kCDensestCmpSync.c:130-136, 342-419).
//
// HBM layout per graph (n nodes, m DAG edges):
//   cd u32[n+1], adj u32[m]          DAG CSR (graph.hip)
//   rho f64[n]                       also read as the pass key: rho >= +0.0,
//                                    so its IEEE bits order like the values
//   cnt u64[n]                       credits of the running pass
//   last u64[n], V u64[n]            last pass's counts, exact totals
//   level u32[n], reordered u32[n], pava_key u64[n], tent u64[n]  (reports)
// fp64 arithmetic in this file is built with -ffp-contract=off: the update
// is t*rho (one rounding), 2*cnt (exact), their sum (one rounding), then a
// correctly rounded division - the reference's operation order.
#include <rocprim/device/device_radix_sort.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "kcl_internal.h"

namespace kcl {
kcl_status count_edges_device(const kcl_graph *g, const uint8_t *d_inc, uint32_t *m_out,
                              hipStream_t st);
uint32_t max_pass_k();
}  // namespace kcl

struct kcl_sim {
  kcl_graph *g = nullptr;
  uint32_t k = 0;
  kcl_sim_opts opts{};
  bool trivial = false;  // no k-clique can exist (k > core + 1): nothing to launch
  kcl::DevicePlan plan;
  kcl::DevBuf<double> rho;
  kcl::DevBuf<unsigned long long> cnt, last, V, stats;
  kcl::DevBuf<uint32_t> level, reordered;
  kcl::DevBuf<unsigned long long> pava_key, tent, sort_keys_in, sort_keys_out;
  kcl::DevBuf<uint32_t> iota;
  kcl::DevBuf<uint8_t> sort_tmp;
  size_t sort_tmp_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  uint64_t last_visits = 0;
  double last_kernel_ms = 0, last_pass_ms = 0;
  kcl_plan_info info{};
  std::vector<uint64_t> root_cliques;  // k-cliques per root (plan filter)
};

namespace kcl {

namespace {

constexpr unsigned long long kExactLimit = 1ull << 53;

uint32_t grid_for(uint64_t work, uint32_t block) {
  uint64_t g = (work + block - 1) / block;
  if (g > 16384) g = 16384;
  return (uint32_t)(g ? g : 1);
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

// rho = (t*rho + 2*cnt)/(t + 2) (kCDensestCmpSync.c:402-403); V += (t+1)*cnt;
// last = cnt; cnt = 0; stats[0] += sum(cnt); stats[1] |= any cnt >= 2^53.
__global__ void rho_update_kernel(uint32_t n, uint32_t t, double *__restrict__ rho,
                                  unsigned long long *__restrict__ cnt,
                                  unsigned long long *__restrict__ last,
                                  unsigned long long *__restrict__ V,
                                  unsigned long long *__restrict__ stats) {
  unsigned long long sum = 0, big = 0;
  const double td = (double)t;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long c = cnt[i];
    const double add = (double)c;
    rho[i] = (td * rho[i] + 2.0 * add) / (td + 2.0);
    V[i] += (unsigned long long)(t + 1) * c;
    last[i] = c;
    cnt[i] = 0;
    sum += c;
    big |= (c >= kExactLimit) ? 1ull : 0ull;
  }
  block_stats(sum, big, stats);
}

// Sort key of CDF_NodeCmp (kCDensestCmpSync.c:334-340): (unsigned long
// long)rho descending; complemented so an ascending stable radix sort of
// (key, id) pairs gives floor(rho) descending, ties by id ascending.
// full != 0: the whole double descending (kCDensestMem.c:314-319; rho >= +0.0
// so its IEEE bits order like the values).
__global__ void level_key_kernel(uint32_t n, const double *__restrict__ rho,
                                 unsigned long long *__restrict__ keys,
                                 uint32_t *__restrict__ iota, uint32_t full) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    keys[i] = full ? ~(unsigned long long)__double_as_longlong(rho[i])
                   : ~(unsigned long long)rho[i];
    iota[i] = i;
  }
}

// level[reordered[i]] = i; PAVA key n-1-level (smaller key wins = larger level).
__global__ void level_scatter_kernel(uint32_t n, const uint32_t *__restrict__ reordered,
                                     uint32_t *__restrict__ level,
                                     unsigned long long *__restrict__ pkey) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = reordered[i];
    level[v] = i;
    pkey[v] = (unsigned long long)(n - 1 - i);
  }
}

// rho_tentative[level[v]] = cnt[v]; cnt = 0.
__global__ void tent_scatter_kernel(uint32_t n, const uint32_t *__restrict__ level,
                                    unsigned long long *__restrict__ cnt,
                                    unsigned long long *__restrict__ tent,
                                    unsigned long long *__restrict__ stats) {
  unsigned long long sum = 0, big = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long c = cnt[i];
    tent[level[i]] = c;
    cnt[i] = 0;
    sum += c;
    big |= (c >= kExactLimit) ? 1ull : 0ull;
  }
  block_stats(sum, big, stats);
}

// Edges with both ends among the first nb levels (CountEdges over the top
// of reordered, Graph.c:158-166).
__global__ void count_top_edges_kernel(const uint32_t *__restrict__ cd,
                                       const uint32_t *__restrict__ adj,
                                       const uint32_t *__restrict__ level, uint32_t n, uint32_t nb,
                                       unsigned long long *__restrict__ out) {
  unsigned long long local = 0;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    if (level[s] >= nb) continue;
    for (uint32_t j = cd[s]; j < cd[s + 1]; ++j) local += (level[adj[j]] < nb) ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(out, local);
}

__global__ void fill_u64_kernel(unsigned long long *p, uint64_t n, unsigned long long v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

kcl_status zero_u64(unsigned long long *p, uint64_t n, hipStream_t st) {
  if (!n) return KCL_OK;
  fill_u64_kernel<<<grid_for(n, 256), 256, 0, st>>>(p, n, 0ull);
  KCL_HIP(hipGetLastError());
  return KCL_OK;
}

kcl_status sim_prepare(kcl_sim *s) {
  kcl_graph *g = s->g;
  const uint32_t n = g->n;
  s->trivial = (n == 0) || (s->k > g->core + 1);
  if (!s->trivial && s->k > max_pass_k())
    return fail(KCL_E_UNSUPPORTED, "k = %u: the enumerator is compiled for k <= %u (graph core %u)",
                s->k, max_pass_k(), g->core);
  KCL_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  for (int i = 0; i < 3; ++i) KCL_HIP(hipEventCreate(&s->ev[i]));
  KCL_TRY(s->rho.alloc(n));
  KCL_TRY(s->cnt.alloc(n));
  KCL_TRY(s->last.alloc(n));
  KCL_TRY(s->V.alloc(n));
  KCL_TRY(s->stats.alloc(4));
  KCL_HIP(hipMemsetAsync(s->rho.p, 0, (size_t)(n ? n : 1) * sizeof(double), s->stream));
  KCL_TRY(zero_u64(s->cnt.p, n, s->stream));
  KCL_TRY(zero_u64(s->last.p, n, s->stream));
  KCL_TRY(zero_u64(s->V.p, n, s->stream));
  if (!s->trivial) {
    const uint32_t split = s->opts.split_work ? s->opts.split_work : (1u << 17);
    PassPlan p;
    KCL_TRY(make_plan(g, s->k, split, s->stream, &p, &s->root_cliques, &s->info,
                      s->opts.schedule, s->opts.model));
    KCL_TRY(upload_plan(p, &s->plan));
  }
  KCL_HIP(hipStreamSynchronize(s->stream));
  return KCL_OK;
}

// One FW pass into cnt, timed; no update.
static kcl_status sim_pass(kcl_sim *s, const uint64_t *d_key, PassMode mode) {
  KCL_HIP(hipEventRecord(s->ev[0], s->stream));
  if (!s->trivial)
    KCL_TRY(launch_pass(s->g, d_key, s->k, mode, s->plan, s->cnt.p,
                        s->stream));
  KCL_HIP(hipEventRecord(s->ev[1], s->stream));
  return KCL_OK;
}

kcl_status sim_step(kcl_sim *s, uint32_t t) {
  const uint32_t n = s->g->n;
  KCL_HIP(hipMemsetAsync(s->stats.p, 0, 4 * sizeof(unsigned long long), s->stream));
  KCL_TRY(sim_pass(s, (const uint64_t *)s->rho.p, kCreditWinner));
  if (n) {
    rho_update_kernel<<<grid_for(n, 256), 256, 0, s->stream>>>(n, t, s->rho.p, s->cnt.p,
                                                               s->last.p, s->V.p, s->stats.p);
    KCL_HIP(hipGetLastError());
  }
  KCL_HIP(hipEventRecord(s->ev[2], s->stream));
  unsigned long long st[2] = {0, 0};
  uint32_t status = 0;
  KCL_HIP(hipMemcpyAsync(st, s->stats.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         s->stream));
  if (!s->trivial)
    KCL_HIP(hipMemcpyAsync(&status, pass_status_word(s->plan), 4, hipMemcpyDeviceToHost,
                           s->stream));
  KCL_HIP(hipStreamSynchronize(s->stream));
  KCL_TRY(pass_status_error(status));
  float k_ms = 0, p_ms = 0;
  KCL_HIP(hipEventElapsedTime(&k_ms, s->ev[0], s->ev[1]));
  KCL_HIP(hipEventElapsedTime(&p_ms, s->ev[0], s->ev[2]));
  s->last_visits = st[0];
  s->last_kernel_ms = k_ms;
  s->last_pass_ms = p_ms;
  if (st[1])
    return fail(KCL_E_OVERFLOW,
                "iteration %u: a node received >= 2^53 cliques; the reference's double "
                "counters are no longer exact",
                t);
  return KCL_OK;
}

// Level order (floor(rho) or full rho, descending, ties by id) and the
// PAVA-preprocess pass: tent[level] = cliques whose last-level member sits
// at that level.  Host copies of tent, reordered and rho.
kcl_status sim_level_pass(kcl_sim *s, bool full_key, std::vector<unsigned long long> &tent,
                          std::vector<uint32_t> &reo, std::vector<double> &rho) {
  kcl_graph *g = s->g;
  const uint32_t n = g->n;
  hipStream_t st = s->stream;
  if (!s->level.p) {
    KCL_TRY(s->level.alloc(n));
    KCL_TRY(s->reordered.alloc(n));
    KCL_TRY(s->pava_key.alloc(n));
    KCL_TRY(s->tent.alloc(n));
    KCL_TRY(s->sort_keys_in.alloc(n));
    KCL_TRY(s->sort_keys_out.alloc(n));
    KCL_TRY(s->iota.alloc(n));
    KCL_HIP(rocprim::radix_sort_pairs(nullptr, s->sort_tmp_bytes, s->sort_keys_in.p,
                                      s->sort_keys_out.p, s->iota.p, s->reordered.p, (size_t)n, 0,
                                      64, st));
    KCL_TRY(s->sort_tmp.alloc(s->sort_tmp_bytes));
  }
  level_key_kernel<<<grid_for(n, 256), 256, 0, st>>>(n, s->rho.p, s->sort_keys_in.p, s->iota.p,
                                                     full_key ? 1u : 0u);
  KCL_HIP(hipGetLastError());
  size_t tb = s->sort_tmp_bytes;
  KCL_HIP(rocprim::radix_sort_pairs(s->sort_tmp.p, tb, s->sort_keys_in.p, s->sort_keys_out.p,
                                    s->iota.p, s->reordered.p, (size_t)n, 0, 64, st));
  level_scatter_kernel<<<grid_for(n, 256), 256, 0, st>>>(n, s->reordered.p, s->level.p,
                                                         s->pava_key.p);
  KCL_HIP(hipGetLastError());
  // PAVA-preprocess pass (kCDensestCmpSync.c:352) into tent by level.
  KCL_HIP(hipMemsetAsync(s->stats.p, 0, 4 * sizeof(unsigned long long), st));
  KCL_TRY(sim_pass(s, (const uint64_t *)s->pava_key.p, kCreditWinner));
  tent_scatter_kernel<<<grid_for(n, 256), 256, 0, st>>>(n, s->level.p, s->cnt.p, s->tent.p,
                                                        s->stats.p);
  KCL_HIP(hipGetLastError());
  tent.resize(n);
  reo.resize(n);
  rho.resize(n);
  unsigned long long stv[2] = {0, 0};
  uint32_t status = 0;
  if (!s->trivial)
    KCL_HIP(hipMemcpyAsync(&status, pass_status_word(s->plan), 4, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(tent.data(), s->tent.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(reo.data(), s->reordered.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(rho.data(), s->rho.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(stv, s->stats.p, 16, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  KCL_TRY(pass_status_error(status));
  if (stv[1]) return fail(KCL_E_OVERFLOW, "a level received >= 2^53 cliques");
  return KCL_OK;
}

kcl_status densest_host(const unsigned long long *tent, const uint32_t *reo, const double *rho,
                        uint32_t n, uint32_t k, kcl_densest_info *out, uint32_t *best_out) {
  // Densest prefix (kCDensestCmpSync.c:354-361): the running sum is a sum
  // of integers below 2^53, so the double is exact and matches.
  double sum = 0, density = -1;
  unsigned long long isum = 0, best_sum = 0;
  uint32_t best = 0;
  for (uint32_t i = 0; i < n; ++i) {
    sum += (double)tent[i];
    isum += tent[i];
    if (sum / (i + 1) > density) {
      density = sum / (i + 1);
      best = i + 1;
      best_sum = isum;
    }
  }
  if (isum >= kExactLimit) return fail(KCL_E_OVERFLOW, "%llu cliques: beyond 2^53", isum);
  // Upper bound (kCDensestCmpSync.c:365-381), sequential in level order.
  double s2 = 0, ub = 0, ip1ck = 1;
  for (uint32_t i = 0; i < n; ++i) {
    s2 += rho[reo[i]];
    if (i + 1 == k)
      ip1ck = 1;
    else if (i + 1 > k)
      ip1ck = (ip1ck * (i + 1)) / (double)(i + 1 - k);
    if (ip1ck < s2)
      ub = ip1ck / (i + 1);
    else {
      if (ub < s2 / (i + 1)) ub = s2 / (i + 1);
      break;
    }
  }
  out->n = best;
  out->density = density;
  out->ub = ub;
  out->sum = best_sum;
  *best_out = best;
  return KCL_OK;
}

kcl_status sim_find_densest(kcl_sim *s, kcl_densest_info *out, uint32_t *h_reordered,
                            uint32_t *h_level, uint64_t *h_tent) {
  kcl_graph *g = s->g;
  const uint32_t n = g->n;
  hipStream_t st = s->stream;
  out->n = 0;
  out->m = 0;
  out->density = -1;
  out->ub = 0;
  out->sum = 0;
  if (n == 0) return KCL_OK;  // the reference reads an unset n here
  std::vector<unsigned long long> tent;
  std::vector<uint32_t> reo;
  std::vector<double> rho;
  KCL_TRY(sim_level_pass(s, false, tent, reo, rho));

  uint32_t best = 0;
  KCL_TRY(densest_host(tent.data(), reo.data(), rho.data(), n, s->k, out, &best));
  // CountEdges over the first `best` nodes of reordered.
  {
    DevBuf<unsigned long long> d_m;
    KCL_TRY(d_m.alloc(1));
    KCL_HIP(hipMemsetAsync(d_m.p, 0, 8, st));
    count_top_edges_kernel<<<grid_for(n, 256), 256, 0, st>>>(g->d_cd.p, g->d_adj.p, s->level.p, n,
                                                             best, d_m.p);
    KCL_HIP(hipGetLastError());
    unsigned long long mm = 0;
    KCL_HIP(hipMemcpyAsync(&mm, d_m.p, 8, hipMemcpyDeviceToHost, st));
    KCL_HIP(hipStreamSynchronize(st));
    out->m = (uint32_t)mm;
  }
  if (h_reordered) memcpy(h_reordered, reo.data(), (size_t)n * 4);
  if (h_tent) memcpy(h_tent, tent.data(), (size_t)n * 8);
  if (h_level) {
    for (uint32_t i = 0; i < n; ++i) h_level[reo[i]] = i;
  }
  return KCL_OK;
}

}  // namespace kcl

using kcl::fail;

static kcl_status check_k(uint32_t k) {
  if (k < 2 || k > 255) return fail(KCL_E_INVALID_ARG, "k = %u: the reference takes 2 <= k <= 255", k);
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_create(kcl_graph *g, uint32_t k, const kcl_sim_opts *opts,
                                     kcl_sim **out) {
  if (!g || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  KCL_TRY(check_k(k));
  kcl_sim *s = new (std::nothrow) kcl_sim();
  if (!s) return fail(KCL_E_OOM_HOST, "out of host memory");
  s->g = g;
  s->k = k;
  if (opts) s->opts = *opts;
  kcl_status st = kcl::sim_prepare(s);
  if (st != KCL_OK) {
    kcl_sim_destroy(s);
    return st;
  }
  *out = s;
  return KCL_OK;
}

extern "C" void kcl_sim_destroy(kcl_sim *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int i = 0; i < 3; ++i)
    if (s->ev[i]) (void)hipEventDestroy(s->ev[i]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

extern "C" kcl_status kcl_sim_step(kcl_sim *s, uint32_t t) {
  if (!s) return fail(KCL_E_INVALID_ARG, "NULL sim");
  if (t == 0) return fail(KCL_E_INVALID_ARG, "iterations count from t = 1");
  return kcl::sim_step(s, t);
}

extern "C" kcl_status kcl_sim_counts(kcl_sim *s, uint64_t *out) {
  if (!s || (!out && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!s->g->n) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(out, s->last.p, (size_t)s->g->n * 8, hipMemcpyDeviceToHost, s->stream));
  KCL_HIP(hipStreamSynchronize(s->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_rho(kcl_sim *s, double *out) {
  if (!s || (!out && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!s->g->n) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(out, s->rho.p, (size_t)s->g->n * 8, hipMemcpyDeviceToHost, s->stream));
  KCL_HIP(hipStreamSynchronize(s->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_set_rho(kcl_sim *s, const double *rho) {
  if (!s || (!rho && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  for (uint32_t i = 0; i < s->g->n; ++i)
    if (!(rho[i] >= 0.0) || rho[i] != rho[i] || (rho[i] == 0.0 && std::signbit(rho[i])))
      return fail(KCL_E_INVALID_ARG, "rho[%u] must be a non-negative number (+0.0 for zero)", i);
  if (!s->g->n) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(s->rho.p, rho, (size_t)s->g->n * 8, hipMemcpyHostToDevice, s->stream));
  KCL_HIP(hipStreamSynchronize(s->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_totals(kcl_sim *s, uint64_t *V) {
  if (!s || (!V && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!s->g->n) return KCL_OK;
  KCL_HIP(hipMemcpyAsync(V, s->V.p, (size_t)s->g->n * 8, hipMemcpyDeviceToHost, s->stream));
  KCL_HIP(hipStreamSynchronize(s->stream));
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_last_pass(kcl_sim *s, uint64_t *visits, double *kernel_ms,
                                        double *pass_ms) {
  if (!s) return fail(KCL_E_INVALID_ARG, "NULL sim");
  if (visits) *visits = s->last_visits;
  if (kernel_ms) *kernel_ms = s->last_kernel_ms;
  if (pass_ms) *pass_ms = s->last_pass_ms;
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_plan_classes(kcl_sim *s, kcl_class_info *out) {
  if (!s || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  static_assert(KCL_NUM_CLASSES == kcl::kNumClasses, "class count of the ABI");
  for (uint32_t grp = 0; grp < 2; ++grp) {
    const uint32_t *begin = grp ? s->plan.eclass_begin : s->plan.class_begin;
    const uint32_t *count = grp ? s->plan.eclass_count : s->plan.class_count;
    const uint32_t kk = grp ? s->k - 1 : s->k;
    for (uint32_t c = 0; c < kcl::kNumClasses; ++c) {
      kcl_class_info &o = out[grp * kcl::kNumClasses + c];
      o = kcl_class_info{};
      o.words = kcl::kClassWords[c];
      o.k = kk;
      o.first = begin[c];
      o.count = s->trivial ? 0u : count[c];
      KCL_TRY(kcl::class_launch_shape(c, kk, false, s->plan.max_per_cu, o.count, &o.waves,
                                      &o.per_cu, &o.blocks, &o.lds_bytes));
    }
  }
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_profile_pass(kcl_sim *s, uint32_t *ntasks, uint32_t *root,
                                           uint32_t *part, uint32_t *nparts, uint64_t *build_cycles,
                                           uint64_t *total_cycles) {
  if (!s || !ntasks) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const uint32_t nt = (uint32_t)s->plan.host_tasks.size();
  const uint32_t cap = *ntasks;
  *ntasks = nt;
  if (!root && !part && !nparts && !build_cycles && !total_cycles) return KCL_OK;  // size query
  if (cap < nt) return fail(KCL_E_INVALID_ARG, "room for %u tasks, the plan has %u", cap, nt);
  if (s->trivial || nt == 0) return KCL_OK;
  kcl::DevBuf<unsigned long long> prof;
  KCL_TRY(prof.alloc(2 * (size_t)nt));
  KCL_HIP(hipMemsetAsync(prof.p, 0, 2 * (size_t)nt * 8, s->stream));
  KCL_TRY(kcl::launch_pass(s->g, (const uint64_t *)s->rho.p, s->k,
                           kcl::kCreditWinner, s->plan, s->cnt.p, s->stream, prof.p));
  std::vector<unsigned long long> h(2 * (size_t)nt);
  KCL_HIP(hipMemcpyAsync(h.data(), prof.p, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
  KCL_TRY(kcl::zero_u64(s->cnt.p, s->g->n, s->stream));  // the pass leaves no credits behind
  KCL_TRY(kcl::pass_status(s->plan, s->stream));
  for (uint32_t i = 0; i < nt; ++i) {
    const kcl::Task &t = s->plan.host_tasks[i];
    if (root) root[i] = t.s;
    const bool et = kcl::task_is_edge(t.part_np);  // edge task: part 0 of 0
    if (part) part[i] = et ? 0u : (t.part_np & 0xffffu);
    if (nparts) nparts[i] = et ? 0u : (t.part_np >> 16);
    if (build_cycles) build_cycles[i] = h[2 * (size_t)i];
    if (total_cycles) total_cycles[i] = h[2 * (size_t)i + 1];
  }
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_root_cliques(kcl_sim *s, uint64_t *out) {
  if (!s || (!out && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const uint32_t n = s->g->n;
  for (uint32_t v = 0; v < n; ++v) out[v] = v < s->root_cliques.size() ? s->root_cliques[v] : 0;
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_root_features(kcl_sim *s, float *x) {
  if (!s || (!x && s->g->n)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  std::vector<float> f;
  KCL_TRY(kcl::root_features(s->g, s->k, s->stream, &f));
  if (!f.empty()) std::memcpy(x, f.data(), f.size() * sizeof(float));
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_cost_samples(kcl_sim *s, uint32_t *nrows, uint32_t *root, float *x,
                                           float *y) {
  if (!s || !nrows) return fail(KCL_E_INVALID_ARG, "NULL argument");
  const std::vector<kcl::Task> &tasks = s->plan.host_tasks;
  std::vector<uint8_t> seen(s->g->n, 0);
  uint32_t roots = 0;
  for (const kcl::Task &t : tasks)
    if (!seen[t.s]) {
      seen[t.s] = 1;
      ++roots;
    }
  const uint32_t cap = *nrows;
  *nrows = roots;
  if (!root && !x && !y) return KCL_OK;  // size query
  if (cap < roots) return fail(KCL_E_INVALID_ARG, "room for %u rows, the plan has %u roots", cap,
                               roots);
  if (roots == 0) return KCL_OK;
  uint32_t nt = (uint32_t)tasks.size();
  std::vector<uint32_t> troot(nt), tparts(nt);
  std::vector<uint64_t> tb(nt), tt(nt);
  KCL_TRY(kcl_sim_profile_pass(s, &nt, troot.data(), nullptr, tparts.data(), tb.data(),
                               tt.data()));
  std::vector<double> bsum(s->g->n, 0.0), csum(s->g->n, 0.0);
  std::vector<uint32_t> parts(s->g->n, 0);
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t v = troot[i];
    bsum[v] += (double)tb[i];
    csum[v] += (double)(tt[i] > tb[i] ? tt[i] - tb[i] : 0);
    parts[v] += 1;
  }
  std::vector<float> feat;
  if (x) KCL_TRY(kcl::root_features(s->g, s->k, s->stream, &feat));
  uint32_t r = 0;
  for (uint32_t v = 0; v < s->g->n; ++v) {  // rows by ascending root id
    if (!parts[v]) continue;
    if (root) root[r] = v;
    if (x)
      std::memcpy(x + (size_t)r * KCL_CM_FEATURES, feat.data() + (size_t)v * KCL_CM_FEATURES,
                  KCL_CM_FEATURES * sizeof(float));
    if (y) {
      y[2 * (size_t)r] = (float)std::log2(1.0 + bsum[v] / parts[v]);
      y[2 * (size_t)r + 1] = (float)std::log2(1.0 + csum[v]);
    }
    ++r;
  }
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_replan(kcl_sim *s, uint32_t schedule, const kcl_cost_model *model) {
  if (!s) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (schedule > KCL_SCHED_DEGREE) return fail(KCL_E_INVALID_ARG, "unknown schedule %u", schedule);
  s->opts.schedule = schedule;
  s->opts.model = model;
  if (s->trivial) return KCL_OK;
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t split = s->opts.split_work ? s->opts.split_work : (1u << 17);
  kcl::PassPlan p;
  KCL_TRY(kcl::plan_for_schedule(s->g, s->k, schedule, model, split, &s->root_cliques, s->stream,
                                 &p));
  KCL_TRY(kcl::upload_plan(p, &s->plan));
  KCL_HIP(hipStreamSynchronize(s->stream));
  s->info.tasks = (uint32_t)(p.tasks.size() + p.edge_tasks.size());
  s->info.plan_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_plan_info(kcl_sim *s, kcl_plan_info *info) {
  if (!s || !info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *info = s->info;
  return KCL_OK;
}

extern "C" kcl_status kcl_sim_find_densest(kcl_sim *s, kcl_densest_info *out,
                                           uint32_t *reordered, uint32_t *level, uint64_t *tent) {
  if (!s || !out) return fail(KCL_E_INVALID_ARG, "NULL argument");
  return kcl::sim_find_densest(s, out, reordered, level, tent);
}

extern "C" kcl_status kcl_cmpsync_run(kcl_graph *g, uint32_t k, uint32_t num_iter,
                                      const kcl_sim_opts *opts, const char *stat_path,
                                      int64_t t0_unix, kcl_record *records, uint32_t *nrec) {
  if (!g) return fail(KCL_E_INVALID_ARG, "NULL graph");
  KCL_TRY(check_k(k));
  if (nrec) *nrec = 0;
  if (t0_unix <= 0) t0_unix = (int64_t)time(nullptr);
  FILE *f = nullptr;
  if (stat_path) {
    f = fopen(stat_path, "w");
    if (!f) return fail(KCL_E_IO, "cannot create %s", stat_path);
    fprintf(f,
            "[Number of Iterations]\t[Number of Nodes]\t[Number of Edges]\t[k-Clique "
            "Density]\t[Upper Bound of Maximum Density]\t[Time (seconds)]\n");
  }
  kcl_sim *s = nullptr;
  kcl_status st = kcl_sim_create(g, k, opts, &s);
  uint32_t r = 0;
  uint32_t t = 1;
  for (uint64_t T = 1; st == KCL_OK && T <= num_iter; T <<= 1) {
    for (; st == KCL_OK && t <= T; ++t) st = kcl::sim_step(s, t);
    if (st != KCL_OK) break;
    kcl_densest_info info;
    st = kcl::sim_find_densest(s, &info, nullptr, nullptr, nullptr);
    if (st != KCL_OK) break;
    const int64_t now = (int64_t)time(nullptr);
    if (f) {
      fprintf(f, "%u\t%u\t%u\t%.12f\t%.12f\t%ld\n", (uint32_t)T, info.n, info.m, info.density,
              info.ub, (long)(now - t0_unix));
      fflush(f);
    }
    if (records && r < 32) {
      records[r] = kcl_record{(uint32_t)T, info.n, info.m, info.density, info.ub, info.sum,
                              (double)(now - t0_unix)};
    }
    ++r;
  }
  if (nrec) *nrec = r < 32 ? r : 32;
  if (f && fclose(f) != 0 && st == KCL_OK) st = fail(KCL_E_IO, "cannot write %s", stat_path);
  kcl_sim_destroy(s);
  return st;
}

extern "C" kcl_status kcl_clique_enum(kcl_graph *g, uint32_t k, kcl_task task, const double *rho,
                                      const uint32_t *level, double *acc, uint64_t *visits) {
  if (!g || !acc) return fail(KCL_E_INVALID_ARG, "NULL argument");
  KCL_TRY(check_k(k));
  if (task != KCL_FRANK_WOLFE && task != KCL_PAVA_PREPROCESS)
    return fail(KCL_E_INVALID_ARG, "task %d: expected 2 (FRANK_WOLFE) or 3 (PAVA_PREPROCESS)",
                (int)task);
  if (task == KCL_FRANK_WOLFE && !rho) return fail(KCL_E_INVALID_ARG, "rho is NULL");
  if (task == KCL_PAVA_PREPROCESS && !level) return fail(KCL_E_INVALID_ARG, "level is NULL");
  const uint32_t n = g->n;
  if (visits) *visits = 0;
  if (n == 0) return KCL_OK;
  std::vector<unsigned long long> key(n);
  if (task == KCL_FRANK_WOLFE) {
    for (uint32_t i = 0; i < n; ++i) {
      if (!(rho[i] >= 0.0))
        return fail(KCL_E_INVALID_ARG, "rho[%u] is negative or NaN", i);
      double r = rho[i] == 0.0 ? 0.0 : rho[i];  // -0.0 orders as +0.0
      memcpy(&key[i], &r, 8);
    }
  } else {
    for (uint32_t i = 0; i < n; ++i) {
      if (level[i] >= n) return fail(KCL_E_INVALID_ARG, "level[%u] = %u >= n", i, level[i]);
      key[i] = (unsigned long long)(n - 1 - level[i]);
    }
  }
  kcl_sim *s = nullptr;
  KCL_TRY(kcl_sim_create(g, k, nullptr, &s));
  kcl_status st = KCL_OK;
  kcl::DevBuf<unsigned long long> d_key;
  std::vector<unsigned long long> cnt(n);
  do {
    if ((st = d_key.alloc(n)) != KCL_OK) break;
    hipError_t e = hipMemcpyAsync(d_key.p, key.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                                  s->stream);
    if (e != hipSuccess) { st = fail(KCL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e)); break; }
    if ((st = kcl::sim_pass(s, (const uint64_t *)d_key.p, kcl::kCreditWinner)) != KCL_OK) break;
    e = hipMemcpyAsync(cnt.data(), s->cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) { st = fail(KCL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e)); break; }
  } while (0);
  kcl_sim_destroy(s);
  if (st != KCL_OK) return st;
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    total += cnt[i];
    if (cnt[i] >= (1ull << 53)) return fail(KCL_E_OVERFLOW, "node %u: >= 2^53 cliques", i);
    if (task == KCL_FRANK_WOLFE) acc[i] += (double)cnt[i];
    else acc[level[i]] += (double)cnt[i];
  }
  if (visits) *visits = total;
  return KCL_OK;
}

extern "C" kcl_status kcl_device_get_info(kcl_device_info *info) {
  if (!info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  int dev = 0;
  KCL_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  KCL_HIP(hipGetDeviceProperties(&p, dev));
  memset(info, 0, sizeof(*info));
  snprintf(info->name, sizeof(info->name), "%s", p.name);
  snprintf(info->arch, sizeof(info->arch), "%s", p.gcnArchName);
  info->cu_count = (uint32_t)p.multiProcessorCount;
  info->hbm_bytes = (uint64_t)p.totalGlobalMem;
  info->lds_bytes_per_cu = (uint32_t)p.maxSharedMemoryPerMultiProcessor;
  info->clock_khz = (uint32_t)p.clockRate;
  return KCL_OK;
}
