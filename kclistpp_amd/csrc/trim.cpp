// trim.cpp - k-clique core trimming (kClistCoreTrim.c).
//
// Pipeline (kClistCoreTrim.c:707-775):
//   1. (k-1)-core by the degree heap, dropping nodes popped while the
//      running core is below k-1; kept nodes get rank n2 - r in pop order
//      (kcore, :242-272)                                   host, sequential
//   2. kept edges relabelled and oriented s > t, input order (:274-294)
//   3. k-clique degree of every kept node (onepass, :394-416) - on the GPU:
//      the enumeration pass with every member credited
//   4. k-clique core peel: min-heap on the k-clique degrees, each pop
//      removes the cliques through the popped node from its alive
//      neighbours (kcliquecore + oneshortpass, :577-666).  The heap is the
//      host's (its tie order defines the results); a pop whose node still
//      lies in at least gpu_peel_min_cliques cliques counts them on the GPU
//      instead: one root (the popped node, members its alive neighbours,
//      split into parts that fill the device) with every member credited,
//      then a gather of the neighbours' counts
//   5. outputs (printckdeg, printdensest, :419-429, :668-705)
// Steps 1 and 4 define the results through their heaps' tie order, so they
// replay the reference's heap discipline on the host.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <vector>

#include "kcl_internal.h"

struct kcl_trim {
  kcl_trim_info info{};
  kcl_trim_opts opts{};
  uint32_t k = 0;
  std::vector<uint32_t> map;     // new id -> input id (n2)
  std::vector<uint64_t> nck;     // k-clique degree per new id (n2)
  std::vector<uint32_t> ic_key;  // peel order (n2)
  std::vector<uint64_t> ic_val;  // running k-clique core at each pop (n2)
  std::vector<uint32_t> es, et;  // kept edges, new ids, input order
  std::vector<uint8_t> keep;     // trimmed node set, by new id
};

namespace kcl {

namespace {

// Pops whose node lies in at least this many alive k-cliques count them on
// the GPU (below it the host's enumeration is faster than a launch).
constexpr uint64_t kDefaultGpuPeelCliques = 1ull << 16;

// Binary min-heap with decrease-by-delta, the reference's tie discipline
// (kClistCoreTrim.c:64-133 and :433-533): inserted in key order, sift-up
// while the parent is strictly larger, sift-down to the right child only if
// strictly smaller than the left one and while strictly smaller.
template <class V>
struct Heap {
  struct Slot {
    uint32_t key;
    V val;
  };
  std::vector<Slot> a;
  std::vector<uint32_t> where;
  uint32_t size = 0;

  explicit Heap(uint32_t n) : a(n ? n : 1), where(n ? n : 1, UINT32_MAX) {}

  void up(uint32_t i) {
    Slot x = a[i];
    while (i > 0) {
      uint32_t p = (i - 1) / 2;
      if (!(a[p].val > x.val)) break;
      a[i] = a[p];
      where[a[i].key] = i;
      i = p;
    }
    a[i] = x;
    where[x.key] = i;
  }
  void down() {
    uint32_t i = 0;
    Slot x = a[0];
    for (;;) {
      uint32_t l = 2 * i + 1;
      if (l >= size) break;
      uint32_t r = l + 1;
      uint32_t c = (r < size && a[r].val < a[l].val) ? r : l;
      if (!(a[c].val < x.val)) break;
      a[i] = a[c];
      where[a[i].key] = i;
      i = c;
    }
    a[i] = x;
    where[x.key] = i;
  }
  void push(uint32_t key, V v) {
    a[size] = Slot{key, v};
    where[key] = size;
    ++size;
    up(size - 1);
  }
  Slot pop() {
    Slot m = a[0];
    where[m.key] = UINT32_MAX;
    --size;
    a[0] = a[size];
    where[a[0].key] = 0;  // also when the heap just emptied, as the reference
    if (size > 0) down();
    return m;
  }
  void decrease(uint32_t key, V by) {
    uint32_t i = where[key];
    if (i == UINT32_MAX) return;
    a[i].val -= by;
    up(i);
  }
  bool alive(uint32_t key) const { return where[key] != UINT32_MAX; }
};

uint32_t merge_lists(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb,
                     uint32_t *out) {
  uint32_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else {
      out[c++] = a[i];
      ++i;
      ++j;
    }
  }
  return c;
}

// Cliques through ck[0..depth) completed from cand, narrowed by DAG rows;
// every member is credited (kClistCoreTrim.c:368-391).
void peel_descend(const std::vector<uint32_t> &cd, const std::vector<uint32_t> &adj,
                  uint32_t kmax, uint32_t depth, uint32_t *ck, const uint32_t *cand, uint32_t nc,
                  std::vector<std::vector<uint32_t>> &buf, uint64_t *nck) {
  if (nc < kmax - depth) return;
  if (depth + 1 == kmax) {
    for (uint32_t i = 0; i < depth; ++i) nck[ck[i]] += nc;
    for (uint32_t i = 0; i < nc; ++i) nck[cand[i]] += 1;
    return;
  }
  uint32_t *nxt = buf[depth].data();
  for (uint32_t i = 0; i < nc; ++i) {
    const uint32_t x = cand[i];
    ck[depth] = x;
    const uint32_t nn = merge_lists(adj.data() + cd[x], cd[x + 1] - cd[x], cand, nc, nxt);
    peel_descend(cd, adj, kmax, depth + 1, ck, nxt, nn, buf, nck);
  }
}

}  // namespace

// out[i] = cnt[list[i]], out[n] = cnt[root]; those counters back to zero.
__global__ void gather_clear_kernel(const uint32_t *__restrict__ list, uint32_t n, uint32_t root,
                                    unsigned long long *__restrict__ cnt,
                                    unsigned long long *__restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
    const uint32_t v = i < n ? list[i] : root;
    out[i] = cnt[v];
    cnt[v] = 0;
  }
}

// One pop's oneshortpass on the GPU: the k-cliques through u among its
// alive neighbours M (ascending, |M| <= kMaxRootDegree), every member
// credited; nck2[M[i]] += cliques through u and M[i].
struct GpuPeel {
  kcl_graph *g = nullptr;
  uint32_t k = 0;
  const unsigned long long *key = nullptr;  // zeros (credit-all ignores keys)
  unsigned long long *cnt = nullptr;        // n2 counters, zero between pops
  DevBuf<uint32_t> list, tickets;
  DevBuf<Task> tasks;
  DevBuf<unsigned long long> out;
  std::vector<Task> host_tasks;
  std::vector<unsigned long long> host_out;
  uint32_t pops = 0;

  kcl_status init() {
    KCL_TRY(list.alloc(kMaxRootDegree));
    KCL_TRY(tickets.alloc(kPassTicketWords));
    KCL_TRY(tasks.alloc(kMaxSplit));
    KCL_TRY(out.alloc(kMaxRootDegree + 1));
    host_out.resize(kMaxRootDegree + 1);
    return KCL_OK;
  }

  static constexpr uint32_t kMaxSplit = 4096;

  kcl_status pop(uint32_t u, const uint32_t *m, uint32_t d, uint64_t cliques,
                 uint64_t *nck2) {
    hipStream_t st = g->stream;
    uint32_t c = 0;
    while (kClassWords[c] * 64 < d) ++c;
    // Parts of about 2^20 cliques, within the items a part can take.
    const uint64_t items = k >= 5 ? (uint64_t)d * (d - 1) / 2 : d;
    uint64_t np = cliques >> 20;
    np = std::max<uint64_t>(1, std::min<uint64_t>({np, items, (uint64_t)kMaxSplit}));
    host_tasks.resize(np);
    for (uint32_t p = 0; p < np; ++p) host_tasks[p] = Task{u, p | ((uint32_t)np << 16)};
    KCL_HIP(hipMemcpyAsync(tasks.p, host_tasks.data(), np * sizeof(Task), hipMemcpyHostToDevice,
                           st));
    KCL_HIP(hipMemcpyAsync(list.p, m, (size_t)d * 4, hipMemcpyHostToDevice, st));
    uint32_t begin[kNumClasses + 1] = {}, count[kNumClasses] = {}, light[kNumClasses] = {};
    const uint32_t zero[kNumClasses + 1] = {};
    for (uint32_t x = c + 1; x <= kNumClasses; ++x) begin[x] = (uint32_t)np;
    count[c] = (uint32_t)np;
    light[c] = (uint32_t)np;
    PassLaunch pl;
    pl.tasks = tasks.p;
    pl.tickets = tickets.p;
    pl.class_begin = begin;
    pl.class_count = count;
    pl.eclass_begin = zero;
    pl.eclass_count = zero;
    pl.light_from = light;
    pl.list = list.p;
    pl.nlist = d;
    KCL_TRY(launch_pass_raw(g, (const uint64_t *)key, k, kCreditAll, pl, cnt, st));
    gather_clear_kernel<<<(d + 256) / 256, 256, 0, st>>>(list.p, d, u, cnt, out.p);
    KCL_HIP(hipGetLastError());
    uint32_t status = 0;
    KCL_HIP(hipMemcpyAsync(host_out.data(), out.p, ((size_t)d + 1) * 8, hipMemcpyDeviceToHost,
                           st));
    KCL_HIP(hipMemcpyAsync(&status, tickets.p + kPassTicketWords - 1, 4, hipMemcpyDeviceToHost,
                           st));
    KCL_HIP(hipStreamSynchronize(st));
    KCL_TRY(pass_status_error(status));
    for (uint32_t i = 0; i < d; ++i) nck2[m[i]] += host_out[i];
    ++pops;
    return KCL_OK;
  }
};

kcl_status trim_run(kcl_trim *T, const uint32_t *pairs, uint64_t m, uint32_t n_in) {
  const uint32_t kmax = T->k;
  kcl_trim_info &info = T->info;
  info.n_in = n_in;
  const auto t0 = std::chrono::steady_clock::now();

  // 1. (k-1)-core and ranks.
  std::vector<uint64_t> off((size_t)n_in + 1, 0);
  for (uint64_t i = 0; i < m; ++i) {
    off[pairs[2 * i] + 1]++;
    off[pairs[2 * i + 1] + 1]++;
  }
  for (uint32_t v = 0; v < n_in; ++v) off[v + 1] += off[v];
  std::vector<uint32_t> rank(n_in, UINT32_MAX);
  {
    std::vector<uint32_t> nbr(2 * m + 1);
    std::vector<uint64_t> fill(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < m; ++i) {
      const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
      nbr[fill[a]++] = b;
      nbr[fill[b]++] = a;
    }
    Heap<uint32_t> h(n_in);
    for (uint32_t v = 0; v < n_in; ++v) h.push(v, (uint32_t)(off[v + 1] - off[v]));
    T->map.assign(n_in, 0);
    uint32_t alive = n_in, r = 0, c = 0;
    for (uint32_t i = 0; i < n_in; ++i) {
      const auto kv = h.pop();
      if (kv.val > c) c = kv.val;
      if (c < kmax - 1) {
        --alive;
      } else {
        ++r;
        T->map[alive - r] = kv.key;
        rank[kv.key] = alive - r;
      }
      for (uint64_t j = off[kv.key]; j < off[kv.key + 1]; ++j) h.decrease(nbr[j], 1u);
    }
    info.core = c;
    info.n2 = alive;
  }
  const uint32_t n2 = info.n2;
  T->map.resize(n2);

  // 2. Kept edges, oriented, in input order.
  T->es.clear();
  T->et.clear();
  for (uint64_t i = 0; i < m; ++i) {
    const uint32_t a = rank[pairs[2 * i]], b = rank[pairs[2 * i + 1]];
    if (a == UINT32_MAX || b == UINT32_MAX) continue;
    T->es.push_back(a > b ? a : b);
    T->et.push_back(a > b ? b : a);
  }
  const uint32_t e2 = (uint32_t)T->es.size();
  info.e2 = e2;
  T->nck.assign(n2, 0);
  T->ic_key.assign(n2, 0);
  T->ic_val.assign(n2, 0);
  T->keep.assign(n2, 0);
  if (n2 == 0) return KCL_OK;  // the reference does not handle this case

  // 3. k-clique degrees on the GPU, over the DAG of the trim's own ranks.
  kcl_graph *g = nullptr;
  {
    std::vector<uint32_t> rel(2 * (size_t)e2);
    for (uint32_t i = 0; i < e2; ++i) {
      rel[2 * i] = T->es[i];
      rel[2 * i + 1] = T->et[i];
    }
    std::vector<uint32_t> ident(n2);
    for (uint32_t v = 0; v < n2; ++v) ident[v] = v;
    KCL_TRY(graph_create(rel.data(), e2, n2, ident.data(), &g));
  }
  kcl_status st = KCL_OK;
  if (g->n > 0 && kmax <= g->core + 1 && kmax > max_pass_k()) {
    kcl_graph_destroy(g);
    return fail(KCL_E_UNSUPPORTED, "k = %u: the enumerator is compiled for k <= %u", kmax,
                max_pass_k());
  }
  DevBuf<unsigned long long> cnt, key;  // also the peel's counters (zeroed)
  if (g->n > 0 && kmax <= g->core + 1) {
    DevicePlan dp;
    PassPlan p;
    do {
      if ((st = make_plan(g, kmax, 1u << 17, g->stream, &p, nullptr, nullptr)) != KCL_OK) break;
      if ((st = upload_plan(p, &dp)) != KCL_OK) break;
      if ((st = cnt.alloc(g->n)) != KCL_OK) break;
      if ((st = key.alloc(g->n)) != KCL_OK) break;
      hipError_t e = hipMemsetAsync(cnt.p, 0, (size_t)g->n * 8, g->stream);
      if (e == hipSuccess) e = hipMemsetAsync(key.p, 0, (size_t)g->n * 8, g->stream);
      if (e != hipSuccess) {
        st = fail(KCL_E_HIP, "hipMemset: %s", hipGetErrorString(e));
        break;
      }
      st = launch_pass(g, (const uint64_t *)key.p, kmax, kCreditAll, dp, cnt.p,
                       g->stream);
      if (st != KCL_OK) break;
      e = hipMemcpyAsync(T->nck.data(), cnt.p, (size_t)g->n * 8, hipMemcpyDeviceToHost, g->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
      if (e != hipSuccess) st = fail(KCL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e));
      else st = pass_status(dp, g->stream);
      if (st == KCL_OK) st = zero_u64(cnt.p, g->n, g->stream);
    } while (0);
  }  // else no k-clique can exist: the degrees stay zero
  std::vector<uint32_t> cd = g->cd, adj = g->adj;  // DAG rows, ascending
  cd.resize((size_t)n2 + 1, cd.empty() ? 0u : cd.back());
  if (st != KCL_OK) {
    kcl_graph_destroy(g);
    return st;
  }
  uint64_t tot = 0;
  for (uint32_t v = 0; v < n2; ++v) tot += T->nck[v];
  tot /= kmax;
  info.ncktot = tot;
  const auto t1 = std::chrono::steady_clock::now();

  // 4. k-clique core peel (kcliquecore, kClistCoreTrim.c:605-666).
  std::vector<uint32_t> cd0((size_t)n2 + 1, 0), adj0(2 * (size_t)e2 + 1);
  for (uint32_t i = 0; i < e2; ++i) {
    cd0[T->es[i] + 1]++;
    cd0[T->et[i] + 1]++;
  }
  for (uint32_t v = 0; v < n2; ++v) cd0[v + 1] += cd0[v];
  {
    std::vector<uint32_t> fill(cd0.begin(), cd0.end() - 1);
    for (uint32_t i = 0; i < e2; ++i) {
      adj0[fill[T->es[i]]++] = T->et[i];
      adj0[fill[T->et[i]]++] = T->es[i];
    }
    for (uint32_t v = 0; v < n2; ++v) std::sort(adj0.begin() + cd0[v], adj0.begin() + cd0[v + 1]);
  }
  Heap<uint64_t> h(n2);
  for (uint32_t v = 0; v < n2; ++v) h.push(v, T->nck[v]);
  uint32_t e = e2;
  uint64_t c = 0, ncktot = tot;
  std::vector<uint64_t> when((size_t)n2 + 1, 0), nck2((size_t)n2 + 1, 0);
  uint32_t maxdeg = 0;
  for (uint32_t v = 0; v < n2; ++v) maxdeg = std::max(maxdeg, cd0[v + 1] - cd0[v]);
  std::vector<uint32_t> alive_nbr(maxdeg + 1), first(maxdeg + 1);
  std::vector<std::vector<uint32_t>> buf(kmax + 1, std::vector<uint32_t>(maxdeg + 1));
  std::vector<uint32_t> ck(kmax + 1);
  // Heavy pops on the GPU (the degrees' graph, counters and zero keys).
  GpuPeel gpu;
  const bool gpu_ok = cnt.p && kmax <= max_pass_k();
  if (gpu_ok) {
    gpu.g = g;
    gpu.k = kmax;
    gpu.key = key.p;
    gpu.cnt = cnt.p;
    if ((st = gpu.init()) != KCL_OK) {
      kcl_graph_destroy(g);
      return st;
    }
  }
  const uint64_t gpu_min = T->opts.gpu_peel_min_cliques ? T->opts.gpu_peel_min_cliques
                                                        : kDefaultGpuPeelCliques;
  double ckrho = (double)ncktot / (double)n2;
  uint32_t size = n2;
  // 32-bit products as in the reference (they wrap for n2 > 65536).
  double erho = (double)(uint32_t)(2u * e) / (double)(uint32_t)(n2 * (n2 - 1u));
  uint32_t i = 0;
  for (; i + 1 < n2; ++i) {
    const auto kv = h.pop();
    const uint32_t u = kv.key;
    if (kv.val > c) c = kv.val;
    T->ic_key[i] = u;
    T->ic_val[i] = c;
    // oneshortpass (kClistCoreTrim.c:577-603).
    uint32_t s = 0;
    for (uint32_t j = cd0[u]; j < cd0[u + 1]; ++j)
      if (h.alive(adj0[j])) alive_nbr[s++] = adj0[j];
    if (s > kmax - 2 && gpu_ok && kv.val >= gpu_min && s <= kMaxRootDegree) {
      if ((st = gpu.pop(u, alive_nbr.data(), s, kv.val, nck2.data())) != KCL_OK) {
        kcl_graph_destroy(g);
        return st;
      }
    } else if (s > kmax - 2) {
      ck[0] = u;
      for (uint32_t j = 0; j < s; ++j) {
        const uint32_t v = alive_nbr[j];
        ck[1] = v;
        const uint32_t nc = merge_lists(alive_nbr.data(), s, adj.data() + cd[v], cd[v + 1] - cd[v],
                                        first.data());
        peel_descend(cd, adj, kmax, 2, ck.data(), first.data(), nc, buf, nck2.data());
      }
    }
    uint64_t tot2 = 0;
    for (uint32_t j = cd0[u]; j < cd0[u + 1]; ++j) {
      const uint32_t v = adj0[j];
      if (h.alive(v)) {
        --e;
        h.decrease(v, nck2[v]);
        tot2 += nck2[v];
        nck2[v] = 0;
      }
    }
    const uint32_t sz = n2 - i - 1;
    when[sz] = tot2 / (uint64_t)(kmax - 1);
    ncktot -= when[sz];
    const double tmp = (double)ncktot / (double)sz;
    if (tmp > ckrho) {
      ckrho = tmp;
      size = sz;
      erho = (double)(uint32_t)(2u * e) / (double)(uint32_t)(sz * (sz - 1u));
    }
  }
  uint64_t ncktrim = tot;
  uint32_t ntrim = 0;  // the reference leaves it unset when no j qualifies
  for (uint32_t j = n2 - 1; j >= 1; --j) {
    if ((double)when[j] >= ckrho) {
      ntrim = j + 1;
      break;
    }
    ncktrim -= when[j];
  }
  {
    const auto kv = h.pop();
    T->ic_key[i] = kv.key;
    T->ic_val[i] = c;
  }
  kcl_graph_destroy(g);
  info.gpu_pops = gpu.pops;
  info.ckcore = c;
  info.ckrho = ckrho;
  info.rho = erho;
  info.size = size;
  info.ntrim = ntrim;
  info.ncktrim = ncktrim;
  for (uint32_t j = 0; j < ntrim; ++j) T->keep[T->ic_key[n2 - 1 - j]] = 1;
  uint32_t mt = 0;
  for (uint32_t j = 0; j < e2; ++j) mt += (T->keep[T->es[j]] && T->keep[T->et[j]]) ? 1u : 0u;
  info.mtrim = mt;
  const auto t2 = std::chrono::steady_clock::now();
  info.seconds_degrees = std::chrono::duration<double>(t1 - t0).count();
  info.seconds_peel = std::chrono::duration<double>(t2 - t1).count();
  return KCL_OK;
}

}  // namespace kcl

using kcl::fail;

extern "C" kcl_status kcl_trim_create(const uint32_t *pairs, uint64_t m, uint32_t n_in,
                                      uint32_t k, kcl_trim **out) {
  return kcl_trim_create_opts(pairs, m, n_in, k, nullptr, out);
}

extern "C" kcl_status kcl_trim_create_opts(const uint32_t *pairs, uint64_t m, uint32_t n_in,
                                           uint32_t k, const kcl_trim_opts *opts,
                                           kcl_trim **out) {
  if (!out) return fail(KCL_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (k < 3 || k > 255)
    return fail(KCL_E_INVALID_ARG, "k = %u: trimming takes 3 <= k <= 255 (kClistCoreTrim.c:27)",
                k);
  if (m && !pairs) return fail(KCL_E_INVALID_ARG, "pairs is NULL");
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) {
    if (pairs[2 * i] == pairs[2 * i + 1])
      return fail(KCL_E_INVALID_ARG, "self-loop at edge %llu", (unsigned long long)i);
    mx = std::max(mx, std::max(pairs[2 * i], pairs[2 * i + 1]));
  }
  if (mx >= 0xFFFFFFFEu) return fail(KCL_E_UNSUPPORTED, "node id %u", mx);
  if (n_in == 0) n_in = mx + 1;
  if (m && n_in <= mx) return fail(KCL_E_INVALID_ARG, "n_in = %u but an edge names node %u",
                                   n_in, mx);
  kcl_trim *T = new (std::nothrow) kcl_trim();
  if (!T) return fail(KCL_E_OOM_HOST, "out of host memory");
  T->k = k;
  if (opts) T->opts = *opts;
  kcl_status s;
  try {
    s = kcl::trim_run(T, pairs, m, n_in);
  } catch (const std::bad_alloc &) {
    s = fail(KCL_E_OOM_HOST, "out of host memory");
  }
  if (s != KCL_OK) {
    delete T;
    return s;
  }
  *out = T;
  return KCL_OK;
}

extern "C" void kcl_trim_destroy(kcl_trim *t) { delete t; }

extern "C" kcl_status kcl_trim_get_info(const kcl_trim *t, kcl_trim_info *info) {
  if (!t || !info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *info = t->info;
  return KCL_OK;
}

extern "C" kcl_status kcl_trim_nodes(const kcl_trim *t, uint32_t *map, uint64_t *nck,
                                     uint32_t *peel_key, uint64_t *peel_core, uint8_t *kept) {
  if (!t) return fail(KCL_E_INVALID_ARG, "NULL trim");
  const size_t n2 = t->info.n2;
  if (map) memcpy(map, t->map.data(), n2 * sizeof(uint32_t));
  if (nck) memcpy(nck, t->nck.data(), n2 * sizeof(uint64_t));
  if (peel_key) memcpy(peel_key, t->ic_key.data(), n2 * sizeof(uint32_t));
  if (peel_core) memcpy(peel_core, t->ic_val.data(), n2 * sizeof(uint64_t));
  if (kept) memcpy(kept, t->keep.data(), n2);
  return KCL_OK;
}

extern "C" kcl_status kcl_trim_edges(const kcl_trim *t, uint32_t *pairs) {
  if (!t || (!pairs && t->info.mtrim)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  uint32_t w = 0;
  for (size_t j = 0; j < t->es.size(); ++j) {
    if (t->keep[t->es[j]] && t->keep[t->et[j]]) {
      pairs[2 * (size_t)w] = t->map[t->es[j]];
      pairs[2 * (size_t)w + 1] = t->map[t->et[j]];
      ++w;
    }
  }
  return KCL_OK;
}

extern "C" kcl_status kcl_trim_write(const kcl_trim *t, const char *ckdeg, const char *ckcore,
                                     const char *ckdens, const char *trimmed, const char *stat,
                                     int64_t t0_unix) {
  if (!t) return fail(KCL_E_INVALID_ARG, "NULL trim");
  const uint32_t n = t->info.n2;
  auto open = [](const char *p, FILE **f) -> kcl_status {
    *f = nullptr;
    if (!p) return KCL_OK;
    *f = fopen(p, "w");
    return *f ? KCL_OK : fail(KCL_E_IO, "cannot create %s", p);
  };
  FILE *f;
  // printckdeg (kClistCoreTrim.c:419-429)
  KCL_TRY(open(ckdeg, &f));
  if (f) {
    for (uint32_t i = 0; i < n; ++i)
      fprintf(f, "%u %llu\n", t->map[i], (unsigned long long)t->nck[i]);
    if (fclose(f)) return fail(KCL_E_IO, "cannot write %s", ckdeg);
  }
  // printdensest (kClistCoreTrim.c:672-684)
  KCL_TRY(open(ckcore, &f));
  if (f) {
    for (uint32_t i = 0; i < n; ++i)
      fprintf(f, "%u %llu\n", t->map[t->ic_key[i]], (unsigned long long)t->ic_val[i]);
    if (fclose(f)) return fail(KCL_E_IO, "cannot write %s", ckcore);
  }
  KCL_TRY(open(ckdens, &f));
  if (f) {
    fprintf(f, "%lf %lf %u", t->info.ckrho, t->info.rho, t->info.size);
    // unsigned loop bounds as in the reference: no ids when size == n
    for (uint32_t i = n - 1; n > 0 && i > n - t->info.size - 1; i--)
      fprintf(f, " %u", t->map[t->ic_key[i]]);
    fprintf(f, "\n");
    if (fclose(f)) return fail(KCL_E_IO, "cannot write %s", ckdens);
  }
  KCL_TRY(open(trimmed, &f));
  if (f) {
    for (size_t j = 0; j < t->es.size(); ++j)
      if (t->keep[t->es[j]] && t->keep[t->et[j]])
        fprintf(f, "%u %u\n", t->map[t->es[j]], t->map[t->et[j]]);
    if (fclose(f)) return fail(KCL_E_IO, "cannot write %s", trimmed);
  }
  KCL_TRY(open(stat, &f));
  if (f) {
    if (t0_unix <= 0) t0_unix = (int64_t)time(nullptr);
    fprintf(f, "[Number of Nodes]\t[Number of Edges]\t[Number of k-Cliques]\t[Time (seconds)]\n");
    fprintf(f, "%u\t%u\t%llu\t%ld\n", t->info.ntrim, t->info.mtrim,
            (unsigned long long)t->info.ncktrim, (long)((int64_t)time(nullptr) - t0_unix));
    if (fclose(f)) return fail(KCL_E_IO, "cannot write %s", stat);
  }
  return KCL_OK;
}
