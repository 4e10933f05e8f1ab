// pass.hip - host side of the enumeration pass: kernel table, launches,
// task plans.  The kernels are in pass_kernel.h (one instantiation unit per
// word class, pass_inst.hip).
#include <algorithm>

#include "pass_kernel.h"

namespace kcl {

namespace {

struct PassTable {
  PassFn fn[kNumClasses][kMaxTemplK + 1] = {};
  PassFn fn_all[kNumClasses][kMaxTemplK + 1] = {};
  PassFn deep[kNumClasses][2] = {};  // large k: credit winner/root, credit all
  size_t lds[kNumClasses] = {};
  PassTable() {
    void (*fill[kNumClasses])(PassFn *, PassFn *, PassFn *, size_t *) = {
        fill_pass_class_0, fill_pass_class_1, fill_pass_class_2, fill_pass_class_3,
        fill_pass_class_4, fill_pass_class_5, fill_pass_class_6, fill_pass_class_7,
        fill_pass_class_8};
    for (uint32_t c = 0; c < kNumClasses; ++c) fill[c](fn[c], fn_all[c], deep[c], &lds[c]);
  }
};

const PassTable &pass_table() {
  static const PassTable table;  // thread-safe static initialisation
  return table;
}

PassFn kernel_for(uint32_t wclass, uint32_t k, bool all) {
  if (wclass >= kNumClasses || k < 2 || k > kMaxDeepK) return nullptr;
  if (k > (uint32_t)kMaxTemplK) return pass_table().deep[wclass][all ? 1 : 0];
  return all ? pass_table().fn_all[wclass][k] : pass_table().fn[wclass][k];
}

// Waves per workgroup of the class-c kernel for k.
uint32_t waves_for(uint32_t wclass, uint32_t k, bool all) {
  return k > (uint32_t)kMaxTemplK ? deep_waves(wclass) : class_waves(wclass, k, all);
}

// Large k: each wavefront's path (and, for k-clique degrees, per-level
// counts) must fit its share of the class's key array.
bool deep_fits(uint32_t wclass, uint32_t k, bool all) {
  if (k <= (uint32_t)kMaxTemplK) return true;
  const uint32_t W = kClassWords[wclass];
  return deep_words(k - 3, all) <= 64u * W / deep_waves(wclass);
}

// The class whose kernel runs class c's tasks: c itself, or for large k the
// narrowest wider class with room for the walk (a task fits any wider
// class); kNumClasses if none has.
uint32_t launch_class(uint32_t c, uint32_t k, bool all) {
  while (c < kNumClasses && !deep_fits(c, k, all)) ++c;
  return c;
}

size_t lds_bytes(uint32_t wclass) { return pass_table().lds[wclass]; }

}  // namespace

uint32_t max_pass_k() { return kMaxDeepK; }

namespace {

// Resident workgroups per CU of a class's kernel: as many as LDS and the 32
// wave slots allow, except that the 1- and 2-word classes of counting passes
// at k <= 5 keep to 16 waves per CU.  Their roots wait on chains of global
// loads while holding wave slots, and at k <= 5 the wider classes that run
// beside them are VALU bound and need those slots: the 1M/16M k = 5 pass
// takes 2.47 ms with 32 light waves per CU, 1.96 ms with 16 or 8.  At k = 6
// the cap cost 2.7 % (1M/16M) and was neutral on 4M/60M.  light_waves is the
// plan's cap (kLightWaves by default; the batch, whose passes are nearly all
// light roots, lifts it and tunes residency itself).
uint32_t blocks_per_cu(uint32_t wclass, const DeviceProps &props, uint32_t waves, uint32_t k,
                       bool all, uint32_t light_waves = kLightWaves) {
  uint32_t per_cu = (uint32_t)(props.lds_per_cu / (lds_bytes(wclass) + 256));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 32 / waves) per_cu = 32 / waves;
  if (!all && k <= 5 && wclass <= 1 && per_cu * waves > light_waves)
    per_cu = std::max(1u, light_waves / waves);
  return per_cu;
}

}  // namespace

namespace {

// Two counters per class (single claims, chunked claims), vertex and edge,
// then the pass's status word.
constexpr uint32_t kTickets = 2 * 2 * kNumClasses;
constexpr uint32_t kTicketWords = kPassTicketWords;
static_assert(kTicketWords == kTickets + 1, "tickets, then the status word");

__global__ void zero_tickets_kernel(uint32_t *t) {
  if (threadIdx.x < kTicketWords) t[threadIdx.x] = 0;
}

}  // namespace

kcl_status class_launch_shape(uint32_t wclass, uint32_t k, bool all, uint32_t max_per_cu,
                              uint32_t count, uint32_t *waves, uint32_t *per_cu,
                              uint32_t *blocks, uint32_t *lds) {
  DeviceProps props;
  KCL_TRY(device_props(&props));
  if (wclass >= kNumClasses) return fail(KCL_E_INVALID_ARG, "word class %u", wclass);
  const uint32_t lc = launch_class(wclass, k, all);
  if (lc >= kNumClasses) return fail(KCL_E_UNSUPPORTED, "k = %u: no word class holds the walk", k);
  wclass = lc;
  const uint32_t w = waves_for(wclass, k, all);
  uint32_t pc = blocks_per_cu(wclass, props, w, k, all);
  if (max_per_cu && max_per_cu < pc) pc = max_per_cu;
  uint64_t b = (uint64_t)props.cu_count * pc;
  if (b > count) b = count;
  *waves = w;
  *per_cu = pc;
  *blocks = (uint32_t)b;
  *lds = (uint32_t)lds_bytes(wclass);
  return KCL_OK;
}

uint32_t class_slots(uint32_t wclass, uint32_t k) {
  DeviceProps props;
  if (wclass >= kNumClasses || device_props(&props) != KCL_OK) return 1;
  return props.cu_count * blocks_per_cu(wclass, props, waves_for(wclass, k, false), k, false);
}

kcl_status launch_pass(const kcl_graph *g, const uint64_t *d_key, uint32_t k, PassMode mode,
                       DevicePlan &plan, unsigned long long *d_cnt, hipStream_t stream,
                       unsigned long long *prof) {
  PassLaunch pl;
  pl.tasks = plan.tasks.p;
  pl.tickets = plan.tickets.p;
  pl.class_begin = plan.class_begin;
  pl.class_count = plan.class_count;
  pl.eclass_begin = plan.eclass_begin;
  pl.eclass_count = plan.eclass_count;
  pl.light_from = plan.light_from;
  pl.max_per_cu = plan.max_per_cu;
  pl.light_waves = plan.light_waves;
  return launch_pass_raw(g, d_key, k, mode, pl, d_cnt, stream, prof);
}

kcl_status launch_pass_raw(const kcl_graph *g, const uint64_t *d_key, uint32_t k, PassMode mode,
                           const PassLaunch &plan, unsigned long long *d_cnt, hipStream_t stream,
                           unsigned long long *prof) {
  DeviceProps props;
  KCL_TRY(device_props(&props));
  // A kernel rather than a memset node, so the pass can be graph-captured.
  static_assert(kTicketWords <= 64, "tickets of one pass fit one wavefront");
  zero_tickets_kernel<<<1, 64, 0, stream>>>(plan.tickets);
  KCL_HIP(hipGetLastError());
  // The class launches go round-robin to the pass's stream and the graph's
  // side streams, so a class's tail overlaps the next class (credits are
  // atomic sums, tickets are per class); the pass's stream joins them all.
  constexpr int kQueues = 1 + kcl_graph::kSideStreams;
  KCL_HIP(hipEventRecord(g->fork, stream));
  for (int i = 0; i < kcl_graph::kSideStreams; ++i)
    KCL_HIP(hipStreamWaitEvent(g->side[i], g->fork, 0));
  uint32_t launches = 0;
  // Vertex tasks with the k kernels, then edge tasks with the (k-1) ones.
  for (uint32_t grp = 0; grp < 2; ++grp) {
    const uint32_t *begin = grp ? plan.eclass_begin : plan.class_begin;
    const uint32_t *cnts = grp ? plan.eclass_count : plan.class_count;
    const uint32_t kk = grp ? k - 1 : k;
    for (int c = (int)kNumClasses - 1; c >= 0; --c) {  // heaviest word class first
      const uint32_t count = cnts[c];
      if (!count) continue;
      const bool all = mode == kCreditAll;
      // Emit passes list cliques with the walk kernels at every k.
      const bool walk = mode == kEmit || kk > (uint32_t)kMaxTemplK;
      const uint32_t lc = launch_class((uint32_t)c, kk, all);
      if (lc >= kNumClasses)
        return fail(KCL_E_UNSUPPORTED, "k = %u: no word class has room for a %u-level %s walk",
                    kk, kk - 3, all ? "k-clique-degree" : "counting");
      PassFn fn = mode == kEmit ? pass_table().deep[lc][0] : kernel_for(lc, kk, all);
      if (!fn) return fail(KCL_E_UNSUPPORTED, "no enumeration kernel for k = %u", kk);
      const uint32_t waves = walk ? deep_waves(lc) : class_waves(lc, kk, all);
      // The walk kernels read the clique size from the bits above the mode.
      const uint32_t mode_arg = walk ? ((uint32_t)mode | (kk << 8)) : (uint32_t)mode;
      // Emit passes: prof is the emit block, shared by every class.
      unsigned long long *prof_arg =
          mode == kEmit ? prof : (prof ? prof + 2 * (size_t)begin[c] : nullptr);
      uint32_t per_cu = blocks_per_cu(lc, props, waves, kk, all, plan.light_waves);
      if (plan.max_per_cu && plan.max_per_cu < per_cu) per_cu = plan.max_per_cu;
      uint64_t blocks = (uint64_t)props.cu_count * per_cu;
      if (blocks > count) blocks = count;
      const int q = (int)(launches++ % kQueues);
      hipLaunchKernelGGL(fn, dim3((uint32_t)blocks), dim3(64 * waves), 0,
                         q ? g->side[q - 1] : stream,
                         (const uint32_t *)g->d_cd.p, (const uint32_t *)g->d_adj.p, d_key,
                         (const unsigned long long *)(g->core_h ? g->d_core.p : nullptr),
                         g->core_h, plan.tasks + begin[c], count,
                         grp ? count : std::min(plan.light_from[c], count), g->dense_items,
                         plan.tickets + 2 * (grp * kNumClasses + c), plan.tickets + kTickets,
                         plan.list, plan.nlist, mode_arg, d_cnt, prof_arg);
      KCL_HIP(hipGetLastError());
    }
  }
  for (int i = 0; i < kcl_graph::kSideStreams; ++i) {
    KCL_HIP(hipEventRecord(g->join[i], g->side[i]));
    KCL_HIP(hipStreamWaitEvent(stream, g->join[i], 0));
  }
  return KCL_OK;
}

uint32_t *pass_status_word(DevicePlan &plan) { return plan.tickets.p + kTickets; }

kcl_status pass_status_error(uint32_t status) {
  if (status & kStatusEdgeOverflow)
    return fail(KCL_E_UNSUPPORTED,
                "an edge task had more common out-neighbours than its word class holds "
                "(the plan's member counts disagree with the graph); the pass's counts are "
                "incomplete");
  if (status) return fail(KCL_E_HIP, "enumeration pass status 0x%x", status);
  return KCL_OK;
}

kcl_status pass_status(DevicePlan &plan, hipStream_t stream) {
  if (!plan.tickets.p) return KCL_OK;
  uint32_t w = 0;
  KCL_HIP(hipMemcpyAsync(&w, pass_status_word(plan), 4, hipMemcpyDeviceToHost, stream));
  KCL_HIP(hipStreamSynchronize(stream));
  return pass_status_error(w);
}

namespace {

// Members of N+(s) & N+(u) per edge task (one wavefront per task: lanes
// walk the shorter list and search the longer).
__global__ void edge_members_kernel(const uint32_t *__restrict__ cd,
                                    const uint32_t *__restrict__ adj,
                                    const Task *__restrict__ tasks, uint32_t n,
                                    uint32_t *__restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n;
       t += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint32_t s = tasks[t].s, u = task_edge_node(tasks[t].part_np);
    const uint32_t bs = cd[s], ls = cd[s + 1] - bs, bu = cd[u], lu = cd[u + 1] - bu;
    const uint32_t *A = adj + (ls <= lu ? bs : bu);
    const uint32_t *B = adj + (ls <= lu ? bu : bs);
    const uint32_t la = ls <= lu ? ls : lu, lb = ls <= lu ? lu : ls;
    uint32_t hits = 0;
    for (uint32_t i = lane; i < la; i += 64) {
      const uint32_t x = A[i];
      uint32_t lo = 0, hi = lb;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (B[mid] < x) lo = mid + 1;
        else hi = mid;
      }
      hits += (lo < lb && B[lo] == x) ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off);
    if (lane == 0) out[t] = hits;
  }
}

}  // namespace

kcl_status add_edge_tasks(const kcl_graph *g, uint32_t k, const std::vector<uint32_t> &big,
                          PassPlan *plan) {
  plan->edge_tasks.clear();
  for (uint32_t c = 0; c <= kNumClasses; ++c) plan->eclass_begin[c] = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) plan->eclass_count[c] = 0;
  if (big.empty()) return KCL_OK;
  if (k < 3)
    return fail(KCL_E_UNSUPPORTED, "node %u has out-degree %u > %u, supported for k >= 3 only",
                big[0], g->cd[big[0] + 1] - g->cd[big[0]], kMaxRootDegree);
  std::vector<Task> cand;
  for (uint32_t s : big)
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j) {
      const uint32_t u = g->adj[j];
      if (u & kEdgeTaskFlag) return fail(KCL_E_UNSUPPORTED, "node id %u >= 2^31 in an edge task", u);
      cand.push_back(Task{s, kEdgeTaskFlag | u});
    }
  std::vector<uint32_t> members(cand.size());
  {
    DevBuf<Task> dt;
    DevBuf<uint32_t> dm;
    KCL_TRY(dt.alloc(cand.size()));
    KCL_TRY(dm.alloc(cand.size()));
    KCL_HIP(hipMemcpyAsync(dt.p, cand.data(), cand.size() * sizeof(Task), hipMemcpyHostToDevice,
                           g->stream));
    uint64_t blocks = (cand.size() + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    edge_members_kernel<<<(uint32_t)blocks, 256, 0, g->stream>>>(g->d_cd.p, g->d_adj.p, dt.p,
                                                                 (uint32_t)cand.size(), dm.p);
    KCL_HIP(hipGetLastError());
    KCL_HIP(hipMemcpyAsync(members.data(), dm.p, members.size() * 4, hipMemcpyDeviceToHost,
                           g->stream));
    KCL_HIP(hipStreamSynchronize(g->stream));
  }
  struct Entry {
    uint32_t c;
    Task t;
  };
  std::vector<Entry> cls[kNumClasses];
  for (size_t i = 0; i < cand.size(); ++i) {
    const uint32_t c = members[i];
    if (c < k - 2) continue;  // no (k-2)-clique among the common members
    if (c > kMaxRootDegree)
      return fail(KCL_E_UNSUPPORTED,
                  "edge (%u, %u) has %u common out-neighbours in the DAG; the enumerator "
                  "handles up to %u",
                  cand[i].s, task_edge_node(cand[i].part_np), c, kMaxRootDegree);
    const uint32_t words = (c + 63) / 64;
    uint32_t w = 0;
    while (kClassWords[w] < words) ++w;
    cls[w].push_back(Entry{c, cand[i]});
  }
  uint32_t at = 0;
  for (uint32_t w = 0; w < kNumClasses; ++w) {
    std::stable_sort(cls[w].begin(), cls[w].end(),
                     [](const Entry &a, const Entry &b) { return a.c > b.c; });
    plan->eclass_begin[w] = at;
    for (const Entry &e : cls[w]) plan->edge_tasks.push_back(e.t);
    plan->eclass_count[w] = (uint32_t)cls[w].size();
    at += (uint32_t)cls[w].size();
  }
  plan->eclass_begin[kNumClasses] = at;
  return KCL_OK;
}

double light_cost(double class_total, uint32_t wclass, uint32_t k) {
  return class_total / ((double)std::max<uint32_t>(1, class_slots(wclass, k)) * kLightShare);
}

kcl_status build_plan(const kcl_graph *g, uint32_t k, const RootCost *rc, uint32_t split_work,
                      PassPlan *plan, const std::vector<uint64_t> *keep) {
  const std::vector<uint32_t> &cd = g->cd;
  const uint32_t n = cd.empty() ? 0 : (uint32_t)(cd.size() - 1);
  std::vector<uint32_t> big;
  struct Entry {
    uint64_t cost, work, build;
    uint32_t s;
    uint32_t np;
  };
  std::vector<Entry> cls[kNumClasses];
  double total[kNumClasses] = {};
  const uint32_t need = k - 1;  // members needed in N+(s)
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t d = cd[s + 1] - cd[s];
    if (d < need || d == 0) continue;
    if (keep && (*keep)[s] == 0) continue;
    if (d > g->edge_above) {
      big.push_back(s);
      continue;
    }
    const uint32_t words = (d + 63) / 64;
    uint32_t c = 0;
    while (kClassWords[c] < words) ++c;  // smallest class that holds d members
    const uint64_t work = rc ? rc->work[s] : (uint64_t)d * d;
    const uint64_t build = rc ? rc->build[s] : (uint64_t)d;
    cls[c].push_back(Entry{work + build, work, build, s, 1});
    total[c] += (double)(work + build);
  }
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    // Counted costs: a part carries at most a quarter of one resident
    // workgroup's fair share of its class (and at least split_work), so
    // neither a heavy root nor the rebuilds of many parts set the class's
    // time.  The plain d+^2 schedule keeps the absolute split_work.
    const double tau = rc ? total[c] / (4.0 * std::max<uint32_t>(1, class_slots(c, k))) : 0.0;
    for (Entry &e : cls[c]) {
      const uint32_t d = cd[e.s + 1] - cd[e.s];
      // Each part rebuilds the root's rows, so a part takes at least
      // kBuildShare times the build's work units of counting.
      const uint64_t part_work =
          std::max<uint64_t>(std::max<uint64_t>(split_work, kBuildShare * e.build), (uint64_t)tau);
      if (split_work && e.work > part_work) {
        uint64_t want = (e.work + part_work - 1) / part_work;
        const uint64_t items = (k >= 5) ? (uint64_t)d * (d - 1) / 2 : d;
        if (want > items) want = items;
        if (want > 4096) want = 4096;
        e.np = (uint32_t)(want ? want : 1);
      }
    }
  }
  plan->tasks.clear();
  uint32_t at = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    std::stable_sort(cls[c].begin(), cls[c].end(),
                     [](const Entry &a, const Entry &b) { return a.cost > b.cost; });
    plan->class_begin[c] = at;
    const double light = light_cost(total[c], c, k);
    uint32_t cnt = 0;
    plan->light_from[c] = UINT32_MAX;
    for (const Entry &e : cls[c])
      for (uint32_t p = 0; p < e.np; ++p) {
        if (plan->light_from[c] == UINT32_MAX && (double)e.cost / e.np < light)
          plan->light_from[c] = cnt;
        plan->tasks.push_back(Task{e.s, p | (e.np << 16)});
        ++cnt;
      }
    plan->class_count[c] = cnt;
    plan->light_from[c] = std::min(plan->light_from[c], cnt);
    at += cnt;
  }
  plan->class_begin[kNumClasses] = at;
  return add_edge_tasks(g, k, big, plan);
}

namespace {

// Tasks of classes below min_class moved into min_class (the floor class's
// own tasks first, then the promoted ones, widest first).
void promote_group(std::vector<Task> &tasks, uint32_t *begin, uint32_t *count,
                   uint32_t min_class) {
  std::vector<Task> out;
  out.reserve(tasks.size());
  uint32_t nb[kNumClasses + 1] = {}, nc[kNumClasses] = {};
  uint32_t at = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    nb[c] = at;
    if (c < min_class) continue;
    for (uint32_t i = begin[c]; i < begin[c + 1]; ++i) out.push_back(tasks[i]);
    if (c == min_class)
      for (uint32_t l = min_class; l-- > 0;)
        for (uint32_t i = begin[l]; i < begin[l + 1]; ++i) out.push_back(tasks[i]);
    nc[c] = (uint32_t)out.size() - at;
    at = (uint32_t)out.size();
  }
  nb[kNumClasses] = at;
  tasks.swap(out);
  for (uint32_t c = 0; c <= kNumClasses; ++c) begin[c] = nb[c];
  for (uint32_t c = 0; c < kNumClasses; ++c) count[c] = nc[c];
}

}  // namespace

void promote_plan(PassPlan *plan, uint32_t min_class) {
  if (min_class == 0 || min_class >= kNumClasses) return;
  // The promoted (lighter) tasks follow the floor class's own: they stay
  // in its light tail.
  for (uint32_t c = 0; c < min_class; ++c) plan->light_from[c] = 0;
  promote_group(plan->tasks, plan->class_begin, plan->class_count, min_class);
  promote_group(plan->edge_tasks, plan->eclass_begin, plan->eclass_count, min_class);
}

kcl_status upload_plan(const PassPlan &plan, DevicePlan *dev) {
  dev->host_tasks = plan.tasks;
  dev->host_tasks.insert(dev->host_tasks.end(), plan.edge_tasks.begin(), plan.edge_tasks.end());
  const uint32_t nv = (uint32_t)plan.tasks.size();
  KCL_TRY(dev->tasks.alloc(dev->host_tasks.size()));
  KCL_TRY(dev->tickets.alloc(kTicketWords));
  if (!dev->host_tasks.empty())
    KCL_HIP(hipMemcpy(dev->tasks.p, dev->host_tasks.data(), dev->host_tasks.size() * sizeof(Task),
                      hipMemcpyHostToDevice));
  for (uint32_t c = 0; c <= kNumClasses; ++c) {
    dev->class_begin[c] = plan.class_begin[c];
    dev->eclass_begin[c] = nv + plan.eclass_begin[c];
  }
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    dev->class_count[c] = plan.class_count[c];
    dev->eclass_count[c] = plan.eclass_count[c];
    dev->light_from[c] = plan.light_from[c];
  }
  return KCL_OK;
}

}  // namespace kcl
