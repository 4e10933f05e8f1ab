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
  size_t lds[kNumClasses] = {};
  PassTable() {
    void (*fill[kNumClasses])(PassFn *, PassFn *, size_t *) = {
        fill_pass_class_0, fill_pass_class_1, fill_pass_class_2, fill_pass_class_3,
        fill_pass_class_4, fill_pass_class_5, fill_pass_class_6, fill_pass_class_7};
    for (uint32_t c = 0; c < kNumClasses; ++c) fill[c](fn[c], fn_all[c], &lds[c]);
  }
};

const PassTable &pass_table() {
  static const PassTable table;  // thread-safe static initialisation
  return table;
}

PassFn kernel_for(uint32_t wclass, uint32_t k, bool all) {
  if (wclass >= kNumClasses || k < 2 || k > (uint32_t)kMaxTemplK) return nullptr;
  return all ? pass_table().fn_all[wclass][k] : pass_table().fn[wclass][k];
}

size_t lds_bytes(uint32_t wclass) { return pass_table().lds[wclass]; }

}  // namespace

uint32_t max_pass_k() { return (uint32_t)kMaxTemplK; }

namespace {

uint32_t blocks_per_cu(uint32_t wclass, const DeviceProps &props) {
  const uint32_t waves = kClassWaves[wclass];
  uint32_t per_cu = (uint32_t)(props.lds_per_cu / (lds_bytes(wclass) + 256));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 32 / waves) per_cu = 32 / waves;
  return per_cu;
}

}  // namespace

namespace {

__global__ void zero_tickets_kernel(uint32_t *t) {
  if (threadIdx.x < kNumClasses) t[threadIdx.x] = 0;
}

}  // namespace

uint32_t class_slots(uint32_t wclass) {
  DeviceProps props;
  if (wclass >= kNumClasses || device_props(&props) != KCL_OK) return 1;
  return props.cu_count * blocks_per_cu(wclass, props);
}

kcl_status launch_pass(const kcl_graph *g, const uint64_t *d_key, uint32_t k, PassMode mode,
                       DevicePlan &plan, unsigned long long *d_cnt, hipStream_t stream,
                       unsigned long long *prof) {
  DeviceProps props;
  KCL_TRY(device_props(&props));
  // A kernel rather than a memset node, so the pass can be graph-captured.
  zero_tickets_kernel<<<1, 64, 0, stream>>>(plan.tickets.p);
  KCL_HIP(hipGetLastError());
  for (int c = (int)kNumClasses - 1; c >= 0; --c) {  // heaviest word class first
    const uint32_t count = plan.class_count[c];
    if (!count) continue;
    PassFn fn = kernel_for((uint32_t)c, k, mode == kCreditAll);
    if (!fn) return fail(KCL_E_UNSUPPORTED, "no enumeration kernel for k = %u", k);
    const uint32_t waves = kClassWaves[c];
    uint32_t per_cu = blocks_per_cu((uint32_t)c, props);
    if (plan.max_per_cu && plan.max_per_cu < per_cu) per_cu = plan.max_per_cu;
    uint64_t blocks = (uint64_t)props.cu_count * per_cu;
    if (blocks > count) blocks = count;
    hipLaunchKernelGGL(fn, dim3((uint32_t)blocks), dim3(64 * waves), 0, stream,
                       (const uint32_t *)g->d_cd.p, (const uint32_t *)g->d_adj.p, d_key,
                       (const unsigned long long *)(g->core_h ? g->d_core.p : nullptr), g->core_h,
                       plan.tasks.p + plan.class_begin[c], count, plan.tickets.p + c,
                       (uint32_t)mode, d_cnt,
                       prof ? prof + 2 * (size_t)plan.class_begin[c] : nullptr);
    KCL_HIP(hipGetLastError());
  }
  return KCL_OK;
}

kcl_status build_plan(const std::vector<uint32_t> &cd, uint32_t k, const RootCost *rc,
                      uint32_t split_work, PassPlan *plan, const std::vector<uint64_t> *keep) {
  const uint32_t n = cd.empty() ? 0 : (uint32_t)(cd.size() - 1);
  struct Entry {
    uint64_t cost;
    uint32_t s;
    uint32_t np;
  };
  std::vector<Entry> cls[kNumClasses];
  const uint32_t need = k - 1;  // members needed in N+(s)
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t d = cd[s + 1] - cd[s];
    if (d < need || d == 0) continue;
    if (keep && (*keep)[s] == 0) continue;
    if (d > kMaxRootDegree)
      return fail(KCL_E_UNSUPPORTED,
                  "node %u has out-degree %u in the DAG; the enumerator handles up to %u", s, d,
                  kMaxRootDegree);
    const uint32_t words = (d + 63) / 64;
    uint32_t c = 0;
    while (kClassWords[c] < words) ++c;  // smallest class that holds d members
    const uint64_t work = rc ? rc->work[s] : (uint64_t)d * d;
    const uint64_t build = rc ? rc->build[s] : (uint64_t)d;
    const uint64_t cost = work + build;
    uint32_t np = 1;
    // Each part rebuilds the root's rows, so a part takes at least
    // kBuildShare times the build's work units of counting.
    const uint64_t part_work = std::max<uint64_t>(split_work, kBuildShare * build);
    if (split_work && work > part_work) {
      uint64_t want = (work + part_work - 1) / part_work;
      const uint64_t items = (k >= 5) ? (uint64_t)d * (d - 1) / 2 : d;
      if (want > items) want = items;
      if (want > 4096) want = 4096;
      np = (uint32_t)(want ? want : 1);
    }
    cls[c].push_back(Entry{cost, s, np});
  }
  plan->tasks.clear();
  uint32_t at = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    std::stable_sort(cls[c].begin(), cls[c].end(),
                     [](const Entry &a, const Entry &b) { return a.cost > b.cost; });
    plan->class_begin[c] = at;
    uint32_t cnt = 0;
    for (const Entry &e : cls[c])
      for (uint32_t p = 0; p < e.np; ++p) {
        plan->tasks.push_back(Task{e.s, p | (e.np << 16)});
        ++cnt;
      }
    plan->class_count[c] = cnt;
    at += cnt;
  }
  plan->class_begin[kNumClasses] = at;
  return KCL_OK;
}

void promote_plan(PassPlan *plan, uint32_t min_class) {
  if (min_class == 0 || min_class >= kNumClasses) return;
  PassPlan out;
  out.tasks.reserve(plan->tasks.size());
  uint32_t at = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    out.class_begin[c] = at;
    if (c < min_class) continue;
    // The floor class first (its own, heavier roots), then the promoted
    // lighter classes, widest first.
    for (uint32_t i = plan->class_begin[c]; i < plan->class_begin[c + 1]; ++i)
      out.tasks.push_back(plan->tasks[i]);
    if (c == min_class)
      for (uint32_t l = min_class; l-- > 0;)
        for (uint32_t i = plan->class_begin[l]; i < plan->class_begin[l + 1]; ++i)
          out.tasks.push_back(plan->tasks[i]);
    out.class_count[c] = (uint32_t)out.tasks.size() - at;
    at = (uint32_t)out.tasks.size();
  }
  out.class_begin[kNumClasses] = at;
  *plan = std::move(out);
}

kcl_status upload_plan(const PassPlan &plan, DevicePlan *dev) {
  dev->host_tasks = plan.tasks;
  KCL_TRY(dev->tasks.alloc(plan.tasks.size()));
  KCL_TRY(dev->tickets.alloc(kNumClasses));
  if (!plan.tasks.empty())
    KCL_HIP(hipMemcpy(dev->tasks.p, plan.tasks.data(), plan.tasks.size() * sizeof(Task),
                      hipMemcpyHostToDevice));
  for (uint32_t c = 0; c <= kNumClasses; ++c) dev->class_begin[c] = plan.class_begin[c];
  for (uint32_t c = 0; c < kNumClasses; ++c) dev->class_count[c] = plan.class_count[c];
  return KCL_OK;
}

}  // namespace kcl
