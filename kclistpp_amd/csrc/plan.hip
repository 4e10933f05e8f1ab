// plan.hip - the pass schedule: which roots get tasks, in which order, and
// how many parts each heavy root is split into.
//
// The reference hands DAG edges to an OpenMP dynamic schedule in edge
// order (kCDensestCmpSync.c:300-305); here every pass runs a fixed task
// list per (graph, k), built once:
//  * a root-credit k-clique counting pass gives each root's k-clique count;
//    roots with none are dropped (exact: they contribute nothing to any
//    pass);
//  * the remaining roots are costed by one of three schedules and split:
//      COUNTED  steps of the k-pass (one per (k-1)-clique with that root,
//               from a root-credit (k-1)-pass) x row words + build entries;
//      DEGREE   d+(s)^2 (plain: no extra pass);
//      LEARNED  cycles predicted by the cost model from eight features.
// The schedule moves time only: counts are integers summed with atomics.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "kcl_internal.h"

namespace kcl {

namespace {

// Per-root clique counts: a root-credit pass of k-cliques over `plan`
// (out[s] = number of k-cliques whose largest id is s).
kcl_status root_counts(kcl_graph *g, uint32_t k, const PassPlan &plan, hipStream_t st,
                       std::vector<uint64_t> *out) {
  out->assign(g->n, 0);
  DevicePlan dp;
  KCL_TRY(upload_plan(plan, &dp));
  DevBuf<unsigned long long> c;
  DevBuf<unsigned long long> key;
  KCL_TRY(c.alloc(g->n));
  KCL_TRY(key.alloc(g->n));
  KCL_TRY(zero_u64(c.p, g->n, st));
  KCL_TRY(zero_u64(key.p, g->n, st));
  KCL_TRY(launch_pass(g, (const uint64_t *)key.p, k, kCreditRoot, dp, c.p, st));
  KCL_HIP(hipMemcpyAsync(out->data(), c.p, (size_t)g->n * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, st));
  return pass_status(dp, st);
}

// Work per root: the steps of its k-pass (one per (k-1)-clique with that
// root, counted by a root-credit (k-1)-pass) times the row width, plus the
// adjacency entries its build scans (so light roots keep a weight and split
// parts amortise their rebuild).
kcl_status root_costs(kcl_graph *g, uint32_t k, hipStream_t st, RootCost *out) {
  PassPlan p;
  KCL_TRY(build_plan(g, k - 1, nullptr, 0, &p));
  KCL_TRY(root_counts(g, k - 1, p, st, &out->work));
  out->build.assign(g->n, 0);
  for (uint32_t s = 0; s < g->n; ++s) {
    uint64_t b = 0;
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j) {
      const uint32_t x = g->adj[j];
      b += g->cd[x + 1] - g->cd[x];
    }
    out->build[s] = b;
    out->work[s] *= (g->cd[s + 1] - g->cd[s] + 63) / 64;
  }
  return KCL_OK;
}

uint32_t class_of(uint32_t d) {
  const uint32_t words = (d + 63) / 64;
  uint32_t c = 0;
  while (c + 1 < kNumClasses && kClassWords[c] < words) ++c;
  return c;
}

// Each slot of a class should see at least this many tasks' worth of work,
// so that no single task sets the class's makespan.
constexpr double kLearnedDepth = 4.0;

// LEARNED tasks: per root, predicted build cycles B and counting cycles C;
// a part of np costs B + C/np.  Within a class with S concurrent task
// slots and total predicted time T, counting is split so that each part's
// counting is at most max(T / (S kLearnedDepth), kBuildShare B).
kcl_status build_plan_learned(const kcl_graph *g, uint32_t k, const std::vector<float> &pred,
                              const std::vector<uint64_t> *keep, PassPlan *plan) {
  const std::vector<uint32_t> &cd = g->cd;
  const uint32_t n = cd.empty() ? 0 : (uint32_t)(cd.size() - 1);
  std::vector<uint32_t> big;
  struct Entry {
    double b, c, cost;
    uint32_t s, np;
  };
  std::vector<Entry> cls[kNumClasses];
  const uint32_t need = k - 1;
  double total[kNumClasses] = {};
  auto cycles = [](float lg) {
    const double e = std::min(60.0, std::max(0.0, (double)lg));
    return std::exp2(e) - 1.0;
  };
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t d = cd[s + 1] - cd[s];
    if (d < need || d == 0) continue;
    if (keep && (*keep)[s] == 0) continue;
    if (d > g->edge_above) {  // edge tasks
      big.push_back(s);
      continue;
    }
    const uint32_t c = class_of(d);
    Entry e{cycles(pred[2 * (size_t)s]), cycles(pred[2 * (size_t)s + 1]), 0.0, s, 1};
    total[c] += e.b + e.c;
    cls[c].push_back(e);
  }
  plan->tasks.clear();
  uint32_t at = 0;
  for (uint32_t c = 0; c < kNumClasses; ++c) {
    const double slots = (double)std::max<uint32_t>(1, class_slots(c, k));
    const double tau = total[c] / (slots * kLearnedDepth);
    for (Entry &e : cls[c]) {
      const double target = std::max(tau, (double)kBuildShare * std::max(e.b, 1.0));
      if (e.c > target) {
        const uint32_t d = cd[e.s + 1] - cd[e.s];
        const double items = (k >= 5) ? (double)d * (d - 1) / 2 : (double)d;
        double want = std::ceil(e.c / target);
        want = std::min(want, std::min(items, 4096.0));
        e.np = (uint32_t)std::max(1.0, want);
      }
      e.cost = e.b + e.c / e.np;
    }
    std::stable_sort(cls[c].begin(), cls[c].end(),
                     [](const Entry &a, const Entry &b) { return a.cost > b.cost; });
    plan->class_begin[c] = at;
    const double light = light_cost(total[c], c, k);
    uint32_t cnt = 0;
    plan->light_from[c] = UINT32_MAX;
    for (const Entry &e : cls[c])
      for (uint32_t p = 0; p < e.np; ++p) {
        if (plan->light_from[c] == UINT32_MAX && e.cost < light) plan->light_from[c] = cnt;
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

}  // namespace

kcl_status root_features(kcl_graph *g, uint32_t k, hipStream_t st, std::vector<float> *out) {
  const uint32_t n = g->n;
  out->assign((size_t)n * KCL_CM_FEATURES, 0.f);
  if (n == 0) return KCL_OK;
  // Edges inside N+(s) = triangles whose largest id is s.
  std::vector<uint64_t> tri(n, 0);
  if (g->core >= 2) {
    PassPlan p3;
    KCL_TRY(build_plan(g, 3, nullptr, 1u << 17, &p3));
    KCL_TRY(root_counts(g, 3, p3, st, &tri));
  }
  const std::vector<uint32_t> &cd = g->cd;
  uint32_t core = 0;
  for (uint32_t s = n; s-- > 0;) {  // degeneracy core: suffix maximum of d+
    const uint32_t d = cd[s + 1] - cd[s];
    core = std::max(core, d);
    uint64_t sum2 = 0;
    uint32_t mx = 0;
    for (uint32_t j = cd[s]; j < cd[s + 1]; ++j) {
      const uint32_t x = g->adj[j];
      const uint32_t dx = cd[x + 1] - cd[x];
      sum2 += dx;
      mx = std::max(mx, dx);
    }
    const double pairs = d >= 2 ? (double)d * (d - 1) / 2 : 1.0;
    float *f = out->data() + (size_t)s * KCL_CM_FEATURES;
    f[0] = (float)std::log2(1.0 + d);
    f[1] = (float)std::log2(1.0 + core);
    f[2] = (float)std::log2(1.0 + (double)sum2);
    f[3] = (float)std::log2(1.0 + (double)tri[s]);
    f[4] = (float)((double)tri[s] / pairs);
    f[5] = (float)std::log2(1.0 + mx);
    f[6] = (float)std::log2((double)kClassWords[class_of(d)]);
    f[7] = (float)k;
  }
  return KCL_OK;
}

namespace {

// log2(1 + cycles) predictions (build, counting) for every node as a root.
kcl_status learned_predictions(kcl_graph *g, uint32_t k, const kcl_cost_model *model,
                               hipStream_t st, std::vector<float> *pred) {
  std::vector<float> feat;
  KCL_TRY(root_features(g, k, st, &feat));
  DevBuf<float> dx, dy;
  KCL_TRY(dx.alloc(feat.size()));
  KCL_TRY(dy.alloc((size_t)g->n * KCL_CM_OUTPUTS));
  KCL_HIP(hipMemcpyAsync(dx.p, feat.data(), feat.size() * sizeof(float), hipMemcpyHostToDevice,
                         st));
  KCL_TRY(cost_model_predict_device(model, dx.p, g->n, dy.p, st));
  pred->assign((size_t)g->n * KCL_CM_OUTPUTS, 0.f);
  KCL_HIP(hipMemcpy(pred->data(), dy.p, pred->size() * sizeof(float), hipMemcpyDeviceToHost));
  return KCL_OK;
}

}  // namespace

kcl_status plan_for_schedule(kcl_graph *g, uint32_t k, uint32_t schedule,
                             const kcl_cost_model *model, uint32_t split,
                             const std::vector<uint64_t> *keep, hipStream_t st, PassPlan *out) {
  if (schedule > KCL_SCHED_DEGREE)
    return fail(KCL_E_INVALID_ARG, "unknown schedule %u", schedule);
  if (schedule == KCL_SCHED_LEARNED && model) {
    std::vector<float> pred;
    KCL_TRY(learned_predictions(g, k, model, st, &pred));
    return build_plan_learned(g, k, pred, keep, out);
  }
  if (schedule == KCL_SCHED_COUNTED && k >= 4) {
    RootCost cost;
    KCL_TRY(root_costs(g, k, st, &cost));
    return build_plan(g, k, &cost, split, out, keep);
  }
  return build_plan(g, k, nullptr, split, out, keep);
}

kcl_status make_plan(kcl_graph *g, uint32_t k, uint32_t split, hipStream_t st, PassPlan *out,
                     std::vector<uint64_t> *root_cliques, kcl_plan_info *info,
                     uint32_t schedule, const kcl_cost_model *model,
                     const std::vector<uint64_t> *candidates) {
  if (schedule > KCL_SCHED_DEGREE)
    return fail(KCL_E_INVALID_ARG, "unknown schedule %u", schedule);
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t n = g->n;
  const bool learned = schedule == KCL_SCHED_LEARNED && model;
  PassPlan all;
  std::vector<float> pred;
  if (learned) {
    KCL_TRY(learned_predictions(g, k, model, st, &pred));
    KCL_TRY(build_plan_learned(g, k, pred, candidates, &all));
  } else {
    KCL_TRY(plan_for_schedule(g, k, schedule, model, split, candidates, st, &all));
  }
  std::vector<uint64_t> ck;
  KCL_TRY(root_counts(g, k, all, st, &ck));
  if (learned) {
    // Same predictions, kept roots only (the class totals change).
    KCL_TRY(build_plan_learned(g, k, pred, &ck, out));
  } else {
    // Reuse the candidate plan's costs: drop roots without k-cliques.
    PassPlan kept;
    kept.tasks.reserve(all.tasks.size());
    uint32_t at = 0;
    for (uint32_t c = 0; c < kNumClasses; ++c) {
      kept.class_begin[c] = at;
      kept.light_from[c] = UINT32_MAX;
      for (uint32_t i = all.class_begin[c]; i < all.class_begin[c + 1]; ++i) {
        if (!ck[all.tasks[i].s]) continue;
        if (kept.light_from[c] == UINT32_MAX && i - all.class_begin[c] >= all.light_from[c])
          kept.light_from[c] = (uint32_t)kept.tasks.size() - at;
        kept.tasks.push_back(all.tasks[i]);
      }
      kept.class_count[c] = (uint32_t)kept.tasks.size() - at;
      kept.light_from[c] = std::min(kept.light_from[c], kept.class_count[c]);
      at = (uint32_t)kept.tasks.size();
    }
    kept.class_begin[kNumClasses] = at;
    at = 0;
    for (uint32_t c = 0; c < kNumClasses; ++c) {
      kept.eclass_begin[c] = at;
      for (uint32_t i = all.eclass_begin[c]; i < all.eclass_begin[c + 1]; ++i)
        if (ck[all.edge_tasks[i].s]) kept.edge_tasks.push_back(all.edge_tasks[i]);
      kept.eclass_count[c] = (uint32_t)kept.edge_tasks.size() - at;
      at = (uint32_t)kept.edge_tasks.size();
    }
    kept.eclass_begin[kNumClasses] = at;
    *out = std::move(kept);
  }
  if (info) {
    uint64_t cliques = 0;
    uint32_t kept = 0, roots_all = 0;
    for (uint32_t v = 0; v < n; ++v) {
      cliques += ck[v];
      kept += ck[v] ? 1u : 0u;
      roots_all += (g->cd[v + 1] - g->cd[v] >= k - 1 && (!candidates || (*candidates)[v]))
                       ? 1u : 0u;
    }
    info->roots_total = roots_all;
    info->roots_kept = kept;
    info->tasks = (uint32_t)(out->tasks.size() + out->edge_tasks.size());
    info->cliques = cliques;
    info->plan_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (root_cliques) root_cliques->swap(ck);
  return KCL_OK;
}

}  // namespace kcl
