// kcl_internal.h - internal declarations of libkclpp (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "kclpp.h"

namespace kcl {

// Thread-local error message; returns s so callers can `return fail(...)`.
kcl_status fail(kcl_status s, const char *fmt, ...);

#define KCL_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      return ::kcl::fail(e_ == hipErrorOutOfMemory ? KCL_E_OOM_DEVICE : KCL_E_HIP,  \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                       \
    }                                                                               \
  } while (0)

#define KCL_TRY(expr)                       \
  do {                                      \
    kcl_status s_ = (expr);                 \
    if (s_ != KCL_OK) return s_;            \
  } while (0)

// Degeneracy ranks by the reference's heap order (host, sequential).
void degeneracy_rank(uint32_t n, uint64_t m, const uint32_t *pairs, uint32_t *rank);

// Device buffer that frees itself.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  kcl_status alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? KCL_E_OOM_DEVICE : KCL_E_HIP,
                  "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    n = count;
    return KCL_OK;
  }
};

// Word classes of the enumeration kernels: a root's out-neighbourhood of d
// nodes is held as d rows of W 64-bit words (d <= 64 W), processed by a
// workgroup of kClassWaves wavefronts.
// Waves per workgroup (the default; class_waves below gives the count per
// class and k): one for the 64-bit class, two for 128 bits, eight wider
// (the wide classes are LDS-limited to few workgroups per CU, so more lanes
// per root pay).
constexpr uint32_t kNumClasses = 9;
constexpr uint32_t kClassWords[kNumClasses] = {1, 2, 3, 4, 6, 8, 10, 12, 16};
constexpr uint32_t kClassWaves[kNumClasses] = {1, 2, 8, 8, 8, 8, 8, 8, 8};
// Waves of the class-c kernel for k (measured per class and k; `all`: the
// k-clique-degree kernels): the narrower classes use eight only for k <= 5;
// the wide ones (one or two workgroups per CU by LDS) use eight up to k = 7
// and four from k = 8, where eight would not leave their deeper loops the
// registers.  Wide means 512+ members for the counting kernels (dense-only
// from k = 6, few registers) and 640+ for the k-clique-degree kernels.  The
// 10-word counting kernels at k = 6, 7 (two workgroups per CU by LDS) run
// twelve: 24 waves per CU, within the SGPR file (4M/60M k = 6 pass 8 %
// faster than with eight).
constexpr uint32_t class_waves(uint32_t c, uint32_t k, bool all = false) {
  return (!all && kClassWords[c] == 10 && k >= 6 && k <= 7) ? 12u
         : kClassWords[c] <= 2 ? kClassWaves[c]
         : kClassWords[c] >= (all ? 10u : 8u) ? (k >= 8 ? 4u : 8u)
                                              : (k <= 5 ? 8u : 4u);
}
constexpr uint32_t kMaxWords = 16;
// Waves per CU of the 1- and 2-word classes in counting passes at k <= 5
// (pass.hip: blocks_per_cu); 32 lifts the cap.
constexpr uint32_t kLightWaves = 16;
constexpr uint32_t kMaxRootDegree = 64 * kMaxWords;

// One root task: root node s, split part p of np (the root's items dealt out
// in blocks, part p taking blocks p, p + np, ...), or an edge task (root edge (s, u), u < 2^31, flagged in the top bit).
struct Task {
  uint32_t s;
  uint32_t part_np;  // part | (np << 16), or kEdgeTaskFlag | u
};
constexpr uint32_t kEdgeTaskFlag = 0x80000000u;
__host__ __device__ inline bool task_is_edge(uint32_t part_np) {
  return (part_np & kEdgeTaskFlag) != 0;
}
__host__ __device__ inline uint32_t task_edge_node(uint32_t part_np) {
  return part_np & ~kEdgeTaskFlag;
}

// Status bits a pass sets in its plan's status word (checked by the host
// after the pass with pass_status).
constexpr uint32_t kStatusEdgeOverflow = 1u;  // an edge task had more common
                                              // members than its class holds

enum PassMode : uint32_t {
  kCreditWinner = 0,  // each clique credits its minimum-key member
  kCreditRoot = 1,    // each clique credits its root (per-root clique counts)
  kCreditAll = 2,     // each clique credits every member (k-clique degrees)
  kEmit = 3,          // each clique kept by the sampler is written out (any k >= 3,
                      // run by the walk kernels; keys must all be equal)
};

// Emit passes: where the kept cliques go.  The walk kernels receive the id
// buffer in place of the count vector and this block in place of the
// timers.  A clique (ids ascending) is kept iff the hash of its ids with
// `seed` is below thr53 / 2^53; `count` counts every kept clique, and only
// the first `cap` are written.
struct EmitCtl {
  unsigned long long count;
  unsigned long long cap;
  unsigned long long thr53;
  unsigned long long seed;
};

// The sampler's hash of a clique (ids in ascending order): splitmix64
// steps over the ids, the top 53 bits compared with thr53.
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct PassPlan {
  // Vertex tasks grouped by word class, heaviest first inside each class.
  std::vector<Task> tasks;
  uint32_t class_begin[kNumClasses + 1] = {};
  uint32_t class_count[kNumClasses] = {};
  // Edge tasks of the roots with d+ > kMaxRootDegree, grouped by the word
  // class of their member count (run by the (k-1) kernels).
  std::vector<Task> edge_tasks;
  uint32_t eclass_begin[kNumClasses + 1] = {};
  uint32_t eclass_count[kNumClasses] = {};
  // Per class, the index (within the class) from which tasks are light
  // enough to be claimed in chunks; class_count means none.
  uint32_t light_from[kNumClasses] = {};
};

struct DevicePlan {
  std::vector<Task> host_tasks;  // vertex tasks, then edge tasks
  DevBuf<Task> tasks;
  DevBuf<uint32_t> tickets;  // one queue head per class, vertex then edge
  uint32_t class_begin[kNumClasses + 1] = {};
  uint32_t class_count[kNumClasses] = {};
  uint32_t eclass_begin[kNumClasses + 1] = {};  // offsets into tasks
  uint32_t eclass_count[kNumClasses] = {};
  uint32_t light_from[kNumClasses] = {};
  uint32_t max_per_cu = 0;  // cap on resident workgroups per CU (0: by LDS)
  uint32_t light_waves = kLightWaves;  // waves per CU of the 1-2 word classes at k <= 5
};

struct DeviceProps {
  int device = 0;
  uint32_t cu_count = 0;
  uint32_t lds_per_cu = 0;
  std::string arch;
};
kcl_status device_props(DeviceProps *out);

// Enumeration pass over a plan: cnt[x] += credits.  key[x] orders members
// (smaller wins, ties to the smaller id).  All launches go to `stream`;
// the ms of the enumeration kernels is recorded between ev0 and ev1.
// prof (nullable): 2 u64 per task, shader-clock cycles until the rows are
// built and of the whole task (the tasks' order is the plan's).
kcl_status launch_pass(const kcl_graph *g, const uint64_t *d_key, uint32_t k, PassMode mode,
                       DevicePlan &plan, unsigned long long *d_cnt, hipStream_t stream,
                       unsigned long long *prof = nullptr);
// A pass over task and ticket buffers the caller owns: tasks in the layout
// of DevicePlan (vertex tasks by class, then edge tasks), tickets of
// kPassTicketWords u32 (the last one is the status word).  The begin/count
// arrays are host arrays of kNumClasses (+1 for begins) entries.
constexpr uint32_t kPassTicketWords = 2 * 2 * kNumClasses + 1;
struct PassLaunch {
  const Task *tasks = nullptr;
  uint32_t *tickets = nullptr;
  const uint32_t *class_begin = nullptr;
  const uint32_t *class_count = nullptr;
  const uint32_t *eclass_begin = nullptr;
  const uint32_t *eclass_count = nullptr;
  const uint32_t *light_from = nullptr;
  uint32_t max_per_cu = 0;
  uint32_t light_waves = kLightWaves;
  // Vertex tasks take their members from this device list (ascending ids,
  // at most the class's 64 W) instead of N+(s); nullptr: N+(s).
  const uint32_t *list = nullptr;
  uint32_t nlist = 0;
};
kcl_status launch_pass_raw(const kcl_graph *g, const uint64_t *d_key, uint32_t k, PassMode mode,
                           const PassLaunch &plan, unsigned long long *d_cnt, hipStream_t stream,
                           unsigned long long *prof = nullptr);
// Device address of the plan's status word (zeroed at the start of every
// pass, set by the kernels; kStatus* bits).
uint32_t *pass_status_word(DevicePlan &plan);
// Error for a status word read back after a pass (KCL_OK when 0).
kcl_status pass_status_error(uint32_t status);
// Reads the status word of the plan's last pass on `stream` (synchronous).
kcl_status pass_status(DevicePlan &plan, hipStream_t stream);

// Largest core bitmap side (ids below it get a dense adjacency bitmap).
constexpr uint32_t kCoreMax = 16384;

// Work estimate per root: counting work (steps x row words) and build work
// (adjacency entries scanned).  A split part counts at least kBuildShare
// times its root's build work.
struct RootCost {
  std::vector<uint64_t> work;
  std::vector<uint64_t> build;
};
constexpr uint64_t kBuildShare = 8;

// Tasks for every root with d+(s) >= k-1 (rc nullable: d^2 estimate).
// keep (nullable): only roots with keep[s] != 0 get tasks.  Roots above
// kMaxRootDegree become edge tasks (their sizes counted on the GPU).
kcl_status build_plan(const kcl_graph *g, uint32_t k, const RootCost *rc, uint32_t split_work,
                      PassPlan *plan, const std::vector<uint64_t> *keep = nullptr);
// Edge tasks for the roots in `big` (every out-edge (s, u) whose common
// out-neighbourhood can hold a (k-2)-clique), grouped by class.
kcl_status add_edge_tasks(const kcl_graph *g, uint32_t k, const std::vector<uint32_t> &big,
                          PassPlan *plan);
kcl_status upload_plan(const PassPlan &plan, DevicePlan *dev);
// Tasks cheaper than a resident workgroup's share of its class divided by
// kLightShare are claimed in chunks (light_cost gives that bound).
constexpr double kLightShare = 64.0;
double light_cost(double class_total, uint32_t wclass, uint32_t k);
// Moves the tasks of classes below min_class into min_class (wider rows:
// a class-c kernel handles any root with d+ <= 64 kClassWords[c]).
void promote_plan(PassPlan *plan, uint32_t min_class);

// Concurrent task slots of a word class's k kernel in one launch
// (workgroups).
uint32_t class_slots(uint32_t wclass, uint32_t k);
// Launch shape of a class kernel in a pass (as launch_pass_raw picks it):
// waves per workgroup, workgroups per CU, workgroups launched for `count`
// tasks, LDS bytes per workgroup.
kcl_status class_launch_shape(uint32_t wclass, uint32_t k, bool all, uint32_t max_per_cu,
                              uint32_t count, uint32_t *waves, uint32_t *per_cu,
                              uint32_t *blocks, uint32_t *lds);

// Plan for (graph, k): root filter (a root-credit k-pass), costs by the
// schedule, split tasks.  root_cliques (nullable) receives each root's
// k-clique count, info (nullable) a summary.  LEARNED without a model is
// DEGREE.
kcl_status make_plan(kcl_graph *g, uint32_t k, uint32_t split, hipStream_t st, PassPlan *out,
                     std::vector<uint64_t> *root_cliques, kcl_plan_info *info,
                     uint32_t schedule = KCL_SCHED_COUNTED,
                     const kcl_cost_model *model = nullptr,
                     const std::vector<uint64_t> *candidates = nullptr);
// Tasks for the roots with keep[s] != 0 under a schedule.
kcl_status plan_for_schedule(kcl_graph *g, uint32_t k, uint32_t schedule,
                             const kcl_cost_model *model, uint32_t split,
                             const std::vector<uint64_t> *keep, hipStream_t st, PassPlan *out);
// Every k-clique (k >= 3) inside the node set `in` (n flags) of g, k global
// ids each, listed on the GPU (sampling.hip).
kcl_status list_cliques_gpu(const kcl_graph *g, uint32_t k, const std::vector<uint8_t> &in,
                            std::vector<uint32_t> *out);
// Cost-model features of every node as a root (n rows of KCL_CM_FEATURES).
kcl_status root_features(kcl_graph *g, uint32_t k, hipStream_t st, std::vector<float> *out);
// Cost-model predictions for device-resident rows (log2(1 + cycles)).
kcl_status cost_model_predict_device(const kcl_cost_model *m, const float *d_x, uint64_t rows,
                                     float *d_y, hipStream_t st);
// Fills n u64 with zero on `st`.
kcl_status zero_u64(unsigned long long *p, uint64_t n, hipStream_t st);

kcl_status sim_step(kcl_sim *s, uint32_t t);
kcl_status sim_level_pass(kcl_sim *s, bool full_key, std::vector<unsigned long long> &tent,
                          std::vector<uint32_t> &reo, std::vector<double> &rho);

// Largest k with compiled enumeration kernels.
uint32_t max_pass_k();

// Densest prefix and upper bound of a report from host copies of tent (by
// level), reordered and rho (kCDensestCmpSync.c:354-381); fills out's n,
// density, ub and sum (not m) and best = n.
kcl_status densest_host(const unsigned long long *tent, const uint32_t *reo, const double *rho,
                        uint32_t n, uint32_t k, kcl_densest_info *out, uint32_t *best);

// A graph from a ready DAG CSR (rows of smaller ids, ascending): no
// relabelling, no core bitmap.  Takes the vectors.
kcl_status graph_from_csr(std::vector<uint32_t> &&cd, std::vector<uint32_t> &&adj,
                          kcl_graph **out);

// Graph creation with the degeneracy heap (rank = nullptr) or given ranks
// (a permutation of 0..n_in-1).
kcl_status graph_create(const uint32_t *pairs, uint64_t m, uint32_t n_in, const uint32_t *rank,
                        kcl_graph **out);

}  // namespace kcl

// Opaque ABI types.
struct kcl_graph;
struct kcl_graph {
  uint32_t n_in = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t core = 0;
  std::vector<uint32_t> rank;  // n_in
  std::vector<uint32_t> cd;    // n + 1 (host copy)
  std::vector<uint32_t> adj;   // m (host copy)
  kcl::DevBuf<uint32_t> d_cd;
  kcl::DevBuf<uint32_t> d_adj;
  // Dense adjacency bitmap of the nodes with id < core_h (the most core
  // part of the degeneracy order): row v, bit u set iff {u, v} is an edge.
  kcl::DevBuf<unsigned long long> d_core;
  uint32_t core_h = 0;
  // Roots with more out-neighbours than this run as edge tasks.
  uint32_t edge_above = kcl::kMaxRootDegree;
  // k >= 6 roots whose expected common out-set per edge item reaches this
  // are counted a wavefront per item.
  float dense_items = 64.0f;
  hipStream_t stream = nullptr;
  // Side streams of a pass's class launches: the pass forks from its
  // stream with `fork` and joins back with `join` (one host thread uses a
  // graph at a time, so one set serves every pass on it).
  static constexpr int kSideStreams = 2;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSideStreams] = {};
};
