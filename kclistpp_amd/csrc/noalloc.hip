// noalloc.hip - kcl_clique_enum_noalloc: one enumeration pass with host
// vectors, like kcl_clique_enum, that allocates nothing.
//
// kcl_clique_enum builds a full plan per call (a counting pass, cost
// estimates, device buffers for the task list, keys and counts).  This
// variant keeps every intermediate in one file-scope device array of
// KCL_NOALLOC_SCRATCH_BYTES and builds a simpler plan on the device: one
// task per root with d+ >= k-1, heavy roots split by the d+^2 rule, tasks
// in arbitrary order inside their word class.  Counts do not depend on the
// plan, so the results equal kcl_clique_enum's.
//
// Scratch layout (byte offsets rounded up to 256):
//   head   tickets + status (kPassTicketWords u32), class counts, class
//          begins, class fill counters, stats (u64 visits, u64 overflow)
//   key    u64[n]  the pass key: rho bits (FW) or n-1-level (PAVA); the
//                  caller's rho is copied here first
//   cnt    u64[n]  the pass's credits (u32 level counts before the pass)
//   aux    u64[n]  the caller's level (u32) before the pass, its acc after
//   tasks  Task[]  vertex tasks by word class
#include <mutex>

#include "kcl_internal.h"

static __device__ __align__(256) unsigned char scratch[KCL_NOALLOC_SCRATCH_BYTES];

namespace kcl {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (looked up once per
// device; host code never passes the symbol itself to a kernel).
kcl_status scratch_address(unsigned char **out) {
  int dev = 0;
  KCL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(KCL_E_HIP, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(KCL_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch[dev] = (unsigned char *)p;
  }
  *out = g_scratch[dev];
  return KCL_OK;
}

constexpr size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// Head words (u32 offsets into the head block).
constexpr uint32_t kHeadTickets = 0;                              // kPassTicketWords
constexpr uint32_t kHeadCount = 64;                               // kNumClasses
constexpr uint32_t kHeadBegin = kHeadCount + 16;                  // kNumClasses + 1
constexpr uint32_t kHeadFill = kHeadBegin + 16;                   // kNumClasses
constexpr uint32_t kHeadFlags = kHeadFill + 16;                   // 1
constexpr size_t kHeadStatsByte = 512;                            // u64[2]
constexpr size_t kHeadBytes = 1024;
static_assert(kPassTicketWords <= kHeadCount, "tickets fit the head");

// Argument errors found on the device (flag bits).
constexpr uint32_t kBadRho = 1u;
constexpr uint32_t kBadLevel = 2u;
constexpr uint32_t kNotPermutation = 4u;

// Split rule of the plain schedule (cost d+^2, parts of at least 2^17
// units or 8 d+): the parts of root s.
__device__ __forceinline__ uint32_t root_parts(uint32_t d, uint32_t k) {
  const uint64_t work = (uint64_t)d * d;
  const uint64_t part = max((uint64_t)1 << 17, 8ull * d);
  uint64_t np = (work + part - 1) / part;
  const uint64_t items = (k >= 5) ? (uint64_t)d * (d - 1) / 2 : d;
  if (np > items) np = items;
  if (np > 4096) np = 4096;
  return (uint32_t)(np ? np : 1);
}

__device__ __forceinline__ uint32_t word_class(uint32_t d) {
  const uint32_t words = (d + 63) / 64;
  uint32_t c = 0;
  while (c + 1 < kNumClasses && kClassWords[c] < words) ++c;
  return c;
}

__global__ void fw_key_kernel(uint32_t n, unsigned long long *__restrict__ key,
                              uint32_t *__restrict__ flags) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double r = __longlong_as_double((long long)key[i]);
    if (!(r >= 0.0)) atomicOr(flags, kBadRho);
    key[i] = r == 0.0 ? 0ull : key[i];  // -0.0 orders as +0.0
  }
}

// key = n-1-level; seen[level] counts how often each level occurs.
__global__ void pava_key_kernel(uint32_t n, const uint32_t *__restrict__ level,
                                unsigned long long *__restrict__ key, uint32_t *__restrict__ seen,
                                uint32_t *__restrict__ flags) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t l = level[i];
    if (l >= n) {
      atomicOr(flags, kBadLevel);
      key[i] = 0;
      continue;
    }
    key[i] = (unsigned long long)(n - 1 - l);
    atomicAdd(&seen[l], 1u);
  }
}

__global__ void permutation_kernel(uint32_t n, const uint32_t *__restrict__ seen,
                                   uint32_t *__restrict__ flags) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (seen[i] != 1u) atomicOr(flags, kNotPermutation);
}

// Tasks per word class (roots with d+ >= k-1, split parts included).
__global__ void count_tasks_kernel(const uint32_t *__restrict__ cd, uint32_t n, uint32_t k,
                                   uint32_t *__restrict__ count) {
  __shared__ uint32_t local[kNumClasses];
  if (threadIdx.x < kNumClasses) local[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t d = cd[s + 1] - cd[s];
    if (d == 0 || d < k - 1) continue;
    atomicAdd(&local[word_class(d)], root_parts(d, k));
  }
  __syncthreads();
  if (threadIdx.x < kNumClasses && local[threadIdx.x])
    atomicAdd(&count[threadIdx.x], local[threadIdx.x]);
}

// Writes each root's tasks into its class's segment (begin[c] + a slot
// reserved from fill[c]).
__global__ void scatter_tasks_kernel(const uint32_t *__restrict__ cd, uint32_t n, uint32_t k,
                                     const uint32_t *__restrict__ begin,
                                     uint32_t *__restrict__ fill, Task *__restrict__ tasks) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t d = cd[s + 1] - cd[s];
    if (d == 0 || d < k - 1) continue;
    const uint32_t c = word_class(d), np = root_parts(d, k);
    const uint32_t at = begin[c] + atomicAdd(&fill[c], np);
    for (uint32_t p = 0; p < np; ++p) tasks[at + p] = Task{s, p | (np << 16)};
  }
}

__global__ void zero_words_kernel(uint32_t *p, uint32_t count) {
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
}

struct Begins {
  uint32_t v[kNumClasses + 1];
};

__global__ void set_begins_kernel(uint32_t *dst, Begins b) {
  if (threadIdx.x <= kNumClasses) dst[threadIdx.x] = b.v[threadIdx.x];
}

// acc (the caller's, copied in) += counts: FW acc[i], PAVA acc[level[i]]
// with level = n-1-key (a permutation, so every slot is written once).
// stats[0] += visits, stats[1] |= any count >= 2^53.
__global__ void accumulate_kernel(uint32_t n, uint32_t pava, const unsigned long long *__restrict__ cnt,
                                  const unsigned long long *__restrict__ key,
                                  double *__restrict__ acc, unsigned long long *__restrict__ stats) {
  unsigned long long sum = 0, big = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long c = cnt[i];
    sum += c;
    big |= (c >= (1ull << 53)) ? 1ull : 0ull;
    const uint32_t at = pava ? (uint32_t)(n - 1 - key[i]) : i;
    acc[at] += (double)c;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off);
    big |= __shfl_xor(big, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (sum) atomicAdd(&stats[0], sum);
    if (big) atomicOr(&stats[1], big);
  }
}

uint32_t grid_of(uint64_t work) {
  uint64_t g = (work + 255) / 256;
  if (g > 8192) g = 8192;
  return (uint32_t)(g ? g : 1);
}

}  // namespace

}  // namespace kcl

using kcl::fail;

extern "C" kcl_status kcl_clique_enum_noalloc(kcl_graph *g, uint32_t k, kcl_task task,
                                              const double *rho, const uint32_t *level,
                                              double *acc, uint64_t *visits) {
  using namespace kcl;
  if (!g || !acc) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (k < 2 || k > 255) return fail(KCL_E_INVALID_ARG, "k = %u: the reference takes 2 <= k <= 255", k);
  if (task != KCL_FRANK_WOLFE && task != KCL_PAVA_PREPROCESS)
    return fail(KCL_E_INVALID_ARG, "task %d: expected 2 (FRANK_WOLFE) or 3 (PAVA_PREPROCESS)",
                (int)task);
  if (task == KCL_FRANK_WOLFE && !rho) return fail(KCL_E_INVALID_ARG, "rho is NULL");
  if (task == KCL_PAVA_PREPROCESS && !level) return fail(KCL_E_INVALID_ARG, "level is NULL");
  const uint32_t n = g->n;
  if (visits) *visits = 0;
  if (n == 0) return KCL_OK;
  const bool pava = task == KCL_PAVA_PREPROCESS;
  const size_t vec = round256((size_t)n * 8);
  if (kHeadBytes + 3 * vec > (size_t)KCL_NOALLOC_SCRATCH_BYTES)
    return fail(KCL_E_UNSUPPORTED,
                "graph of %u nodes: the allocation-free pass holds 24 bytes per node plus its "
                "tasks in %u bytes of scratch (at most about %u nodes)",
                n, (unsigned)KCL_NOALLOC_SCRATCH_BYTES,
                (unsigned)(((size_t)KCL_NOALLOC_SCRATCH_BYTES - kHeadBytes) / 32));
  if (g->core > g->edge_above)
    return fail(KCL_E_UNSUPPORTED,
                "a root has %u DAG out-neighbours (> %u): edge tasks need plan buffers; use "
                "kcl_clique_enum",
                g->core, g->edge_above);
  const bool trivial = k > g->core + 1;  // no k-clique: acc stays, visits = 0
  if (!trivial && k > max_pass_k())
    return fail(KCL_E_UNSUPPORTED, "k = %u: the enumerator is compiled for k <= %u (graph core %u)",
                k, max_pass_k(), g->core);

  unsigned char *base = nullptr;
  KCL_TRY(scratch_address(&base));
  uint32_t *head = (uint32_t *)base;
  uint32_t *tickets = head + kHeadTickets;
  uint32_t *d_count = head + kHeadCount;
  uint32_t *d_begin = head + kHeadBegin;
  uint32_t *d_fill = head + kHeadFill;
  uint32_t *d_flags = head + kHeadFlags;
  unsigned long long *d_stats = (unsigned long long *)(base + kHeadStatsByte);
  unsigned long long *d_key = (unsigned long long *)(base + kHeadBytes);
  unsigned long long *d_cnt = (unsigned long long *)(base + kHeadBytes + vec);
  unsigned long long *d_aux = (unsigned long long *)(base + kHeadBytes + 2 * vec);
  Task *d_tasks = (Task *)(base + kHeadBytes + 3 * vec);
  const size_t task_room = ((size_t)KCL_NOALLOC_SCRATCH_BYTES - kHeadBytes - 3 * vec) / sizeof(Task);
  hipStream_t st = g->stream;

  // Every call shares this device's scratch: clear what is read before it
  // is written (kernels, not memset nodes).
  zero_words_kernel<<<1, 256, 0, st>>>(head, (uint32_t)(kHeadStatsByte / 4) + 4);
  KCL_HIP(hipGetLastError());

  // Keys (and, for PAVA, the permutation check).
  if (!pava) {
    KCL_HIP(hipMemcpyAsync(d_key, rho, (size_t)n * 8, hipMemcpyHostToDevice, st));
    fw_key_kernel<<<grid_of(n), 256, 0, st>>>(n, d_key, d_flags);
  } else {
    KCL_TRY(zero_u64(d_cnt, ((size_t)n + 1) / 2, st));  // seen counts, u32[n]
    KCL_HIP(hipMemcpyAsync(d_aux, level, (size_t)n * 4, hipMemcpyHostToDevice, st));
    pava_key_kernel<<<grid_of(n), 256, 0, st>>>(n, (const uint32_t *)d_aux, d_key,
                                                (uint32_t *)d_cnt, d_flags);
    KCL_HIP(hipGetLastError());
    permutation_kernel<<<grid_of(n), 256, 0, st>>>(n, (const uint32_t *)d_cnt, d_flags);
  }
  KCL_HIP(hipGetLastError());

  // Task counts per word class.
  uint32_t counts[kNumClasses] = {}, flags = 0;
  if (!trivial) {
    count_tasks_kernel<<<grid_of(n), 256, 0, st>>>(g->d_cd.p, n, k, d_count);
    KCL_HIP(hipGetLastError());
    KCL_HIP(hipMemcpyAsync(counts, d_count, sizeof(counts), hipMemcpyDeviceToHost, st));
  }
  KCL_HIP(hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  if (flags & kBadRho) return fail(KCL_E_INVALID_ARG, "rho has a negative or NaN entry");
  if (flags & kBadLevel) return fail(KCL_E_INVALID_ARG, "level has an entry >= n = %u", n);
  if (flags & kNotPermutation)
    return fail(KCL_E_INVALID_ARG,
                "level is not a permutation of 0..n-1 (the allocation-free pass scatters "
                "rho_tentative in parallel; kcl_clique_enum accepts repeated levels)");

  if (!trivial) {
    uint32_t begin[kNumClasses + 1] = {}, light[kNumClasses] = {};
    uint64_t total = 0;
    for (uint32_t c = 0; c < kNumClasses; ++c) {
      begin[c] = (uint32_t)total;
      total += counts[c];
      // The one-word class (light roots) is claimed in chunks, the others
      // one task at a time.
      light[c] = c == 0 ? 0u : counts[c];
    }
    begin[kNumClasses] = (uint32_t)total;
    if (total > task_room)
      return fail(KCL_E_UNSUPPORTED,
                  "%llu tasks do not fit the allocation-free scratch (room for %zu with %u nodes)",
                  (unsigned long long)total, task_room, n);
    // begin[] goes to the device as a kernel argument, not a copy.
    Begins bv;
    for (uint32_t c = 0; c <= kNumClasses; ++c) bv.v[c] = begin[c];
    set_begins_kernel<<<1, 64, 0, st>>>(d_begin, bv);
    KCL_HIP(hipGetLastError());
    scatter_tasks_kernel<<<grid_of(n), 256, 0, st>>>(g->d_cd.p, n, k, d_begin, d_fill, d_tasks);
    KCL_HIP(hipGetLastError());
    KCL_TRY(zero_u64(d_cnt, n, st));
    const uint32_t zero[kNumClasses + 1] = {};
    PassLaunch pl;
    pl.tasks = d_tasks;
    pl.tickets = tickets;
    pl.class_begin = begin;
    pl.class_count = counts;
    pl.eclass_begin = zero;
    pl.eclass_count = zero;
    pl.light_from = light;
    KCL_TRY(launch_pass_raw(g, (const uint64_t *)d_key, k, kCreditWinner, pl, d_cnt, st));
  } else {
    KCL_TRY(zero_u64(d_cnt, n, st));
  }

  // acc += counts on the device (the caller's acc copied in and back).
  KCL_HIP(hipMemcpyAsync(d_aux, acc, (size_t)n * 8, hipMemcpyHostToDevice, st));
  accumulate_kernel<<<grid_of(n), 256, 0, st>>>(n, pava ? 1u : 0u, d_cnt, d_key,
                                                (double *)d_aux, d_stats);
  KCL_HIP(hipGetLastError());
  unsigned long long stats[2] = {0, 0};
  uint32_t status = 0;
  KCL_HIP(hipMemcpyAsync(stats, d_stats, sizeof(stats), hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(&status, tickets + kPassTicketWords - 1, 4, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  KCL_TRY(pass_status_error(status));
  if (stats[1]) return fail(KCL_E_OVERFLOW, "a node received >= 2^53 cliques");
  KCL_HIP(hipMemcpyAsync(acc, d_aux, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  if (visits) *visits = stats[0];
  return KCL_OK;
}
