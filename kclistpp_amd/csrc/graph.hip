// graph.hip - kcl_graph: degeneracy ranks (host heap) + device DAG CSR.
//
// Replaces SortByCore + Relabel + MakeGraph (Graph.c:57-156).  The ranks
// come from the host heap replay (their order is sequential by
// definition); relabelling, orientation (s = larger rank), the (s,t) sort
// and the CSR are built on the GPU:
//   key[i] = rank_hi << 32 | rank_lo         (relabel kernel)
//   radix sort of the 64-bit keys            (rocPRIM)
//   adj[i] = low word, cd from row boundaries (csr kernel)
// Rows come out ascending because the keys are sorted (Graph.c:125).
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "kcl_internal.h"

namespace kcl {

namespace {

__global__ void relabel_kernel(const uint32_t *__restrict__ pairs,
                               const uint32_t *__restrict__ rank, uint64_t m,
                               uint64_t *__restrict__ keys, uint32_t *__restrict__ max_s) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t local = 0;
  for (; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = rank[pairs[2 * i]], b = rank[pairs[2 * i + 1]];
    uint32_t s = a > b ? a : b, t = a > b ? b : a;
    keys[i] = ((uint64_t)s << 32) | t;
    local = s > local ? s : local;
  }
  // Wave maximum, one atomic per wave.
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(local, off);
    local = o > local ? o : local;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(max_s, local);
}

// cd[v] = index of the first edge with s >= v; each entry written once.
__global__ void csr_kernel(const uint64_t *__restrict__ keys, uint64_t m, uint32_t n,
                           uint32_t *__restrict__ cd, uint32_t *__restrict__ adj) {
  // Edge i owns the rows v in (s(i-1), s(i)], with s(-1) = -1 and s(m) = n.
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i <= m; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = (i == 0) ? 0u : (uint32_t)(keys[i - 1] >> 32) + 1u;
    uint32_t hi = (i == m) ? n : (uint32_t)(keys[i] >> 32);
    if (i < m) adj[i] = (uint32_t)keys[i];
    for (uint32_t v = lo; v <= hi; ++v) cd[v] = (uint32_t)i;
  }
}

__global__ void max_degree_kernel(const uint32_t *__restrict__ cd, uint32_t n,
                                  uint32_t *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t local = 0;
  for (; i < n; i += gridDim.x * blockDim.x) {
    uint32_t d = cd[i + 1] - cd[i];
    local = d > local ? d : local;
  }
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(local, off);
    local = o > local ? o : local;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, local);
}

__global__ void count_edges_kernel(const uint32_t *__restrict__ cd,
                                   const uint32_t *__restrict__ adj,
                                   const uint8_t *__restrict__ inc, uint32_t n,
                                   unsigned long long *__restrict__ out) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long local = 0;
  for (; s < n; s += gridDim.x * blockDim.x) {
    if (!inc[s]) continue;
    for (uint32_t j = cd[s]; j < cd[s + 1]; ++j) local += inc[adj[j]];
  }
  for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(out, local);
}

// Dense bitmap of the edges among nodes < h: one thread per row s < h sets
// (s, t) and (t, s) for its out-neighbours t (t < s).
__global__ void core_bits_kernel(const uint32_t *__restrict__ cd, const uint32_t *__restrict__ adj,
                                 uint32_t rows, uint32_t width,
                                 unsigned long long *__restrict__ bits) {
  const uint32_t stride = width >> 6;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < rows;
       s += gridDim.x * blockDim.x) {
    for (uint32_t j = cd[s]; j < cd[s + 1]; ++j) {
      const uint32_t t = adj[j];
      atomicOr(&bits[(size_t)s * stride + (t >> 6)], 1ull << (t & 63));
      atomicOr(&bits[(size_t)t * stride + (s >> 6)], 1ull << (s & 63));
    }
  }
}

uint32_t grid_for(uint64_t work, uint32_t block) {
  uint64_t g = (work + block - 1) / block;
  if (g > 65536) g = 65536;
  return (uint32_t)(g ? g : 1);
}

}  // namespace

kcl_status device_props(DeviceProps *out) {
  // Queried once per device (launches ask for it, also during stream capture).
  static std::mutex mu;
  static std::vector<DeviceProps> cache;
  int dev = 0;
  KCL_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  for (const DeviceProps &c : cache)
    if (c.device == dev) {
      *out = c;
      return KCL_OK;
    }
  hipDeviceProp_t p;
  KCL_HIP(hipGetDeviceProperties(&p, dev));
  DeviceProps d;
  d.device = dev;
  d.cu_count = (uint32_t)p.multiProcessorCount;
  d.lds_per_cu = (uint32_t)p.maxSharedMemoryPerMultiProcessor;
  d.arch = p.gcnArchName;
  cache.push_back(d);
  *out = d;
  return KCL_OK;
}

static kcl_status build_device_csr(kcl_graph *g, const uint32_t *pairs) {
  const uint64_t m = g->m;
  hipStream_t st = g->stream;
  DevBuf<uint32_t> d_pairs, d_rank, d_max;
  DevBuf<uint64_t> d_keys, d_sorted;
  KCL_TRY(d_pairs.alloc(2 * m));
  KCL_TRY(d_rank.alloc(g->n_in));
  KCL_TRY(d_max.alloc(1));
  KCL_TRY(d_keys.alloc(m));
  KCL_TRY(d_sorted.alloc(m));
  KCL_HIP(hipMemcpyAsync(d_pairs.p, pairs, 2 * m * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  KCL_HIP(hipMemcpyAsync(d_rank.p, g->rank.data(), g->n_in * sizeof(uint32_t),
                         hipMemcpyHostToDevice, st));
  KCL_HIP(hipMemsetAsync(d_max.p, 0, sizeof(uint32_t), st));
  relabel_kernel<<<grid_for(m, 256), 256, 0, st>>>(d_pairs.p, d_rank.p, m, d_keys.p, d_max.p);
  KCL_HIP(hipGetLastError());
  d_pairs.release();
  d_rank.release();

  uint32_t max_s = 0;
  KCL_HIP(hipMemcpyAsync(&max_s, d_max.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  g->n = max_s + 1;

  unsigned end_bit = 32;
  while (end_bit < 64 && (1ull << (end_bit - 32)) <= max_s) ++end_bit;
  size_t tmp_bytes = 0;
  KCL_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, d_keys.p, d_sorted.p, (size_t)m, 0,
                                   end_bit, st));
  DevBuf<uint8_t> tmp;
  KCL_TRY(tmp.alloc(tmp_bytes));
  KCL_HIP(rocprim::radix_sort_keys(tmp.p, tmp_bytes, d_keys.p, d_sorted.p, (size_t)m, 0,
                                   end_bit, st));
  tmp.release();
  d_keys.release();

  KCL_TRY(g->d_cd.alloc((size_t)g->n + 1));
  KCL_TRY(g->d_adj.alloc(m));
  csr_kernel<<<grid_for(m + 1, 256), 256, 0, st>>>(d_sorted.p, m, g->n, g->d_cd.p, g->d_adj.p);
  KCL_HIP(hipGetLastError());
  KCL_HIP(hipMemsetAsync(d_max.p, 0, sizeof(uint32_t), st));
  max_degree_kernel<<<grid_for(g->n, 256), 256, 0, st>>>(g->d_cd.p, g->n, d_max.p);
  KCL_HIP(hipGetLastError());
  uint32_t core = 0;
  KCL_HIP(hipMemcpyAsync(&core, d_max.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  g->cd.resize((size_t)g->n + 1);
  g->adj.resize(m);
  KCL_HIP(hipMemcpyAsync(g->cd.data(), g->d_cd.p, ((size_t)g->n + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, st));
  KCL_HIP(hipMemcpyAsync(g->adj.data(), g->d_adj.p, m * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         st));
  KCL_HIP(hipStreamSynchronize(st));
  g->core = core;
  // Core bitmap: nodes [0, h) with h = min(kCoreMax, n rounded up to 64).
  const uint32_t h = std::min<uint32_t>(kCoreMax, (g->n + 63) & ~63u);
  if (h > 0) {
    KCL_TRY(g->d_core.alloc((size_t)h * (h / 64)));
    KCL_HIP(hipMemsetAsync(g->d_core.p, 0, (size_t)h * (h / 64) * 8, st));
    const uint32_t rows = std::min(h, g->n);
    core_bits_kernel<<<grid_for(rows, 256), 256, 0, st>>>(g->d_cd.p, g->d_adj.p, rows, h,
                                                          g->d_core.p);
    KCL_HIP(hipGetLastError());
    KCL_HIP(hipStreamSynchronize(st));
    g->core_h = h;
  }
  return KCL_OK;
}

}  // namespace kcl

using kcl::fail;

namespace {

// The graph's stream, its side streams and the pass fork/join events (all
// default priority, non-blocking).
hipError_t create_streams(kcl_graph *g) {
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  for (int i = 0; e == hipSuccess && i < kcl_graph::kSideStreams; ++i) {
    e = hipStreamCreateWithFlags(&g->side[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
  return e;
}

}  // namespace

extern "C" kcl_status kcl_graph_create(const uint32_t *pairs, uint64_t m, uint32_t n_in,
                                       kcl_graph **out) {
  return kcl::graph_create(pairs, m, n_in, nullptr, out);
}

kcl_status kcl::graph_create(const uint32_t *pairs, uint64_t m, uint32_t n_in,
                             const uint32_t *given_rank, kcl_graph **out) {
  if (!out) return fail(KCL_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (m && !pairs) return fail(KCL_E_INVALID_ARG, "pairs is NULL");
  if (m >= 0xFFFFFFFFull) return fail(KCL_E_UNSUPPORTED, "m = %llu: the limit is 2^32-2",
                                      (unsigned long long)m);
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
    if (a == b) return fail(KCL_E_INVALID_ARG, "self-loop at edge %llu (node %u)",
                            (unsigned long long)i, a);
    mx = std::max(mx, std::max(a, b));
  }
  if (mx >= 0xFFFFFFFEu) return fail(KCL_E_UNSUPPORTED, "node id %u: ids must be < 2^32-2", mx);
  if (n_in == 0) n_in = mx + 1;
  if (m && n_in <= mx) return fail(KCL_E_INVALID_ARG, "n_in = %u but an edge names node %u",
                                   n_in, mx);
  kcl_graph *g = new (std::nothrow) kcl_graph();
  if (!g) return fail(KCL_E_OOM_HOST, "out of host memory");
  g->n_in = n_in;
  g->m = m;
  try {
    g->rank.resize(n_in);
  } catch (...) {
    delete g;
    return fail(KCL_E_OOM_HOST, "out of host memory for %u ranks", n_in);
  }
  if (given_rank)
    memcpy(g->rank.data(), given_rank, (size_t)n_in * sizeof(uint32_t));
  else
    kcl::degeneracy_rank(n_in, m, pairs, g->rank.data());
  hipError_t e = create_streams(g);
  if (e != hipSuccess) {
    kcl_graph_destroy(g);
    return fail(KCL_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  if (m == 0) {
    g->n = 0;
    g->core = 0;
    g->cd.assign(1, 0);
    kcl_status s = g->d_cd.alloc(1);
    if (s == KCL_OK) s = g->d_adj.alloc(1);
    if (s == KCL_OK) {
      e = hipMemsetAsync(g->d_cd.p, 0, sizeof(uint32_t), g->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
      if (e != hipSuccess) s = fail(KCL_E_HIP, "hipMemset: %s", hipGetErrorString(e));
    }
    if (s != KCL_OK) {
      kcl_graph_destroy(g);
      return s;
    }
    *out = g;
    return KCL_OK;
  }
  kcl_status s = kcl::build_device_csr(g, pairs);
  if (s != KCL_OK) {
    kcl_graph_destroy(g);
    return s;
  }
  *out = g;
  return KCL_OK;
}

kcl_status kcl::graph_from_csr(std::vector<uint32_t> &&cd, std::vector<uint32_t> &&adj,
                               kcl_graph **out) {
  *out = nullptr;
  if (cd.empty() || cd.back() != adj.size())
    return fail(KCL_E_INVALID_ARG, "CSR offsets do not match the adjacency");
  kcl_graph *g = new (std::nothrow) kcl_graph();
  if (!g) return fail(KCL_E_OOM_HOST, "out of host memory");
  g->n = (uint32_t)(cd.size() - 1);
  g->n_in = g->n;
  g->m = adj.size();
  for (uint32_t v = 0; v < g->n; ++v) g->core = std::max(g->core, cd[v + 1] - cd[v]);
  g->cd = std::move(cd);
  g->adj = std::move(adj);
  kcl_status st = KCL_OK;
  hipError_t e = create_streams(g);
  if (e != hipSuccess) st = fail(KCL_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  if (st == KCL_OK) st = g->d_cd.alloc(g->n + 1);
  if (st == KCL_OK) st = g->d_adj.alloc(g->m);
  if (st == KCL_OK) {
    e = hipMemcpy(g->d_cd.p, g->cd.data(), (g->n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && g->m)
      e = hipMemcpy(g->d_adj.p, g->adj.data(), g->m * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) st = fail(KCL_E_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  if (st != KCL_OK) {
    kcl_graph_destroy(g);
    return st;
  }
  *out = g;
  return KCL_OK;
}

extern "C" void kcl_graph_destroy(kcl_graph *g) {
  if (!g) return;
  g->d_cd.release();
  g->d_adj.release();
  if (g->stream) (void)hipStreamDestroy(g->stream);
  for (int i = 0; i < kcl_graph::kSideStreams; ++i) {
    if (g->side[i]) (void)hipStreamDestroy(g->side[i]);
    if (g->join[i]) (void)hipEventDestroy(g->join[i]);
  }
  if (g->fork) (void)hipEventDestroy(g->fork);
  delete g;
}

extern "C" kcl_status kcl_graph_set_edge_split(kcl_graph *g, uint32_t max_vertex_degree) {
  if (!g) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (max_vertex_degree < 1 || max_vertex_degree > kcl::kMaxRootDegree)
    return fail(KCL_E_INVALID_ARG, "max_vertex_degree = %u: 1 <= value <= %u", max_vertex_degree,
                kcl::kMaxRootDegree);
  g->edge_above = max_vertex_degree;
  return KCL_OK;
}

extern "C" kcl_status kcl_graph_set_dense_threshold(kcl_graph *g, float min_items) {
  if (!g) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (!(min_items >= 0.0f) || min_items > 1e9f)
    return fail(KCL_E_INVALID_ARG, "min_items = %g: 0 <= value <= 1e9", (double)min_items);
  g->dense_items = min_items;
  return KCL_OK;
}

extern "C" kcl_status kcl_graph_get_info(const kcl_graph *g, kcl_graph_info *info) {
  if (!g || !info) return fail(KCL_E_INVALID_ARG, "NULL argument");
  info->n_in = g->n_in;
  info->n = g->n;
  info->m = g->m;
  info->core = g->core;
  return KCL_OK;
}

extern "C" kcl_status kcl_graph_rank(const kcl_graph *g, uint32_t *rank) {
  if (!g || !rank) return fail(KCL_E_INVALID_ARG, "NULL argument");
  memcpy(rank, g->rank.data(), (size_t)g->n_in * sizeof(uint32_t));
  return KCL_OK;
}

extern "C" kcl_status kcl_graph_csr(const kcl_graph *g, uint32_t *cd, uint32_t *adj) {
  if (!g || !cd || (!adj && g->m)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  memcpy(cd, g->cd.data(), ((size_t)g->n + 1) * sizeof(uint32_t));
  if (g->m) memcpy(adj, g->adj.data(), g->m * sizeof(uint32_t));
  return KCL_OK;
}

extern "C" kcl_status kcl_graph_edges(const kcl_graph *g, uint32_t *st) {
  if (!g || (!st && g->m)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  for (uint32_t s = 0; s < g->n; ++s)
    for (uint32_t j = g->cd[s]; j < g->cd[s + 1]; ++j) {
      st[2 * (uint64_t)j] = s;
      st[2 * (uint64_t)j + 1] = g->adj[j];
    }
  return KCL_OK;
}

namespace kcl {
kcl_status count_edges_device(const kcl_graph *g, const uint8_t *d_inc, uint32_t *m_out,
                              hipStream_t st) {
  DevBuf<unsigned long long> d_out;
  KCL_TRY(d_out.alloc(1));
  KCL_HIP(hipMemsetAsync(d_out.p, 0, sizeof(unsigned long long), st));
  if (g->n) {
    count_edges_kernel<<<grid_for(g->n, 256), 256, 0, st>>>(g->d_cd.p, g->d_adj.p, d_inc, g->n,
                                                            d_out.p);
    KCL_HIP(hipGetLastError());
  }
  unsigned long long v = 0;
  KCL_HIP(hipMemcpyAsync(&v, d_out.p, sizeof(v), hipMemcpyDeviceToHost, st));
  KCL_HIP(hipStreamSynchronize(st));
  *m_out = (uint32_t)v;  // the reference's counter is unsigned (Graph.c:159)
  return KCL_OK;
}
}  // namespace kcl

extern "C" kcl_status kcl_count_edges(const kcl_graph *g, uint32_t num_nodes,
                                      const uint32_t *nodes, uint32_t *m_out) {
  if (!g || !m_out || (num_nodes && !nodes)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  std::vector<uint8_t> inc((size_t)g->n + 1, 0);
  for (uint32_t i = 0; i < num_nodes; ++i) {
    if (nodes[i] >= g->n) return fail(KCL_E_INVALID_ARG, "node %u >= n = %u", nodes[i], g->n);
    inc[nodes[i]] = 1;
  }
  kcl::DevBuf<uint8_t> d_inc;
  KCL_TRY(d_inc.alloc(inc.size()));
  KCL_HIP(hipMemcpyAsync(d_inc.p, inc.data(), inc.size(), hipMemcpyHostToDevice, g->stream));
  return kcl::count_edges_device(g, d_inc.p, m_out, g->stream);
}
