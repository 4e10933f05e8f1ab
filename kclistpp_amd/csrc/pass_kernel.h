// pass_kernel.h - the k-clique enumeration pass on gfx950 (device code).
//
// Replaces CDF_CliqueEnum + MakeSubgraph + CDF_CliqueEnumThread +
// CDF_FrankWolfeUpdateRates (kCDensestCmpSync.c:71-117, 138-325).
//
// Formulation.  Every k-clique Q has a unique largest id s (its root);
// Q \ {s} is a (k-1)-clique inside L = N+(s), the root's out-neighbours.
// The reference credits Q to its member of smallest (key, id), where key
// is rho (Frank-Wolfe) or the reversed level (PAVA).  Orient the graph
// induced on L by that order (u -> v when (key_u, id_u) < (key_v, id_v));
// then the winner among Q \ {s} is its source u, and
//     winner(Q) = u if (key_u, id_u) < (key_s, s) else s.
// Since every member of L has a smaller id than s, that test is
// key_u <= key_s.  So the pass needs, for every u in L, the number of
// (k-2)-cliques inside A'[u] (u's out-set in the oriented graph), and
// credits it to u or to s.  The same count with all credit sent to s
// gives per-root clique counts (the cost pre-pass and trimming).
//
// Kernel shape.  One wavefront per root task, persistent, taking tasks
// from a global ticket (tasks are sorted heaviest first).  L (<= 64*W
// nodes) is staged in LDS with the members' keys; the oriented rows A'
// are W 64-bit words each, built from the members' adjacency slices
// (coalesced reads, binary search in L, ds_or into the row of the
// earlier endpoint).  Counting is a per-lane state machine over items:
//   k = 2: item u, one clique {s, u}
//   k = 3: item u, popcount(A'[u])
//   k = 4: item u,       P = A'[u],          steps w in P: popcount(P & A'[w])
//   k >= 5: item (u, v),  P = A'[u] & A'[v], steps w in P: cliques of size
//           k-4 inside P & A'[w] (nested loops, compile-time depth)
// At k <= 5, lanes that finish their items draw new ones together once
// enough of them wait (steps are short, so rounds of steps then keep most
// lanes busy); at k >= 6 a lane draws at once.  Credits gather in LDS (u64 per
// member) and are flushed with one global atomicAdd per non-zero member,
// plus one wave-reduced atomicAdd for the root.
//
// Heavy roots are split into parts, each rebuilding A' (the rebuild is cheap
// next to the counting it splits).  Items are dealt out in blocks of
// kPartBlock(NW) consecutive items (member items: fewer, so that min(np, 8)
// parts get some): part p of np takes blocks p, p + np,
// p + 2 np, ...  Dealt-out blocks give every part the same mix of items (in
// ranked kernels an item's cost grows steeply with its members' ranks, so
// contiguous ranges leave one part with most of a root's work), and a block
// of neighbours keeps every wavefront of the part busy on the heavy items
// at its start.
//
// Dense roots (k >= 6).  The k >= 6 kernels of rows of 4+ words number
// the members by their (key, id) rank before building A', so A'[x] holds only ranks above x
// (its words below x's own are zero) and the winners are the ranks below
// nwin.  A root whose expected common out-set per item is large is counted
// a wavefront per item: the sets of the outer levels are the same in
// every lane (scalar registers), and in the last two levels lane l takes
// member 64j + l of each word j and counts its row over words >= j.  The
// level above those walks its members word by word with the word index a
// compile-time constant: a member's row is zero below its own word, so the
// next set is formed from that word on (row words moved to scalar
// registers, one scalar AND each) and its edges counted from there.  Rows
// of 8 or more words (roots inside a dense core) are always counted so.
//
// Large k (K = 0 instantiation: the clique size is a launch argument, for
// k above the compiled ones).  Ranked rows as above; a wavefront per item u
// walks the cliques of A'[u] depth first with the current set spread over
// the lanes (lane j holds word j), choosing each level's next member with a
// ballot.  Only the chosen members are kept (a u16 path per level in the
// LDS the keys used before ranking); a level's remaining set is recomputed
// from the path when the walk returns to it.  The last two levels are the
// edges of the final set, shared by the lanes as in the dense counting.
// Values one lane holds reach the others through readlane, never through
// LDS: the compiler orders a lane's LDS accesses only against its own.
// Every lane writes the same path entry, so each reads back its own write.
//
// The walk also lists cliques (mode kEmit, any k >= 3, all keys equal so
// that ranks follow ids): every clique found is hashed and, when the
// sampler keeps it, written out with its ids in ascending order.
#pragma once

#include "kcl_internal.h"

namespace kcl {

namespace {

// Words between consecutive rows in LDS: odd for the even classes below 16
// words, so that 64 lanes reading one word of consecutive rows hit distinct
// banks (an even stride of 2^a * odd words repeats banks every 16 / 2^a
// lanes).  The 16-word class stays unpadded (it would not fit).
template <int W>
constexpr int row_stride() {
  return (W % 2 == 0 && W < 16) ? W + 1 : W;
}

template <int W>
struct Lds {
  static constexpr int D = 64 * W;
  unsigned long long rows[D * row_stride<W>()];  // oriented rows A'
  uint64_t key[D];                 // member keys
  unsigned long long cred[D];   // credits gathered per member
  uint32_t node[D];             // member ids, ascending
  uint32_t beg[D];              // cd[member]
  uint32_t pre[D + 1];          // prefix sums (member degrees, later |A'[u]|)
  unsigned long long root_sum;  // the root's credits, summed over the waves
  uint32_t task;                // ticket broadcast
  uint32_t next_item;           // counting items handed out so far
  uint32_t wcount[16];          // edge tasks: members found per wavefront
  uint32_t ecount;              // edge tasks: members found so far
  uint32_t chunk;               // 0 while claiming single tasks, else the chunk size
  uint32_t nwin;                // ranked kernels: members with key <= the root's key
};

__device__ __forceinline__ uint32_t select_bit(unsigned long long x, uint32_t r) {
  // Position of the r-th (0-based) set bit of x; r < popcount(x).
  uint32_t pos = 0, c;
  c = __popcll(x & 0xffffffffull);
  if (r >= c) { r -= c; x >>= 32; pos += 32; }
  c = __popc((uint32_t)x & 0xffffu);
  if (r >= c) { r -= c; x >>= 16; pos += 16; }
  c = __popc((uint32_t)x & 0xffu);
  if (r >= c) { r -= c; x >>= 8; pos += 8; }
  c = __popc((uint32_t)x & 0xfu);
  if (r >= c) { r -= c; x >>= 4; pos += 4; }
  c = __popc((uint32_t)x & 0x3u);
  if (r >= c) { r -= c; x >>= 2; pos += 2; }
  c = (uint32_t)x & 1u;
  if (r >= c) pos += 1;
  return pos;
}

template <int W>
__device__ __forceinline__ uint32_t popc_set(const unsigned long long (&p)[W]) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) c += __popcll(p[i]);
  return c;
}

// Number of J-cliques inside P (P is a set of members; cliques follow the
// orientation, so each is counted once, from its source).
template <int W, int J>
__device__ __forceinline__ uint64_t cliques_in(const unsigned long long (&P)[W],
                                               const unsigned long long *rows) {
  if constexpr (J == 1) {
    return popc_set<W>(P);
  } else {
    uint64_t total = 0;
#pragma unroll
    for (int wi = 0; wi < W; ++wi) {
      uint64_t bits = P[wi];
      while (bits) {
        const uint32_t b = (uint32_t)__builtin_ctzll(bits);
        bits &= bits - 1;
        const unsigned long long *r = rows + (uint32_t)(wi * 64 + b) * row_stride<W>();
        unsigned long long Q[W];
#pragma unroll
        for (int j = 0; j < W; ++j) Q[j] = P[j] & r[j];
        if constexpr (J >= 3) {
          if (popc_set<W>(Q) < (uint32_t)(J - 1)) continue;
        }
        total += cliques_in<W, J - 1>(Q, rows);
      }
    }
    return total;
  }
}

// Number of J-cliques inside P, crediting each clique to every one of its
// members in P (LDS atomics; the caller credits the members outside P).
template <int W, int J>
__device__ __forceinline__ uint64_t cliques_all(const unsigned long long (&P)[W],
                                                const unsigned long long *rows,
                                                unsigned long long *cred) {
  uint64_t total = 0;
#pragma unroll
  for (int wi = 0; wi < W; ++wi) {
    uint64_t bits = P[wi];
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctzll(bits);
      bits &= bits - 1;
      const uint32_t x = (uint32_t)(wi * 64) + b;
      if constexpr (J == 1) {
        atomicAdd(&cred[x], 1ull);
        ++total;
      } else {
        const unsigned long long *r = rows + x * row_stride<W>();
        unsigned long long Q[W];
#pragma unroll
        for (int j = 0; j < W; ++j) Q[j] = P[j] & r[j];
        if (popc_set<W>(Q) < (uint32_t)(J - 1)) continue;
        const uint64_t c = cliques_all<W, J - 1>(Q, rows, cred);
        if (c) atomicAdd(&cred[x], (unsigned long long)c);
        total += c;
      }
    }
  }
  return total;
}

// Dense counting (ranked kernels).  Members are numbered by (key, id)
// rank, so the row of member x holds only members ranked after x and, for
// x in word j, its words below j are zero.  A wavefront works on one set S
// at a time, the same in every lane (kept in scalar registers), and the
// lanes share the last level.

template <int W>
__device__ __forceinline__ void to_sgpr(unsigned long long (&S)[W]) {
#pragma unroll
  for (int j = 0; j < W; ++j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)S[j]);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(S[j] >> 32));
    S[j] = ((unsigned long long)hi << 32) | lo;
  }
}

// v if this lane's bit of mask is set, else 0: the set word itself is the
// lane mask of v_cndmask (lane l holds member 64 j + l of word j).
// (One v_cndmask with the scalar mask operand; the shift-and-compare form
// takes four instructions.  mask must be wave-uniform: a scalar register
// pair.)
__device__ __forceinline__ uint32_t keep_if_member(uint32_t v, unsigned long long mask) {
  uint32_t r;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(mask));
  return r;
}

// A row word that is the same in every lane, moved to scalar registers.
__device__ __forceinline__ unsigned long long uniform_word(unsigned long long x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// This lane's share of the edges inside S, from word J0 on (the words of S
// below J0 are zero): for each non-zero word j of S, lane l counts the row
// of member 64j + l inside S (words j and above only) and keeps the count
// if that member is in S.  Every lane loads (rows past the root's members
// hold stale bits and are masked out the same way), so the loads of
// successive words need no branch between them.
template <int W, int J0>
__device__ __forceinline__ uint32_t wave_edges_from(const unsigned long long (&S)[W],
                                                    const unsigned long long *rows,
                                                    uint32_t lane) {
  uint32_t c = 0;
#pragma unroll
  for (int j = J0; j < W; ++j) {
    if (S[j] == 0) continue;
    const unsigned long long *r = rows + (uint32_t)(64 * j + lane) * row_stride<W>();
    uint32_t cj = 0;
#pragma unroll
    for (int jj = j; jj < W; ++jj) cj += (uint32_t)__popcll(S[jj] & r[jj]);
    c += keep_if_member(cj, S[j]);
  }
  return c;
}

// wave_edges_from<W, 0> without the scalar-mask select, for callers whose
// set the compiler may keep in vector registers (the walk kernels).
template <int W>
__device__ __forceinline__ uint32_t wave_edges_plain(const unsigned long long (&S)[W],
                                                     const unsigned long long *rows,
                                                     uint32_t lane) {
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (S[j] == 0) continue;
    const unsigned long long *r = rows + (uint32_t)(64 * j + lane) * row_stride<W>();
    uint32_t cj = 0;
#pragma unroll
    for (int jj = j; jj < W; ++jj) cj += (uint32_t)__popcll(S[jj] & r[jj]);
    c += ((S[j] >> lane) & 1ull) ? cj : 0u;
  }
  return c;
}

// This lane's share of the triangles inside S, for the members w of S in
// words WI and above: the set of w's later neighbours, S & A'[w], has no
// member below w's own word, so it is formed (and moved to scalar
// registers) from word WI on, and its edges are counted from there.
template <int W, int WI>
__device__ __forceinline__ uint32_t wave_triangles_from(const unsigned long long (&S)[W],
                                                        const unsigned long long *rows,
                                                        uint32_t lane) {
  if constexpr (WI >= W) {
    return 0;
  } else {
    uint32_t total = 0;
    unsigned long long bits = S[WI];
    while (bits) {
      const uint32_t w = (uint32_t)WI * 64 + (uint32_t)__builtin_ctzll(bits);
      bits &= bits - 1;
      const unsigned long long *r = rows + w * row_stride<W>();
      unsigned long long Q[W];
      uint32_t size = 0;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (j < WI) {
          Q[j] = 0;
        } else {
          // The row word moves to scalar registers first, so the
          // intersection is one scalar AND.
          Q[j] = S[j] & uniform_word(r[j]);
          size += (uint32_t)__popcll(Q[j]);
        }
      }
      if (size < 2) continue;
      total += wave_edges_from<W, WI>(Q, rows, lane);
    }
    return total + wave_triangles_from<W, WI + 1>(S, rows, lane);
  }
}

// This lane's share of the J-cliques inside S (J >= 2): the members of S
// are walked in the same order by every lane, the last two levels go to
// wave_edges.
template <int W, int J>
__device__ __forceinline__ uint64_t wave_cliques(const unsigned long long (&S)[W],
                                                 const unsigned long long *rows, uint32_t lane) {
  if constexpr (J == 2) {
    return wave_edges_from<W, 0>(S, rows, lane);
  } else if constexpr (J == 3) {
    return wave_triangles_from<W, 0>(S, rows, lane);
  } else {
    uint64_t total = 0;
#pragma unroll 1
    for (int wi = 0; wi < W; ++wi) {
      unsigned long long bits = 0;
#pragma unroll
      for (int j = 0; j < W; ++j) bits = (j == wi) ? S[j] : bits;
      while (bits) {
        const uint32_t w = (uint32_t)wi * 64 + (uint32_t)__builtin_ctzll(bits);
        bits &= bits - 1;
        const unsigned long long *r = rows + w * row_stride<W>();
        // Words below w's own are zero in its row (ranked members).
        unsigned long long Q[W];
#pragma unroll
        for (int j = 0; j < W; ++j) Q[j] = (j >= wi) ? (S[j] & uniform_word(r[j])) : 0ull;
        if (popc_set<W>(Q) < (uint32_t)(J - 1)) continue;
        total += wave_cliques<W, J - 1>(Q, rows, lane);
      }
    }
    return total;
  }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Lane `src`'s 64-bit value in every lane (src wave-uniform).
__device__ __forceinline__ unsigned long long read_lane64(unsigned long long x, uint32_t src) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)src);
  return ((unsigned long long)hi << 32) | lo;
}

// Word j of the set of members ranked after p.
__device__ __forceinline__ unsigned long long after_mask(uint32_t p, uint32_t j) {
  const uint32_t pw = p >> 6, pb = p & 63;
  if (j > pw) return ~0ull;
  if (j < pw || pb == 63) return 0ull;
  return ~0ull << (pb + 1);
}

// Large-k layout of a wavefront's share of the key array (used once the
// members are ranked), in u64 words: (k-clique degrees only) one count per
// level, then one u16 member per level.
__host__ __device__ constexpr uint32_t deep_words(uint32_t levels, bool all) {
  return (all ? levels : 0u) + (levels + 3u) / 4u;
}

// Emit passes: the hash of the ids ahead of the last two (the walk's
// prefix, ascending) and of the whole clique.
__device__ __forceinline__ unsigned long long emit_hash(unsigned long long h, uint32_t id) {
  return mix64(h ^ id);
}

// Emit passes: the slot of this lane's clique (keep = the sampler kept
// it), one atomic per wavefront; ~0 when it is not to be written (not kept,
// or past the buffer's capacity).  Called by every lane of the wavefront.
__device__ __forceinline__ unsigned long long emit_slot(bool keep, uint32_t lane,
                                                        EmitCtl *__restrict__ ctl,
                                                        unsigned long long cap) {
  const unsigned long long m = __ballot(keep);
  if (m == 0ull) return ~0ull;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(&ctl->count, (unsigned long long)__popcll(m));
  base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), (int)leader) << 32) |
         (uint32_t)__shfl((int)(uint32_t)base, (int)leader);
  if (!keep) return ~0ull;
  const unsigned long long slot = base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
  return slot < cap ? slot : ~0ull;
}

// In-place inclusive scan of a[0..d) by one wavefront.
__device__ __forceinline__ void wave_scan(uint32_t *a, uint32_t d, uint32_t lane) {
  const uint32_t chunk = (d + 63) >> 6;
  const uint32_t lo = min(d, lane * chunk), hi = min(d, lo + chunk);
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += a[i];
  uint32_t inc = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t v = __shfl_up(inc, off);
    if (lane >= (uint32_t)off) inc += v;
  }
  uint32_t run = inc - sum;
  for (uint32_t i = lo; i < hi; ++i) {
    run += a[i];
    a[i] = run;
  }
}

// Items per dealt-out block of a split root: 64 per wavefront.  Learned
// passes (ms) on 1M/16M k = 6, 4M/60M k = 6, R-MAT scales 15 and 16 k = 7
// at one state, by block size: 1 (items dealt one by one) 16.8, 3621, 2271,
// 6830; 4 NW 15.7, 3614, 1828, 6057; 16 NW 15.3, 3608, 1640, 5659;
// 64 NW 15.2, 3610, 1437, 5562; 256 NW 16.2, 3605, 1667, 6428; one
// contiguous range per part 15.7, 3644, 1708, 6506.
constexpr uint32_t kPartBlock(uint32_t nw) { return 64u * nw; }

// ALL: credit every member of every clique (k-clique degrees) instead of
// the winner or the root.
template <int W, int K, int NW, bool ALL>
__global__ __launch_bounds__(64 * NW) void pass_kernel(const uint32_t *__restrict__ cd,
                                                       const uint32_t *__restrict__ adj,
                                                       const uint64_t *__restrict__ key,
                                                       const unsigned long long *__restrict__ core,
                                                       uint32_t core_h,
                                                       const Task *__restrict__ tasks,
                                                       uint32_t ntasks, uint32_t light_from,
                                                       float dense_items,
                                                       uint32_t *__restrict__ ticket,
                                                       uint32_t *__restrict__ status,
                                                       const uint32_t *__restrict__ list,
                                                       uint32_t nlist,
                                                       uint32_t mode,
                                                       unsigned long long *__restrict__ cnt,
                                                       unsigned long long *__restrict__ prof) {
  static_assert(K == 0 || K >= 2, "k >= 2, or 0: k from the launch (large k)");
  static_assert(NW >= 1 && NW <= 16, "at most 16 wavefronts per workgroup");
  // Large-k kernels take the clique size in the bits above the mode.
  uint32_t kdeep = 0;
  unsigned long long *timers = prof;
  if constexpr (K == 0) {
    kdeep = mode >> 8;
    mode &= 0xffu;
    if (mode == kEmit) timers = nullptr;  // prof carries the emit block
  }
  constexpr uint32_t NT = 64 * NW;  // threads of the workgroup sharing one root
  constexpr uint32_t RS = (uint32_t)row_stride<W>();
  __shared__ Lds<W> sh;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const bool to_root = !ALL && (mode == kCreditRoot);
  // Ranked kernels (k >= 6, rows of 4+ words) number the members by
  // (key, id) rank before building the rows; their dense roots are counted
  // a wavefront per item.  Narrower rows (d <= 192) hardly ever reach the
  // dense threshold, so they skip the ranking.
  constexpr bool kRanked = (K >= 6 && W >= 4) || K == 0;
  constexpr bool kDense = kRanked && !ALL && K != 0;
  // Rows of 8+ words belong to roots inside a dense core (d+ > 448): those
  // kernels count every task densely and leave out the per-lane path, whose
  // registers would halve their occupancy.
  constexpr bool kDenseOnly = kDense && W >= 8;

  // Tickets: tasks [0, light_from) one at a time from ticket[0] (heavy
  // first), then the light tail [light_from, ntasks) in chunks of
  // kLightChunk from ticket[1] (one same-address atomic per chunk).
  constexpr uint32_t kLightChunk = 8;
  uint32_t chunk_base = 0, chunk_left = 0, chunk_size = 1;
  if (tid == 0) sh.chunk = 0;
  __syncthreads();
  for (;;) {
    if (chunk_left == 0) {
      if (tid == 0) {
        uint32_t t = light_from, c = 0;
        if (sh.chunk == 0) {  // still in the heavy part
          t = atomicAdd(&ticket[0], 1u);
          c = 1;
        }
        if (t >= light_from) {
          t = light_from + atomicAdd(&ticket[1], kLightChunk);
          c = kLightChunk;
        }
        sh.task = t;
        sh.chunk = c == 1 ? 0u : c;
      }
      __syncthreads();
      chunk_base = sh.task;
      chunk_size = sh.chunk == 0 ? 1u : sh.chunk;
      chunk_left = chunk_size;
    }
    const uint32_t ti = chunk_base + (chunk_size - chunk_left);
    --chunk_left;
    if (ti >= ntasks) break;  // same value in every thread of the workgroup
    // Optional per-task timing (shader clock): timers[2 ti] = cycles until
    // the rows are built, timers[2 ti + 1] = cycles of the whole task.
    const unsigned long long t_start = timers ? __builtin_amdgcn_s_memtime() : 0ull;
    const Task tk = tasks[ti];
    const uint32_t s = tk.s;
    // A vertex task (root s, part of np) or an edge task (root edge (s, u),
    // members N+(s) & N+(u), run by the (k-1) kernel: its cliques are the
    // k-cliques whose two largest ids are s and u).  In an edge task the
    // clique's winner among {s, u} is the one of smaller (key, id); members
    // have smaller ids than both, so "member key <= that key" decides.
    const bool edge_task = task_is_edge(tk.part_np);
    const uint32_t part = edge_task ? 0u : (tk.part_np & 0xffffu);
    const uint32_t np = edge_task ? 1u : (tk.part_np >> 16);
    // The j-th item of this part (increasing in j; np = 1: item j).  Edge
    // items (k >= 5) are dealt out in blocks of kPartBlock(NW); parts past
    // the last block rebuild the rows and stop.  Member items (k <= 4 and the
    // walk kernels) come in blocks of 2^sh, the largest power of two up to
    // kPartBlock(NW) that still gives min(np, 8) parts items (part_shift;
    // every part computes it from the same n): a root of at most 64 W members
    // would otherwise sit in one block.  Shrinking the edge-item blocks the
    // same way made the 1M/16M k = 5 pass 13 % slower, and shrinking them
    // until every part had items cost 5 % on R-MAT scale 15 at k = 7.
    constexpr uint32_t kBlk = kPartBlock(NW);
    auto part_edge_item = [&](uint32_t j) {
      return np == 1 ? j : ((j / kBlk) * np + part) * kBlk + j % kBlk;
    };
    auto part_item = [&](uint32_t j, uint32_t sh2) {
      return np == 1 ? j : ((((j >> sh2) * np + part) << sh2) | (j & ((1u << sh2) - 1u)));
    };
    auto part_shift = [&](uint32_t n) {
      const uint32_t useful = min(np, 8u);
      const uint32_t b = max(1u, min(kPartBlock(NW), (n + useful - 1) / useful));
      return 31u - (uint32_t)__builtin_clz(b);
    };
    uint32_t d;
    uint64_t ks;
    uint32_t eu = s;  // edge tasks: the second root node u
    if (!edge_task) {
      // Members: N+(s), or the launch's explicit list (ascending ids, at
      // most 64 W: the k-clique-core peel's alive neighbours of s).
      const uint32_t b0 = list ? 0u : cd[s];
      const uint32_t *mlist = list ? list : adj + b0;
      d = list ? nlist : cd[s + 1] - b0;
      ks = key[s];
      // Stage the members, their keys and adjacency extents.
      for (uint32_t i = tid; i < d; i += NT) {
        const uint32_t x = mlist[i];
        sh.node[i] = x;
        sh.key[i] = key[x];
        const uint32_t xb = cd[x];
        sh.beg[i] = xb;
        sh.pre[i + 1] = cd[x + 1] - xb;
        sh.cred[i] = 0;
      }
    } else {
      eu = task_edge_node(tk.part_np);
      const uint64_t k_s = key[s], k_u = key[eu];
      ks = (k_u <= k_s) ? k_u : k_s;
      // Members: N+(s) & N+(u) in ascending order.  Each thread takes
      // entries of the shorter list and searches the longer one; the hits
      // are compacted in order (per-wavefront ballots, then wavefront
      // offsets).  The plan sized the class by this intersection.
      const uint32_t bs = cd[s], ls = cd[s + 1] - bs, bu = cd[eu], lu = cd[eu + 1] - bu;
      const uint32_t *A = adj + (ls <= lu ? bs : bu);
      const uint32_t *B = adj + (ls <= lu ? bu : bs);
      const uint32_t la = ls <= lu ? ls : lu, lb = ls <= lu ? lu : ls;
      const uint32_t wave = tid >> 6;
      if (tid == 0) sh.ecount = 0;
      __syncthreads();
      for (uint32_t base = 0; base < la; base += NT) {
        const uint32_t i = base + tid;
        bool hit = false;
        uint32_t x = 0;
        if (i < la) {
          x = A[i];
          uint32_t lo = 0, hi = lb;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (B[mid] < x) lo = mid + 1;
            else hi = mid;
          }
          hit = lo < lb && B[lo] == x;
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) sh.wcount[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t at = sh.ecount;
        for (uint32_t w = 0; w < wave; ++w) at += sh.wcount[w];
        at += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (hit && at < (uint32_t)Lds<W>::D) sh.node[at] = x;
        __syncthreads();
        if (tid == 0) {
          uint32_t tot = 0;
          for (uint32_t w = 0; w < NW; ++w) tot += sh.wcount[w];
          sh.ecount += tot;
        }
        __syncthreads();
      }
      // The plan put this task in a class that holds its common members
      // (edge_members_kernel); a larger intersection means the plan and the
      // graph disagree.  Count what fits, and flag the pass so the host
      // turns it into an error instead of returning wrong counts.
      if (tid == 0 && sh.ecount > (uint32_t)Lds<W>::D) atomicOr(status, kStatusEdgeOverflow);
      d = min(sh.ecount, (uint32_t)Lds<W>::D);
      for (uint32_t i = tid; i < d; i += NT) {
        const uint32_t x = sh.node[i];
        sh.key[i] = key[x];
        const uint32_t xb = cd[x];
        sh.beg[i] = xb;
        sh.pre[i + 1] = cd[x + 1] - xb;
        sh.cred[i] = 0;
      }
    }
    if (tid == 0) {
      sh.pre[0] = 0;
      sh.root_sum = 0;
      sh.nwin = 0;
    }
    if constexpr (K >= 3 || K == 0) {
      for (uint32_t i = tid; i < d * RS; i += NT) sh.rows[i] = 0;
    }
    __syncthreads();

    unsigned long long root_acc = 0;
    if constexpr (K == 2) {
      // Cliques {s, u}: items are the members themselves.
      const uint32_t bsh = part_shift(d);
      for (uint32_t j = tid;; j += NT) {
        const uint32_t u = part_item(j, bsh);
        if (u >= d) break;
        if constexpr (ALL) {
          sh.cred[u] += 1;
          root_acc += 1;
        } else {
          if (!to_root && sh.key[u] <= ks) sh.cred[u] += 1;
          else root_acc += 1;
        }
      }
    } else {
      // Oriented rows.  Members are ascending, so those inside the core
      // bitmap (id < core_h) are a prefix of h members; the edges among them
      // come from h(h-1)/2 bit tests when that is cheaper than scanning their
      // adjacency slices.  The other members' slices are scanned: an edge
      // (x, y) with y < x has x outside the core unless both ends are inside.
      if (tid < 64) wave_scan(sh.pre + 1, d, lane);
      if constexpr (kRanked) {
        // rank(i) = #{j : (key_j, j) < (key_i, i)} (index order is id order),
        // kept in cred[i] until the rows are built.  Members with
        // key <= key_s are the first nwin ranks.
        uint32_t wins = 0;
        for (uint32_t i = tid; i < d; i += NT) {
          const uint64_t ki = sh.key[i];
          uint32_t r = 0;
          for (uint32_t j = 0; j < d; ++j) {
            const uint64_t kj = sh.key[j];
            r += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
          }
          sh.cred[i] = r;
          wins += ki <= ks ? 1u : 0u;
        }
        if (wins) atomicAdd(&sh.nwin, wins);
      }
      __syncthreads();
      const uint32_t total = sh.pre[d];
      uint32_t h = 0;
      if (core) {
        uint32_t a = 0, b = d;
        while (a < b) {
          const uint32_t mid = (a + b) >> 1;
          if (sh.node[mid] < core_h) a = mid + 1;
          else b = mid;
        }
        h = a;
        if ((uint64_t)h * (h - 1) / 2 >= (uint64_t)sh.pre[h]) h = 0;  // slices are cheaper
      }
      // kBuildBatch loads per thread are issued before any is used, so their
      // latencies overlap instead of adding up.
      constexpr uint32_t kBuildBatch = 8;
      auto set_edge = [&](uint32_t a, uint32_t i) {  // members a < i are adjacent
        uint32_t from, to;
        if constexpr (kRanked) {  // bit positions are ranks
          const uint32_t ra = (uint32_t)sh.cred[a], ri = (uint32_t)sh.cred[i];
          from = min(ra, ri);
          to = max(ra, ri);
        } else {
          const bool a_first = sh.key[a] <= sh.key[i];  // ties: smaller id (a) first
          from = a_first ? a : i;
          to = a_first ? i : a;
        }
        atomicOr(&sh.rows[from * RS + (to >> 6)], 1ull << (to & 63));
      };
      if (h > 1) {
        const uint32_t pairs = h * (h - 1) / 2;
        const uint32_t stride = core_h >> 6;  // words per bitmap row
        for (uint32_t p0 = tid; p0 < pairs; p0 += NT * kBuildBatch) {
          unsigned long long wv[kBuildBatch];
          uint32_t iv[kBuildBatch], jv[kBuildBatch];
#pragma unroll
          for (uint32_t b = 0; b < kBuildBatch; ++b) {
            const uint32_t pi = p0 + b * NT;
            iv[b] = UINT32_MAX;
            if (pi < pairs) {
              // pair index pi = i(i-1)/2 + j, 0 <= j < i
              uint32_t i = (uint32_t)((1.0f + sqrtf(1.0f + 8.0f * (float)pi)) * 0.5f);
              while ((uint64_t)i * (i - 1) / 2 > pi) --i;
              while ((uint64_t)(i + 1) * i / 2 <= pi) ++i;
              const uint32_t j = pi - i * (i - 1) / 2;
              iv[b] = i;
              jv[b] = j;
              const uint32_t y = sh.node[j];
              wv[b] = core[(size_t)sh.node[i] * stride + (y >> 6)];
            }
          }
#pragma unroll
          for (uint32_t b = 0; b < kBuildBatch; ++b) {
            if (iv[b] == UINT32_MAX) continue;
            if ((wv[b] >> (sh.node[jv[b]] & 63)) & 1ull) set_edge(jv[b], iv[b]);
          }
        }
      }
      uint32_t cur = h;
      for (uint32_t e0 = sh.pre[h] + tid; e0 < total; e0 += NT * kBuildBatch) {
        uint32_t yv[kBuildBatch], iv[kBuildBatch];
#pragma unroll
        for (uint32_t b = 0; b < kBuildBatch; ++b) {
          const uint32_t e = e0 + b * NT;
          iv[b] = UINT32_MAX;
          if (e < total) {
            uint32_t lo = cur, hi = d;  // member owning element e: pre[i] <= e < pre[i+1]
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (sh.pre[mid] <= e) lo = mid;
              else hi = mid;
            }
            cur = lo;
            iv[b] = lo;
            yv[b] = adj[sh.beg[lo] + (e - sh.pre[lo])];
          }
        }
#pragma unroll
        for (uint32_t b = 0; b < kBuildBatch; ++b) {
          const uint32_t i = iv[b];
          if (i == UINT32_MAX) continue;
          const uint32_t y = yv[b];
          uint32_t a = 0, c = i;  // y among members [0, i) (all smaller ids)
          while (a < c) {
            const uint32_t mid = (a + c) >> 1;
            if (sh.node[mid] < y) a = mid + 1;
            else c = mid;
          }
          if (a < i && sh.node[a] == y) set_edge(a, i);
        }
      }
      if constexpr (kRanked) {
        __syncthreads();
        // From here on members are ranks: beg[r] = the node of rank r.
        for (uint32_t i = tid; i < d; i += NT) {
          const uint32_t r = (uint32_t)sh.cred[i];
          sh.beg[r] = sh.node[i];
          sh.cred[i] = 0;
        }
      }
      __syncthreads();
      if (timers && tid == 0) timers[2 * (size_t)ti] = __builtin_amdgcn_s_memtime() - t_start;

      if constexpr (K == 0) {
        // Large k, and the emit passes: a wavefront per item u (ranks of
        // this part), the (kdeep-2)-cliques of A'[u] walked depth first.
        // c = members still to choose inside the current set; at c = 3 the
        // chosen member's set is final and its edges are counted (or
        // listed) by the lanes.  Items of kdeep <= 4 (emit passes only)
        // stop at the item's own set.
        if (tid == 0) sh.next_item = 0;  // counts this part's items (part_item)
        const uint32_t bsh = part_shift(d);
        __syncthreads();
        const uint32_t wid = tid >> 6;
        const uint32_t levels = kdeep > 3 ? kdeep - 3 : 1u;  // path / count slots, 1-based
        unsigned long long *ws =
            reinterpret_cast<unsigned long long *>(sh.key) + wid * (Lds<W>::D / NW);
        unsigned long long *lacc = ws;  // [levels] (k-clique degrees; lane 0 only)
        uint16_t *path = reinterpret_cast<uint16_t *>(ws + (ALL ? levels : 0u));  // [levels]
        const bool word_lane = lane < (uint32_t)W;
        const uint32_t nwin = sh.nwin;
        // Emit passes: the id buffer and the sampler (cnt and prof carry them).
        const bool emit = !ALL && mode == kEmit;
        uint32_t *eout = reinterpret_cast<uint32_t *>(cnt);
        EmitCtl *ectl = reinterpret_cast<EmitCtl *>(prof);
        unsigned long long ecap = 0, ethr = 0, eseed = 0;
        uint32_t ksz = kdeep + (edge_task ? 1u : 0u);  // ids per listed clique
        if (emit) {
          ecap = ectl->cap;
          ethr = ectl->thr53;
          eseed = ectl->seed;
        }
        // The hash of a clique ends with the root pair's smaller node (edge
        // tasks) and the root.
        auto tail_hash = [&](unsigned long long h) {
          if (edge_task) h = emit_hash(h, eu);
          return emit_hash(h, s);
        };
        // Emit: write this lane's clique (prefix ranks u, path[1..np-1],
        // then the ranks x, y when given) if kept.
        auto write_clique = [&](unsigned long long slot, uint32_t u0, uint32_t npath, uint32_t wl,
                                bool has_w, uint32_t x, bool has_x, uint32_t y, bool has_y) {
          if (slot == ~0ull) return;
          uint32_t *o = eout + slot * ksz;
          uint32_t at = 0;
          o[at++] = sh.beg[u0];
          for (uint32_t i = 1; i < npath; ++i) o[at++] = sh.beg[path[i]];
          if (has_w) o[at++] = sh.beg[wl];
          if (has_x) o[at++] = sh.beg[x];
          if (has_y) o[at++] = sh.beg[y];
          if (edge_task) o[at++] = eu;
          o[at] = s;
        };
        // The edges of the final set F (scalar registers), whose cliques
        // also hold the prefix hashed in hp: counted, or listed.
        auto final_edges = [&](const unsigned long long (&F)[W], unsigned long long hp,
                               uint32_t u0, uint32_t npath, uint32_t wl, bool has_w) -> uint32_t {
          if (!emit) return wave_edges_plain<W>(F, sh.rows, lane);
#pragma unroll
          for (int j = 0; j < W; ++j) {
            if (F[j] == 0ull) continue;
            const bool mine = (F[j] >> lane) & 1ull;
            const uint32_t y = 64 * (uint32_t)j + lane;
            const unsigned long long hy = emit_hash(hp, mine ? sh.beg[y] : 0u);
#pragma unroll
            for (int jj = j; jj < W; ++jj) {
              unsigned long long e = mine ? (F[jj] & sh.rows[y * RS + jj]) : 0ull;
              while (__ballot(e != 0ull)) {
                const bool has = e != 0ull;
                const uint32_t z = 64 * (uint32_t)jj + (has ? (uint32_t)__builtin_ctzll(e) : 0u);
                if (has) e &= e - 1ull;
                const unsigned long long h = has ? tail_hash(emit_hash(hy, sh.beg[z])) : 0ull;
                const bool keep = has && (h >> 11) < ethr;
                write_clique(emit_slot(keep, lane, ectl, ecap), u0, npath, wl, has_w, y, true, z,
                             true);
              }
            }
          }
          return 0u;
        };
        for (;;) {
          uint32_t it = 0;
          if (lane == 0) it = atomicAdd(&sh.next_item, 1u);
          it = (uint32_t)__builtin_amdgcn_readfirstlane((int)it);
          const uint32_t u = part_item(it, bsh);
          if (u >= d) break;
          unsigned long long R = word_lane ? sh.rows[u * RS + lane] : 0ull;
          uint32_t l = 1, c = kdeep - 2;
          unsigned long long item_total = 0;
          if constexpr (ALL) {
            if (lane == 0) lacc[1] = 0;
          }
          if constexpr (!ALL) {
            if (c <= 2) {
              // The item's own set is final (kdeep <= 4, emit passes).
              const unsigned long long hu = emit ? emit_hash(eseed, sh.beg[u]) : 0ull;
              if (c == 0) {
                if (emit) {
                  const bool keep = lane == 0 && (tail_hash(hu) >> 11) < ethr;
                  write_clique(emit_slot(keep, lane, ectl, ecap), u, 1, 0, false, 0, false, 0,
                               false);
                } else {
                  item_total = 1;
                }
              } else {
                unsigned long long F[W];
#pragma unroll
                for (int j = 0; j < W; ++j) F[j] = read_lane64(R, (uint32_t)j);
                if (c == 1) {
                  if (emit) {
#pragma unroll
                    for (int j = 0; j < W; ++j) {
                      if (F[j] == 0ull) continue;
                      const bool mine = (F[j] >> lane) & 1ull;
                      const uint32_t z = 64 * (uint32_t)j + lane;
                      const bool keep =
                          mine && (tail_hash(emit_hash(hu, sh.beg[z])) >> 11) < ethr;
                      write_clique(emit_slot(keep, lane, ectl, ecap), u, 1, 0, false, z, true, 0,
                                   false);
                    }
                  } else {
                    item_total = popc_set<W>(F);
                  }
                } else {
                  item_total = wave_sum((unsigned long long)final_edges(F, hu, u, 1, 0, false));
                }
              }
              if (lane == 0 && item_total) {
                if (!to_root && u < nwin) atomicAdd(&sh.cred[u], item_total);
                else root_acc += item_total;
              }
              continue;
            }
          }
          for (;;) {
            const unsigned long long nz = __ballot(R != 0ull);
            if (nz == 0ull) {
              if (l == 1) break;
              // Level l is done: back to level l-1, whose remaining set is
              // A'[u] & A'[path[1]] & ... & A'[path[l-2]], after path[l-1].
              --l;
              ++c;
              if constexpr (ALL) {
                if (lane == 0) {
                  const unsigned long long v = lacc[l + 1];
                  if (v) {
                    atomicAdd(&sh.cred[path[l]], v);
                    lacc[l] += v;
                  }
                }
              }
              unsigned long long S = word_lane ? sh.rows[u * RS + lane] : 0ull;
              for (uint32_t i = 1; i < l; ++i)
                S &= word_lane ? sh.rows[(uint32_t)path[i] * RS + lane] : 0ull;
              R = S & after_mask(path[l], lane);
              continue;
            }
            const uint32_t j0 = (uint32_t)__builtin_ctzll(nz);
            const unsigned long long word = read_lane64(R, j0);
            const uint32_t w = 64 * j0 + (uint32_t)__builtin_ctzll(word);
            if (lane == j0) R &= R - 1ull;
            const unsigned long long N = word_lane ? (R & sh.rows[w * RS + lane]) : 0ull;
            const uint32_t cn = c - 1;  // members still to choose inside N
            if (wave_sum32((uint32_t)__popcll(N)) < cn) continue;
            if (cn == 2) {
              // The edges of N, its words moved to scalar registers: lane l
              // takes member 64 j + l of each word j.
              unsigned long long F[W];
#pragma unroll
              for (int j = 0; j < W; ++j) F[j] = read_lane64(N, (uint32_t)j);
              uint32_t ce = 0;
              if constexpr (ALL) {
#pragma unroll
                for (int j = 0; j < W; ++j) {
                  if (F[j] == 0ull || !((F[j] >> lane) & 1ull)) continue;
                  const unsigned long long *ry = sh.rows + (64 * (uint32_t)j + lane) * RS;
                  uint32_t cy = 0;
#pragma unroll
                  for (int jj = j; jj < W; ++jj) {
                    unsigned long long e = F[jj] & ry[jj];
                    cy += (uint32_t)__popcll(e);
                    while (e) {
                      atomicAdd(&sh.cred[64 * (uint32_t)jj + (uint32_t)__builtin_ctzll(e)], 1ull);
                      e &= e - 1ull;
                    }
                  }
                  if (cy) atomicAdd(&sh.cred[64 * (uint32_t)j + lane], (unsigned long long)cy);
                  ce += cy;
                }
              } else {
                unsigned long long hp = 0;
                if (emit) {
                  hp = emit_hash(eseed, sh.beg[u]);
                  for (uint32_t i = 1; i < l; ++i) hp = emit_hash(hp, sh.beg[path[i]]);
                  hp = emit_hash(hp, sh.beg[w]);
                }
                ce = final_edges(F, hp, u, l, w, true);
              }
              const unsigned long long cb = wave_sum((unsigned long long)ce);
              if constexpr (ALL) {
                if (lane == 0 && cb) {
                  atomicAdd(&sh.cred[w], cb);
                  lacc[l] += cb;
                }
              } else {
                item_total += cb;
              }
              continue;
            }
            path[l] = (uint16_t)w;  // the same value from every lane
            ++l;
            c = cn;
            if constexpr (ALL) {
              if (lane == 0) lacc[l] = 0;
            }
            R = N;
          }
          if (lane == 0) {
            if constexpr (ALL) {
              const unsigned long long v = lacc[1];
              if (v) {
                atomicAdd(&sh.cred[u], v);
                root_acc += v;
              }
            } else if (item_total) {
              if (!to_root && u < nwin) atomicAdd(&sh.cred[u], item_total);
              else root_acc += item_total;
            }
          }
        }
      } else if constexpr (K == 3) {
        const uint32_t bsh = part_shift(d);
        for (uint32_t j = tid;; j += NT) {
          const uint32_t u = part_item(j, bsh);
          if (u >= d) break;
          uint32_t c = 0;
#pragma unroll
          for (int j = 0; j < W; ++j) c += __popcll(sh.rows[u * RS + j]);
          if constexpr (ALL) {
            if (c) {
              atomicAdd(&sh.cred[u], (unsigned long long)c);
              root_acc += c;
#pragma unroll
              for (int j = 0; j < W; ++j) {
                unsigned long long bits = sh.rows[u * RS + j];
                while (bits) {
                  const uint32_t x = (uint32_t)j * 64 + (uint32_t)__builtin_ctzll(bits);
                  bits &= bits - 1;
                  atomicAdd(&sh.cred[x], 1ull);
                }
              }
            }
          } else if (c) {
            if (!to_root && sh.key[u] <= ks) sh.cred[u] += c;
            else root_acc += c;
          }
        }
      } else {
        // Items: u (k = 4) or DAG edges (u, v) (k >= 5).
        constexpr bool kEdgeItems = (K >= 5);
        constexpr int kStepDepth = kEdgeItems ? K - 4 : 1;  // cliques counted per step
        constexpr uint32_t kMinSet = kEdgeItems ? (uint32_t)(K - 3) : 2u;
        uint32_t nitems = d;
        if constexpr (kEdgeItems) {
          for (uint32_t u = tid; u < d; u += NT) {
            uint32_t c = 0;
#pragma unroll
            for (int j = 0; j < W; ++j) c += __popcll(sh.rows[u * RS + j]);
            sh.pre[u + 1] = c;
          }
          __syncthreads();
          if (tid < 64) wave_scan(sh.pre + 1, d, lane);
          __syncthreads();
          nitems = sh.pre[d];
        }
        // next_item counts this part's items (part_item).
        const uint32_t ibsh = part_shift(nitems);
        if (tid == 0) sh.next_item = 0;
        __syncthreads();

        // Dense roots (ranked kernels, expected |P| of an item ~ rho^2 d / 3
        // >= dense_items, rho the edge density of the oriented graph on L): a
        // wavefront per item, the lanes sharing the last level.  Items are
        // handed out one per wavefront from next_item, so a wavefront's items
        // are increasing.
        bool dense = kDenseOnly;
        if constexpr (kDense && !kDenseOnly) {
          const float dd = (float)d, ni = (float)nitems;
          dense = d >= 2 && 4.0f * ni * ni >= 3.0f * dense_items * dd * (dd - 1.0f) * (dd - 1.0f);
        }
        if (dense) {
          if constexpr (kDense) {
            const uint32_t nwin = sh.nwin;
            uint32_t ucur = 0;
            // Per-lane sums of the items of member upend (credited to it when
            // u changes); items credited to the root go to root_acc at once.
            uint32_t upend = UINT32_MAX;
            unsigned long long pend = 0;
            auto flush = [&]() {
              const unsigned long long c = wave_sum(pend);
              if (lane == 0 && c) atomicAdd(&sh.cred[upend], c);
              pend = 0;
            };
            for (;;) {
              uint32_t it = 0;
              if (lane == 0) it = atomicAdd(&sh.next_item, 1u);
              it = (uint32_t)__builtin_amdgcn_readfirstlane((int)it);
              it = kEdgeItems ? part_edge_item(it) : part_item(it, ibsh);
              if (it >= nitems) break;
              uint32_t lo = ucur, hi = d;  // pre[u] <= it < pre[u+1]
              while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sh.pre[mid] <= it) lo = mid;
                else hi = mid;
              }
              ucur = lo;
              const uint32_t u = lo;
              uint32_t r = it - sh.pre[u];
              uint32_t v = 0;
              bool found = false;
#pragma unroll
              for (int j = 0; j < W; ++j) {
                const unsigned long long word = sh.rows[u * RS + j];
                const uint32_t c = __popcll(word);
                if (!found) {
                  if (r < c) {
                    v = (uint32_t)j * 64 + select_bit(word, r);
                    found = true;
                  } else {
                    r -= c;
                  }
                }
              }
              unsigned long long P[W];
#pragma unroll
              for (int j = 0; j < W; ++j) P[j] = sh.rows[u * RS + j] & sh.rows[v * RS + j];
              to_sgpr<W>(P);
              if (popc_set<W>(P) < (uint32_t)(K - 3)) continue;
              const uint64_t c = wave_cliques<W, K - 3>(P, sh.rows, lane);
              if (to_root || u >= nwin) {
                root_acc += c;
              } else {
                if (u != upend) {
                  if (upend != UINT32_MAX) flush();
                  upend = u;
                }
                pend += c;
              }
            }
            if (upend != UINT32_MAX) flush();
          }
        } else if constexpr (!kDenseOnly) {
          uint32_t u = 0, vsec = 0, wi = 0;
          unsigned long long P[W];
          unsigned long long bits = 0;
          unsigned long long acc = 0;
          bool live = false;
          uint32_t ucur = 0;

          // Next item with |P| >= kMinSet, or live = false.  For k <= 5 items
          // are handed out on demand (one LDS atomic per wavefront and round,
          // each lane taking the next index), so a lane that drew light items
          // keeps drawing while another works through a heavy one; for k >= 6,
          // where an item's own inner loops dominate, lanes take every NT-th
          // item (measured faster there).  A lane's items are increasing,
          // which the search for u below relies on.
          constexpr bool kOnDemand = (K <= 5);
          uint32_t e = tid;  // static items: the lane's next index within the part
          auto fetch = [&]() {
            live = false;
            for (;;) {
              uint32_t it;
              if constexpr (kOnDemand) {
                const unsigned long long act = __ballot(1);
                const uint32_t leader = (uint32_t)__builtin_ctzll(act);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&sh.next_item, (uint32_t)__popcll(act));
                base = __shfl(base, (int)leader);
                it = base + (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
              } else {
                it = e;
                e += NT;
              }
              it = kEdgeItems ? part_edge_item(it) : part_item(it, ibsh);
              if (it >= nitems) return;
              if constexpr (kEdgeItems) {
                uint32_t lo = ucur, hi = d;  // pre[u] <= it < pre[u+1]
                while (hi - lo > 1) {
                  const uint32_t mid = (lo + hi) >> 1;
                  if (sh.pre[mid] <= it) lo = mid;
                  else hi = mid;
                }
                ucur = lo;
                u = lo;
                uint32_t r = it - sh.pre[u];
                uint32_t v = 0;
                bool found = false;
#pragma unroll
                for (int j = 0; j < W; ++j) {
                  const unsigned long long word = sh.rows[u * RS + j];
                  const uint32_t c = __popcll(word);
                  if (!found) {
                    if (r < c) {
                      v = (uint32_t)j * 64 + select_bit(word, r);
                      found = true;
                    } else {
                      r -= c;
                    }
                  }
                }
#pragma unroll
                for (int j = 0; j < W; ++j) P[j] = sh.rows[u * RS + j] & sh.rows[v * RS + j];
                vsec = v;
              } else {
                u = it;
#pragma unroll
                for (int j = 0; j < W; ++j) P[j] = sh.rows[u * RS + j];
              }
              if (popc_set<W>(P) >= kMinSet) {
                wi = 0;
                bits = P[0];
                live = true;
                return;
              }
            }
          };

          // One step: the next w of the lane's set P.
          auto step = [&]() {
            const uint32_t w = wi * 64 + (uint32_t)__builtin_ctzll(bits);
            bits &= bits - 1;
            const unsigned long long *r = sh.rows + w * RS;
            unsigned long long Q[W];
#pragma unroll
            for (int j = 0; j < W; ++j) Q[j] = P[j] & r[j];
            if constexpr (ALL) {
              const uint64_t c = cliques_all<W, kStepDepth>(Q, sh.rows, sh.cred);
              if (c) atomicAdd(&sh.cred[w], (unsigned long long)c);
              acc += c;
            } else {
              acc += cliques_in<W, kStepDepth>(Q, sh.rows);
            }
          };
          auto next_word = [&]() {
            ++wi;
            unsigned long long nb = 0;
#pragma unroll
            for (int j = 0; j < W; ++j) nb = (j == (int)wi) ? P[j] : nb;
            bits = nb;
          };
          // The item's cliques: to its first member or to the root.
          auto finish_item = [&]() {
            if (acc) {
              if constexpr (ALL) {
                atomicAdd(&sh.cred[u], acc);
                if constexpr (kEdgeItems) atomicAdd(&sh.cred[vsec], acc);
                root_acc += acc;
              } else {
                const bool wins = kRanked ? (u < sh.nwin) : (sh.key[u] <= ks);
                if (!to_root && wins) atomicAdd(&sh.cred[u], acc);
                else root_acc += acc;
              }
              acc = 0;
            }
          };

          fetch();
          if constexpr (kOnDemand) {
            // Steps are a few instructions and a fetch (item search, bit
            // select, the set P) is several steps long, so lanes that run
            // out of work wait until kRefillAt of them can fetch together:
            // a round of steps then keeps most lanes busy instead of paying
            // a near-empty fetch branch in every round (measured at k = 5 on
            // hub roots: 8, 16, 32, 48 waiting lanes give passes of 2.73,
            // 2.57, 2.50, 2.55 ms).  A lane that drew past the last item is
            // done (the queue only grows).
            constexpr uint32_t kRefillAt = 32;
            bool done = !live;
            for (;;) {
              for (;;) {
                if (live && !bits) {
                  while (!bits && wi + 1 < (uint32_t)W) next_word();
                  if (!bits) {
                    finish_item();
                    live = false;
                  }
                }
                const bool waiting = !live && !done;
                if (__ballot(live) == 0ull ||
                    (uint32_t)__popcll(__ballot(waiting)) >= kRefillAt)
                  break;
                if (live) step();
              }
              const bool waiting = !live && !done;
              if (__ballot(waiting) == 0ull) break;
              if (waiting) {
                fetch();
                done = !live;
              }
            }
          }
          else {
            // k >= 6: a lane refills at once.  Its steps are nested loops of
            // their own, of uneven length, so lanes waiting for a batch
            // would idle longer than a lone fetch costs (batched at 32:
            // 1M/16M k = 6 pass 16.8 -> 17.6 ms, R-MAT scale 15 k = 7
            // 2310 -> 2340 ms, 4M/60M k = 6 3621 -> 3600 ms).
            while (live) {
              if (bits) {
                step();
              } else if (wi + 1 < (uint32_t)W) {
                next_word();
              } else {
                finish_item();
                fetch();
              }
            }
          }
        }
      }
    }

    // Flush: the root's share (reduced per wave, then per workgroup), then
    // each member's.
    root_acc = wave_sum(root_acc);
    if (lane == 0 && root_acc) atomicAdd(&sh.root_sum, root_acc);
    __syncthreads();
    if (tid == 0 && sh.root_sum) {
      // The root's share: to s (root counts, and every clique's root in
      // k-clique degrees); in an edge task to the winner of {s, u} and, for
      // k-clique degrees, to u as well.
      const uint32_t r = (edge_task && !ALL && !to_root && key[eu] <= ks) ? eu : s;
      atomicAdd(&cnt[r], sh.root_sum);
      if (ALL && edge_task) atomicAdd(&cnt[eu], sh.root_sum);
    }
    if (timers && tid == 0) timers[2 * (size_t)ti + 1] = __builtin_amdgcn_s_memtime() - t_start;
    for (uint32_t i = tid; i < d; i += NT) {
      const unsigned long long c = sh.cred[i];
      if (c) atomicAdd(&cnt[kRanked ? sh.beg[i] : sh.node[i]], c);
    }
    __syncthreads();
  }
}

}  // namespace

using PassFn = void (*)(const uint32_t *, const uint32_t *, const uint64_t *,
                        const unsigned long long *, uint32_t, const Task *, uint32_t, uint32_t,
                        float, uint32_t *, uint32_t *, const uint32_t *, uint32_t, uint32_t,
                        unsigned long long *, unsigned long long *);

constexpr int kMaxTemplK = 10;  // k handled by compiled kernels: 2..kMaxTemplK
constexpr uint32_t kMaxDeepK = 255;  // larger k (the reference's unsigned char k)

// Waves per workgroup of the large-k kernels: one for rows of up to 128
// members, two up to 256, four wider (each wavefront keeps its path in its
// share of the key array).
constexpr uint32_t deep_waves(uint32_t c) {
  return kClassWords[c] <= 2 ? 1u : kClassWords[c] <= 4 ? 2u : 4u;
}

// Kernel pointers and LDS size of word class <c> (one translation unit per
// class, pass_inst.hip, so the instantiations compile in parallel): the
// k = 2..kMaxTemplK kernels, their k-clique-degree versions, and the
// large-k kernel and its degree version.
void fill_pass_class_0(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_1(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_2(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_3(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_4(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_5(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_6(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_7(PassFn *, PassFn *, PassFn *, size_t *);
void fill_pass_class_8(PassFn *, PassFn *, PassFn *, size_t *);
static_assert(kNumClasses == 9, "one fill_pass_class_<c> per word class");

}  // namespace kcl
