/*
 * kcl_oracle.c - CPU restatement of the reference's Sim-kClist++ path
 * (kCDensestCmpSync.c, Graph.c, BinaryHeap.c) and of its k-clique core
 * trimming (kClistCoreTrim.c).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links or calls it.
 *
 * The enumeration here is deliberately the plain edge-rooted kClist
 * recursion over sorted lists with one argmin per clique, a different
 * formulation from the GPU path (vertex-rooted bitmaps with bulk
 * popcounts), so that a mistake in either one shows up as a mismatch.
 *
 * Parity pins (checked in tests/test_oracle.py): golden vectors written
 * by the compiled reference (ranks, per-iteration counts, rho bits, stat
 * records, trimming outputs), brute force over all k-subsets on small
 * graphs, and closed forms (K_n, disjoint cliques, clique-free graphs).
 *
 * Built with -O2 -ffp-contract=off: fp64 expressions are evaluated as
 * written, one IEEE rounding per operation, like gcc's x86-64 output for
 * the reference.
 */
#include "kcl_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------ */
/* Edge-list reader (Graph.c:34-53). */

int orc_read_edge_list(const char *path, uint32_t **pairs, uint64_t *m, uint32_t *n) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  uint64_t cap = 1 << 16, cnt = 0;
  uint32_t *buf = (uint32_t *)malloc(cap * 2 * sizeof(uint32_t));
  uint32_t a, b, mx = 0;
  while (fscanf(f, "%u %u", &a, &b) == 2) {
    if (cnt == cap) {
      cap *= 2;
      buf = (uint32_t *)realloc(buf, cap * 2 * sizeof(uint32_t));
    }
    buf[2 * cnt] = a;
    buf[2 * cnt + 1] = b;
    if (a > mx) mx = a;
    if (b > mx) mx = b;
    ++cnt;
  }
  fclose(f);
  *pairs = buf;
  *m = cnt;
  *n = mx + 1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Min-heap with decrease-by-one (BinaryHeap.c:6-84).  Ties are settled
 * by layout only: bubble-up moves while the parent is strictly larger,
 * bubble-down picks the right child only when strictly smaller than the
 * left one and moves while strictly smaller than the parent. */

typedef struct {
  uint32_t size;
  uint32_t *key;   /* slot -> node */
  uint64_t *val;   /* slot -> value */
  uint32_t *slot;  /* node -> slot, UINT32_MAX once popped */
} heap_t;

static void heap_init(heap_t *h, uint32_t n) {
  h->size = 0;
  h->key = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
  h->val = (uint64_t *)malloc((size_t)(n ? n : 1) * sizeof(uint64_t));
  h->slot = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) h->slot[i] = UINT32_MAX;
}

static void heap_release(heap_t *h) {
  free(h->key);
  free(h->val);
  free(h->slot);
}

static void heap_exchange(heap_t *h, uint32_t a, uint32_t b) {
  uint32_t ka = h->key[a], kb = h->key[b];
  uint64_t va = h->val[a];
  h->key[a] = kb;
  h->val[a] = h->val[b];
  h->key[b] = ka;
  h->val[b] = va;
  h->slot[kb] = a;
  h->slot[ka] = b;
}

static void heap_sift_up(heap_t *h, uint32_t i) {
  while (i > 0) {
    uint32_t p = (i - 1) >> 1;
    if (!(h->val[p] > h->val[i])) break;
    heap_exchange(h, i, p);
    i = p;
  }
}

static void heap_sift_down(heap_t *h) {
  uint32_t i = 0;
  for (;;) {
    uint32_t l = 2 * i + 1, r = l + 1, c;
    if (l >= h->size) break;
    c = (r < h->size && h->val[r] < h->val[l]) ? r : l;
    if (!(h->val[c] < h->val[i])) break;
    heap_exchange(h, i, c);
    i = c;
  }
}

static void heap_push(heap_t *h, uint32_t node, uint64_t v) {
  uint32_t at = h->size++;
  h->key[at] = node;
  h->val[at] = v;
  h->slot[node] = at;
  heap_sift_up(h, at);
}

static void heap_pop(heap_t *h, uint32_t *node, uint64_t *v) {
  *node = h->key[0];
  *v = h->val[0];
  h->slot[*node] = UINT32_MAX;
  --h->size;
  h->key[0] = h->key[h->size];
  h->val[0] = h->val[h->size];
  /* The reference writes the moved node's slot unconditionally, also when
   * the heap just became empty (then it is the popped node itself). */
  h->slot[h->key[0]] = 0;
  heap_sift_down(h);
}

static void heap_decrease(heap_t *h, uint32_t node, uint64_t by) {
  uint32_t at = h->slot[node];
  if (at == UINT32_MAX) return;
  h->val[at] -= by;
  heap_sift_up(h, at);
}

/* Undirected adjacency in edge order (Graph.c:58-74, kClistCoreTrim.c:198-220). */
static void undirected_csr(uint32_t n, uint64_t m, const uint32_t *pairs, uint64_t **off_out,
                           uint32_t **nbr_out) {
  uint64_t *off = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  uint32_t *nbr = (uint32_t *)malloc((size_t)(2 * m + 1) * sizeof(uint32_t));
  for (uint64_t i = 0; i < m; ++i) {
    off[pairs[2 * i] + 1]++;
    off[pairs[2 * i + 1] + 1]++;
  }
  for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
  uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  memcpy(fill, off, ((size_t)n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
    nbr[fill[a]++] = b;
    nbr[fill[b]++] = a;
  }
  free(fill);
  *off_out = off;
  *nbr_out = nbr;
}

void orc_degeneracy_rank(uint32_t n, uint64_t m, const uint32_t *pairs, uint32_t *rank) {
  uint64_t *off;
  uint32_t *nbr;
  undirected_csr(n, m, pairs, &off, &nbr);
  heap_t h;
  heap_init(&h, n);
  for (uint32_t v = 0; v < n; ++v) heap_push(&h, v, off[v + 1] - off[v]);
  for (uint32_t popped = 0; popped < n; ++popped) {
    uint32_t u;
    uint64_t du;
    heap_pop(&h, &u, &du);
    rank[u] = n - (popped + 1);
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) heap_decrease(&h, nbr[j], 1);
  }
  heap_release(&h);
  free(off);
  free(nbr);
}

/* ------------------------------------------------------------------ */
/* Relabel + CSR (Graph.c:95-156). */

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x > y) - (x < y);
}

uint32_t orc_build_dag(uint32_t n, uint64_t m, const uint32_t *pairs, const uint32_t *rank,
                       uint32_t *es, uint32_t *et, uint32_t *cd, uint32_t *adj,
                       uint32_t *core) {
  (void)n;
  uint64_t *keys = (uint64_t *)malloc((size_t)(m ? m : 1) * sizeof(uint64_t));
  uint32_t n2 = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = rank[pairs[2 * i]], b = rank[pairs[2 * i + 1]];
    uint32_t s = a > b ? a : b, t = a > b ? b : a;
    if (s + 1 > n2) n2 = s + 1;
    keys[i] = ((uint64_t)s << 32) | t;
  }
  qsort(keys, (size_t)m, sizeof(uint64_t), cmp_u64);
  memset(cd, 0, ((size_t)n2 + 1) * sizeof(uint32_t));
  for (uint64_t i = 0; i < m; ++i) {
    es[i] = (uint32_t)(keys[i] >> 32);
    et[i] = (uint32_t)keys[i];
    adj[i] = et[i];
    cd[es[i] + 1]++;
  }
  uint32_t mx = 0;
  for (uint32_t v = 0; v < n2; ++v) {
    if (cd[v + 1] > mx) mx = cd[v + 1];
    cd[v + 1] += cd[v];
  }
  *core = mx;
  free(keys);
  return n2;
}

/* ------------------------------------------------------------------ */
/* kClist enumeration with the per-clique update (kCDensestCmpSync.c:138-325). */

typedef struct {
  const uint32_t *cd, *adj;
  uint32_t k;
  int task;
  const double *rho;
  const uint32_t *level;
  uint64_t *acc;
  uint32_t *members; /* members[0] = s, [1] = t, then descending ids */
  uint32_t **buf;    /* candidate lists per depth */
  const uint8_t *mask;
  uint64_t visits;
} enum_ctx;

/* Sorted-list intersection. */
static uint32_t intersect(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb,
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

static void credit(enum_ctx *c) {
  uint32_t k = c->k;
  ++c->visits;
  if (c->acc == NULL) return;
  if (c->task == 2) {
    /* Scan in ascending id order, switch on strictly smaller rho
     * (kCDensestCmpSync.c:139-143); ascending order is members[] reversed. */
    uint32_t w = c->members[k - 1];
    for (int i = (int)k - 2; i >= 0; --i) {
      uint32_t x = c->members[i];
      if (c->rho[w] > c->rho[x]) w = x;
    }
    c->acc[w] += 1;
  } else {
    uint32_t w = c->members[0];
    for (uint32_t i = 1; i < k; ++i) {
      uint32_t x = c->members[i];
      if (c->level[x] > c->level[w]) w = x;
    }
    c->acc[c->level[w]] += 1;
  }
}

static void descend(enum_ctx *c, const uint32_t *cand, uint32_t nc, uint32_t depth) {
  /* depth = number of members chosen so far (>= 2); need k - depth more. */
  uint32_t need = c->k - depth;
  if (nc < need) return;
  for (uint32_t i = 0; i < nc; ++i) {
    uint32_t x = cand[i];
    c->members[depth] = x;
    if (need == 1) {
      credit(c);
      continue;
    }
    uint32_t *nxt = c->buf[depth + 1];
    uint32_t nn = intersect(cand, i, c->adj + c->cd[x], c->cd[x + 1] - c->cd[x], nxt);
    descend(c, nxt, nn, depth + 1);
  }
}

static uint64_t run_enum(uint32_t n, const uint32_t *cd, const uint32_t *adj, uint64_t m,
                         const uint32_t *es, const uint32_t *et, uint32_t k, int task,
                         const double *rho, const uint32_t *level, uint64_t *acc,
                         const uint8_t *mask) {
  (void)n;
  enum_ctx c;
  uint32_t core = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (cd[v + 1] - cd[v] > core) core = cd[v + 1] - cd[v];
  c.cd = cd;
  c.adj = adj;
  c.k = k;
  c.task = task;
  c.rho = rho;
  c.level = level;
  c.acc = acc;
  c.mask = mask;
  c.visits = 0;
  c.members = (uint32_t *)malloc((k + 1) * sizeof(uint32_t));
  c.buf = (uint32_t **)malloc((k + 2) * sizeof(uint32_t *));
  for (uint32_t d = 0; d < k + 2; ++d) c.buf[d] = (uint32_t *)malloc((core + 1) * sizeof(uint32_t));
  uint32_t *tmp = (uint32_t *)malloc((core + 1) * sizeof(uint32_t));
  for (uint64_t e = 0; e < m; ++e) {
    uint32_t s = es[e], t = et[e];
    if (mask && (!mask[s] || !mask[t])) continue;
    c.members[0] = s;
    c.members[1] = t;
    if (k == 2) {
      credit(&c);
      continue;
    }
    uint32_t nc = intersect(adj + cd[s], cd[s + 1] - cd[s], adj + cd[t], cd[t + 1] - cd[t], tmp);
    if (mask) {
      uint32_t w = 0;
      for (uint32_t i = 0; i < nc; ++i)
        if (mask[tmp[i]]) tmp[w++] = tmp[i];
      nc = w;
    }
    memcpy(c.buf[2], tmp, nc * sizeof(uint32_t));
    descend(&c, c.buf[2], nc, 2);
  }
  for (uint32_t d = 0; d < k + 2; ++d) free(c.buf[d]);
  free(c.buf);
  free(c.members);
  free(tmp);
  return c.visits;
}

uint64_t orc_clique_enum(uint32_t n, const uint32_t *cd, const uint32_t *adj, uint64_t m,
                         const uint32_t *es, const uint32_t *et, uint32_t k, int task,
                         const double *rho, const uint32_t *level, uint64_t *acc) {
  if (k < 2) return 0;
  return run_enum(n, cd, adj, m, es, et, k, task, rho, level, acc, NULL);
}

uint64_t orc_count_in_set(uint32_t n, const uint32_t *cd, const uint32_t *adj, uint64_t m,
                          const uint32_t *es, const uint32_t *et, uint32_t k,
                          const uint8_t *mask) {
  if (k < 2) return 0;
  return run_enum(n, cd, adj, m, es, et, k, 2, NULL, NULL, NULL, mask);
}

void orc_rho_update(uint32_t n, uint32_t t, double *rho, const uint64_t *cnt) {
  for (uint32_t i = 0; i < n; ++i) {
    double add = (double)cnt[i];
    rho[i] = ((double)t * rho[i] + 2.0 * add) / ((double)t + 2.0);
  }
}

/* ------------------------------------------------------------------ */
/* Densest prefix, edge count and upper bound (kCDensestCmpSync.c:334-383). */

static const double *g_sort_rho;

/* Key (unsigned long long)rho descending (kCDensestCmpSync.c:334-340);
 * glibc's stable merge sort leaves ties in index order. */
static int cmp_level(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  unsigned long long kx = (unsigned long long)g_sort_rho[x];
  unsigned long long ky = (unsigned long long)g_sort_rho[y];
  if (kx > ky) return -1;
  if (kx < ky) return 1;
  return (x > y) - (x < y);
}

void orc_find_densest(uint32_t n, const uint32_t *cd, const uint32_t *adj, uint64_t m,
                      const uint32_t *es, const uint32_t *et, uint32_t k, const double *rho,
                      uint32_t *reordered, uint32_t *level, uint64_t *tent, orc_densest *out) {
  for (uint32_t i = 0; i < n; ++i) reordered[i] = i;
  g_sort_rho = rho;
  qsort(reordered, n, sizeof(uint32_t), cmp_level);
  for (uint32_t i = 0; i < n; ++i) level[reordered[i]] = i;
  memset(tent, 0, (size_t)n * sizeof(uint64_t));
  orc_clique_enum(n, cd, adj, m, es, et, k, 3, rho, level, tent);

  double sum = 0, density = -1;
  uint32_t best = 0;
  uint64_t isum = 0, best_sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    sum += (double)tent[i];
    isum += tent[i];
    if (sum / (i + 1) > density) {
      density = sum / (i + 1);
      best = i + 1;
      best_sum = isum;
    }
  }
  uint8_t *inc = (uint8_t *)calloc((size_t)n + 1, 1);
  for (uint32_t i = 0; i < best; ++i) inc[reordered[i]] = 1;
  uint32_t me = 0;
  for (uint64_t e = 0; e < m; ++e) me += (inc[es[e]] && inc[et[e]]);
  free(inc);

  double ub = 0, s2 = 0, ip1ck = 1;
  for (uint32_t i = 0; i < n; ++i) {
    s2 += rho[reordered[i]];
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
  out->m = me;
  out->density = density;
  out->ub = ub;
  out->sum = best_sum;
}

int orc_cmpsync(uint32_t n, const uint32_t *cd, const uint32_t *adj, uint64_t m,
                const uint32_t *es, const uint32_t *et, uint32_t k, uint32_t num_iter,
                orc_record *records, uint64_t *counts_out, double *rho_out) {
  double *rho = (double *)calloc((size_t)n + 1, sizeof(double));
  uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  uint32_t *reordered = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  uint32_t *level = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  uint64_t *tent = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  int nrec = 0;
  uint32_t t = 1;
  for (uint64_t T = 1; T <= num_iter; T <<= 1) {
    for (; t <= T; ++t) {
      memset(cnt, 0, (size_t)n * sizeof(uint64_t));
      orc_clique_enum(n, cd, adj, m, es, et, k, 2, rho, NULL, cnt);
      if (counts_out) memcpy(counts_out + (size_t)(t - 1) * n, cnt, (size_t)n * sizeof(uint64_t));
      orc_rho_update(n, t, rho, cnt);
    }
    orc_densest d;
    orc_find_densest(n, cd, adj, m, es, et, k, rho, reordered, level, tent, &d);
    records[nrec].T = (uint32_t)T;
    records[nrec].n = d.n;
    records[nrec].m = d.m;
    records[nrec].density = d.density;
    records[nrec].ub = d.ub;
    records[nrec].sum = d.sum;
    ++nrec;
  }
  if (rho_out) memcpy(rho_out, rho, (size_t)n * sizeof(double));
  free(rho);
  free(cnt);
  free(reordered);
  free(level);
  free(tent);
  return nrec;
}

/* ------------------------------------------------------------------ */
/* k-clique core trimming (kClistCoreTrim.c:242-705). */

/* All k-cliques containing ck[0..depth) plus kmax-depth more from cand,
 * candidates taken in list order and narrowed by the DAG rows
 * (kClistCoreTrim.c:368-391).  Credits every member. */
static void trim_descend(const uint32_t *cd, const uint32_t *adj, uint32_t kmax, uint32_t depth,
                         uint32_t *ck, const uint32_t *cand, uint32_t nc, uint32_t **buf,
                         uint64_t *nck) {
  if (nc < kmax - depth) return;
  if (depth + 1 == kmax) {
    for (uint32_t i = 0; i < depth; ++i) nck[ck[i]] += nc;
    for (uint32_t i = 0; i < nc; ++i) nck[cand[i]] += 1;
    return;
  }
  for (uint32_t i = 0; i < nc; ++i) {
    uint32_t x = cand[i];
    ck[depth] = x;
    uint32_t *nxt = buf[depth];
    uint32_t nn = intersect(adj + cd[x], cd[x + 1] - cd[x], cand, nc, nxt);
    trim_descend(cd, adj, kmax, depth + 1, ck, nxt, nn, buf, nck);
  }
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

int orc_trim(uint32_t n_in, uint64_t m, const uint32_t *pairs, uint32_t kmax, uint32_t *map,
             uint64_t *nck, uint32_t *ic_key, uint64_t *ic_val, uint32_t *trim_pairs,
             orc_trim_info *info) {
  if (kmax < 3) return -1;
  memset(info, 0, sizeof(*info));
  info->n_in = n_in;

  /* kcore (kClistCoreTrim.c:242-272): peel by degree; nodes popped while
   * the running core value is below kmax-1 are dropped (rank -1). */
  uint64_t *off;
  uint32_t *nbr;
  undirected_csr(n_in, m, pairs, &off, &nbr);
  uint32_t *rank = (uint32_t *)malloc((size_t)n_in * sizeof(uint32_t));
  heap_t h;
  heap_init(&h, n_in);
  for (uint32_t v = 0; v < n_in; ++v) heap_push(&h, v, off[v + 1] - off[v]);
  uint32_t alive = n_in, r = 0;
  uint64_t c = 0;
  for (uint32_t i = 0; i < n_in; ++i) {
    uint32_t u;
    uint64_t du;
    heap_pop(&h, &u, &du);
    if (du > c) c = du;
    if (c < kmax - 1) {
      rank[u] = UINT32_MAX;
      --alive;
    } else {
      ++r;
      map[alive - r] = u;
      rank[u] = alive - r;
    }
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) heap_decrease(&h, nbr[j], 1);
  }
  heap_release(&h);
  free(off);
  free(nbr);
  uint32_t n2 = alive;
  info->core = (uint32_t)c;
  info->n2 = n2;

  /* relabelnodes (kClistCoreTrim.c:274-294): keep input edge order. */
  uint32_t *es = (uint32_t *)malloc((size_t)(m ? m : 1) * sizeof(uint32_t));
  uint32_t *et = (uint32_t *)malloc((size_t)(m ? m : 1) * sizeof(uint32_t));
  uint32_t e2 = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t a = rank[pairs[2 * i]], b = rank[pairs[2 * i + 1]];
    if (a == UINT32_MAX || b == UINT32_MAX) continue;
    es[e2] = a > b ? a : b;
    et[e2] = a > b ? b : a;
    ++e2;
  }
  info->e2 = e2;
  free(rank);

  if (n2 == 0) { /* the reference does not guard this case */
    free(es);
    free(et);
    return 0;
  }

  /* mkspecial (kClistCoreTrim.c:306-331): DAG rows, ascending. */
  uint32_t *cd = (uint32_t *)calloc((size_t)n2 + 1, sizeof(uint32_t));
  uint32_t *adj = (uint32_t *)malloc((size_t)(e2 ? e2 : 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < e2; ++i) cd[es[i] + 1]++;
  uint32_t dmax = 0;
  for (uint32_t v = 0; v < n2; ++v) {
    if (cd[v + 1] > dmax) dmax = cd[v + 1];
    cd[v + 1] += cd[v];
  }
  {
    uint32_t *fill = (uint32_t *)malloc(((size_t)n2 + 1) * sizeof(uint32_t));
    memcpy(fill, cd, ((size_t)n2 + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < e2; ++i) adj[fill[es[i]]++] = et[i];
    free(fill);
    for (uint32_t v = 0; v < n2; ++v) qsort(adj + cd[v], cd[v + 1] - cd[v], sizeof(uint32_t), cmp_u32);
  }
  /* Lists are bounded by the degeneracy; undirected rows can be longer. */
  uint32_t cap = (uint32_t)c + 2;
  uint32_t **buf = (uint32_t **)malloc((kmax + 1) * sizeof(uint32_t *));
  for (uint32_t d = 0; d < kmax + 1; ++d) buf[d] = (uint32_t *)malloc((size_t)(cap + dmax + 1) * sizeof(uint32_t));
  uint32_t *ck = (uint32_t *)malloc((kmax + 1) * sizeof(uint32_t));
  uint32_t *first = (uint32_t *)malloc((size_t)(cap + dmax + 1) * sizeof(uint32_t));

  /* onepass (kClistCoreTrim.c:394-416). */
  memset(nck, 0, (size_t)n2 * sizeof(uint64_t));
  for (uint32_t i = 0; i < e2; ++i) {
    ck[0] = es[i];
    ck[1] = et[i];
    uint32_t nc = intersect(adj + cd[es[i]], cd[es[i] + 1] - cd[es[i]], adj + cd[et[i]],
                            cd[et[i] + 1] - cd[et[i]], first);
    trim_descend(cd, adj, kmax, 2, ck, first, nc, buf, nck);
  }
  uint64_t tot = 0;
  for (uint32_t v = 0; v < n2; ++v) tot += nck[v];
  tot /= kmax;
  info->ncktot = tot;

  /* mkgraph2 (kClistCoreTrim.c:547-573): undirected rows, ascending. */
  uint32_t *cd0 = (uint32_t *)calloc((size_t)n2 + 1, sizeof(uint32_t));
  uint32_t *adj0 = (uint32_t *)malloc((size_t)(2 * (uint64_t)e2 + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < e2; ++i) {
    cd0[es[i] + 1]++;
    cd0[et[i] + 1]++;
  }
  for (uint32_t v = 0; v < n2; ++v) cd0[v + 1] += cd0[v];
  {
    uint32_t *fill = (uint32_t *)malloc(((size_t)n2 + 1) * sizeof(uint32_t));
    memcpy(fill, cd0, ((size_t)n2 + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < e2; ++i) {
      adj0[fill[es[i]]++] = et[i];
      adj0[fill[et[i]]++] = es[i];
    }
    free(fill);
    for (uint32_t v = 0; v < n2; ++v) qsort(adj0 + cd0[v], cd0[v + 1] - cd0[v], sizeof(uint32_t), cmp_u32);
  }

  /* kcliquecore (kClistCoreTrim.c:605-666). */
  heap_init(&h, n2);
  for (uint32_t v = 0; v < n2; ++v) heap_push(&h, v, nck[v]);
  uint32_t e = e2;
  uint64_t cc = 0, ncktot = tot;
  uint64_t *when = (uint64_t *)calloc((size_t)n2 + 1, sizeof(uint64_t));
  uint64_t *nck2 = (uint64_t *)calloc((size_t)n2 + 1, sizeof(uint64_t));
  uint32_t *alive_nbr = (uint32_t *)malloc(((size_t)cd0[n2] + 1) * sizeof(uint32_t));
  double ckrho = (double)ncktot / (double)n2;
  uint32_t size = n2;
  /* uint32 products as in the reference (they wrap for n2 > 65536). */
  double erho = (double)(uint32_t)(2u * e) / (double)(uint32_t)(n2 * (n2 - 1u));
  uint32_t i;
  for (i = 0; i + 1 < n2; ++i) {
    uint32_t u;
    uint64_t vu;
    heap_pop(&h, &u, &vu);
    if (vu > cc) cc = vu;
    ic_key[i] = u;
    ic_val[i] = cc;
    /* oneshortpass (kClistCoreTrim.c:577-603). */
    uint32_t s = 0;
    for (uint32_t j = cd0[u]; j < cd0[u + 1]; ++j)
      if (h.slot[adj0[j]] != UINT32_MAX) alive_nbr[s++] = adj0[j];
    if (s > kmax - 2) {
      ck[0] = u;
      for (uint32_t j = 0; j < s; ++j) {
        uint32_t v = alive_nbr[j];
        ck[1] = v;
        uint32_t nc = intersect(alive_nbr, s, adj + cd[v], cd[v + 1] - cd[v], first);
        trim_descend(cd, adj, kmax, 2, ck, first, nc, buf, nck2);
      }
    }
    uint64_t tot2 = 0;
    for (uint32_t j = cd0[u]; j < cd0[u + 1]; ++j) {
      uint32_t v = adj0[j];
      if (h.slot[v] != UINT32_MAX) {
        --e;
        heap_decrease(&h, v, nck2[v]);
        tot2 += nck2[v];
        nck2[v] = 0;
      }
    }
    uint32_t sz = n2 - i - 1;
    when[sz] = tot2 / (uint64_t)(kmax - 1);
    ncktot -= when[sz];
    double tmp = (double)ncktot / (double)sz;
    if (tmp > ckrho) {
      ckrho = tmp;
      size = sz;
      erho = (double)(uint32_t)(2u * e) / (double)(uint32_t)(sz * (sz - 1u));
    }
  }
  uint64_t ncktrim = tot;
  uint32_t ntrim = 0; /* the reference leaves this unset if no j qualifies */
  for (uint32_t j = n2 - 1; j >= 1; --j) {
    if ((double)when[j] >= ckrho) {
      ntrim = j + 1;
      break;
    }
    ncktrim -= when[j];
  }
  {
    uint32_t u;
    uint64_t vu;
    heap_pop(&h, &u, &vu);
    ic_key[i] = u;
    ic_val[i] = cc;
  }
  heap_release(&h);
  info->ckcore = cc;
  info->ckrho = ckrho;
  info->rho = erho;
  info->size = size;
  info->ntrim = ntrim;
  info->ncktrim = ncktrim;

  /* Trimmed graph (kClistCoreTrim.c:686-699): the last ntrim peeled
   * nodes, edges in relabel order, endpoints in old ids. */
  uint8_t *keep = (uint8_t *)calloc((size_t)n2 + 1, 1);
  for (uint32_t j = 0; j < ntrim; ++j) keep[ic_key[n2 - 1 - j]] = 1;
  uint32_t mt = 0;
  for (uint32_t j = 0; j < e2; ++j) {
    if (keep[es[j]] && keep[et[j]]) {
      trim_pairs[2 * mt] = map[es[j]];
      trim_pairs[2 * mt + 1] = map[et[j]];
      ++mt;
    }
  }
  info->mtrim = mt;

  free(keep);
  free(when);
  free(nck2);
  free(alive_nbr);
  free(cd0);
  free(adj0);
  for (uint32_t d = 0; d < kmax + 1; ++d) free(buf[d]);
  free(buf);
  free(ck);
  free(first);
  free(cd);
  free(adj);
  free(es);
  free(et);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Dinic max-flow (value only; the max-flow value is unique). */

typedef struct {
  uint32_t to;
  uint64_t cap;
} arc_t;

uint64_t orc_max_flow(uint32_t nv, uint64_t ne, const uint32_t *from, const uint32_t *to,
                      const uint64_t *cap, uint32_t s, uint32_t t) {
  uint64_t na = 2 * ne;
  uint64_t *head = (uint64_t *)calloc((size_t)nv + 1, sizeof(uint64_t));
  arc_t *arc = (arc_t *)malloc((size_t)(na ? na : 1) * sizeof(arc_t));
  uint64_t *rev = (uint64_t *)malloc((size_t)(na ? na : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < ne; ++i) {
    head[from[i] + 1]++;
    head[to[i] + 1]++;
  }
  for (uint32_t v = 0; v < nv; ++v) head[v + 1] += head[v];
  uint64_t *fill = (uint64_t *)malloc(((size_t)nv + 1) * sizeof(uint64_t));
  memcpy(fill, head, ((size_t)nv + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < ne; ++i) {
    uint64_t a = fill[from[i]]++, b = fill[to[i]]++;
    arc[a].to = to[i];
    arc[a].cap = cap[i];
    arc[b].to = from[i];
    arc[b].cap = 0;
    rev[a] = b;
    rev[b] = a;
  }
  free(fill);
  int64_t *dist = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
  uint32_t *queue = (uint32_t *)malloc((size_t)nv * sizeof(uint32_t));
  uint64_t *it = (uint64_t *)malloc((size_t)nv * sizeof(uint64_t));
  uint32_t *stack_v = (uint32_t *)malloc((size_t)nv * sizeof(uint32_t));
  uint64_t *stack_a = (uint64_t *)malloc((size_t)nv * sizeof(uint64_t));
  uint64_t flow = 0;
  for (;;) {
    for (uint32_t v = 0; v < nv; ++v) dist[v] = -1;
    uint32_t qh = 0, qt = 0;
    dist[s] = 0;
    queue[qt++] = s;
    while (qh < qt) {
      uint32_t v = queue[qh++];
      for (uint64_t a = head[v]; a < head[v + 1]; ++a)
        if (arc[a].cap > 0 && dist[arc[a].to] < 0) {
          dist[arc[a].to] = dist[v] + 1;
          queue[qt++] = arc[a].to;
        }
    }
    if (dist[t] < 0) break;
    for (uint32_t v = 0; v < nv; ++v) it[v] = head[v];
    /* Iterative blocking-flow DFS. */
    for (;;) {
      uint32_t depth = 0, v = s;
      int found = 0;
      for (;;) {
        if (v == t) {
          found = 1;
          break;
        }
        uint64_t a;
        for (a = it[v]; a < head[v + 1]; ++a)
          if (arc[a].cap > 0 && dist[arc[a].to] == dist[v] + 1) break;
        it[v] = a;
        if (a == head[v + 1]) {
          if (depth == 0) break;
          dist[v] = -1;
          --depth;
          v = stack_v[depth];
          it[v]++;
          continue;
        }
        stack_v[depth] = v;
        stack_a[depth] = a;
        ++depth;
        v = arc[a].to;
      }
      if (!found) break;
      uint64_t push = UINT64_MAX;
      for (uint32_t d = 0; d < depth; ++d)
        if (arc[stack_a[d]].cap < push) push = arc[stack_a[d]].cap;
      for (uint32_t d = 0; d < depth; ++d) {
        arc[stack_a[d]].cap -= push;
        arc[rev[stack_a[d]]].cap += push;
      }
      flow += push;
    }
  }
  free(head);
  free(arc);
  free(rev);
  free(dist);
  free(queue);
  free(it);
  free(stack_v);
  free(stack_a);
  return flow;
}
