/*
 * ref_harness.c - drives the reference's own Sim-kClist++ code, compiled
 * from its sources (kCDensestCmpSync.c, Graph.c, BinaryHeap.c) into
 * oracle/_ref/libkclref.so by oracle/Makefile.
 *
 * TEST INFRASTRUCTURE ONLY: used to write golden vectors in the build
 * container and as the timed CPU baseline.  It calls the reference's
 * non-static symbols and reads its extern globals; the per-iteration loop
 * below is kCDensestCmpSync.c:392-413 with the vectors captured between
 * steps.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#include "Graph.h"

extern double *rho;
extern double *rho_add;
extern double *rho_tentative;
extern unsigned *level;
extern unsigned *reordered;

/* kCDensestCmpSync.c:128 and :327-332 */
enum { REF_FRANK_WOLFE = 2, REF_PAVA_PREPROCESS = 3 };
typedef struct {
  unsigned n;
  unsigned m;
  double density;
  double ub;
} ref_info;

void AllocCdf(Graph *g, unsigned k);
void CDF_CliqueEnum(Graph *g, unsigned char k, int task);
ref_info CDF_FindDensestSubset(Graph *g, unsigned char k);

typedef struct {
  Graph *g;
  unsigned n_in;
  unsigned *rank;
  unsigned k;
} kclref;

/* ReadEdgeList -> SortByCore -> Relabel -> MakeGraph (kCDensestCmpSync.c:433-443),
 * keeping a copy of the rank vector that MakeGraph frees. */
void *kclref_load(const char *path) {
  EdgeList *el = ReadEdgeList((char *)path);
  kclref *h = (kclref *)calloc(1, sizeof(kclref));
  h->n_in = el->n;
  SortByCore(el);
  h->rank = (unsigned *)malloc(el->n * sizeof(unsigned));
  memcpy(h->rank, el->rank, el->n * sizeof(unsigned));
  Relabel(el);
  h->g = MakeGraph(el);
  return h;
}

/* The same graph from uint32 pairs in memory: the EdgeList ReadEdgeList
 * would build from their text (n = max id + 1, Graph.c:34-53), then the
 * reference's own SortByCore / Relabel / MakeGraph. */
void *kclref_load_pairs(const unsigned *pairs, unsigned m) {
  EdgeList *el = (EdgeList *)malloc(sizeof(EdgeList));
  el->e = m;
  el->n = 0;
  el->edges = (Edge *)malloc((m ? m : 1) * sizeof(Edge));
  for (unsigned i = 0; i < m; ++i) {
    el->edges[i].s = pairs[2 * (size_t)i];
    el->edges[i].t = pairs[2 * (size_t)i + 1];
    if (el->edges[i].s > el->n) el->n = el->edges[i].s;
    if (el->edges[i].t > el->n) el->n = el->edges[i].t;
  }
  ++el->n;
  kclref *h = (kclref *)calloc(1, sizeof(kclref));
  h->n_in = el->n;
  SortByCore(el);
  h->rank = (unsigned *)malloc(el->n * sizeof(unsigned));
  memcpy(h->rank, el->rank, el->n * sizeof(unsigned));
  Relabel(el);
  h->g = MakeGraph(el);
  return h;
}

/* The reference's Graph filled from a DAG CSR the caller holds (rows of
 * smaller ids, ascending: the layout MakeGraph builds, Graph.c:120-156),
 * with edges in (s, t) order.  The benchmark's CPU baseline uses it on a
 * graph whose CSR the tests show equal to the reference's, to skip a second
 * degeneracy ordering of a large input.  No ranks are kept. */
void *kclref_load_csr(unsigned n, const unsigned *cd, const unsigned *adj) {
  Graph *g = (Graph *)malloc(sizeof(Graph));
  g->n = n;
  g->e = cd[n];
  g->cd = (unsigned *)malloc((n + 1) * sizeof(unsigned));
  memcpy(g->cd, cd, (n + 1) * sizeof(unsigned));
  g->adj = (unsigned *)malloc((g->e ? g->e : 1) * sizeof(unsigned));
  memcpy(g->adj, adj, g->e * sizeof(unsigned));
  g->edges = (Edge *)malloc((g->e ? g->e : 1) * sizeof(Edge));
  g->core = 0;
  for (unsigned s = 0; s < n; ++s) {
    unsigned d = cd[s + 1] - cd[s];
    if (d > g->core) g->core = d;
    for (unsigned j = cd[s]; j < cd[s + 1]; ++j) {
      g->edges[j].s = s;
      g->edges[j].t = adj[j];
    }
  }
  kclref *h = (kclref *)calloc(1, sizeof(kclref));
  h->n_in = n;
  h->g = g;
  return h;
}

unsigned kclref_n_in(void *p) { return ((kclref *)p)->n_in; }
unsigned kclref_n(void *p) { return ((kclref *)p)->g->n; }
unsigned kclref_e(void *p) { return ((kclref *)p)->g->e; }
unsigned kclref_core(void *p) { return ((kclref *)p)->g->core; }

void kclref_rank(void *p, unsigned *out) {
  kclref *h = (kclref *)p;
  memcpy(out, h->rank, h->n_in * sizeof(unsigned));
}

void kclref_edges(void *p, unsigned *out) {
  kclref *h = (kclref *)p;
  for (unsigned i = 0; i < h->g->e; ++i) {
    out[2 * i] = h->g->edges[i].s;
    out[2 * i + 1] = h->g->edges[i].t;
  }
}

void kclref_csr(void *p, unsigned *cd, unsigned *adj) {
  kclref *h = (kclref *)p;
  memcpy(cd, h->g->cd, (h->g->n + 1) * sizeof(unsigned));
  memcpy(adj, h->g->adj, h->g->e * sizeof(unsigned));
}

void kclref_threads(int p) { omp_set_num_threads(p); }

/* AllocCdf + rho = 0 (kCDensestCmpSync.c:388-393). */
void kclref_start(void *p, unsigned k) {
  kclref *h = (kclref *)p;
  h->k = k;
  AllocCdf(h->g, k);
  for (unsigned i = 0; i < h->g->n; ++i) rho[i] = 0;
}

/* One Frank-Wolfe step t (kCDensestCmpSync.c:399-403); copies the pass's
 * rho_add (before the update) to add_out and the updated rho to rho_out. */
void kclref_step(void *p, unsigned t, double *add_out, double *rho_out) {
  kclref *h = (kclref *)p;
  Graph *g = h->g;
  for (unsigned i = 0; i < g->n; ++i) rho_add[i] = 0;
  CDF_CliqueEnum(g, (unsigned char)h->k, REF_FRANK_WOLFE);
  if (add_out) memcpy(add_out, rho_add, g->n * sizeof(double));
  for (unsigned i = 0; i < g->n; ++i) rho[i] = (t * rho[i] + 2.0 * rho_add[i]) / (t + 2.0);
  if (rho_out) memcpy(rho_out, rho, g->n * sizeof(double));
}

/* Report step (kCDensestCmpSync.c:406-408): returns n, m, density, ub and
 * copies level, reordered and rho_tentative. */
void kclref_report(void *p, unsigned *nm, double *dens_ub, unsigned *level_out,
                   unsigned *reordered_out, double *tent_out) {
  kclref *h = (kclref *)p;
  Graph *g = h->g;
  for (unsigned i = 0; i < g->n; ++i) rho_tentative[i] = 0;
  ref_info info = CDF_FindDensestSubset(g, (unsigned char)h->k);
  nm[0] = info.n;
  nm[1] = info.m;
  dens_ub[0] = info.density;
  dens_ub[1] = info.ub;
  if (level_out) memcpy(level_out, level, g->n * sizeof(unsigned));
  if (reordered_out) memcpy(reordered_out, reordered, g->n * sizeof(unsigned));
  if (tent_out) memcpy(tent_out, rho_tentative, g->n * sizeof(double));
}

/* One CDF_CliqueEnum(PAVA_PREPROCESS) with the caller's level vector
 * (kCDensestCmpSync.c:217-246 read the global level[]); copies
 * rho_tentative (indexed by level) to tent_out. */
void kclref_pava_level(void *p, const unsigned *level_in, double *tent_out) {
  kclref *h = (kclref *)p;
  Graph *g = h->g;
  memcpy(level, level_in, g->n * sizeof(unsigned));
  for (unsigned i = 0; i < g->n; ++i) rho_tentative[i] = 0;
  CDF_CliqueEnum(g, (unsigned char)h->k, REF_PAVA_PREPROCESS);
  memcpy(tent_out, rho_tentative, g->n * sizeof(double));
}

/* Set rho from the caller (to time or check a pass at a given state). */
void kclref_set_rho(void *p, const double *src) {
  kclref *h = (kclref *)p;
  memcpy(rho, src, h->g->n * sizeof(double));
}

/* Times one CDF_CliqueEnum(FRANK_WOLFE) over a subset of the root edges
 * (indices into the sorted edge array), with the current rho.  Returns the
 * wall seconds; *visits gets the k-cliques credited (sum of rho_add). */
double kclref_time_fw_roots(void *p, const unsigned *roots, unsigned nroots, double *visits) {
  kclref *h = (kclref *)p;
  Graph *g = h->g;
  Edge *all = g->edges;
  unsigned e_all = g->e;
  Edge *sub = (Edge *)malloc((nroots ? nroots : 1) * sizeof(Edge));
  for (unsigned i = 0; i < nroots; ++i) sub[i] = all[roots[i]];
  for (unsigned i = 0; i < g->n; ++i) rho_add[i] = 0;
  g->edges = sub;
  g->e = nroots;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  CDF_CliqueEnum(g, (unsigned char)h->k, REF_FRANK_WOLFE);
  clock_gettime(CLOCK_MONOTONIC, &b);
  g->edges = all;
  g->e = e_all;
  free(sub);
  double s = 0;
  for (unsigned i = 0; i < g->n; ++i) s += rho_add[i];
  *visits = s;
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* Thread-seconds of each root edge alone: the sample's edges are handed to
 * the host threads one at a time (dynamic schedule), and each thread runs
 * the reference's CDF_CliqueEnum(FRANK_WOLFE) over a one-edge view of the
 * graph (its own parallel region then has one thread: nesting is off) under
 * its own clock.  sec_out[i] receives edge i's seconds; returns the wall
 * seconds of the whole sample, *visits the cliques credited. */
double kclref_time_fw_each(void *p, const unsigned *roots, unsigned nroots, double *sec_out,
                           double *visits) {
  kclref *h = (kclref *)p;
  Graph *g = h->g;
  for (unsigned i = 0; i < g->n; ++i) rho_add[i] = 0;
  int levels = omp_get_max_active_levels();
  omp_set_max_active_levels(1);
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  #pragma omp parallel for schedule(dynamic, 1)
  for (unsigned i = 0; i < nroots; ++i) {
    Graph one = *g;
    one.edges = g->edges + roots[i];
    one.e = 1;
    struct timespec x, y;
    clock_gettime(CLOCK_MONOTONIC, &x);
    CDF_CliqueEnum(&one, (unsigned char)h->k, REF_FRANK_WOLFE);
    clock_gettime(CLOCK_MONOTONIC, &y);
    sec_out[i] = (double)(y.tv_sec - x.tv_sec) + 1e-9 * (double)(y.tv_nsec - x.tv_nsec);
  }
  clock_gettime(CLOCK_MONOTONIC, &b);
  omp_set_max_active_levels(levels);
  double s = 0;
  for (unsigned i = 0; i < g->n; ++i) s += rho_add[i];
  *visits = s;
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

void kclref_free(void *p) {
  kclref *h = (kclref *)p;
  FreeGraph(h->g);
  free(h->rank);
  free(h);
}
