/*
 * kCDensestSamplingKclistpp - Sampling++ on one MI355X, drop-in for the
 * reference program of the same name (kCDensestSamplingKclistpp.c:491-536).
 *
 *   ./kCDensestSamplingKclistpp p T k edgeListFileName [seed]
 *
 * p is accepted for compatibility; the counting, listing and rounds run on
 * the GPU.  After rounds 1, 2, 4, ... up to T it prints the reference's
 * line "Approximate densest subgraph: N nodes, M edges, edge density = ...,
 * k-clique density = ..., upper bound = .... X milliseconds."  The kept
 * cliques depend on `seed` (default 0) instead of the clock, so a run can
 * be repeated.  Unlike the reference, bad arguments and failures exit
 * non-zero with a message.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kclpp.h"

static int parse_u64(const char *s, unsigned long long max, unsigned long long *out) {
  char *end = NULL;
  unsigned long long v = strtoull(s, &end, 10);
  if (!s[0] || *end || v > max) return -1;
  *out = v;
  return 0;
}

static void print_elapsed(time_t a, time_t b) {
  long d = (long)(b - a);
  printf("- Time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
}

static int die(const char *prog, kcl_status st) {
  fprintf(stderr, "%s: %s: %s\n", prog, kcl_status_name(st), kcl_last_error());
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "usage: %s p T k edgeListFileName [seed]\n", argv[0]);
    return 2;
  }
  unsigned long long p, num_iter, k, seed = 0;
  if (parse_u64(argv[1], 1u << 20, &p) || parse_u64(argv[2], 0xffffffffu, &num_iter) ||
      parse_u64(argv[3], 255, &k) || (argc == 6 && parse_u64(argv[5], ~0ull, &seed))) {
    fprintf(stderr, "%s: p, T, k and seed must be unsigned integers (k <= 255)\n", argv[0]);
    return 2;
  }
  if (k < 3) {
    fprintf(stderr, "%s: k must be at least 3\n", argv[0]);
    return 2;
  }
  (void)p;
  const char *file_name = argv[4];
  struct timespec ts0;
  clock_gettime(CLOCK_MONOTONIC, &ts0);
  time_t t1 = time(NULL), t2;

  printf("Reading edgelist from file %s\n", file_name);
  uint32_t *pairs = NULL, n_in = 0;
  uint64_t m = 0;
  kcl_status st = kcl_read_edge_list(file_name, &pairs, &m, &n_in);
  if (st != KCL_OK) return die(argv[0], st);
  printf("Number of nodes = %u\n", n_in);
  printf("Number of edges = %llu\n", (unsigned long long)m);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;

  printf("Building the graph structure\n");
  kcl_graph *g = NULL;
  st = kcl_graph_create(pairs, m, n_in, &g);
  kcl_free(pairs);
  if (st != KCL_OK) return die(argv[0], st);
  kcl_graph_info info;
  kcl_graph_get_info(g, &info);
  printf("Number of nodes (degree > 0) = %u\n", info.n);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;

  kcl_sampling_opts o;
  memset(&o, 0, sizeof(o));
  o.seed = seed;
  kcl_sampling *s = NULL;
  st = kcl_sampling_create(g, (uint32_t)k, &o, &s);
  if (st != KCL_OK) return die(argv[0], st);
  kcl_sampling_info si;
  kcl_sampling_get_info(s, &si);
  printf("Number of %u-cliques: %llu\n", (unsigned)k, (unsigned long long)si.total);
  printf("Number of sampled %u-cliques: %llu\n", (unsigned)k, (unsigned long long)si.sampled);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;

  for (unsigned long long T = 1; T <= num_iter; T <<= 1) {
    while (si.rounds < T) {
      if ((st = kcl_sampling_step(s)) != KCL_OK) return die(argv[0], st);
      kcl_sampling_get_info(s, &si);
    }
    kcl_sampling_record r;
    if ((st = kcl_sampling_report(s, &r, NULL)) != KCL_OK) return die(argv[0], st);
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    long ms = (long)((ts.tv_sec - ts0.tv_sec) * 1000 + (ts.tv_nsec - ts0.tv_nsec) / 1000000);
    printf("Approximate densest subgraph: %u nodes, %u edges, edge density = %f, k-clique "
           "density = %f, upper bound = %f. %ld milliseconds.\n",
           r.n, r.m, r.edge_density, r.density, r.ub, ms);
    fflush(stdout);
  }
  t2 = time(NULL);
  print_elapsed(t1, t2);
  kcl_sampling_destroy(s);
  kcl_graph_destroy(g);
  return 0;
}
