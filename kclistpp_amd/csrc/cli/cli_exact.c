/*
 * kCDensestMem - exact k-clique densest subgraph on one MI355X; drop-in for
 * the reference program of the same name (kCDensestMem.c:547-600).
 *
 *   ./kCDensestMem p k edgeListFileName tag
 *
 * Writes stat_exact_<tag>_<p>_<k>.txt: the reference's header, one row per
 * round that did not certify, then the final record (iterations, stopping
 * condition, nodes, edges, density, max-flow calls, seconds).  The final
 * subgraph is the maximal k-clique densest subgraph.  Exit 1 on failure or
 * when no certificate was reached within the iteration limit.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kclpp.h"

static void print_elapsed(time_t a, time_t b) {
  long d = (long)(b - a);
  printf("- Time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s p k edgeListFileName tag\n", argv[0]);
    return 2;
  }
  char *end = NULL;
  unsigned long k = strtoul(argv[2], &end, 10);
  if (!argv[2][0] || *end || k < 2 || k > 255) {
    fprintf(stderr, "%s: k must be an integer in 2..255\n", argv[0]);
    return 2;
  }
  strtoul(argv[1], &end, 10);
  if (!argv[1][0] || *end) {
    fprintf(stderr, "%s: p must be an unsigned integer\n", argv[0]);
    return 2;
  }
  time_t t0 = time(NULL), t1 = t0, t2;
  printf("Reading edgelist from file %s\n", argv[3]);
  uint32_t *pairs = NULL, n_in = 0;
  uint64_t m = 0;
  kcl_status st = kcl_read_edge_list(argv[3], &pairs, &m, &n_in);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  printf("Number of nodes = %u\n", n_in);
  printf("Number of edges = %llu\n", (unsigned long long)m);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;
  printf("Building the graph structure\n");
  kcl_graph *g = NULL;
  st = kcl_graph_create(pairs, m, n_in, &g);
  kcl_free(pairs);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  kcl_graph_info info;
  kcl_graph_get_info(g, &info);
  printf("Number of nodes (degree > 0) = %u\n", info.n);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;
  printf("Iterate over all cliques\n");
  size_t len = strlen(argv[4]) + strlen(argv[1]) + strlen(argv[2]) + 32;
  char *name = (char *)malloc(len);
  if (!name) {
    kcl_graph_destroy(g);
    return 1;
  }
  snprintf(name, len, "stat_exact_%s_%s_%s.txt", argv[4], argv[1], argv[2]);
  kcl_exact_result r;
  st = kcl_exact_run(g, (uint32_t)k, NULL, name, (int64_t)t0, &r, NULL);
  free(name);
  kcl_graph_destroy(g);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  printf("Number of %lu-cliques in the densest subgraph: %llu\n", k,
         (unsigned long long)r.cliques);
  printf("Densest subgraph: %u nodes, %u edges, density = %.12f (%s, %u max-flow calls, %u "
         "iterations)\n",
         r.n, r.m, r.density, r.certified ? "certified" : "NOT certified", r.max_flow_calls,
         r.num_iter);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  long d = (long)(t2 - t0);
  printf("- Overall time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
  return r.certified ? 0 : 1;
}
