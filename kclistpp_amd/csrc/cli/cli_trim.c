/*
 * kClistCoreTrim - k-clique core trimming with the k-clique degrees on one
 * MI355X; drop-in for the reference program (kClistCoreTrim.c:707-775).
 *
 *   ./kClistCoreTrim k edgelist.txt ckdeg.txt ckcore.txt ckdens.txt tag
 *
 * Writes ckdeg, ckcore and ckdens in the reference's formats, plus
 * trimmed_<tag>_<k>.txt (the trimmed graph, input ids) and
 * stat_trim_<tag>_<k>.txt.  k >= 3.  Failures exit non-zero.
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
  if (argc != 7) {
    fprintf(stderr, "usage: %s k edgelist.txt ckdeg.txt ckcore.txt ckdens.txt tag\n", argv[0]);
    return 2;
  }
  char *end = NULL;
  unsigned long k = strtoul(argv[1], &end, 10);
  if (!argv[1][0] || *end || k < 3 || k > 255) {
    fprintf(stderr, "%s: k must be an integer in 3..255\n", argv[0]);
    return 2;
  }
  time_t t0 = time(NULL), t1 = t0, t2;
  printf("Reading edgelist from file %s\n", argv[2]);
  uint32_t *pairs = NULL, n_in = 0;
  uint64_t m = 0;
  kcl_status st = kcl_read_edge_list(argv[2], &pairs, &m, &n_in);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;
  printf("Number of nodes: %u\n", n_in);
  printf("Number of edges: %llu\n", (unsigned long long)m);
  printf("Computing %lu-clique degrees and the %lu-clique core decomposition\n", k, k);
  kcl_trim *tr = NULL;
  st = kcl_trim_create(pairs, m, n_in, (uint32_t)k, &tr);
  kcl_free(pairs);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  kcl_trim_info info;
  kcl_trim_get_info(tr, &info);
  printf("Number of nodes (with core value > %lu): %u\n", k - 2, info.n2);
  printf("Number of edges (between nodes with core value > %lu): %u\n", k - 2, info.e2);
  printf("Core number = %u\n", info.core);
  printf("Number of %lu-cliques: %llu\n", k, (unsigned long long)info.ncktot);
  printf("%lu-clique core number of the graph = %llu\n", k, (unsigned long long)info.ckcore);
  printf("%lu-clique density of the dense subgraph = %lf\n", k, info.ckrho);
  printf("Edge density of the found dense subgraph = %lf\n", info.rho);
  printf("Size of the found dense subgraph = %u\n", info.size);
  const char *tag = argv[6];
  size_t len = strlen(tag) + strlen(argv[1]) + 32;
  char *trim_name = (char *)malloc(len), *stat_name = (char *)malloc(len);
  if (!trim_name || !stat_name) {
    fprintf(stderr, "%s: out of memory\n", argv[0]);
    kcl_trim_destroy(tr);
    return 1;
  }
  snprintf(trim_name, len, "trimmed_%s_%s.txt", tag, argv[1]);
  snprintf(stat_name, len, "stat_trim_%s_%s.txt", tag, argv[1]);
  st = kcl_trim_write(tr, argv[3], argv[4], argv[5], trim_name, stat_name, (int64_t)t0);
  if (st == KCL_OK)
    printf("The trimmed graph has %u nodes, %u edges and %llu k-cliques.\n", info.ntrim,
           info.mtrim, (unsigned long long)info.ncktrim);
  free(trim_name);
  free(stat_name);
  kcl_trim_destroy(tr);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    return 1;
  }
  t2 = time(NULL);
  print_elapsed(t1, t2);
  long d = (long)(t2 - t0);
  printf("- Overall time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
  return 0;
}
