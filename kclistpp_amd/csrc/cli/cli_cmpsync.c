/*
 * kCDensestCmpSync - Sim-kClist++ on one MI355X, drop-in for the
 * reference program of the same name (kCDensestCmpSync.c:421-468).
 *
 *   ./kCDensestCmpSync p T k edgeListFileName tag
 *
 * p is accepted for compatibility (it names the output file as in the
 * reference); the enumeration runs on the GPU.  T is rounded down to a
 * power of two; one record per power of two goes to
 * stat_cmpsync_<tag>_<p>_<T>_<k>.txt with the reference's header and
 * "%u\t%u\t%u\t%.12f\t%.12f\t%ld" rows.  Unlike the reference, bad
 * arguments and failures exit non-zero with a message.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kclpp.h"

static int parse_uint(const char *s, unsigned long max, unsigned long *out) {
  char *end = NULL;
  unsigned long v = strtoul(s, &end, 10);
  if (!s[0] || *end || v > max) return -1;
  *out = v;
  return 0;
}

static void print_elapsed(time_t a, time_t b) {
  long d = (long)(b - a);
  printf("- Time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
}

int main(int argc, char **argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s p T k edgeListFileName tag\n", argv[0]);
    return 2;
  }
  unsigned long p, num_iter, k;
  if (parse_uint(argv[1], 1u << 20, &p) || parse_uint(argv[2], 0xffffffffu, &num_iter) ||
      parse_uint(argv[3], 255, &k)) {
    fprintf(stderr, "%s: p, T and k must be unsigned integers (k <= 255)\n", argv[0]);
    return 2;
  }
  if (k < 2) {
    fprintf(stderr, "%s: k must be at least 2\n", argv[0]);
    return 2;
  }
  const char *file_name = argv[4];
  time_t t0 = time(NULL), t1 = t0, t2;

  printf("Reading edgelist from file %s\n", file_name);
  uint32_t *pairs = NULL, n_in = 0;
  uint64_t m = 0;
  kcl_status st = kcl_read_edge_list(file_name, &pairs, &m, &n_in);
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
  printf("Core value (max truncated degree) = %u\n", info.core);
  printf("Number of nodes (degree > 0) = %u\n", info.n);
  t2 = time(NULL);
  print_elapsed(t1, t2);
  t1 = t2;

  printf("Iterate over all cliques\n");
  const char *tag = argv[5];
  size_t len = strlen(tag) + strlen(argv[1]) + strlen(argv[2]) + strlen(argv[3]) + 64;
  char *out_name = (char *)malloc(len);
  if (!out_name) {
    fprintf(stderr, "%s: out of memory\n", argv[0]);
    kcl_graph_destroy(g);
    return 1;
  }
  snprintf(out_name, len, "stat_cmpsync_%s_%s_%s_%s.txt", tag, argv[1], argv[2], argv[3]);
  kcl_record rec[32];
  uint32_t nrec = 0;
  st = kcl_cmpsync_run(g, (uint32_t)k, (uint32_t)num_iter, NULL, out_name, (int64_t)t0, rec,
                       &nrec);
  for (uint32_t i = 0; i < nrec; ++i)
    fprintf(stderr,
            "Approximate densest subgraph: %u nodes, density = %f, upper bound = %f. %ld "
            "seconds.\n",
            rec[i].n, rec[i].density, rec[i].ub, (long)rec[i].seconds);
  free(out_name);
  if (st != KCL_OK) {
    fprintf(stderr, "%s: %s: %s\n", argv[0], kcl_status_name(st), kcl_last_error());
    kcl_graph_destroy(g);
    return 1;
  }
  t2 = time(NULL);
  print_elapsed(t1, t2);
  kcl_graph_destroy(g);
  long d = (long)(t2 - t0);
  printf("- Overall time = %ldh%ldm%lds\n", d / 3600, (d % 3600) / 60, d % 60);
  return 0;
}
