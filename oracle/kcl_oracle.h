/*
 * kcl_oracle.h - CPU restatement of the reference's Sim-kClist++ path,
 * k-clique core trimming and exact-density checks.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product library (libkclpp.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Plain C,
 * single-threaded, no FMA contraction (built with -ffp-contract=off), so
 * fp64 results follow the reference's IEEE order bit for bit.
 */
#ifndef KCL_ORACLE_H
#define KCL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Text edge list, "%u %u" per line; n = max id + 1 (Graph.c:34-53).
 * *pairs is malloc'd (2*m uint32), free with orc_free. Returns 0 or -1. */
int orc_read_edge_list(const char *path, uint32_t **pairs, uint64_t *m, uint32_t *n);
void orc_free(void *p);

/* Degeneracy order by the reference's binary heap (Graph.c:57-91,
 * BinaryHeap.c:6-84): rank[v] = n - (pop position + 1). */
void orc_degeneracy_rank(uint32_t n, uint64_t m, const uint32_t *pairs, uint32_t *rank);

/* Relabel + orient + CSR (Graph.c:95-156).  Outputs (caller-allocated):
 *   es/et[m]   edges sorted by (s,t), s > t, in relabelled ids
 *   cd[n2+1], adj[m]  out-neighbours (smaller ids), ascending
 * Returns n2 = max relabelled s + 1 (0 for m = 0); *core = max out-degree. */
uint32_t orc_build_dag(uint32_t n, uint64_t m, const uint32_t *pairs, const uint32_t *rank,
                       uint32_t *es, uint32_t *et, uint32_t *cd, uint32_t *adj,
                       uint32_t *core);

/* One enumeration pass (kCDensestCmpSync.c:138-146, 148-296, 298-325).
 * task 2 (FRANK_WOLFE): every k-clique gives one unit to its member with
 *   the smallest rho, ties to the smallest id; acc[winner] += 1.
 * task 3 (PAVA_PREPROCESS): one unit to the member with the largest
 *   level; acc[level[winner]] += 1.
 * acc is accumulated into (the caller zeroes it, as the reference does).
 * Returns the number of k-cliques visited.  k >= 2. */
uint64_t orc_clique_enum(uint32_t n, const uint32_t *cd, const uint32_t *adj,
                         uint64_t m, const uint32_t *es, const uint32_t *et,
                         uint32_t k, int task, const double *rho, const uint32_t *level,
                         uint64_t *acc);

/* rho = (t*rho + 2*cnt)/(t + 2) per node (kCDensestCmpSync.c:402-403). */
void orc_rho_update(uint32_t n, uint32_t t, double *rho, const uint64_t *cnt);

/* CDF_FindDensestSubset (kCDensestCmpSync.c:342-383): fills reordered,
 * level and tent (= rho_tentative, by level) and the four report fields. */
typedef struct {
  uint32_t n;
  uint32_t m;
  double density;
  double ub;
  uint64_t sum;       /* exact numerator: cliques credited to the top n levels */
} orc_densest;

void orc_find_densest(uint32_t n, const uint32_t *cd, const uint32_t *adj,
                      uint64_t m, const uint32_t *es, const uint32_t *et,
                      uint32_t k, const double *rho, uint32_t *reordered,
                      uint32_t *level, uint64_t *tent, orc_densest *out);

/* Whole CDF_Main loop (kCDensestCmpSync.c:385-419) without the file:
 * runs iterations t = 1..Tmax (Tmax = largest power of two <= num_iter)
 * and fills one record per power of two (T, n, m, density, ub).
 * counts_out (nullable) receives the per-iteration count vectors,
 * Tmax * n uint64, row t-1 for iteration t.  rho_out (nullable) gets the
 * final rho.  Returns the number of records written (<= 32). */
typedef struct {
  uint32_t T;
  uint32_t n;
  uint32_t m;
  double density;
  double ub;
  uint64_t sum;
} orc_record;

int orc_cmpsync(uint32_t n, const uint32_t *cd, const uint32_t *adj,
                uint64_t m, const uint32_t *es, const uint32_t *et,
                uint32_t k, uint32_t num_iter, orc_record *records,
                uint64_t *counts_out, double *rho_out);

/* ---- k-clique core trimming (kClistCoreTrim.c) ---- */

typedef struct {
  uint32_t n_in;        /* nodes of the input (max id + 1) */
  uint32_t n2;          /* nodes with running core >= k-1 */
  uint32_t e2;          /* edges between them */
  uint32_t core;        /* degeneracy of the input */
  uint64_t ncktot;      /* k-cliques in the (k-1)-core */
  uint64_t ckcore;      /* k-clique core number */
  double ckrho;         /* best k-clique density of the peel */
  double rho;           /* its edge density */
  uint32_t size;        /* its size */
  uint32_t ntrim;       /* nodes kept by the trim */
  uint64_t ncktrim;     /* k-cliques in the trimmed graph */
  uint32_t mtrim;       /* edges in the trimmed graph */
} orc_trim_info;

/* Runs kcore -> relabelnodes -> mkspecial -> onepass -> kcliquecore and
 * the output selection of printdensest (kClistCoreTrim.c:242-705).
 * Outputs (caller-allocated, sizes from the input):
 *   map[n_in]     new -> old id for the first n2 entries
 *   nck[n_in]     k-clique degree per new id (first n2 entries)
 *   ic_key[n_in], ic_val[n_in]  peel order and running core value
 *   trim_pairs[2*m] edges of the trimmed graph in old ids (mtrim rows)
 * Returns 0, or -1 for k < 3. */
int orc_trim(uint32_t n_in, uint64_t m, const uint32_t *pairs, uint32_t k,
             uint32_t *map, uint64_t *nck, uint32_t *ic_key, uint64_t *ic_val,
             uint32_t *trim_pairs, orc_trim_info *info);

/* ---- exact checks ---- */

/* Max-flow value by Dinic on a capacity list (integer, unique value).
 * Edges (from[i], to[i], cap[i]); returns the s-t max-flow. */
uint64_t orc_max_flow(uint32_t nv, uint64_t ne, const uint32_t *from, const uint32_t *to,
                      const uint64_t *cap, uint32_t s, uint32_t t);

/* Number of k-cliques with all members in the vertex set S (mask[v] != 0)
 * of the relabelled DAG. */
uint64_t orc_count_in_set(uint32_t n, const uint32_t *cd, const uint32_t *adj,
                          uint64_t m, const uint32_t *es, const uint32_t *et,
                          uint32_t k, const uint8_t *mask);

#ifdef __cplusplus
}
#endif

#endif
