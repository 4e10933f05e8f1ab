/*
 * kclpp.h - C ABI of the MI355X Sim-kClist++ library (libkclpp.so).
 *
 * The library replaces the reference's data-parallel hot path: kClist
 * k-clique enumeration over the degeneracy-ordered DAG fused with the
 * Sim-kClist++ update (every k-clique gives one unit to its member of
 * smallest rho, ties to the smallest id), the PAVA-preprocess pass, the
 * densest-prefix report, and the k-clique degree pass of core trimming.
 * The enumeration runs as hand-written HIP kernels on one gfx950 GPU; the
 * sequential parts whose order defines the results (the degeneracy heap,
 * the upper-bound scan, the trimming peel, max-flow) run on the host.
 *
 * Conventions: plain pointers and sizes, no globals.  Output arrays are
 * caller-allocated host memory.  Every call returns a kcl_status; on
 * failure kcl_last_error() holds a message for the calling thread.  A
 * kcl_graph / kcl_sim is used by one host thread at a time; separate
 * handles are independent.  Calls are synchronous.
 *
 * Reference citations are to btsun/kclistpp.
 */
#ifndef KCLPP_H
#define KCLPP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kcl_status {
  KCL_OK = 0,
  KCL_E_IO = 1,          /* a file cannot be opened, read or written */
  KCL_E_PARSE = 2,       /* malformed input */
  KCL_E_INVALID_ARG = 3, /* NULL handle, k out of range, self-loop, ... */
  KCL_E_OOM_HOST = 4,
  KCL_E_OOM_DEVICE = 5,
  KCL_E_HIP = 6,         /* HIP runtime failure or no usable GPU */
  KCL_E_OVERFLOW = 7,    /* a count reached 2^53, where the reference's
                            double counters stop being exact */
  KCL_E_UNSUPPORTED = 8  /* input outside the library's limits */
} kcl_status;

/* Task of one enumeration pass (kCDensestCmpSync.c:128). */
typedef enum kcl_task {
  KCL_FRANK_WOLFE = 2,     /* credit the member of smallest (rho, id) */
  KCL_PAVA_PREPROCESS = 3  /* credit the member of largest level */
} kcl_task;

const char *kcl_last_error(void);
const char *kcl_status_name(kcl_status s);

/* Largest k the GPU enumerator is compiled for.  The reference takes any
 * k in 2..255; here k above KCL_MAX_K is accepted when the graph cannot
 * hold a k-clique (k > core + 1: every count is zero) and otherwise
 * returns KCL_E_UNSUPPORTED. */
#define KCL_MAX_K 10

/* ---------------------------------------------------------------------
 * Edge lists.  Replaces ReadEdgeList (Graph.h:34, Graph.c:34-53): text,
 * two unsigned integers per edge, read until the first token that does
 * not parse; n_in = max id + 1.  *pairs is allocated by the library
 * (2*m uint32, s0 t0 s1 t1 ...), released with kcl_free.  A path ending
 * in ".u32" is read as raw little-endian uint32 pairs instead.
 * ------------------------------------------------------------------- */
kcl_status kcl_read_edge_list(const char *path, uint32_t **pairs, uint64_t *m,
                              uint32_t *n_in);
kcl_status kcl_write_edge_list(const char *path, const uint32_t *pairs, uint64_t m);
void kcl_free(void *p);

/* ---------------------------------------------------------------------
 * Graph.  Replaces SortByCore + Relabel + MakeGraph (Graph.h:38-46,
 * Graph.c:57-156): degeneracy ranks from the reference's heap order
 * (host), relabel + orientation s > t + sorted DAG CSR (device).
 * The input must be a simple graph (no self-loops or repeated pairs), as
 * the reference assumes; self-loops are rejected, repeats are not checked.
 * n_in = 0 means max id + 1.  Limits: m < 2^32.
 * ------------------------------------------------------------------- */
typedef struct kcl_graph kcl_graph;

typedef struct kcl_graph_info {
  uint32_t n_in; /* ids in the input (max id + 1) */
  uint32_t n;    /* relabelled nodes: max rank over edge endpoints + 1 */
  uint64_t m;    /* edges */
  uint32_t core; /* max out-degree of the DAG ("core value", Graph.c:139) */
} kcl_graph_info;

kcl_status kcl_graph_create(const uint32_t *pairs, uint64_t m, uint32_t n_in, kcl_graph **out);
void kcl_graph_destroy(kcl_graph *g); /* FreeGraph (Graph.h:28) */
kcl_status kcl_graph_get_info(const kcl_graph *g, kcl_graph_info *info);
/* rank[v] for v < n_in (Graph.c:82). */
kcl_status kcl_graph_rank(const kcl_graph *g, uint32_t *rank);
/* cd[n+1], adj[m]: out-neighbours ascending (Graph.c:131-145). */
kcl_status kcl_graph_csr(const kcl_graph *g, uint32_t *cd, uint32_t *adj);
/* st[2m]: edges sorted by (s, t) (Graph.c:125). */
kcl_status kcl_graph_edges(const kcl_graph *g, uint32_t *st);
/* Roots with more DAG out-neighbours than max_vertex_degree (1..1024,
 * default 1024) are enumerated as edge tasks: one task per out-edge (s, u)
 * over N+(s) & N+(u), the reference's own root unit (kCDensestCmpSync.c:
 * 300-305).  Above 1024 there is no vertex task, so this is also how graphs
 * whose degeneracy exceeds 1024 run (each common out-neighbourhood must stay
 * <= 1024).  Changes timings only; plans made after the call use it. */
kcl_status kcl_graph_set_edge_split(kcl_graph *g, uint32_t max_vertex_degree);
/* Counting of roots with dense out-neighbourhoods (k >= 6): a root whose
 * expected common out-set per edge item, rho^2 d / 3 (d = d+(s), rho the
 * edge density of G[N+(s)]), is at least min_items is counted a wavefront
 * per item instead of a lane per item.  Roots of more than 448 members are
 * always counted that way.  0 = every root; default 64.  Changes timings
 * only (no reference counterpart); passes launched after the call use it. */
kcl_status kcl_graph_set_dense_threshold(kcl_graph *g, float min_items);
/* CountEdges (Graph.h:51, Graph.c:158-166): edges with both ends in nodes[]. */
kcl_status kcl_count_edges(const kcl_graph *g, uint32_t num_nodes, const uint32_t *nodes,
                           uint32_t *m_out);

/* ---------------------------------------------------------------------
 * One enumeration pass with host vectors.  Replaces
 * CDF_CliqueEnum(Graph *g, unsigned char k, task_t task)
 * (kCDensestCmpSync.c:298-325) including the globals it reads and writes:
 *   KCL_FRANK_WOLFE:     reads rho[n], acc[w] += 1.0 per clique
 *                        (rho_add, kCDensestCmpSync.c:144-145)
 *   KCL_PAVA_PREPROCESS: reads level[n], acc[level[w]] += 1.0
 *                        (rho_tentative, kCDensestCmpSync.c:239-240)
 * acc is accumulated into, like the reference's (callers zero it).
 * visits (nullable) receives the number of k-cliques.  2 <= k <= 255.
 * ------------------------------------------------------------------- */
kcl_status kcl_clique_enum(kcl_graph *g, uint32_t k, kcl_task task, const double *rho,
                           const uint32_t *level, double *acc, uint64_t *visits);

/* The same pass without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): every intermediate
 * (keys, counts, the task list, tickets) lives in one static device array
 * of KCL_NOALLOC_SCRATCH_BYTES per GPU, held from library load until the
 * process ends, and the call makes no hipMalloc/hipFree or other allocating
 * call.  Arguments and results as kcl_clique_enum, with these limits:
 * 24 bytes per node plus 8 per task must fit the scratch (about 2 million
 * nodes), no root may need edge tasks (core <= the graph's
 * kcl_graph_set_edge_split bound), and for KCL_PAVA_PREPROCESS level must
 * be a permutation of 0..n-1 (as the reference's level always is).
 * Sizes beyond the scratch return KCL_E_UNSUPPORTED.  Calls on one GPU
 * share its scratch: calls that overlap on one GPU (two host threads, or
 * another call before one returns) race and are not supported.  The
 * scratch keeps the last call's contents; each call clears what it reads. */
#define KCL_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
kcl_status kcl_clique_enum_noalloc(kcl_graph *g, uint32_t k, kcl_task task, const double *rho,
                                   const uint32_t *level, double *acc, uint64_t *visits);

/* ---------------------------------------------------------------------
 * Device-resident Sim-kClist++ state: rho (fp64), the per-iteration
 * counts (uint64) and the exact totals V (uint64, rho_t = 2 V_t /
 * ((t+1)(t+2))) stay in HBM across iterations.  Replaces AllocCdf +
 * the loop body of CDF_Main (kCDensestCmpSync.c:130-136, 385-413).
 * ------------------------------------------------------------------- */
typedef struct kcl_sim kcl_sim;

/* How a pass orders and splits its roots.  The schedule changes only the
 * time of a pass, never its counts.  The reference's counterpart is the
 * OpenMP dynamic schedule over edges (kCDensestCmpSync.c:300-305). */
typedef enum kcl_schedule {
  KCL_SCHED_COUNTED = 0, /* cost from a root-credit (k-1)-clique counting
                            pass (steps x row words + build entries) */
  KCL_SCHED_LEARNED = 1, /* cost predicted by a kcl_cost_model from eight
                            per-root features (falls back to
                            KCL_SCHED_DEGREE when no model is given) */
  KCL_SCHED_DEGREE = 2   /* plain: cost d+(s)^2, heavy first */
} kcl_schedule;

typedef struct kcl_cost_model kcl_cost_model;

typedef struct kcl_sim_opts {
  uint32_t schedule;      /* kcl_schedule */
  uint32_t split_work;    /* COUNTED/DEGREE: work units per split part; 0 = default */
  const kcl_cost_model *model; /* LEARNED: the model (nullable; read only while
                                  the plan is built, in kcl_sim_create) */
  uint32_t reserved[4];
} kcl_sim_opts;

kcl_status kcl_sim_create(kcl_graph *g, uint32_t k, const kcl_sim_opts *opts, kcl_sim **out);
void kcl_sim_destroy(kcl_sim *s);
/* Iteration t >= 1: FW pass, then rho = (t*rho + 2*cnt)/(t + 2)
 * (kCDensestCmpSync.c:399-403), fp64 in the reference's operation order. */
kcl_status kcl_sim_step(kcl_sim *s, uint32_t t);
/* Counts of the last pass (n uint64; rho_add of kCDensestCmpSync.c:401). */
kcl_status kcl_sim_counts(kcl_sim *s, uint64_t *out);
kcl_status kcl_sim_rho(kcl_sim *s, double *out);
kcl_status kcl_sim_set_rho(kcl_sim *s, const double *rho);
kcl_status kcl_sim_totals(kcl_sim *s, uint64_t *V);
/* k-cliques visited by the last pass and its GPU time (HIP events on the
 * pass's stream, enumeration kernels only / whole pass). */
kcl_status kcl_sim_last_pass(kcl_sim *s, uint64_t *visits, double *kernel_ms, double *pass_ms);

/* The enumeration schedule, fixed per graph and k.  Roots are DAG nodes
 * with out-degree >= k-1; those with no k-clique are dropped at set-up by
 * a counting pass (exact); heavy ones are split into several tasks. */
typedef struct kcl_plan_info {
  uint32_t roots_total; /* nodes with out-degree >= k-1 */
  uint32_t roots_kept;  /* of those, roots of at least one k-clique */
  uint32_t tasks;       /* wavefront tasks per pass (after splitting) */
  uint32_t reserved;
  uint64_t cliques;     /* k-cliques in the graph (= visits of every pass) */
  double plan_ms;       /* set-up time of the plan (counting passes included) */
} kcl_plan_info;

kcl_status kcl_sim_plan_info(kcl_sim *s, kcl_plan_info *info);
/* k-cliques per root (n entries): the cliques whose largest id is that
 * node; zero for nodes the plan does not schedule. */
kcl_status kcl_sim_root_cliques(kcl_sim *s, uint64_t *out);

/* One Frank-Wolfe pass at the current rho with per-task timing, leaving the
 * state unchanged.  *ntasks: in = room in the arrays, out = tasks of the
 * plan (all arrays NULL: size query).  Per task (plan order, nullable):
 * root node, part, parts of that root, shader-clock cycles until the root's
 * rows were built, cycles of the whole task. */
kcl_status kcl_sim_profile_pass(kcl_sim *s, uint32_t *ntasks, uint32_t *root, uint32_t *part,
                                uint32_t *nparts, uint64_t *build_cycles, uint64_t *total_cycles);

/* The launch shape of the plan's word classes (no counterpart in the
 * reference, whose OpenMP loop has one kind of task).  Entries 0..8 are
 * the vertex-task classes, 9..17 the edge-task classes (run by the (k-1)
 * kernels); a class's tasks are the profile arrays' entries
 * [first, first + count).  Used to price a pass's measured task cycles
 * against the GPU's resident workgroups (a makespan lower bound). */
#define KCL_NUM_CLASSES 9
typedef struct kcl_class_info {
  uint32_t words;     /* 64-bit words per bitmap row (members <= 64 words) */
  uint32_t k;         /* clique size of the kernel (k, or k-1 for edge tasks) */
  uint32_t first;     /* first task of the class in plan order */
  uint32_t count;     /* tasks of the class */
  uint32_t waves;     /* wavefronts per workgroup */
  uint32_t per_cu;    /* resident workgroups per CU the launch assumes */
  uint32_t blocks;    /* workgroups launched (min(CUs x per_cu, count)) */
  uint32_t lds_bytes; /* LDS per workgroup */
} kcl_class_info;
kcl_status kcl_sim_plan_classes(kcl_sim *s, kcl_class_info out[2 * KCL_NUM_CLASSES]);

/* ---------------------------------------------------------------------
 * Learned root-cost model (no counterpart in the reference, whose OpenMP
 * loop has no cost estimate).  A BF16 MLP 8 -> 64 -> 64 -> 2 (ReLU) run
 * and trained on the GPU by the library's own MFMA kernels (fp32 master
 * weights, fp32 accumulation, Adam).  Inputs are eight features of a root
 * s (kcl_sim_root_features); outputs are log2(1 + cycles) of building the
 * root's rows and of counting its cliques, which the LEARNED schedule
 * turns into task order and split counts.
 *
 * Parameter vector (KCL_CM_PARAMS floats): W1[8][64], b1[64], W2[64][64],
 * b2[64], W3[64][2], b3[2] (row = input unit).  Normalisation
 * (KCL_CM_NORM floats): feature mean[8], std[8], target mean[2], std[2];
 * the network sees (x - mean)/std and predicts (y - mean)/std.
 * ------------------------------------------------------------------- */
#define KCL_CM_FEATURES 8
#define KCL_CM_HIDDEN 64
#define KCL_CM_OUTPUTS 2
#define KCL_CM_PARAMS 4866
#define KCL_CM_NORM 20

/* Fresh model: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from a
 * splitmix64 stream of `seed`, zero biases, identity normalisation. */
kcl_status kcl_cost_model_create(uint64_t seed, kcl_cost_model **out);
void kcl_cost_model_destroy(kcl_cost_model *m);
/* Binary file: "KCLCM001", then KCL_CM_NORM and KCL_CM_PARAMS
 * little-endian floats. */
kcl_status kcl_cost_model_load(const char *path, kcl_cost_model **out);
kcl_status kcl_cost_model_save(const kcl_cost_model *m, const char *path);
kcl_status kcl_cost_model_get(const kcl_cost_model *m, float *params, float *norm);
kcl_status kcl_cost_model_set(kcl_cost_model *m, const float *params, const float *norm);
/* y[rows][2] = de-normalised predictions for x[rows][8] (GPU). */
kcl_status kcl_cost_model_predict(kcl_cost_model *m, const float *x, uint64_t rows, float *y);
/* Loss L = sum over rows and outputs of (out - y_norm)^2 / (2 rows) and its
 * gradient with respect to the parameter vector (GPU forward + backward). */
kcl_status kcl_cost_model_gradient(kcl_cost_model *m, const float *x, const float *y,
                                   uint64_t rows, float *loss, float *grad);

typedef struct kcl_cm_train_opts {
  uint32_t epochs;   /* 0 = 60 */
  uint32_t batch;    /* rows per step, 0 = 2048 */
  float lr;          /* Adam step, 0 = 2e-3 */
  uint32_t fit_norm; /* 1: set the normalisation from x and y first */
  uint64_t seed;     /* row shuffle */
  uint32_t reserved[4];
} kcl_cm_train_opts;

/* Adam on minibatches of shuffled rows (GPU).  loss_per_epoch (nullable,
 * room for epochs floats) receives each epoch's mean loss. */
kcl_status kcl_cost_model_train(kcl_cost_model *m, const float *x, const float *y, uint64_t rows,
                                const kcl_cm_train_opts *opts, float *loss_per_epoch);

/* Features of every node as a root of this sim's k (n rows of 8 floats):
 * log2(1+d+), log2(1+core), log2(1 + sum of the out-neighbours' d+),
 * log2(1 + edges inside N+(s)), edges inside N+(s) / C(d+, 2),
 * log2(1 + largest out-neighbour d+), log2(row words), k. */
kcl_status kcl_sim_root_features(kcl_sim *s, float *x);
/* Training rows from one timed pass (kcl_sim_profile_pass) of the current
 * plan: one row per scheduled root, with its features and labels
 * log2(1 + build cycles) (mean over the root's parts) and
 * log2(1 + counting cycles) (summed over its parts).  *nrows: in = room,
 * out = rows (all arrays NULL: size query). */
kcl_status kcl_sim_cost_samples(kcl_sim *s, uint32_t *nrows, uint32_t *root, float *x, float *y);
/* Rebuild the plan with another schedule (model: for KCL_SCHED_LEARNED). */
kcl_status kcl_sim_replan(kcl_sim *s, uint32_t schedule, const kcl_cost_model *model);

typedef struct kcl_densest_info {
  uint32_t n;       /* nodes of the densest prefix */
  uint32_t m;       /* edges inside it */
  double density;   /* k-cliques / n as the reference computes it */
  double ub;        /* the reference's upper bound */
  uint64_t sum;     /* exact k-clique count of the prefix (density = sum/n) */
} kcl_densest_info;

/* CDF_FindDensestSubset (kCDensestCmpSync.c:342-383): level order by
 * floor(rho) descending then id, PAVA-preprocess pass, densest prefix,
 * CountEdges, upper bound.  reordered/level/tent (n each) are nullable. */
kcl_status kcl_sim_find_densest(kcl_sim *s, kcl_densest_info *out, uint32_t *reordered,
                                uint32_t *level, uint64_t *tent);

/* ---------------------------------------------------------------------
 * Whole Sim-kClist++ run: CDF_Main (kCDensestCmpSync.c:385-419).
 * Iterations run up to the largest power of two <= num_iter, with one
 * record per power of two.  stat_path (nullable) receives the reference's
 * file: header + "%u\t%u\t%u\t%.12f\t%.12f\t%ld\n" (T, n, m, density, ub,
 * whole seconds since t0_unix).  records (nullable) gets up to 32 rows.
 * ------------------------------------------------------------------- */
typedef struct kcl_record {
  uint32_t T;
  uint32_t n;
  uint32_t m;
  double density;
  double ub;
  uint64_t sum;
  double seconds;
} kcl_record;

kcl_status kcl_cmpsync_run(kcl_graph *g, uint32_t k, uint32_t num_iter, const kcl_sim_opts *opts,
                           const char *stat_path, int64_t t0_unix, kcl_record *records,
                           uint32_t *nrec);

/* ---------------------------------------------------------------------
 * Batched Sim-kClist++: many graphs, each with its own k, iterated
 * together (no counterpart in the reference, which runs one graph per
 * process; per graph the results are exactly those of kcl_sim /
 * kcl_cmpsync_run on that graph alone).  The graphs are copied into one
 * disjoint union in HBM; one iteration of every graph is one replay of a
 * captured hipGraph (the enumeration launches of all distinct k plus one
 * rho update).  The graph handles may be destroyed after create.
 * ------------------------------------------------------------------- */
typedef struct kcl_batch kcl_batch;

typedef struct kcl_batch_opts {
  uint32_t schedule;            /* kcl_schedule of every plan */
  uint32_t split_work;          /* as kcl_sim_opts */
  const kcl_cost_model *model;  /* for KCL_SCHED_LEARNED (nullable) */
  uint32_t tune;      /* 1: pick the row-width floor and the resident
                         workgroups per CU by timing one pass of each */
  uint32_t row_bits;  /* tune = 0: bitmap row width floor 64/128/256 (0 = 64) */
  uint32_t per_cu;    /* tune = 0: workgroups per CU cap (0 = LDS-limited) */
  uint32_t reserved[3];
} kcl_batch_opts;

typedef struct kcl_batch_info {
  uint32_t graphs;
  uint32_t distinct_k;
  uint64_t nodes;     /* relabelled nodes over all graphs */
  uint64_t edges;
  uint64_t cliques;   /* k-cliques over all graphs (each with its own k) */
  uint32_t tasks;     /* workgroup tasks per iteration */
  uint32_t row_bits;  /* chosen row-width floor (tune = 1: each k's plan is tuned on its
                         own; this is the choice for the plan whose pass is longest) */
  uint32_t per_cu;    /* chosen cap (0: LDS-limited), the same plan's */
  uint32_t reserved;
  double plan_ms;     /* union + plans */
  double tune_ms;     /* autotuning */
} kcl_batch_info;

kcl_status kcl_batch_create(kcl_graph *const *graphs, const uint32_t *ks, uint32_t count,
                            const kcl_batch_opts *opts, kcl_batch **out);
void kcl_batch_destroy(kcl_batch *b);
kcl_status kcl_batch_get_info(const kcl_batch *b, kcl_batch_info *info);
/* k-cliques per graph (count entries). */
kcl_status kcl_batch_graph_cliques(const kcl_batch *b, uint64_t *out);
/* Iteration t >= 1 of every graph (one hipGraph replay). */
kcl_status kcl_batch_step(kcl_batch *b, uint32_t t);
/* Back to the state after create: rho, counts and totals zero. */
kcl_status kcl_batch_reset(kcl_batch *b);
/* Visits of the last iteration (all graphs) and its GPU time. */
kcl_status kcl_batch_last_pass(const kcl_batch *b, uint64_t *visits, double *ms);
/* Per graph i (n_i entries): last counts, rho, exact totals V. */
kcl_status kcl_batch_counts(kcl_batch *b, uint32_t i, uint64_t *out);
kcl_status kcl_batch_rho(kcl_batch *b, uint32_t i, double *out);
kcl_status kcl_batch_totals(kcl_batch *b, uint32_t i, uint64_t *out);
/* CDF_FindDensestSubset of every graph (count entries). */
kcl_status kcl_batch_find_densest(kcl_batch *b, kcl_densest_info *out);
/* CDF_Main of every graph: iterations 1 .. the largest power of two <=
 * num_iter from the current state (a fresh or reset batch gives the
 * reference's run), a record per power of two; records: count x 32,
 * nrec: count. */
kcl_status kcl_batch_run(kcl_batch *b, uint32_t num_iter, kcl_record *records, uint32_t *nrec);

/* ---------------------------------------------------------------------
 * k-clique core trimming: kClistCoreTrim (kClistCoreTrim.c:707-775).
 * (k-1)-core by the reference's degree heap (nodes popped while the running
 * core is below k-1 are dropped; kept nodes get new ids n2-1, n2-2, ... in
 * pop order), k-clique degree of every kept node (GPU), then the k-clique
 * core peel by the reference's min-heap; the trimmed graph keeps the last
 * ntrim peeled nodes.  3 <= k <= 255 (k <= KCL_MAX_K when cliques exist).
 * ------------------------------------------------------------------- */
typedef struct kcl_trim kcl_trim;

typedef struct kcl_trim_info {
  uint32_t n_in;    /* input nodes (max id + 1) */
  uint32_t n2;      /* nodes of the (k-1)-core */
  uint32_t e2;      /* edges among them */
  uint32_t core;    /* degeneracy of the input */
  uint64_t ncktot;  /* k-cliques of the (k-1)-core */
  uint64_t ckcore;  /* k-clique core number */
  double ckrho;     /* best k-clique density along the peel */
  double rho;       /* edge density of that subgraph (reference's formula) */
  uint32_t size;    /* its size */
  uint32_t ntrim;   /* nodes kept by the trim */
  uint64_t ncktrim; /* k-cliques of the trimmed graph */
  uint32_t mtrim;   /* edges of the trimmed graph */
  uint32_t gpu_pops; /* peel pops whose cliques were counted on the GPU */
  double seconds_degrees; /* core + relabel + GPU k-clique degrees */
  double seconds_peel;    /* the peel (host heap, heavy pops on the GPU) */
} kcl_trim_info;

kcl_status kcl_trim_create(const uint32_t *pairs, uint64_t m, uint32_t n_in, uint32_t k,
                           kcl_trim **out);
/* The same with options (no counterpart in the reference, whose peel is one
 * host thread).  The peel's heap stays on the host (its tie order defines
 * ckcore, ckdens and the trimmed set); a popped node that still lies in at
 * least gpu_peel_min_cliques alive k-cliques (0 = 2^16) has them counted
 * on the GPU - its alive neighbours as one root's members, every member
 * credited - instead of by the host's enumeration (oneshortpass,
 * kClistCoreTrim.c:577-603).  Results do not depend on it. */
typedef struct kcl_trim_opts {
  uint64_t gpu_peel_min_cliques;
  uint32_t reserved[4];
} kcl_trim_opts;
kcl_status kcl_trim_create_opts(const uint32_t *pairs, uint64_t m, uint32_t n_in, uint32_t k,
                                const kcl_trim_opts *opts, kcl_trim **out);
void kcl_trim_destroy(kcl_trim *t);
kcl_status kcl_trim_get_info(const kcl_trim *t, kcl_trim_info *info);
/* Per new id (n2 entries each, every pointer nullable): input id, k-clique
 * degree (ckdeg), peel order and running core (ckcore), kept flag. */
kcl_status kcl_trim_nodes(const kcl_trim *t, uint32_t *map, uint64_t *nck, uint32_t *peel_key,
                          uint64_t *peel_core, uint8_t *kept);
/* Trimmed edges (2*mtrim input ids, in input order, larger new id first). */
kcl_status kcl_trim_edges(const kcl_trim *t, uint32_t *pairs);
/* The reference's five output files (each path nullable): ckdeg, ckcore,
 * ckdens, trimmed_<tag>_<k>.txt and stat_trim_<tag>_<k>.txt. */
kcl_status kcl_trim_write(const kcl_trim *t, const char *ckdeg, const char *ckcore,
                          const char *ckdens, const char *trimmed, const char *stat,
                          int64_t t0_unix);

/* ---------------------------------------------------------------------
 * Exact k-clique densest subgraph: the result of kCDensestMem
 * (kCDensestMem.c:480-545), computed without storing every clique:
 * Sim-kClist++ rounds on the GPU (1, 2, 4, ... iterations), the top PAVA
 * block (kCDensestMem.c:329-364) as the candidate of density D = p/q, the
 * D-core by k-clique degree as the region that must contain every densest
 * subgraph.  The candidate is stable when the region equals it (the
 * storage-free counterpart of CDF_CheckStability, kCDensestMem.c:366-403);
 * a stable candidate is first tested by the reference's closed-form
 * criterion A (CDF_CheckDensestGoldberg, :405-432) over the allocation
 * V_T / (T(T+3)/2), then the region's cliques go to a host max-flow of
 * Goldberg's network (criterion B, :434-467; from round T = 8 on also for
 * candidates that are not stable) that either certifies D - returning the
 * maximal densest subgraph - or yields a denser candidate.
 * Rounds stop at max_iter iterations; regions whose clique count exceeds
 * max_network_cliques are not sent to the max-flow.  Without a certificate
 * the result is the densest subgraph found, with upper_bound = the largest
 * average of the top-i normalised rho values (rho_T / (T(T+3)/((T+1)(T+2)))
 * is a convex combination of clique-to-member allocations, so every set S
 * has cliques(S) <= sum over S of it).
 * stat_path (nullable): the reference's stat_exact file: a row per round,
 * "%u\t%u\t%.12f\t%ld\t%s" with Info NOT STABLE (the region is larger than
 * the candidate), STABLE BUT NOT DENSEST (it equals the candidate and
 * neither criterion certified it) or OVER BUDGET (the candidate alone holds
 * more than max_network_cliques cliques: not tested; the reference, which
 * stores every clique, has no such row), then the final record with
 * Stopping Condition Goldberg or Max Flow.  nodes_out (nullable, room for n
 * entries of the graph): input ids of the subgraph, ascending.
 * ------------------------------------------------------------------- */
typedef struct kcl_exact_opts {
  uint32_t max_iter;             /* 0 = 65536 */
  uint32_t max_seconds;          /* wall-clock budget, 0 = none */
  uint64_t max_network_cliques;  /* 0 = 100,000,000 (about (20 k + 8) bytes of
                                    host memory per clique; at most
                                    (2^32 - 1) / k) */
  uint32_t schedule;             /* kcl_schedule of the Sim iterations */
  uint32_t gpu_list_min;         /* a certificate region holding at least this many
                                    cliques has them listed on the GPU, a smaller
                                    one on the host; 0 = 1,048,576 */
  const kcl_cost_model *model;   /* for KCL_SCHED_LEARNED (nullable) */
  uint32_t reserved[2];
} kcl_exact_opts;

typedef struct kcl_exact_result {
  uint32_t certified;      /* 1: optimality proved by the max-flow */
  uint32_t num_iter;       /* Sim iterations run */
  uint32_t n;              /* nodes of the (maximal) densest subgraph */
  uint32_t m;              /* its edges */
  uint64_t cliques;        /* its k-cliques: density = cliques / n exactly */
  double density;
  uint32_t max_flow_calls;
  uint32_t region;         /* nodes of the last region */
  double seconds;
  double upper_bound;      /* bound on the maximum density from the last
                              allocation (equals density when certified) */
  uint32_t stop;           /* certified by: 1 criterion A (Goldberg),
                              2 max-flow; 0 not certified */
  uint32_t degree_passes;  /* k-clique degree passes of the region tests */
  double host_seconds;     /* host time of the certificates (clique
                              listing, max-flow, criterion A) */
} kcl_exact_result;

kcl_status kcl_exact_run(kcl_graph *g, uint32_t k, const kcl_exact_opts *opts,
                         const char *stat_path, int64_t t0_unix, kcl_exact_result *res,
                         uint32_t *nodes_out);

/* ---------------------------------------------------------------------
 * Sampling++ (kCDensestSamplingKclistpp.c, "Seq-Sampling++"): replaces
 * SampleCliques + Solve (:438-489).  The count is the enumeration's; each
 * k-clique is kept with probability p = min(1, target / count) by a seeded
 * hash of its ids (the reference uses rand() seeded with the clock, :492),
 * so a (graph, k, target, seed) gives one kept set on any schedule.  The
 * "++" rounds are synchronous: in a round every kept clique gives one unit
 * to its member of smallest load after the previous round, ties to the
 * smaller id (the reference runs them one clique at a time in a shuffled
 * order, :345-359).  Reports follow ExtractDensest and the line printed at
 * :475-479: the densest prefix by decreasing load, its induced subgraph
 * rebuilt and its k-cliques counted exactly, the scaled upper bound.
 * ------------------------------------------------------------------- */
typedef struct kcl_sampling kcl_sampling;

typedef struct kcl_sampling_opts {
  uint64_t target; /* cliques to keep on average; 0 = 1e7 (CLIQUES_TO_SAMPLE, :138) */
  uint64_t seed;   /* the sampler's seed */
  uint32_t reserved[4];
} kcl_sampling_opts;

typedef struct kcl_sampling_info {
  uint64_t total;   /* k-cliques of the graph */
  uint64_t sampled; /* cliques kept */
  double prob;      /* keep probability */
  uint32_t rounds;  /* rounds run */
} kcl_sampling_info;

typedef struct kcl_sampling_record {
  uint32_t T;              /* rounds so far */
  uint32_t n;              /* nodes of the approximate densest subgraph (after Relabel) */
  uint32_t m;              /* its edges */
  uint32_t prefix;         /* densest prefix of the load order (ExtractDensest's n) */
  uint64_t prefix_sampled; /* kept cliques credited to the prefix (ExtractDensest's m) */
  uint64_t cliques;        /* k-cliques of the subgraph, exact */
  double edge_density;     /* 2 m / n / (n - 1) */
  double density;          /* cliques / n */
  double ub_sample;        /* ExtractDensest's upper bound over the sample */
  double ub;               /* ub_sample / p / (1 - sqrt(6 ln(graph n) / ub_sample)) */
  double seconds;          /* kcl_sampling_run: since its start */
} kcl_sampling_record;

/* Counts, then lists the kept cliques into device memory. */
kcl_status kcl_sampling_create(kcl_graph *g, uint32_t k, const kcl_sampling_opts *opts,
                               kcl_sampling **out);
void kcl_sampling_destroy(kcl_sampling *s);
kcl_status kcl_sampling_get_info(kcl_sampling *s, kcl_sampling_info *info);
/* The kept cliques (sampled x k ids, each clique ascending; order unspecified). */
kcl_status kcl_sampling_cliques(kcl_sampling *s, uint32_t *out);
/* One synchronous "++" round. */
kcl_status kcl_sampling_step(kcl_sampling *s);
/* Loads per node (n entries): units received so far. */
kcl_status kcl_sampling_loads(kcl_sampling *s, uint64_t *out);
/* The report after the rounds run so far; order (nullable, n entries): the
 * nodes by decreasing load. */
kcl_status kcl_sampling_report(kcl_sampling *s, kcl_sampling_record *rec, uint32_t *order);
/* Solve (:446-489): rounds up to num_iter with a report after rounds 1, 2,
 * 4, ...; *nrec: in = room in recs, out = reports made. */
kcl_status kcl_sampling_run(kcl_sampling *s, uint32_t num_iter, kcl_sampling_record *recs,
                            uint32_t *nrec);

/* ---------------------------------------------------------------------
 * Library information.
 * ------------------------------------------------------------------- */
typedef struct kcl_device_info {
  char name[64];
  char arch[32];
  uint32_t cu_count;
  uint64_t hbm_bytes;
  uint32_t lds_bytes_per_cu;
  uint32_t clock_khz; /* peak shader clock (the shader-clock timers' rate at most) */
} kcl_device_info;

kcl_status kcl_device_get_info(kcl_device_info *info);
const char *kcl_version(void);

#ifdef __cplusplus
}
#endif

#endif /* KCLPP_H */
