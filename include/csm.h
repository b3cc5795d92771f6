/*
 * csm.h -- C ABI of the MI355X-native ClusterSM matching path (libcsm.so).
 *
 * The reference has no library API: its boundaries are the ClusterSM CLI
 * (main.cu:24-175) and the host functions that main() calls.  Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *   - every function returns a csm_status; nothing calls exit();
 *   - output buffers are owned by the caller;
 *   - vertex ids are uint32_t, counts uint64_t;
 *   - an embedding is num_v uint32_t in query-vertex-id order, the row
 *     layout of the reference's result table (join.cu:163-165);
 *   - a csm_ctx is used by one host thread at a time, except
 *     csm_serve_submit / csm_serve_poll, which are thread-safe.
 */
#ifndef CSM_H
#define CSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSM_MAX_QV 32        /* reference MAX_VQ (defs.h:11) */
#define CSM_MAX_QE 512       /* reference MAX_EQ is 64 (defs.h:12); 512 >= 32*31/2 */
#define CSM_MAX_CLUSTERS 1024 /* reference MAX_CLUSTERS is 150 (defs.h:15), unchecked */
#define CSM_MAX_COMPACT 4096
#define CSM_NO_PARENT 0xFFFFFFFFu

typedef enum csm_status {
  CSM_OK = 0,
  CSM_E_ARG = 1,        /* invalid argument */
  CSM_E_IO = 2,         /* file cannot be opened (io.cu:29-33 exits 1) */
  CSM_E_FORMAT = 3,     /* malformed t/v/e text (io.cu:36-42 exits 1) */
  CSM_E_QUERY = 4,      /* query outside the envelope: |V|<2 or >32, not simple, disconnected */
  CSM_E_PLAN = 5,       /* planner limits exceeded (clusters, compact size) */
  CSM_E_DEVICE = 6,     /* HIP runtime error (cuchk exits with the code, cuda_helpers.h:16-33) */
  CSM_E_NOMEM = 7,      /* device arena exhausted, or host memory (an entry point that cannot allocate on the host) */
  CSM_E_NO_GPU = 8,     /* no usable gfx950 device */
  CSM_E_BUSY = 9,       /* serving admission queue full */
  CSM_E_NOT_FOUND = 10, /* unknown serving ticket */
  CSM_E_PENDING = 11,   /* serving result not ready yet */
  CSM_E_LIMIT = 12,     /* data graph beyond the int32 CSR layout (2|E| >= 2^32) */
  CSM_E_TIMEOUT = 13,   /* result status only: the time limit ended the search; count is a lower bound */
  CSM_E_BUDGET = 14     /* result status only: the caller's item budget ended the search; count is a lower bound */
} csm_status;

typedef struct csm_host_graph csm_host_graph; /* parsed or generated graph in host memory */
typedef struct csm_ctx csm_ctx;               /* device + arena */
typedef struct csm_graph csm_graph;           /* data graph resident in HBM */
typedef struct csm_query csm_query;           /* prepared query (plan + candidate bitmaps) */

/* ---------------------------------------------------------------- host graph
 * Replaces cpuGraph (cpuGraph.h:13-45) and readGraphToCPU2 (io.cu:24-108):
 * "t N M", "v id label deg" for ids 0..N-1 in order, then one line per
 * undirected edge; any line type other than 'v' is an edge line.  The file
 * order of edges is kept (it fixes the query CSR order, io.cu:86-93).
 * Malformed input returns CSM_E_FORMAT instead of the reference's undefined
 * behaviour (declared degree != edge lines, ids out of order or range). */
csm_status csm_host_graph_read(const char *path, csm_host_graph **out);
csm_status csm_host_graph_from_edges(uint32_t num_v, const uint32_t *labels,
                                     uint64_t num_e, const uint32_t *edges /* 2*num_e, file order */,
                                     csm_host_graph **out);
void csm_host_graph_free(csm_host_graph *g);
csm_status csm_host_graph_info(const csm_host_graph *g, uint32_t *num_v, uint64_t *num_e);
/* Borrowed views, valid until csm_host_graph_free. */
csm_status csm_host_graph_views(const csm_host_graph *g, const uint32_t **labels,
                                const uint32_t **edges);
/* Write in the t/v/e format (declared degree = number of edge lines). */
csm_status csm_host_graph_write(const csm_host_graph *g, const char *path);
/* Binary graph file: labels and the file-order edge list as raw u32 arrays
 * behind a 24-byte header ("CSMGBIN1", u32 N, u32 0, u64 M).  No counterpart
 * in the reference, whose only input is the text format (io.cu:24-108);
 * for graphs too large to re-parse as text on every run.  Reading checks
 * the header, the file size and every endpoint < N (CSM_E_FORMAT). */
csm_status csm_host_graph_write_bin(const csm_host_graph *g, const char *path);
csm_status csm_host_graph_read_bin(const char *path, csm_host_graph **out);
/* Message of the last parser / planner failure on the calling thread. */
const char *csm_host_last_error(void);

/* ------------------------------------------------------------------- planner
 * Replaces copyMeta's query half (io.cu:265-270), decycle (decycle.cpp:6-134),
 * getVertexCover/clustering (filter.cu:23-401), encodingMeta::init
 * (structure.cu:219-239), the root choice (main.cu:144) and getBFSEdgeOrder
 * (order.cpp:150-205).  The matching order is the edge order as the join
 * consumes it (join.cu:322-420): order_src[0] is matched first, an edge whose
 * second endpoint is unmatched extends the embedding, any other edge is a
 * non-tree ("decycle") check, attached to its later-matched endpoint. */
typedef struct csm_query_plan {
  uint32_t num_v, num_e;
  uint32_t labels[CSM_MAX_QV];
  uint32_t degree[CSM_MAX_QV];                       /* full-query degree */
  uint32_t edge_src[CSM_MAX_QE], edge_dst[CSM_MAX_QE]; /* file order */
  uint32_t tree_root;
  uint32_t tree_degree[CSM_MAX_QV];
  uint32_t tree_offsets[CSM_MAX_QV + 1];
  uint32_t tree_adj[2 * CSM_MAX_QV];
  uint32_t cover_size;
  uint32_t cover[CSM_MAX_QV];                        /* selection order */
  uint32_t num_clusters, num_cluster_us, num_layers;
  uint32_t cluster_offsets[CSM_MAX_CLUSTERS + 1];
  uint32_t cluster_us[CSM_MAX_COMPACT];              /* query_us_compact_ */
  uint8_t cluster_valid[CSM_MAX_CLUSTERS];
  uint32_t root;
  uint32_t order_len;
  uint32_t order_src[CSM_MAX_QE], order_dst[CSM_MAX_QE]; /* (u_matched, u) */
  uint32_t match_vertex[CSM_MAX_QV];                 /* vertex matched at position i */
  uint32_t match_parent[CSM_MAX_QV];                 /* parent position, CSM_NO_PARENT at 0 */
  uint32_t num_back[CSM_MAX_QV];
  uint32_t back_pos[CSM_MAX_QV][CSM_MAX_QV];         /* earlier positions with a non-tree edge */
} csm_query_plan;

csm_status csm_plan_build(const csm_host_graph *query, csm_query_plan *out);

/* ------------------------------------------------------------ device context
 * Replaces device selection (main.cu:36-43) and the per-step cudaMalloc /
 * cudaMemset churn (join.cu:234-264,313-319; memManag.cuh:9-16) with one
 * device arena.  arena_bytes = 0 picks 85% of the free HBM. */
csm_status csm_open(int device, uint64_t arena_bytes, csm_ctx **out);
csm_status csm_close(csm_ctx *ctx);
const char *csm_last_error(const csm_ctx *ctx);
const char *csm_status_string(csm_status s);
typedef struct csm_ctx_info {
  int device;
  uint32_t num_cus;
  uint64_t arena_bytes;
  uint64_t arena_used;
  uint64_t arena_peak;
  char arch[32];
} csm_ctx_info;
csm_status csm_ctx_query(const csm_ctx *ctx, csm_ctx_info *out);

/* -------------------------------------------------------------- data graph
 * Replaces gpuGraph + allocateMemGPU/copyGraphToGPU (gpuGraph.h:9-22,
 * io.cu:220-249).  The CSR is built on the device from the edge list: both
 * directions, radix sort, self-loops and duplicates removed, adjacency sorted
 * ascending, uint32 offsets. */
csm_status csm_graph_upload(csm_ctx *ctx, const csm_host_graph *g, csm_graph **out);
csm_status csm_graph_upload_edges(csm_ctx *ctx, uint32_t num_v, const uint32_t *labels,
                                  uint64_t num_e, const uint32_t *edges, csm_graph **out);
csm_status csm_graph_free(csm_ctx *ctx, csm_graph *g);
csm_status csm_graph_info(const csm_graph *g, uint32_t *num_v, uint64_t *num_e,
                          uint32_t *max_degree);
/* offsets: num_v+1, adj: 2*num_e, labels: num_v (any may be NULL). */
csm_status csm_graph_download(csm_ctx *ctx, const csm_graph *g, uint32_t *offsets,
                              uint32_t *adj, uint32_t *labels);
/* Host graph (sorted, simple) from the device CSR, for query generation. */
csm_status csm_graph_to_host(csm_ctx *ctx, const csm_graph *g, csm_host_graph **out);

/* ------------------------------------------------------------------- query
 * Replaces copyMeta (io.cu:265-303, __constant__ globals globals.cu:25-38)
 * with a per-query descriptor so that many queries can be resident. */
csm_status csm_query_prepare(csm_ctx *ctx, const csm_graph *g, const csm_query_plan *plan,
                             csm_query **out);
csm_status csm_query_free(csm_ctx *ctx, csm_query *q);

/* ------------------------------------------------------------------ filter
 * Replaces clusterFilter (filter.cuh:130-144, filter.cu:1234-1396): the
 * label / degree / neighbour-label-count predicate (filter.cu:403-539) into
 * per-query-vertex candidate bitmaps of ceil(num_v/64) uint64 words, then
 * (refine != 0) the cluster refinement, in cluster order:
 *   core   Cur(k)   <- { v in Cur(k) : N(v) meets Cur(m) for every member m }
 *   member Cur(m)   <- { w in Cur(m) : N(w) meets Cur(k) }
 * counts: num_v entries, |C(u)| per query vertex (may be NULL). */
csm_status csm_filter(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine,
                      uint64_t *counts);
/* bits: plan.num_v rows of ceil(graph num_v / 64) words. */
csm_status csm_candidates_download(csm_ctx *ctx, const csm_query *q, uint64_t *bits);

/* ------------------------------------------------------------------- match
 * Replaces join (join.cuh:41-55, join.cu:290-435) and the num_res output
 * (main.cu:172).  Semantics: every injective, label-preserving map f with
 * f(u) in C(u) and every query edge mapped to a data edge (non-induced).
 * cap: count = min(total, cap), capped = total > cap.  With a list, the rows
 * are the first min(cap, list_capacity) embeddings in lexicographic order of
 * the matching-order tuple, returned sorted in query-id order. */
typedef struct csm_match_opts {
  uint64_t cap;            /* 0 = unlimited */
  uint32_t refine;         /* run the cluster refinement (default 1) */
  uint32_t reuse_filter;   /* 1: candidates already computed by csm_filter */
  uint32_t *list;          /* NULL = count only */
  uint64_t list_capacity;  /* rows */
  uint64_t chunk_items;    /* 0 = autotuned */
  uint32_t time_limit_ms;  /* 0 = none; else the search stops after this and the result's status is
                              CSM_E_TIMEOUT with the count so far (the call itself returns CSM_OK) */
  uint32_t reorder;        /* 1: candidate-aware matching order after the filter, as csm_query_reorder
                              (same counts; a capped list holds the first embeddings of that order;
                              the query keeps the new order) */
  uint64_t item_budget;    /* 0 = none; else the search stops once about this many items (row x candidate
                              tests) are examined: the result's status is CSM_E_BUDGET with the count so
                              far (the call itself returns CSM_OK) */
  uint32_t symmetry;       /* 1 (count only): count one embedding per Aut(Q) class, as csm_query_symmetry's
                              conditions select; the result's automorphisms holds |Aut(Q)|, so the full
                              count is count * automorphisms (when not capped) */
  uint32_t lds_lists;      /* 1 (count only): the count level's rows look up every anchor's candidate list
                              and copy the short lists that many of their items search into LDS, where the
                              items' membership searches run (default 0: hub bitmaps and adjacency
                              searches, which measured faster on the synthetic workloads) */
} csm_match_opts;

/* Kernel classes for per-kernel accounting. */
enum {
  CSM_K_FILTER = 0,      /* label/degree/NLC filter */
  CSM_K_REFINE = 1,      /* cluster refinement (core + member steps) */
  CSM_K_EXPAND = 2,      /* join extension, survivor marking */
  CSM_K_MATERIALIZE = 3, /* join extension, row write-out */
  CSM_K_COUNT = 4,       /* last level, count only */
  CSM_K_AUX = 5,         /* scans, compaction, chunk selection */
  CSM_K_BATCH = 6,       /* persistent batched kernel */
  CSM_K_NUM = 7
};
typedef struct csm_kernel_stat {
  uint64_t launches;
  uint64_t ns;             /* summed device time (only with csm_set_profiling) */
  uint64_t bytes;          /* algorithmic bytes: the logical reads and writes of the kernels' algorithm */
  uint64_t bytes_ref;      /* join classes: the reference algorithm's bytes for the same rows (extension by
                              a scan of each row's parent adjacency, join.cu:99-177), which the candidate
                              index avoids; the other classes: equal to bytes */
} csm_kernel_stat;

typedef struct csm_result {
  uint64_t count;
  uint32_t capped;
  uint32_t status;         /* csm_status of this query */
  uint64_t listed;
  uint64_t partial_rows;   /* rows materialised over all levels */
  uint64_t items;          /* (row, neighbour) pairs examined by the join */
  double ms_filter;
  double ms_join;
  double ms_total;
  double ms_latency;       /* serving: arrival to result */
  uint32_t path;           /* csm_path: which method produced the result */
  uint32_t anomalies;      /* batch paths: task records the persistent kernel read late or found bad in
                              this call (never expected; any makes the call recount its unfinished
                              queries on the plain path and is also reported on stderr) */
  csm_kernel_stat kernels[CSM_K_NUM];
  uint64_t automorphisms;  /* with the symmetry option: |Aut(Q)| (1 when the search was too costly and
                              every embedding was counted); otherwise 1 */
} csm_result;

/* csm_result.path */
enum {
  CSM_PATH_PLAIN = 0,       /* csm_match */
  CSM_PATH_HOST = 1,        /* answered without a join (two vertices, or an empty candidate set) */
  CSM_PATH_BATCH = 2,       /* the persistent batch kernel */
  CSM_PATH_BATCH_BUDGET = 3,/* not produced any more (an item budget stops a query with CSM_E_BUDGET) */
  CSM_PATH_BATCH_ABORT = 4  /* the batch kernel gave up (time limit, pool): finished by the plain path */
};

/* Per-launch device timing with HIP events on the context stream. */
csm_status csm_set_profiling(csm_ctx *ctx, int on);

/* Diagnostics, no reference counterpart: a streaming copy kernel (16-byte
 * loads and stores) over `bytes` of arena memory, `iters` launches timed with
 * HIP events.  Gives the HBM bandwidth this device sustains and a kernel of
 * known traffic for calibrating hardware byte counters. */
csm_status csm_copy_probe(csm_ctx *ctx, uint64_t bytes, uint32_t iters, double *ms_per_copy);

csm_status csm_match(csm_ctx *ctx, const csm_graph *g, csm_query *q,
                     const csm_match_opts *opts, csm_result *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): csm_match in count-only mode (opts->list must be NULL)
 * whose filter scratch, candidate index and join workspace all live in the
 * library's static device scratch of CSM_NOALLOC_SCRATCH_BYTES per GPU,
 * resolved at csm_open.  It calls no hipMalloc, hipFree or allocating library
 * routine and takes no arena block.  A query whose workspace does not fit
 * returns CSM_E_NOMEM with csm_last_error naming the limit; the context stays
 * usable.  Every call on one GPU shares that GPU's scratch: calls that overlap
 * on one GPU (from several contexts or threads) race and must be serialised by
 * the caller.  Results equal csm_match's. */
#define CSM_NOALLOC_SCRATCH_BYTES (128ull << 20)
csm_status csm_match_noalloc(csm_ctx *ctx, const csm_graph *g, csm_query *q,
                             const csm_match_opts *opts, csm_result *out);

/* Candidate-aware matching order for q, from its filtered candidate counts
 * (the filter runs first if it has not): the root has the fewest candidates
 * per query degree, then each position takes the unmatched vertex with the
 * most matched neighbours, the fewest candidates, the highest degree.  No
 * counterpart in the reference, which always matches in the BFS edge order
 * from the last cluster's core (main.cu:144-146, order.cpp:150-205; its
 * candidate-count choice is commented out at order.cpp:160-173).  Counts are
 * unchanged; a capped list then holds the first embeddings of the new order,
 * which is why the default order stays the reference's. */
csm_status csm_query_reorder(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine);

/* Symmetry-breaking conditions for counting, no counterpart in the reference
 * (which counts every automorphic image), behind an option: the labelled automorphism group Aut(Q) walked as a
 * stabiliser chain, each vertex's orbit under the automorphisms fixing the
 * vertices before it giving f(v) < f(w) for the orbit's other members
 * (Grochow & Kellis, RECOMB 2007).  Exactly one embedding of every Aut(Q)
 * class satisfies the conditions.  pairs (may be NULL): room for
 * 2 * CSM_MAX_QV * CSM_MAX_QV ids, (v, w) per condition.  CSM_E_LIMIT (with
 * *automorphisms = 1 and no pairs) when the search is too costly. */
csm_status csm_query_symmetry(const csm_query_plan *plan, uint64_t *automorphisms, uint32_t *num_pairs,
                              uint32_t *pairs);

/* ----------------------------------------------------------------- batched
 * No reference counterpart (the reference runs one process per query,
 * runClusterSM.sh:9-13).  All queries run in one persistent kernel fed by a
 * device work queue of (query, level, row chunk) tasks.  out: n results. */
typedef struct csm_batch_opts {
  uint64_t cap;            /* per-query cap, 0 = unlimited */
  uint32_t refine;         /* cluster refinement before the join (opts == NULL: off; the
                              candidate-neighbour lists already end dead branches early) */
  uint32_t use_graph;      /* capture the launch sequence in a hipGraph */
  uint32_t chunk_rows;     /* root ids per root task (multiple of 64), 0 = default 1024 */
  uint32_t wgs_per_cu;     /* persistent workgroups per CU (1..8), 0 = default 4 */
  uint64_t pool_bytes;     /* frontier pool, 0 = auto */
  uint32_t block_kb;       /* pool block size in KiB (32..4096), 0 = default (512, smaller when the pool is small) */
  uint32_t max_out;        /* full output blocks a task holds before it yields (1..8), 0 = default 1 */
  uint32_t ilp;            /* items per thread per tile (1..4), 0 = default 4 */
  uint32_t time_limit_ms;  /* 0: none (the kernel still stops after 120 s and the plain path finishes
                              its queries, exact); else per query: a query unfinished after this has
                              status CSM_E_TIMEOUT and its count so far, as with csm_match */
  uint64_t item_budget;    /* 0 = none; else, as in csm_match_opts, a query over about this many items
                              stops with status CSM_E_BUDGET and its count so far */
  uint32_t reorder;        /* 1: candidate-aware matching order for every query, from the batch's own
                              filter counts (as csm_query_reorder; same counts; the queries keep it) */
  uint32_t reserved;
} csm_batch_opts;
csm_status csm_batch_match(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                           const csm_batch_opts *opts, csm_result *out);
/* Autotuner, no reference counterpart: times csm_batch_match on up to
 * `sample` of the queries (evenly spaced; cap, refine and limits from `base`)
 * for a small grid of workgroups per CU, output blocks per task, block size
 * and items per thread, and writes the fastest setting into *out (the other
 * fields copied from `base`).  Results of the trial runs are discarded. */
csm_status csm_batch_autotune(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                              const csm_batch_opts *base, uint32_t sample, csm_batch_opts *out);

/* ----------------------------------------------------------------- serving
 * No reference counterpart (runClusterSM.sh:9-13 runs one process per query).
 * The graph and one persistent join kernel stay resident; requests are
 * admitted in arrival order while query slots are free and their estimated
 * footprints fit the budget, prepared beside the running kernel and handed to
 * it through a doorbell in pinned memory; each result is published as soon
 * as its query's last task ends. */
typedef struct csm_serve_opts {
  uint64_t budget_bytes;   /* admission budget for intermediate results */
  uint32_t max_inflight;   /* query slots of the resident kernel (0 = 256, at most 256) */
  uint32_t queue_capacity; /* pending submissions before CSM_E_BUSY */
  uint64_t cap;            /* per-query cap */
  uint32_t refine;
  uint32_t chunk_rows;
  uint32_t wgs_per_cu;     /* resident join workgroups per CU (1..4), 0 = 2: the rest of each CU runs the
                              arriving queries' filter and index builds beside it */
  uint32_t reserved;
} csm_serve_opts;
csm_status csm_serve_start(csm_ctx *ctx, const csm_graph *g, const csm_serve_opts *opts);
csm_status csm_serve_submit(csm_ctx *ctx, const csm_query_plan *plan, uint64_t *ticket);
csm_status csm_serve_poll(csm_ctx *ctx, uint64_t ticket, int wait, csm_result *out);
csm_status csm_serve_stop(csm_ctx *ctx);

/* -------------------------------------------------------------- generators
 * No reference counterpart: deterministic synthetic inputs (the reference
 * reads external datasets, runClusterSM.sh:8-13).  Labels are drawn per
 * vertex, uniform (zipf_s = 0) or Zipf(zipf_s) over num_labels labels. */
csm_status csm_gen_erdos_renyi(uint32_t num_v, uint64_t num_e, uint32_t num_labels,
                               double zipf_s, uint64_t seed, csm_host_graph **out);
csm_status csm_gen_rmat(uint32_t scale, uint64_t num_samples, double a, double b, double c,
                        uint32_t num_labels, double zipf_s, uint64_t seed, csm_host_graph **out);
csm_status csm_gen_chung_lu(uint32_t num_v, uint64_t num_samples, double exponent,
                            uint32_t max_degree, uint32_t num_labels, double zipf_s,
                            uint64_t seed, csm_host_graph **out);
/* Random-walk query from a uniformly chosen start vertex.  edges_mode:
 *   0 induced (all data edges among the visited vertices)
 *   1 tree (first-visit edges only)
 *   2 tree + extra_edges induced non-tree edges
 *   3 tree + induced edges up to avg_degree * num_v / 2 edges */
csm_status csm_gen_query(const csm_host_graph *data, uint32_t num_v, uint32_t edges_mode,
                         uint32_t extra_edges, double avg_degree, uint64_t seed,
                         csm_host_graph **out);

#ifdef __cplusplus
}
#endif
#endif /* CSM_H */
