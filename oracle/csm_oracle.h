/*
 * csm_oracle.h -- CPU oracle for the ClusterSM subgraph-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libcsm.so, the ClusterSM CLI) never links or calls it.
 *
 * It is a plain-C restatement of the reference pipeline, written
 * independently of the product's C++ planner so that the two check each
 * other:
 *   - query parsing           io.cu:24-108            (readGraphToCPU2)
 *   - decycle                 decycle.cpp:6-134
 *   - tree vertex cover       filter.cu:23-112        (dfs_cvc, getVertexCover)
 *   - clustering + flattening filter.cu:127-401, structure.cu:219-239
 *   - root + BFS edge order   main.cu:144-146, order.cpp:150-205
 *   - NLC candidate predicate filter.cu:403-539, query NLC filter.cu:555-566
 *   - join semantics          join.cu:290-435 (extend / non-tree check,
 *                             injectivity join.cu:152-155)
 * plus the build's own cluster refinement (sound; see oc_refine) and a
 * backtracking matcher that enumerates in the same matching order.
 *
 * Parity pinning of each part is listed in the header of csm_oracle.c.
 */
#ifndef CSM_ORACLE_H
#define CSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAX_QV 32          /* query vertices (reference MAX_VQ, defs.h:11) */
#define OC_MAX_QE 512         /* query edges: >= 32*31/2 */
#define OC_MAX_CLUSTERS 4096
#define OC_MAX_COMPACT 16384

enum {
  OC_OK = 0,
  OC_E_IO = 1,
  OC_E_FORMAT = 2,
  OC_E_RANGE = 3,
  OC_E_QUERY = 4,
  OC_E_NOMEM = 5,
  OC_E_PLAN = 6,
  OC_E_ARG = 7
};

/* Query graph with the reference's file-order CSR and edge ids. */
typedef struct oc_query {
  uint32_t n, m;
  uint32_t label[OC_MAX_QV];
  uint32_t deg[OC_MAX_QV];          /* full-query degree */
  uint32_t off[OC_MAX_QV + 1];
  uint32_t adj[2 * OC_MAX_QE];
  uint32_t eid[2 * OC_MAX_QE];      /* eid of direction (v, adj[off]) */
  uint32_t esrc[OC_MAX_QE], edst[OC_MAX_QE];  /* file edge k = (esrc, edst) */
} oc_query;

/* Data graph: simple, undirected, adjacency sorted ascending. */
typedef struct oc_data {
  uint32_t n;
  uint64_t m;                       /* undirected edges after dedup */
  uint32_t *label;
  uint64_t *off;
  uint32_t *adj;
} oc_data;

typedef struct oc_plan {
  uint32_t n, m;
  /* decycle (decycle.cpp) */
  uint32_t tree_root;
  uint32_t tree_deg[OC_MAX_QV];
  uint32_t tree_off[OC_MAX_QV + 1];
  uint32_t tree_adj[2 * OC_MAX_QV];
  uint32_t n_dropped;
  uint32_t dropped_src[OC_MAX_QE], dropped_dst[OC_MAX_QE];
  /* vertex cover (filter.cu:43-112), in selection order */
  uint32_t n_cover;
  uint32_t cover[OC_MAX_QV];
  /* clusters flattened as encodingMeta::init (structure.cu:219-239) */
  uint32_t n_clusters;
  uint32_t n_compact;
  uint32_t n_layers;
  uint32_t cluster_off[OC_MAX_CLUSTERS + 1];
  uint8_t cluster_valid[OC_MAX_CLUSTERS];
  uint32_t compact[OC_MAX_COMPACT];
  /* root (main.cu:144) and BFS edge order (order.cpp:150-205) */
  uint32_t root;
  uint32_t order_len;
  uint32_t order_src[OC_MAX_QE], order_dst[OC_MAX_QE];  /* (u_matched, u) */
  /* matching order derived from the edge order (join.cu:348-420) */
  uint32_t mvert[OC_MAX_QV];        /* vertex matched at position i */
  uint32_t mpos[OC_MAX_QV];         /* position of vertex u */
  uint32_t parent[OC_MAX_QV];       /* parent position of position i (i>0) */
  uint32_t nback[OC_MAX_QV];
  uint32_t back[OC_MAX_QV][OC_MAX_QV];  /* earlier positions joined by non-tree edges */
} oc_plan;

typedef struct oc_result {
  uint64_t count;      /* min(total, cap) */
  uint32_t capped;     /* 1 if more than cap embeddings exist */
  uint32_t timed_out;  /* 1 if the time budget stopped the run */
  uint64_t listed;     /* rows written to the list buffer */
  double seconds;
} oc_result;

int oc_query_read(const char *path, oc_query *q);
int oc_query_from_arrays(uint32_t n, const uint32_t *labels, uint32_t m,
                         const uint32_t *edges, oc_query *q);

int oc_data_read(const char *path, oc_data **out);
int oc_data_from_arrays(uint32_t n, const uint32_t *labels, uint64_t m,
                        const uint32_t *edges, oc_data **out);
void oc_data_free(oc_data *g);
uint64_t oc_data_num_edges(const oc_data *g);

int oc_plan_build(const oc_query *q, oc_plan *p);

/* Candidate bitmaps, layout [u][ceil(n/64)] u64 words. */
int oc_filter_nlc(const oc_data *g, const oc_query *q, uint64_t *bits);
int oc_refine(const oc_data *g, const oc_plan *p, uint64_t *bits);

/*
 * Enumerate embeddings in matching order.  cap: result cap; the list (if
 * non-null, capacity list_cap rows of n u32 in query-id order) receives the
 * first min(cap, list_cap) embeddings in lexicographic order of the
 * matching-order tuple.  nthreads>1 is only allowed without a list.
 * budget_s <= 0 means no time budget.
 */
int oc_match(const oc_data *g, const oc_query *q, const oc_plan *p,
             uint64_t cap, uint32_t *list, uint64_t list_cap, int nthreads,
             double budget_s, int use_refine, oc_result *res);

#ifdef __cplusplus
}
#endif
#endif
