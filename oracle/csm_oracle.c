/*
 * csm_oracle.c -- CPU oracle (checker) for the ClusterSM matching path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product never links it.
 *
 * Every function cites the reference file:line it restates.  How each part
 * is pinned (tests/test_oracle.py, tests/test_host.py):
 *   oc_query_read / oc_data_read  bundled example files; declared-degree sum
 *                                 = 2M; eid rule 2k / 2k+1 (io.cu:82-94)
 *   decycle                       the reference's own decycle.cpp compiled
 *                                 into oracle/_ref (bit-equal tree CSR on
 *                                 random queries); invariants |E_T|=|V|-1
 *   vertex cover                  brute-force minimum cover size on random
 *                                 trees; worked example (cover [1,3])
 *   clustering / root / order     seven traces made by hand from the
 *                                 reference source (tests/golden/
 *                                 planner_traces.json and the bundled pair:
 *                                 merges through a core and through shared
 *                                 members, combines that create a cluster or
 *                                 keep a holder, several layers); the product
 *                                 planner (independent C++) agrees with them
 *                                 and on thousands of random queries
 *   oc_filter_nlc                 brute-force predicate per (u, v) in Python
 *   oc_refine                     the spec recomputed with Python sets
 *                                 (exact); soundness against brute force;
 *                                 the worked example's refined sets
 *   oc_match                      brute-force enumeration and networkx
 *                                 monomorphism goldens (tests/golden), worked
 *                                 example count 4 and its sorted list
 */
#include "csm_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* tokenizer for the t/v/e text format                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  const char *p, *end;
} oc_tok;

static int tok_char(oc_tok *t, char *c) {
  while (t->p < t->end && (unsigned char)*t->p <= ' ') t->p++;
  if (t->p >= t->end) return 0;
  *c = *t->p++;
  return 1;
}

static int tok_u64(oc_tok *t, uint64_t *v) {
  while (t->p < t->end && (unsigned char)*t->p <= ' ') t->p++;
  if (t->p >= t->end || *t->p < '0' || *t->p > '9') return 0;
  uint64_t x = 0;
  while (t->p < t->end && *t->p >= '0' && *t->p <= '9') {
    x = x * 10u + (uint64_t)(*t->p - '0');
    t->p++;
  }
  *v = x;
  return 1;
}

static char *slurp(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = (char *)malloc((size_t)n + 1);
  if (!buf) {
    fclose(f);
    return NULL;
  }
  size_t got = fread(buf, 1, (size_t)n, f);
  fclose(f);
  buf[got] = 0;
  *len = got;
  return buf;
}

/*
 * Parse "t N M", then "v id label deg" lines (ids 0..N-1 in order) and edge
 * lines.  As readGraphToCPU2 (io.cu:63-96) any line type other than 'v' is an
 * edge.  Rejected (the reference has undefined behaviour there): vertex lines
 * out of order or after edges, ids out of range, declared degrees that do not
 * match the edge lines, more or fewer than M edge lines.
 * Outputs labels/declared degrees and the edge list in file order.
 */
static int parse_tve(const char *buf, size_t len, uint32_t *n_out, uint64_t *m_out,
                     uint32_t **labels, uint32_t **degs, uint32_t **edges,
                     uint64_t max_n, uint64_t max_m) {
  oc_tok t = {buf, buf + len};
  char c;
  uint64_t n, m;
  if (!tok_char(&t, &c) || c != 't') return OC_E_FORMAT;  /* io.cu:36-42 */
  if (!tok_u64(&t, &n) || !tok_u64(&t, &m)) return OC_E_FORMAT;
  if (n > max_n || m > max_m || n > 0xFFFFFFFFull) return OC_E_RANGE;
  uint32_t *lab = (uint32_t *)calloc(n ? n : 1, 4);
  uint32_t *deg = (uint32_t *)calloc(n ? n : 1, 4);
  uint32_t *cnt = (uint32_t *)calloc(n ? n : 1, 4);
  uint32_t *ed = (uint32_t *)malloc((m ? m : 1) * 8);
  if (!lab || !deg || !cnt || !ed) {
    free(lab); free(deg); free(cnt); free(ed);
    return OC_E_NOMEM;
  }
  uint64_t nv = 0, ne = 0;
  int rc = OC_OK;
  while (tok_char(&t, &c)) {
    if (c == 'v') {
      uint64_t id, l, d;
      if (!tok_u64(&t, &id) || !tok_u64(&t, &l) || !tok_u64(&t, &d)) { rc = OC_E_FORMAT; break; }
      if (ne != 0 || id != nv || id >= n || l > 0xFFFFFFFFull || d > 0xFFFFFFFFull) { rc = OC_E_FORMAT; break; }
      lab[id] = (uint32_t)l;
      deg[id] = (uint32_t)d;
      nv++;
    } else {
      uint64_t s, d;
      if (!tok_u64(&t, &s) || !tok_u64(&t, &d)) { rc = OC_E_FORMAT; break; }
      if (s >= n || d >= n || ne >= m) { rc = OC_E_FORMAT; break; }
      ed[2 * ne] = (uint32_t)s;
      ed[2 * ne + 1] = (uint32_t)d;
      cnt[s]++;
      cnt[d]++;
      ne++;
    }
  }
  if (rc == OC_OK && (nv != n || ne != m)) rc = OC_E_FORMAT;
  if (rc == OC_OK)
    for (uint64_t v = 0; v < n; v++)
      if (cnt[v] != deg[v]) { rc = OC_E_FORMAT; break; }  /* io.cu:75 trusts it */
  free(cnt);
  if (rc != OC_OK) {
    free(lab); free(deg); free(ed);
    return rc;
  }
  *n_out = (uint32_t)n;
  *m_out = m;
  *labels = lab;
  *degs = deg;
  *edges = ed;
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* query graph: file-order CSR with eids (io.cu:82-94)                       */
/* ------------------------------------------------------------------------ */

int oc_query_from_arrays(uint32_t n, const uint32_t *labels, uint32_t m,
                         const uint32_t *edges, oc_query *q) {
  if (n < 2 || n > OC_MAX_QV || m > OC_MAX_QE || m < n - 1) return OC_E_QUERY;
  memset(q, 0, sizeof(*q));
  q->n = n;
  q->m = m;
  uint8_t seen[OC_MAX_QV][OC_MAX_QV];
  memset(seen, 0, sizeof(seen));
  for (uint32_t k = 0; k < m; k++) {
    uint32_t s = edges[2 * k], d = edges[2 * k + 1];
    if (s >= n || d >= n || s == d || seen[s][d]) return OC_E_QUERY;
    seen[s][d] = seen[d][s] = 1;
    q->deg[s]++;
    q->deg[d]++;
    q->esrc[k] = s;
    q->edst[k] = d;
  }
  for (uint32_t v = 0; v < n; v++) {
    q->label[v] = labels[v];
    q->off[v + 1] = q->off[v] + q->deg[v];
  }
  /* placement in file order, io.cu:86-93 */
  uint32_t fill[OC_MAX_QV];
  memset(fill, 0, sizeof(fill));
  for (uint32_t k = 0; k < m; k++) {
    uint32_t s = q->esrc[k], d = q->edst[k];
    uint32_t o = q->off[s] + fill[s]++;
    q->adj[o] = d;
    q->eid[o] = 2 * k;
    o = q->off[d] + fill[d]++;
    q->adj[o] = s;
    q->eid[o] = 2 * k + 1;
  }
  return OC_OK;
}

int oc_query_read(const char *path, oc_query *q) {
  size_t len;
  char *buf = slurp(path, &len);
  if (!buf) return OC_E_IO;
  uint32_t n, *lab, *deg, *ed;
  uint64_t m;
  int rc = parse_tve(buf, len, &n, &m, &lab, &deg, &ed, OC_MAX_QV, OC_MAX_QE);
  free(buf);
  if (rc != OC_OK) return rc;
  rc = oc_query_from_arrays(n, lab, (uint32_t)m, ed, q);
  free(lab); free(deg); free(ed);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* data graph: simple, sorted adjacency                                      */
/* ------------------------------------------------------------------------ */

/* CSR build, one thread per vertex range: each thread streams the whole
 * edge list and places the entries whose owner falls in its range, then
 * sorts and deduplicates its lists; a final pass compacts them. */
typedef struct {
  const uint32_t *edges;
  uint64_t m;
  uint32_t lo, hi;           /* vertex range [lo, hi) */
  const uint64_t *off;       /* raw list starts (duplicates included) */
  uint32_t *adj;
  uint64_t *len;             /* per vertex: entries after dedup */
  uint64_t *deg;             /* per vertex: raw degree (pass 0) */
  int pass;
} oc_csr_job;

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

static void *oc_csr_worker(void *arg) {
  oc_csr_job *j = (oc_csr_job *)arg;
  const uint32_t lo = j->lo, hi = j->hi;
  if (j->pass == 0) {  /* degrees of the range */
    for (uint64_t k = 0; k < j->m; k++) {
      const uint32_t s = j->edges[2 * k], d = j->edges[2 * k + 1];
      if (s == d) continue;
      if (s >= lo && s < hi) j->deg[s]++;
      if (d >= lo && d < hi) j->deg[d]++;
    }
    return NULL;
  }
  for (uint32_t v = lo; v < hi; v++) j->len[v] = 0;
  for (uint64_t k = 0; k < j->m; k++) {
    const uint32_t s = j->edges[2 * k], d = j->edges[2 * k + 1];
    if (s == d) continue;
    if (s >= lo && s < hi) j->adj[j->off[s] + j->len[s]++] = d;
    if (d >= lo && d < hi) j->adj[j->off[d] + j->len[d]++] = s;
  }
  for (uint32_t v = lo; v < hi; v++) {
    uint32_t *a = j->adj + j->off[v];
    const uint64_t e = j->len[v];
    qsort(a, (size_t)e, 4, cmp_u32);
    uint64_t w = 0;
    for (uint64_t i = 0; i < e; i++)
      if (i == 0 || a[i] != a[i - 1]) a[w++] = a[i];
    j->len[v] = w;
  }
  return NULL;
}

int oc_data_from_arrays(uint32_t n, const uint32_t *labels, uint64_t m,
                        const uint32_t *edges, oc_data **out) {
  for (uint64_t k = 0; k < 2 * m; k++)
    if (edges[k] >= n) return OC_E_RANGE;
  oc_data *g = (oc_data *)calloc(1, sizeof(oc_data));
  if (!g) return OC_E_NOMEM;
  g->n = n;
  g->label = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
  g->off = (uint64_t *)calloc((size_t)n + 1, 8);
  uint64_t *deg = (uint64_t *)calloc((size_t)n + 1, 8);
  uint64_t *len = (uint64_t *)calloc((size_t)n + 1, 8);
  if (!g->label || !g->off || !deg || !len) { free(deg); free(len); oc_data_free(g); return OC_E_NOMEM; }
  memcpy(g->label, labels, (size_t)n * 4);
  int nt = 1;
  if (m >= (1u << 20)) {
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    nt = c < 1 ? 1 : c > 16 ? 16 : (int)c;
  }
  oc_csr_job jobs[16];
  pthread_t th[16];
  for (int t = 0; t < nt; t++) {
    jobs[t].edges = edges;
    jobs[t].m = m;
    jobs[t].lo = (uint32_t)((uint64_t)n * (uint64_t)t / (uint64_t)nt);
    jobs[t].hi = (uint32_t)((uint64_t)n * (uint64_t)(t + 1) / (uint64_t)nt);
    jobs[t].off = g->off;
    jobs[t].len = len;
    jobs[t].deg = deg;
  }
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (uint32_t v = 0; v < n; v++) g->off[v + 1] = g->off[v] + deg[v];
      g->adj = (uint32_t *)malloc((size_t)(g->off[n] ? g->off[n] : 1) * 4);
      if (!g->adj) { free(deg); free(len); oc_data_free(g); return OC_E_NOMEM; }
    }
    for (int t = 0; t < nt; t++) {
      jobs[t].pass = pass;
      jobs[t].adj = g->adj;
    }
    if (nt == 1) {
      oc_csr_worker(&jobs[0]);
    } else {
      for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, oc_csr_worker, &jobs[t]);
      for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    }
  }
  /* compact the deduplicated lists */
  uint64_t w = 0;
  for (uint32_t v = 0; v < n; v++) {
    const uint64_t b = g->off[v];
    if (w != b) memmove(g->adj + w, g->adj + b, (size_t)len[v] * 4);
    g->off[v] = w;
    w += len[v];
  }
  g->off[n] = w;
  g->m = w / 2;
  free(deg);
  free(len);
  *out = g;
  return OC_OK;
}

int oc_data_read(const char *path, oc_data **out) {
  size_t len;
  char *buf = slurp(path, &len);
  if (!buf) return OC_E_IO;
  uint32_t n, *lab, *deg, *ed;
  uint64_t m;
  int rc = parse_tve(buf, len, &n, &m, &lab, &deg, &ed, 0xFFFFFFFFull, 1ull << 40);
  free(buf);
  if (rc != OC_OK) return rc;
  rc = oc_data_from_arrays(n, lab, m, ed, out);
  free(lab); free(deg); free(ed);
  return rc;
}

void oc_data_free(oc_data *g) {
  if (!g) return;
  free(g->label);
  free(g->off);
  free(g->adj);
  free(g);
}

uint64_t oc_data_num_edges(const oc_data *g) { return g->m; }

/* ------------------------------------------------------------------------ */
/* planner                                                                   */
/* ------------------------------------------------------------------------ */

/* decycle.cpp:6-134: BFS spanning tree from the first max-degree vertex. */
static int oc_decycle(const oc_query *q, oc_plan *p) {
  uint32_t n = q->n;
  uint32_t root = 0, maxdeg = q->deg[0];
  for (uint32_t u = 1; u < n; u++)           /* decycle.cpp:15-22 */
    if (q->deg[u] > maxdeg) { root = u; maxdeg = q->deg[u]; }
  uint8_t vis[OC_MAX_QV] = {0};
  uint8_t keep[2 * OC_MAX_QE];
  memset(keep, 0, sizeof(keep));
  uint32_t queue[OC_MAX_QV], qh = 0, qt = 0;
  queue[qt++] = root;
  vis[root] = 1;
  while (qh < qt) {                          /* decycle.cpp:27-43 */
    uint32_t top = queue[qh++];
    for (uint32_t o = q->off[top]; o < q->off[top + 1]; o++) {
      uint32_t nb = q->adj[o];
      if (!vis[nb]) {
        vis[nb] = 1;
        queue[qt++] = nb;
        keep[q->eid[o]] = 1;                 /* vve[{top,nbr}] of a simple graph */
      }
    }
  }
  if (qt != n) return OC_E_QUERY;            /* disconnected query */
  for (uint32_t u = 0; u < n; u++) p->tree_deg[u] = q->deg[u];
  p->n_dropped = 0;
  for (uint32_t e = 0; e < 2 * q->m; e += 2) {  /* decycle.cpp:45-60 */
    if (keep[e] || keep[e + 1]) {
      keep[e] = keep[e + 1] = 1;
    } else {
      uint32_t s = q->esrc[e / 2], d = q->edst[e / 2];
      p->dropped_src[p->n_dropped] = s;
      p->dropped_dst[p->n_dropped] = d;
      p->n_dropped++;
      p->tree_deg[s]--;
      p->tree_deg[d]--;
    }
  }
  uint32_t w = 0;                            /* decycle.cpp:93-118 */
  p->tree_off[0] = 0;
  for (uint32_t u = 0; u < n; u++) {
    for (uint32_t o = q->off[u]; o < q->off[u + 1]; o++)
      if (keep[q->eid[o]]) p->tree_adj[w++] = q->adj[o];
    p->tree_off[u + 1] = w;
  }
  p->tree_root = root;
  return OC_OK;
}

/* filter.cu:23-41 */
static void oc_dfs_cvc(const oc_plan *p, uint32_t v, uint32_t father, int f[][2]) {
  f[v][0] = 0;
  f[v][1] = 1;
  for (uint32_t o = p->tree_off[v]; o < p->tree_off[v + 1]; o++) {
    uint32_t nb = p->tree_adj[o];
    if (nb == father) continue;
    oc_dfs_cvc(p, nb, v, f);
    f[v][0] += f[nb][1];
    f[v][1] += f[nb][0] < f[nb][1] ? f[nb][0] : f[nb][1];
  }
}

/* filter.cu:43-112: DP minimum vertex cover, top-down with an explicit stack. */
static int oc_vertex_cover(const oc_plan *tp, uint32_t n, uint32_t *cover, uint32_t *ncover) {
  int f[OC_MAX_QV][2];
  uint32_t root = 0;
  for (uint32_t u = 1; u < n; u++)
    if (tp->tree_deg[u] > tp->tree_deg[root]) root = u;
  oc_dfs_cvc(tp, root, 0xFFFFFFFFu, f);
  uint32_t best = (uint32_t)(f[root][0] < f[root][1] ? f[root][0] : f[root][1]);
  uint32_t st_v[OC_MAX_QV * 2], st_fa[OC_MAX_QV * 2];
  uint8_t st_sel[OC_MAX_QV * 2];
  int sp = 0;
  uint32_t k = 0;
  st_v[sp] = root; st_fa[sp] = 0xFFFFFFFFu; st_sel[sp] = 0; sp++;
  while (sp > 0) {
    sp--;
    uint32_t v = st_v[sp], fa = st_fa[sp];
    int fsel = st_sel[sp];
    int push_sel;
    if (f[v][1] <= f[v][0]) {
      cover[k++] = v;
      push_sel = 1;
    } else {
      if (fsel || fa == 0xFFFFFFFFu) {
        push_sel = 0;
      } else {
        cover[k++] = v;
        push_sel = 1;
      }
    }
    for (uint32_t o = tp->tree_off[v]; o < tp->tree_off[v + 1]; o++) {
      uint32_t nb = tp->tree_adj[o];
      if (nb == fa) continue;
      st_v[sp] = nb; st_fa[sp] = v; st_sel[sp] = (uint8_t)push_sel; sp++;
    }
  }
  if (k != best) return OC_E_PLAN;  /* reconstruction must be optimal */
  *ncover = k;
  return OC_OK;
}

typedef struct {
  uint32_t size;
  uint32_t us[OC_MAX_QV + 1];
} oc_cluster;

/* filter.cu:127-401 + structure.cu:219-239 */
static int oc_clustering(oc_plan *p, uint32_t n) {
  oc_cluster *cl = (oc_cluster *)calloc(OC_MAX_CLUSTERS, sizeof(oc_cluster));
  uint8_t *valid = (uint8_t *)malloc(OC_MAX_CLUSTERS);
  if (!cl || !valid) { free(cl); free(valid); return OC_E_NOMEM; }
  memset(valid, 1, OC_MAX_CLUSTERS);          /* structure.cu:41-42 */
  uint32_t num_clusters = p->n_cover;
  for (uint32_t i = 0; i < num_clusters; i++) {  /* filter.cu:150-162 */
    uint32_t core = p->cover[i];
    cl[i].size = p->tree_deg[core] + 1;
    cl[i].us[0] = core;
    for (uint32_t t = 0; t < p->tree_deg[core]; t++) cl[i].us[1 + t] = p->tree_adj[p->tree_off[core] + t];
  }
  uint32_t layer = 0, i = 0, num_new, num_actual;
  int rc = OC_OK;
  do {                                          /* filter.cu:174-371 */
    num_new = 0;
    num_actual = 0;
    for (; i < num_clusters; ++i) {
      if (!valid[i]) continue;
      uint32_t core_i = cl[i].us[0];
      for (uint32_t j = i + 1; j < num_clusters; ++j) {
        if (!valid[j]) continue;
        uint32_t core_j = cl[j].us[0];
        for (uint32_t ip = 1; ip < cl[i].size; ++ip) {
          uint32_t ui = cl[i].us[ip];
          for (uint32_t jp = 0; jp < cl[j].size; ++jp) {
            if (ui != cl[j].us[jp]) continue;
            num_new++;
            uint32_t idx = num_clusters + num_new - 1;
            if (idx >= OC_MAX_CLUSTERS) { rc = OC_E_PLAN; goto done; }
            valid[idx] = 1;
            oc_cluster *nc = &cl[idx];
            if (jp) {                           /* both are non-core */
              nc->size = core_i == core_j ? 2 : 3;
              nc->us[0] = ui;
              nc->us[1] = core_i;
              if (core_i != core_j) nc->us[2] = core_j;
            } else {                            /* u_j is core_j */
              nc->size = 2;
              nc->us[0] = ui;
              nc->us[1] = core_i;
            }
          }
        }
      }
    }
    layer++;
    num_actual = num_new;
    for (uint32_t out = num_clusters; out < num_clusters + num_new; ++out) {  /* filter.cu:270-366 */
      if (!valid[out]) continue;
      uint32_t core = cl[out].us[0];
      int dup = 0;
      uint32_t grp[OC_MAX_CLUSTERS > 4096 ? 4096 : OC_MAX_CLUSTERS];
      uint32_t ng = 0;
      uint64_t nbrset = 0;   /* members as a bitset: std::set iterates ascending */
      grp[ng++] = out;
      for (uint32_t t = 1; t < cl[out].size; t++) nbrset |= 1ull << cl[out].us[t];
      for (uint32_t in = out + 1; in < num_clusters + num_new; ++in) {
        if (!valid[in]) continue;
        if (cl[in].us[0] == core) {
          dup = 1;
          grp[ng++] = in;
          for (uint32_t t = 1; t < cl[in].size; t++) nbrset |= 1ull << cl[in].us[t];
        }
      }
      if (!dup) continue;
      uint32_t setsz = (uint32_t)__builtin_popcountll(nbrset);
      /* grp is ascending already (out, then increasing in) like std::set<int> */
      int create_new = 1;
      uint32_t it = 0;
      while (create_new && it < ng) {
        create_new = create_new && (cl[grp[it]].size != setsz + 1);
        it++;
      }
      if (!create_new) {                      /* a cluster contains all */
        uint32_t largest = grp[it - 1];
        uint32_t w = 0;
        for (uint32_t t = 0; t < ng; t++) if (grp[t] != largest) grp[w++] = grp[t];
        ng = w;
      } else {
        num_new++;
        num_actual++;
        uint32_t idx = num_clusters + num_new - 1;
        if (idx >= OC_MAX_CLUSTERS) { rc = OC_E_PLAN; goto done; }
        valid[idx] = 1;
        cl[idx].size = setsz + 1;
        cl[idx].us[0] = core;
        uint32_t t = 1;
        for (uint32_t b = 0; b < 64; b++) if (nbrset >> b & 1) cl[idx].us[t++] = b;
      }
      num_actual -= ng;
      for (uint32_t t = 0; t < ng; t++) valid[grp[t]] = 0;
    }
    num_clusters += num_new;
    if (layer > 1000) { rc = OC_E_PLAN; goto done; }
  } while (num_actual > 1);

  /* encodingMeta::init (structure.cu:219-239) */
  {
    uint32_t tot = 0;
    for (uint32_t c = 0; c < num_clusters; c++) {
      p->cluster_off[c] = tot;
      p->cluster_valid[c] = valid[c];
      if (tot + cl[c].size > OC_MAX_COMPACT) { rc = OC_E_PLAN; goto done; }
      for (uint32_t t = 0; t < cl[c].size; t++) p->compact[tot + t] = cl[c].us[t];
      tot += cl[c].size;
    }
    p->cluster_off[num_clusters] = tot;
    p->n_clusters = num_clusters;
    p->n_compact = tot;
    p->n_layers = layer + 1;
  }
done:
  (void)n;
  free(cl);
  free(valid);
  return rc;
}

/* order.cpp:150-205 on the full query, starting at the root. */
static int oc_edge_order(const oc_query *q, oc_plan *p) {
  uint32_t start = p->root;
  uint32_t start_nbr = q->adj[q->off[start]];
  uint8_t visited[OC_MAX_QV][OC_MAX_QV];
  memset(visited, 0, sizeof(visited));
  uint32_t k = 0;
  p->order_src[k] = start;
  p->order_dst[k] = start_nbr;
  k++;
  visited[start][start_nbr] = 1;
  uint8_t vis[OC_MAX_QV] = {0};
  uint32_t queue[4 * OC_MAX_QE + OC_MAX_QV], qh = 0, qt = 0;
  queue[qt++] = start;
  while (qh < qt) {
    uint32_t u = queue[qh++];
    vis[u] = 1;
    for (uint32_t o = q->off[u]; o < q->off[u + 1]; o++) {
      uint32_t nb = q->adj[o];
      if (!vis[nb]) queue[qt++] = nb;
      if (!visited[nb][u] && !visited[u][nb]) {
        if (k >= OC_MAX_QE) return OC_E_PLAN;
        p->order_src[k] = u;
        p->order_dst[k] = nb;
        k++;
        visited[u][nb] = 1;
      }
    }
  }
  p->order_len = k;
  if (k != q->m) return OC_E_QUERY;
  return OC_OK;
}

/* Matching order from the edge order, as the join consumes it (join.cu:322-420):
 * order[0].first is matched first; an edge whose second endpoint is unmatched
 * extends the rows, otherwise it is a non-tree check.  The check is attached
 * to the later-matched endpoint (same result: checks are filters). */
static int oc_matching_order(oc_plan *p, uint32_t n) {
  uint8_t mapped[OC_MAX_QV] = {0};
  uint32_t k = 0;
  p->mvert[k] = p->order_src[0];
  p->mpos[p->order_src[0]] = k;
  p->parent[k] = 0xFFFFFFFFu;
  p->nback[k] = 0;
  mapped[p->order_src[0]] = 1;
  k++;
  for (uint32_t e = 0; e < p->order_len; e++) {
    uint32_t a = p->order_src[e], b = p->order_dst[e];
    if (!mapped[a]) return OC_E_PLAN;
    if (!mapped[b]) {
      p->mvert[k] = b;
      p->mpos[b] = k;
      p->parent[k] = p->mpos[a];
      p->nback[k] = 0;
      mapped[b] = 1;
      k++;
    } else {
      uint32_t pa = p->mpos[a], pb = p->mpos[b];
      uint32_t late = pa > pb ? pa : pb, early = pa > pb ? pb : pa;
      p->back[late][p->nback[late]++] = early;
    }
  }
  if (k != n) return OC_E_PLAN;
  return OC_OK;
}

int oc_plan_build(const oc_query *q, oc_plan *p) {
  memset(p, 0, sizeof(*p));
  p->n = q->n;
  p->m = q->m;
  int rc = oc_decycle(q, p);
  if (rc) return rc;
  rc = oc_vertex_cover(p, q->n, p->cover, &p->n_cover);
  if (rc) return rc;
  rc = oc_clustering(p, q->n);
  if (rc) return rc;
  /* main.cu:144: core of the last cluster */
  p->root = p->compact[p->cluster_off[p->n_clusters - 1]];
  rc = oc_edge_order(q, p);
  if (rc) return rc;
  return oc_matching_order(p, q->n);
}

/* ------------------------------------------------------------------------ */
/* filter + refinement                                                       */
/* ------------------------------------------------------------------------ */

static inline int bit_get(const uint64_t *row, uint64_t v) { return (int)((row[v >> 6] >> (v & 63)) & 1); }
static inline void bit_set(uint64_t *row, uint64_t v) { row[v >> 6] |= 1ull << (v & 63); }

/*
 * filter.cu:403-539 predicate, query NLC from filter.cu:555-566:
 *   label(v) == label(u)  and  deg(v) >= deg_Q(u)  and
 *   for every label l < NUM_VLQ: |{w in N(v): label(w)=l}| >= |{u' in N_Q(u): label(u')=l}|
 * with NUM_VLQ = max query label + 1 (io.cu:269).
 */
int oc_filter_nlc(const oc_data *g, const oc_query *q, uint64_t *bits) {
  uint64_t words = ((uint64_t)g->n + 63) / 64;
  memset(bits, 0, (size_t)(words * q->n * 8));
  uint32_t nvlq = 0;
  for (uint32_t u = 0; u < q->n; u++) if (q->label[u] + 1 > nvlq) nvlq = q->label[u] + 1;
  uint32_t *qnlc = (uint32_t *)calloc((size_t)q->n * nvlq, 4);
  uint32_t *dnlc = (uint32_t *)calloc(nvlq, 4);
  if (!qnlc || !dnlc) { free(qnlc); free(dnlc); return OC_E_NOMEM; }
  for (uint32_t u = 0; u < q->n; u++)
    for (uint32_t o = q->off[u]; o < q->off[u + 1]; o++) qnlc[u * nvlq + q->label[q->adj[o]]]++;
  for (uint32_t v = 0; v < g->n; v++) {
    uint32_t lv = g->label[v];
    if (lv >= nvlq) continue;
    int any = 0;
    for (uint32_t u = 0; u < q->n; u++) any |= q->label[u] == lv;
    if (!any) continue;
    memset(dnlc, 0, nvlq * 4);
    for (uint64_t o = g->off[v]; o < g->off[v + 1]; o++) {
      uint32_t l = g->label[g->adj[o]];
      if (l < nvlq) dnlc[l]++;
    }
    uint64_t dv = g->off[v + 1] - g->off[v];
    for (uint32_t u = 0; u < q->n; u++) {
      if (q->label[u] != lv || q->deg[u] > dv) continue;
      int ok = 1;
      for (uint32_t l = 0; l < nvlq && ok; l++) ok = dnlc[l] >= qnlc[u * nvlq + l];
      if (ok) bit_set(bits + (uint64_t)u * words, v);
    }
  }
  free(qnlc);
  free(dnlc);
  return OC_OK;
}

/*
 * The build's cluster refinement (its own specification; the reference's
 * encode kernels, filter.cu:584-1155, have undefined behaviour and cannot be
 * restated bit for bit).  Clusters are visited in index order; for cluster
 * (core k; members m_1..m_r), every member being a tree neighbour of k:
 *   core step:   Cur(k)   <- { v in Cur(k) : for all i, N(v) meets Cur(m_i) }
 *   member step: Cur(m_i) <- Cur(m_i) intersected with N(Cur(k))
 * Each step keeps f(u) for every embedding f, so the final sets are sound.
 * The final Cur(u) equals the set after u's last occurrence in the flattened
 * cluster array, the position the reference reads (filter.cu:1290-1307).
 */
int oc_refine(const oc_data *g, const oc_plan *p, uint64_t *bits) {
  uint64_t words = ((uint64_t)g->n + 63) / 64;
  uint64_t *keep = (uint64_t *)malloc((size_t)words * 8);
  uint64_t *nb = (uint64_t *)malloc((size_t)words * 8);
  if (!keep || !nb) { free(keep); free(nb); return OC_E_NOMEM; }
  for (uint32_t c = 0; c < p->n_clusters; c++) {
    uint32_t b = p->cluster_off[c], e = p->cluster_off[c + 1];
    uint32_t k = p->compact[b];
    uint64_t *ck = bits + (uint64_t)k * words;
    memset(keep, 0, (size_t)words * 8);
    memset(nb, 0, (size_t)words * 8);
    for (uint64_t v = 0; v < g->n; v++) {
      if (!bit_get(ck, v)) continue;
      int ok = 1;
      for (uint32_t t = b + 1; t < e && ok; t++) {
        const uint64_t *cm = bits + (uint64_t)p->compact[t] * words;
        int found = 0;
        for (uint64_t o = g->off[v]; o < g->off[v + 1] && !found; o++) found = bit_get(cm, g->adj[o]);
        ok = found;
      }
      if (ok) {
        bit_set(keep, v);
        for (uint64_t o = g->off[v]; o < g->off[v + 1]; o++) bit_set(nb, g->adj[o]);
      }
    }
    memcpy(ck, keep, (size_t)words * 8);
    for (uint32_t t = b + 1; t < e; t++) {
      uint64_t *cm = bits + (uint64_t)p->compact[t] * words;
      for (uint64_t w = 0; w < words; w++) cm[w] &= nb[w];
    }
  }
  free(keep);
  free(nb);
  return OC_OK;
}

/* ------------------------------------------------------------------------ */
/* backtracking matcher                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  const oc_data *g;
  const oc_plan *p;
  const uint64_t *bits;
  uint64_t words;
  uint32_t n;
  uint64_t cap;
  uint32_t *list;
  uint64_t list_lim;
  double deadline;
  /* shared */
  _Atomic uint64_t total;
  _Atomic uint64_t next_root;
  _Atomic int stop;
  _Atomic int timed_out;
  uint64_t listed;
  const uint32_t *roots;
  uint64_t nroots;
} oc_mctx;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int has_edge(const oc_data *g, uint32_t a, uint32_t b) {
  uint64_t lo = g->off[a], hi = g->off[a + 1];
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    uint32_t x = g->adj[mid];
    if (x == b) return 1;
    if (x < b) lo = mid + 1; else hi = mid;
  }
  return 0;
}

/* join.cu:144-166 extension test + join.cu:51-52 non-tree check */
static inline int accept(const oc_mctx *m, uint32_t pos, const uint32_t *f, uint32_t w) {
  const oc_plan *p = m->p;
  if (!bit_get(m->bits + (uint64_t)p->mvert[pos] * m->words, w)) return 0;
  for (uint32_t j = 0; j < pos; j++) if (f[j] == w) return 0;
  for (uint32_t t = 0; t < p->nback[pos]; t++)
    if (!has_edge(m->g, f[p->back[pos][t]], w)) return 0;
  return 1;
}

static void emit(oc_mctx *m, const uint32_t *f) {
  if (m->list && m->listed < m->list_lim) {
    uint32_t *row = m->list + m->listed * m->n;
    for (uint32_t i = 0; i < m->n; i++) row[m->p->mvert[i]] = f[i];
    m->listed++;
  }
}

static void dfs(oc_mctx *m, uint32_t pos, uint32_t *f, uint64_t *local, uint64_t *tick) {
  if (atomic_load_explicit(&m->stop, memory_order_relaxed)) return;
  const oc_plan *p = m->p;
  const oc_data *g = m->g;
  uint32_t par = f[p->parent[pos]];
  if (pos + 1 == m->n) {
    uint64_t c = 0;
    for (uint64_t o = g->off[par]; o < g->off[par + 1]; o++) {
      uint32_t w = g->adj[o];
      if (accept(m, pos, f, w)) {
        c++;
        if (m->list) {
          f[pos] = w;
          emit(m, f);
          uint64_t t = atomic_fetch_add(&m->total, 1) + 1;
          if (t > m->cap) { atomic_store(&m->stop, 1); return; }
        }
      }
    }
    if (!m->list && c) {
      uint64_t t = atomic_fetch_add(&m->total, c) + c;
      if (t > m->cap) atomic_store(&m->stop, 1);
    }
    *local += c;
    return;
  }
  for (uint64_t o = g->off[par]; o < g->off[par + 1]; o++) {
    uint32_t w = g->adj[o];
    if (!accept(m, pos, f, w)) continue;
    f[pos] = w;
    dfs(m, pos + 1, f, local, tick);
    if (atomic_load_explicit(&m->stop, memory_order_relaxed)) return;
    if (m->deadline > 0 && ((++*tick) & 1023) == 0 && now_s() > m->deadline) {
      atomic_store(&m->timed_out, 1);
      atomic_store(&m->stop, 1);
      return;
    }
  }
}

static void *worker(void *arg) {
  oc_mctx *m = (oc_mctx *)arg;
  uint32_t f[OC_MAX_QV];
  uint64_t local = 0, tick = 0;
  for (;;) {
    uint64_t i = atomic_fetch_add(&m->next_root, 1);
    if (i >= m->nroots || atomic_load(&m->stop)) break;
    f[0] = m->roots[i];
    dfs(m, 1, f, &local, &tick);
  }
  return NULL;
}

int oc_match(const oc_data *g, const oc_query *q, const oc_plan *p,
             uint64_t cap, uint32_t *list, uint64_t list_cap, int nthreads,
             double budget_s, int use_refine, oc_result *res) {
  if (nthreads < 1) nthreads = 1;
  if (list && nthreads != 1) return OC_E_ARG;
  double t0 = now_s();
  uint64_t words = ((uint64_t)g->n + 63) / 64;
  uint64_t *bits = (uint64_t *)malloc((size_t)(words * q->n * 8));
  if (!bits) return OC_E_NOMEM;
  int rc = oc_filter_nlc(g, q, bits);
  if (!rc && use_refine) rc = oc_refine(g, p, bits);
  if (rc) { free(bits); return rc; }
  /* root candidates ascending (firstJoinKernel, join.cu:70-97) */
  const uint64_t *rb = bits + (uint64_t)p->mvert[0] * words;
  uint64_t nroots = 0;
  for (uint64_t w = 0; w < words; w++) nroots += (uint64_t)__builtin_popcountll(rb[w]);
  uint32_t *roots = (uint32_t *)malloc((size_t)(nroots ? nroots : 1) * 4);
  if (!roots) { free(bits); return OC_E_NOMEM; }
  uint64_t k = 0;
  for (uint64_t v = 0; v < g->n; v++) if (bit_get(rb, v)) roots[k++] = (uint32_t)v;

  oc_mctx *m = (oc_mctx *)calloc(1, sizeof(oc_mctx));
  if (!m) { free(bits); free(roots); return OC_E_NOMEM; }
  m->g = g; m->p = p; m->bits = bits; m->words = words; m->n = q->n;
  m->cap = cap; m->list = list; m->list_lim = list ? (list_cap < cap ? list_cap : cap) : 0;
  m->deadline = budget_s > 0 ? t0 + budget_s : 0;
  m->roots = roots; m->nroots = nroots;
  atomic_store(&m->total, 0);
  atomic_store(&m->next_root, 0);
  atomic_store(&m->stop, 0);
  atomic_store(&m->timed_out, 0);
  if (nthreads == 1) {
    worker(m);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, m);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
  }
  uint64_t total = atomic_load(&m->total);
  res->count = total > cap ? cap : total;
  res->capped = total > cap;
  res->timed_out = (uint32_t)atomic_load(&m->timed_out);
  res->listed = m->listed;
  res->seconds = now_s() - t0;
  free(m);
  free(bits);
  free(roots);
  return OC_OK;
}
