// Test harness around the reference's decycle() (decycle.cpp:6-134), built
// from the reference source tree by oracle/Makefile into oracle/_ref/.
// It fills a cpuGraph the way readGraphToCPU2 does (io.cu:44-107) and copies
// the decycled tree out.  cpuGraph's constructor lives in cpuGraph.cpp, which
// needs <cuda.h>, so the object's members are constructed here directly.
#include <cstring>
#include <new>

#include "cpuGraph.h"
#include "decycle.h"

extern "C" int ref_decycle(uint32_t n, const uint32_t *labels, uint32_t m,
                           const uint32_t *edges, uint32_t *tree_off,
                           uint32_t *tree_adj, uint32_t *tree_deg,
                           uint32_t *tree_m) {
  alignas(cpuGraph) unsigned char storage[sizeof(cpuGraph)];
  std::memset(storage, 0, sizeof(storage));
  cpuGraph *g = reinterpret_cast<cpuGraph *>(storage);
  new (&g->vve) std::map<std::pair<vtype, vtype>, etype>();
  new (&g->evv) std::map<etype, std::pair<vtype, vtype>>();
  g->num_v = n;
  g->num_e = m;
  g->isQuery = true;
  g->vertexIDs_ = new vtype[n];
  g->offsets_ = new offtype[n + 1];
  g->neighbors_ = new vtype[2 * m];
  g->edgeIDs_ = new etype[2 * m];
  g->outdeg_ = new degtype[n];
  g->vLabels_ = new vltype[n];
  std::memset(g->outdeg_, 0, sizeof(degtype) * n);
  for (uint32_t k = 0; k < m; ++k) {
    g->outdeg_[edges[2 * k]]++;
    g->outdeg_[edges[2 * k + 1]]++;
  }
  g->offsets_[0] = 0;
  for (uint32_t v = 0; v < n; ++v) {
    g->vertexIDs_[v] = v;
    g->vLabels_[v] = labels[v];
    g->offsets_[v + 1] = g->offsets_[v] + g->outdeg_[v];
  }
  offtype *offs = new offtype[n]();
  etype eid = 0;
  for (uint32_t k = 0; k < m; ++k) {
    vtype s = edges[2 * k], d = edges[2 * k + 1];
    g->vve.insert({std::pair<vtype, vtype>(s, d), eid});
    g->evv.insert({eid, std::pair<vtype, vtype>(s, d)});
    g->vve.insert({std::pair<vtype, vtype>(d, s), eid + 1});
    g->evv.insert({eid + 1, std::pair<vtype, vtype>(d, s)});
    offtype o = g->offsets_[s] + offs[s]++;
    g->neighbors_[o] = d;
    g->edgeIDs_[o] = eid;
    o = g->offsets_[d] + offs[d]++;
    g->neighbors_[o] = s;
    g->edgeIDs_[o] = eid + 1;
    eid += 2;
  }
  delete[] offs;
  g->keep = new bool[2 * m]();

  decycle(g);

  for (uint32_t v = 0; v <= n; ++v) tree_off[v] = g->offsets_[v];
  for (uint32_t i = 0; i < g->offsets_[n]; ++i) tree_adj[i] = g->neighbors_[i];
  for (uint32_t v = 0; v < n; ++v) tree_deg[v] = g->outdeg_[v];
  *tree_m = g->num_e;

  delete[] g->vertexIDs_;
  delete[] g->offsets_;
  delete[] g->neighbors_;
  delete[] g->edgeIDs_;
  delete[] g->outdeg_;
  delete[] g->vLabels_;
  delete[] g->keep;
  g->vve.~map();
  g->evv.~map();
  return 0;
}
