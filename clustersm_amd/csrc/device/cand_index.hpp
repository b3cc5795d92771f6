// Candidate adjacency index of one query: for every query edge, oriented from
// the earlier-matched endpoint s to the later one t, and every x in C(s), the
// sorted list N(x) ∩ C(t).  Lists of a query are concatenated in one array;
// the list of x is found through x's rank among the set bits of C(s).
#pragma once

#include "common.hpp"

namespace csm {

// Position k of the matching order as the join kernels read it.
struct LevelJoin {
  uint32_t u;                         // query vertex matched at this position
  uint32_t parent;                    // parent position
  uint32_t nanchor;                   // parent + non-tree neighbours
  uint32_t ninj;                      // earlier positions with the same label
  uint32_t anchor[CSM_MAX_QV + 1];    // anchor positions (anchor[0] = parent)
  uint32_t inj[CSM_MAX_QV];
  uint32_t src_u[CSM_MAX_QV + 1];     // query vertex at each anchor
  uint64_t off_base[CSM_MAX_QV + 1];  // start of the anchor's slot in cn_off
};

// Two-load list locator (built on request): a candidate bitmap word with the
// number of set bits before it in its row, then per (slot, candidate) the
// list's start, length and the candidate's hub index, each one 16-byte load.
struct alignas(16) RankLoc {
  uint64_t pref;  // set bits of the row before this word
  uint64_t bits;  // the word
};
struct alignas(16) CnLoc {
  uint64_t start;  // in cn_adj
  uint32_t len;
  int32_t hub;     // the candidate's hub index, -1 for none
};

struct CandIndex {
  const uint64_t *bits;       // nq x words candidate bitmaps
  const uint64_t *rank_pref;  // nq x (words + 1): exclusive popcount prefix, global over all rows
  const uint64_t *cn_off;     // per slot and x in C(s): start of x's list in cn_adj
  const uint32_t *cn_len;     // and its length
  const uint32_t *cn_adj;
  uint32_t words;
  const RankLoc *rloc;        // nq x words (null unless built)
  const CnLoc *loc;           // one per cn_off entry (null unless built)

  // rank of x among the set bits of C(u); x must be a candidate of u
  __device__ inline uint64_t rank(uint32_t u, uint32_t x) const {
    const uint64_t *pref = rank_pref + static_cast<uint64_t>(u) * (words + 1);
    const uint32_t b = x >> 6;
    const uint64_t w = bits[static_cast<uint64_t>(u) * words + b];
    return pref[b] - pref[0] + static_cast<uint64_t>(__popcll(w & ((1ull << (x & 63)) - 1)));
  }
  // [begin, end) of N(x) ∩ C(u_k) for anchor a of position k
  __device__ inline void list(const LevelJoin &L, uint32_t a, uint32_t x, uint64_t &beg, uint64_t &end) const {
    const uint64_t i = L.off_base[a] + rank(L.src_u[a], x);
    beg = cn_off[i];
    end = beg + cn_len[i];
  }
  // the same list, with x's hub index, through the locators (two dependent
  // 16-byte loads)
  __device__ inline CnLoc locate(const LevelJoin &L, uint32_t a, uint32_t x) const {
    const RankLoc r = rloc[static_cast<uint64_t>(L.src_u[a]) * words + (x >> 6)];
    const uint64_t rk = r.pref + static_cast<uint64_t>(__popcll(r.bits & ((1ull << (x & 63)) - 1)));
    return loc[L.off_base[a] + rk];
  }
};

struct CandIndexHost {
  uint64_t *rank_pref = nullptr;
  uint64_t *cn_off = nullptr;
  uint32_t *cn_len = nullptr;
  uint32_t *cn_adj = nullptr;
  RankLoc *rloc = nullptr;    // locators, when requested
  CnLoc *loc = nullptr;
  uint64_t total_lists = 0;   // entries of cn_adj
  uint64_t build_bytes = 0;   // algorithmic bytes of the build
  LevelJoin level[CSM_MAX_QV];
};

// Builds the index of q (candidates must be filtered) in memory from `mem`
// (held by the caller).  Launches on ctx->stream and synchronises.
csm_status build_cand_index(csm_ctx *ctx, const csm_graph *g, const csm_query *q, CandIndexHost &ci,
                            WorkMem &mem, KernelTimer &kt);

// Total length of the n lists of a slot starting at entry `base` (host copy
// of their lengths on ctx->stream; synchronises).
csm_status slot_total(csm_ctx *ctx, const CandIndexHost &ci, uint64_t base, uint64_t n, uint64_t &total);

// The plain path's join of one filtered query (join.hip); its workspace comes
// from `mem` (nullptr: the context arena).
csm_status run_join(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts &o, csm_result &res,
                    WorkMem *mem = nullptr);

// Candidate-aware matching order from q's filtered counts: rewrites q's plan
// order and level descriptors (join.hip).
csm_status apply_reorder(csm_ctx *ctx, const csm_graph *g, csm_query *q);

// All queries' indexes with one launch per pass (one shared list array).
// With `locators`, also the two-load locators (rloc, loc) the persistent
// kernel reads.
csm_status build_cand_index_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<const csm_query *> &qs,
                                  std::vector<CandIndexHost> &out, WorkMem &mem, KernelTimer &kt,
                                  bool locators = false);

}  // namespace csm
