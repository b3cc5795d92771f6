// Build of the candidate adjacency index (see cand_index.hpp).
//
// For every slot (query edge s -> t) and x in C(s): a bound on |N(x) ∩ C(t)|
// from the neighbour-label table (x's neighbours of label(t); x's degree
// when the byte saturates or the label is beyond the table) goes to x's rank
// in the slot's offset array; one exclusive scan over all slots at once turns
// the bounds into list starts in cn_adj; one pass over the adjacency then
// writes every list, sorted, and its exact length.  Candidates of degree at
// least kHeavyDegree are left out of that pass (one wave walking a word whose
// vertices hold hundreds of thousands of neighbours would run long after the
// rest): the bound pass lists them, and one workgroup per (slot, vertex)
// writes each of their lists.
#include <algorithm>
#include <vector>

#include "cand_index.hpp"

namespace csm {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;

struct SlotDesc {
  const uint64_t *bits;       // the slot's query
  const uint64_t *rank_pref;
  uint32_t src, dst;
  uint64_t base;    // first entry of the slot in cn_off
  uint64_t ncand;   // |C(src)|
  uint32_t dst_label;
  uint32_t pad;
};

// A (slot, candidate) pair whose list is built by cn_heavy_kernel.
struct HeavyPair {
  uint32_t slot;   // index into the slot array
  uint32_t x;
  uint64_t entry;  // the candidate's entry in cn_off / cn_len
};

struct RankJob {
  const uint64_t *bits;
  uint64_t *out;    // nq x (words + 1)
  uint32_t nq;
};

// Per-word popcounts of every candidate bitmap (blockIdx.y = query); the
// entry after each row's last word is 0, so one exclusive scan over the
// concatenation gives every row's prefix.
__global__ void rank_words_kernel(const RankJob *jobs, uint32_t words) {
  const RankJob j = jobs[blockIdx.y];
  const uint64_t total = static_cast<uint64_t>(j.nq) * (words + 1);
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += 256ull * gridDim.x) {
    const uint64_t u = i / (words + 1), b = i % (words + 1);
    j.out[i] = b < words ? static_cast<uint64_t>(__popcll(j.bits[u * words + b])) : 0;
  }
}

__device__ inline bool bit_of(const uint64_t *bits, uint32_t v) { return (bits[v >> 6] >> (v & 63)) & 1ull; }

// List bounds: per slot and x in C(s) (blockIdx.y = slot), the number of
// x's neighbours labelled label(t), else deg(x); a zero closes each slot.
// Candidates of degree >= heavy_deg are appended to `heavy` (at most
// heavy_cap pairs; *heavy_n counts every one, so the host sees an overflow).
__global__ __launch_bounds__(kThreads) void bound_kernel(const uint32_t *__restrict__ off,
                                                         const uint8_t *__restrict__ nlc, uint32_t nlc_labels,
                                                         uint32_t words, const SlotDesc *__restrict__ slots,
                                                         uint64_t *__restrict__ cn_off, uint32_t slot0,
                                                         uint32_t heavy_deg, HeavyPair *__restrict__ heavy,
                                                         unsigned long long *heavy_n, uint64_t heavy_cap) {
  const SlotDesc sd = slots[blockIdx.y];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t *xrow = sd.bits + static_cast<uint64_t>(sd.src) * words;
  const uint64_t *pref = sd.rank_pref + static_cast<uint64_t>(sd.src) * (words + 1);
  const uint64_t pref0 = pref[0];
  const bool table = nlc != nullptr && sd.dst_label < nlc_labels;
  if (blockIdx.x == 0 && threadIdx.x == 0) cn_off[sd.base + sd.ncand] = 0;
  for (uint32_t b = blockIdx.x * kWaves + wave; b < words; b += gridDim.x * kWaves) {  // wave-uniform
    const uint64_t xw = xrow[b];
    if (xw == 0) continue;
    const bool act = (xw >> lane) & 1ull;
    const uint32_t x = b * 64u + static_cast<uint32_t>(lane);
    const uint64_t entry = sd.base + (pref[b] - pref0) + lanes_below(xw);
    uint32_t deg = 0;
    if (act) {
      deg = off[x + 1] - off[x];
      uint32_t bound = deg;
      if (table) {
        const uint32_t c = nlc[static_cast<uint64_t>(x) * kNlcStride + sd.dst_label];
        if (c < 255) bound = min(bound, c);
      }
      cn_off[entry] = bound;
    }
    const bool hv = act && deg >= heavy_deg;
    const uint64_t hm = ballot(hv);
    if (hm) {  // one counter update per wave
      unsigned long long at = 0;
      if (lane == 0) at = atomicAdd(heavy_n, static_cast<unsigned long long>(__popcll(hm)));
      at = __shfl(at, 0, kWave);
      const uint64_t idx = at + lanes_below(hm);
      if (hv && idx < heavy_cap) heavy[idx] = HeavyPair{slot0 + blockIdx.y, x, entry};
    }
  }
}

// A wave takes one 64-bit word of C(s): the adjacency ranges of its set
// lanes are contiguous in the CSR order of the concatenation [lane 0's
// neighbours, lane 1's, ...], which the wave walks 64 entries at a time
// (coalesced loads, no divergence on skewed degrees).  Each entry finds its
// owner lane by binary search over the degree prefix.  An owner's hits are
// written at its list start plus the hits before them (earlier chunks: a
// running count in LDS; this chunk: a ballot restricted to the owner's
// lanes), so every list stays sorted; the final count is the list's length.
__global__ __launch_bounds__(kThreads) void cn_kernel(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                                                      uint32_t words, const SlotDesc *__restrict__ slots,
                                                      const uint64_t *__restrict__ cn_off, uint32_t *__restrict__ cn_len,
                                                      uint32_t *__restrict__ cn_adj, uint32_t heavy_deg) {
  __shared__ uint32_t s_pre[kWaves][kWave + 1];  // exclusive prefix of the lanes' degrees (0 for unset lanes)
  __shared__ uint32_t s_o0[kWaves][kWave];       // adjacency start per lane
  __shared__ uint32_t s_run[kWaves][kWave];      // hits so far per lane
  __shared__ uint64_t s_start[kWaves][kWave];    // list start per lane
  const SlotDesc sd = slots[blockIdx.y];
  const uint64_t *bits = sd.bits;
  const uint64_t *rank_pref = sd.rank_pref;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t *xrow = bits + static_cast<uint64_t>(sd.src) * words;
  const uint64_t *yrow = bits + static_cast<uint64_t>(sd.dst) * words;
  const uint64_t *pref = rank_pref + static_cast<uint64_t>(sd.src) * (words + 1);
  const uint64_t pref0 = pref[0];
  uint32_t *pre = s_pre[wave];
  for (uint32_t b = blockIdx.x * kWaves + wave; b < words; b += gridDim.x * kWaves) {
    const uint64_t xw = xrow[b];
    if (xw == 0) continue;
    const bool act = (xw >> lane) & 1ull;
    const uint32_t v = b * 64u + static_cast<uint32_t>(lane);
    const uint64_t slot = sd.base + (pref[b] - pref0) + lanes_below(xw);
    uint32_t o0 = 0, dv = 0;
    bool hv = false;  // heavy: its list is cn_heavy_kernel's, its range empty here
    if (act) {
      o0 = off[v];
      dv = off[v + 1] - o0;
      if (dv >= heavy_deg) {
        hv = true;
        dv = 0;
      }
    }
    uint32_t incl = dv;
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(incl, d, kWave);
      if (lane >= d) incl += y;
    }
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    pre[lane] = incl - dv;
    if (lane == kWave - 1) pre[kWave] = total;
    s_o0[wave][lane] = o0;
    s_run[wave][lane] = 0;
    s_start[wave][lane] = act ? cn_off[slot] : 0;
    wave_sync();
    for (uint32_t base = 0; base < total; base += kWave) {
      const uint32_t e = base + static_cast<uint32_t>(lane);
      bool hit = false;
      uint32_t owner = 0, w = 0;
      if (e < total) {
        uint32_t lo = 0, hi = kWave;  // last lane with pre <= e (unset lanes have empty ranges)
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (pre[mid] <= e) lo = mid; else hi = mid;
        }
        owner = lo;
        w = adj[s_o0[wave][owner] + (e - pre[owner])];
        hit = bit_of(yrow, w);
      }
      const uint64_t m = ballot(hit);
      if (hit) {
        const uint32_t first = max(base, pre[owner]) - base;  // owner's first lane in this chunk
        const uint64_t seg = m & ~((1ull << first) - 1ull);
        cn_adj[s_start[wave][owner] + s_run[wave][owner] + lanes_below(seg)] = w;
      }
      wave_sync();
      if (hit) atomicAdd(&s_run[wave][owner], 1u);
      wave_sync();
    }
    if (act && !hv) cn_len[slot] = s_run[wave][lane];
    wave_sync();
  }
}

// One workgroup per heavy (slot, x) pair: x's adjacency 256 entries at a
// time, hits compacted in order (wave ballots, then the waves' counts in
// LDS), so the list is sorted like every other; then its length.
__global__ __launch_bounds__(kThreads) void cn_heavy_kernel(const uint32_t *__restrict__ off,
                                                            const uint32_t *__restrict__ adj, uint32_t words,
                                                            const SlotDesc *__restrict__ slots,
                                                            const HeavyPair *__restrict__ pairs, uint64_t npairs,
                                                            const uint64_t *__restrict__ cn_off,
                                                            uint32_t *__restrict__ cn_len, uint32_t *__restrict__ cn_adj) {
  __shared__ uint32_t s_cnt[kWaves];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (uint64_t p = blockIdx.x; p < npairs; p += gridDim.x) {  // uniform over the workgroup
    const HeavyPair hp = pairs[p];
    const SlotDesc &sd = slots[hp.slot];
    const uint64_t *yrow = sd.bits + static_cast<uint64_t>(sd.dst) * words;
    const uint32_t o0 = off[hp.x], deg = off[hp.x + 1] - o0;
    const uint64_t start = cn_off[hp.entry];
    uint32_t run = 0;
    for (uint32_t base = 0; base < deg; base += kThreads) {
      const uint32_t e = base + threadIdx.x;
      uint32_t w = 0;
      bool hit = false;
      if (e < deg) {
        w = adj[o0 + e];
        hit = bit_of(yrow, w);
      }
      const uint64_t m = ballot(hit);
      if (lane == 0) s_cnt[wave] = static_cast<uint32_t>(__popcll(m));
      __syncthreads();
      uint32_t before = 0, tot = 0;
      for (int i = 0; i < kWaves; ++i) {
        const uint32_t c = s_cnt[i];
        before += i < wave ? c : 0u;
        tot += c;
      }
      if (hit) cn_adj[start + run + before + lanes_below(m)] = w;
      run += tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) cn_len[hp.entry] = run;
  }
}

// Rank locators (blockIdx.y = query): each candidate word beside the number
// of set bits before it in its row.
__global__ void rank_loc_kernel(const RankJob *jobs, uint32_t words, RankLoc *const *outs) {
  const RankJob j = jobs[blockIdx.y];
  RankLoc *out = outs[blockIdx.y];
  const uint64_t total = static_cast<uint64_t>(j.nq) * words;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += 256ull * gridDim.x) {
    const uint64_t u = i / words, b = i % words;
    const uint64_t *pref = j.out + u * (words + 1);
    out[i] = RankLoc{pref[b] - pref[0], j.bits[i]};
  }
}

// List locators (blockIdx.y = slot): per candidate x of the slot's source,
// its list start and length and x's hub index in one 16-byte entry, at the
// same entry as cn_off / cn_len.
__global__ __launch_bounds__(kThreads) void loc_kernel(uint32_t words, const SlotDesc *__restrict__ slots,
                                                       const uint64_t *__restrict__ cn_off,
                                                       const uint32_t *__restrict__ cn_len,
                                                       const int32_t *__restrict__ hub_id, CnLoc *__restrict__ loc) {
  const SlotDesc sd = slots[blockIdx.y];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t *xrow = sd.bits + static_cast<uint64_t>(sd.src) * words;
  const uint64_t *pref = sd.rank_pref + static_cast<uint64_t>(sd.src) * (words + 1);
  const uint64_t pref0 = pref[0];
  for (uint32_t b = blockIdx.x * kWaves + wave; b < words; b += gridDim.x * kWaves) {  // wave-uniform
    const uint64_t xw = xrow[b];
    if (xw == 0 || !((xw >> lane) & 1ull)) continue;
    const uint32_t x = b * 64u + static_cast<uint32_t>(lane);
    const uint64_t entry = sd.base + (pref[b] - pref0) + lanes_below(xw);
    loc[entry] = CnLoc{cn_off[entry], cn_len[entry], hub_id ? hub_id[x] : -1};
  }
}

uint32_t grid_for(uint64_t n) { return static_cast<uint32_t>(std::min<uint64_t>(65536, (n + 255) / 256 + 1)); }

}  // namespace

csm_status build_cand_index_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<const csm_query *> &qs,
                                  std::vector<CandIndexHost> &out, WorkMem &mem, KernelTimer &kt, bool locators) {
  hipStream_t s = ctx->stream;
  const uint32_t nqs = static_cast<uint32_t>(qs.size());
  if (nqs == 0) return CSM_OK;
  const uint32_t words = qs[0]->desc.words;
  out.assign(nqs, CandIndexHost());
  auto get = [&](uint64_t bytes) { return mem.get(bytes); };
  // rank prefixes of every query in one buffer
  std::vector<uint64_t> rank_base(nqs);
  uint64_t rank_n = 0;
  for (uint32_t i = 0; i < nqs; ++i) {
    rank_base[i] = rank_n;
    rank_n += static_cast<uint64_t>(qs[i]->desc.nq) * (words + 1);
  }
  uint64_t *rank_all = static_cast<uint64_t *>(get(8 * rank_n));
  if (!rank_all) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (candidate index ranks)");
  // slots: one per (query, position k >= 1, anchor a)
  std::vector<SlotDesc> slots;
  uint64_t entries = 0;
  for (uint32_t i = 0; i < nqs; ++i) {
    const csm_query *q = qs[i];
    CandIndexHost &ci = out[i];
    ci.rank_pref = rank_all + rank_base[i];
    std::memset(&ci.level[0], 0, sizeof(LevelJoin));
    ci.level[0].u = q->level[0].u;
    for (uint32_t k = 1; k < q->desc.nq; ++k) {
      const LevelDesc &L = q->level[k];
      LevelJoin &J = ci.level[k];
      std::memset(&J, 0, sizeof(J));
      J.u = L.u;
      J.parent = L.parent;
      J.nanchor = L.nanchor;
      J.ninj = L.ninj;
      for (uint32_t t = 0; t < L.ninj; ++t) J.inj[t] = L.inj[t];
      for (uint32_t a = 0; a < L.nanchor; ++a) {
        const uint32_t src = q->level[L.anchor[a]].u;
        J.anchor[a] = L.anchor[a];
        J.src_u[a] = src;
        J.off_base[a] = entries;
        slots.push_back({q->bits, ci.rank_pref, src, L.u, entries, q->counts[src], q->desc.label[L.u], 0});
        entries += q->counts[src] + 1;
      }
    }
  }
  const uint64_t scan_n = std::max<uint64_t>(rank_n, entries);
  uint64_t *cn_off = static_cast<uint64_t *>(get(8 * std::max<uint64_t>(entries, 1)));
  uint32_t *cn_len = static_cast<uint32_t *>(get(4 * std::max<uint64_t>(entries, 1)));
  uint64_t *partials = static_cast<uint64_t *>(get(8 * (scan_n / kScanTile + 2)));
  uint64_t *total_d = static_cast<uint64_t *>(get(64));
  SlotDesc *d_slots = static_cast<SlotDesc *>(get(sizeof(SlotDesc) * std::max<size_t>(slots.size(), 1)));
  RankJob *d_jobs = static_cast<RankJob *>(get(sizeof(RankJob) * nqs));
  if (!cn_off || !cn_len || !partials || !total_d || !d_slots || !d_jobs)
    return set_error(ctx, CSM_E_NOMEM, "arena exhausted (candidate index)");
  // heavy pairs: room for every heavy vertex in every slot, within 1/64 of
  // the free arena; when the pairs do not fit, every list goes through the
  // per-word walk (the bounds are the same either way)
  uint64_t heavy_cap = std::min<uint64_t>(static_cast<uint64_t>(g->num_heavy) * slots.size(),
                                          std::max<uint64_t>(mem.fixed() ? 0 : 1ull << 16, mem.largest_free() / 64 / sizeof(HeavyPair)));
  HeavyPair *heavy = heavy_cap ? static_cast<HeavyPair *>(get(sizeof(HeavyPair) * heavy_cap)) : nullptr;
  if (!heavy) heavy_cap = 0;
  std::vector<RankJob> jobs(nqs);
  for (uint32_t i = 0; i < nqs; ++i) jobs[i] = {qs[i]->bits, out[i].rank_pref, qs[i]->desc.nq};
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_jobs, jobs.data(), sizeof(RankJob) * nqs, hipMemcpyHostToDevice, s));
  kt.begin(CSM_K_AUX);
  rank_words_kernel<<<dim3(std::min<uint32_t>(1024, grid_for(rank_n / nqs + 1)), nqs), 256, 0, s>>>(d_jobs, words);
  scan_u64_exclusive(rank_all, rank_n, partials, nullptr, s);
  kt.end(CSM_K_AUX, 16ull * rank_n);
  if (locators) {
    uint64_t rloc_n = 0;
    for (uint32_t i = 0; i < nqs; ++i) rloc_n += static_cast<uint64_t>(qs[i]->desc.nq) * words;
    RankLoc *rloc_all = static_cast<RankLoc *>(get(sizeof(RankLoc) * std::max<uint64_t>(rloc_n, 1)));
    RankLoc **d_outs = static_cast<RankLoc **>(get(sizeof(RankLoc *) * nqs));
    if (!rloc_all || !d_outs) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (rank locators)");
    std::vector<RankLoc *> outs(nqs);
    for (uint64_t i = 0, o = 0; i < nqs; o += static_cast<uint64_t>(qs[i]->desc.nq) * words, ++i) {
      outs[i] = rloc_all + o;
      out[i].rloc = outs[i];
    }
    CSM_HIP_TRY(ctx, hipMemcpyAsync(d_outs, outs.data(), sizeof(RankLoc *) * nqs, hipMemcpyHostToDevice, s));
    kt.begin(CSM_K_AUX);
    rank_loc_kernel<<<dim3(std::min<uint32_t>(1024, grid_for(rloc_n / nqs + 1)), nqs), 256, 0, s>>>(d_jobs, words,
                                                                                                    d_outs);
    kt.end(CSM_K_AUX, 32ull * rloc_n);
  }
  if (slots.empty()) return CSM_OK;
  for (uint32_t i = 0; i < nqs; ++i) {
    out[i].cn_off = cn_off;
    out[i].cn_len = cn_len;
  }
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_slots, slots.data(), sizeof(SlotDesc) * slots.size(), hipMemcpyHostToDevice, s));
  const uint32_t gx = std::max<uint32_t>(1, std::min<uint32_t>(std::max<uint32_t>(8, 1024 / nqs), ceil_div(words, kWaves)));
  uint64_t cand_words = 0;
  for (auto &sd : slots) cand_words += sd.ncand;
  unsigned long long *heavy_n = reinterpret_cast<unsigned long long *>(total_d + 1);
  CSM_HIP_TRY(ctx, hipMemsetAsync(heavy_n, 0, 8, s));
  for (size_t s0 = 0; s0 < slots.size(); s0 += 65535) {  // grid.y limit
    const uint32_t ns = static_cast<uint32_t>(std::min<size_t>(65535, slots.size() - s0));
    kt.begin(CSM_K_AUX);
    bound_kernel<<<dim3(gx, ns), kThreads, 0, s>>>(g->off, g->nlc, g->nlc_labels, words, d_slots + s0, cn_off,
                                                   static_cast<uint32_t>(s0), heavy ? kHeavyDegree : ~0u, heavy,
                                                   heavy_n, heavy_cap);
    kt.end(CSM_K_AUX, 8ull * words * ns + 9ull * cand_words * ns / slots.size());
  }
  scan_u64_exclusive(cn_off, entries, partials, total_d, s);
  CSM_HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, total_d, 16, hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  const uint64_t total = ctx->pinned[0];
  const bool use_heavy = heavy && ctx->pinned[1] <= heavy_cap;
  const uint64_t npairs = use_heavy ? ctx->pinned[1] : 0;
  const uint32_t heavy_deg = use_heavy ? kHeavyDegree : ~0u;
  uint32_t *cn_adj = static_cast<uint32_t *>(get(4 * std::max<uint64_t>(total, 1)));
  if (!cn_adj) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (candidate lists)");
  for (size_t s0 = 0; s0 < slots.size(); s0 += 65535) {
    const uint32_t ns = static_cast<uint32_t>(std::min<size_t>(65535, slots.size() - s0));
    kt.begin(CSM_K_AUX);
    cn_kernel<<<dim3(gx, ns), kThreads, 0, s>>>(g->off, g->adj, words, d_slots + s0, cn_off, cn_len, cn_adj,
                                                heavy_deg);
    kt.end(CSM_K_AUX, 8ull * words * ns + 4ull * total * ns / slots.size());
  }
  if (npairs) {
    kt.begin(CSM_K_AUX);
    cn_heavy_kernel<<<static_cast<uint32_t>(std::min<uint64_t>(npairs, 65536)), kThreads, 0, s>>>(
        g->off, g->adj, words, d_slots, heavy, npairs, cn_off, cn_len, cn_adj);
    kt.end(CSM_K_AUX, 16ull * npairs);
  }
  if (locators) {
    CnLoc *loc = static_cast<CnLoc *>(get(sizeof(CnLoc) * std::max<uint64_t>(entries, 1)));
    if (!loc) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (list locators)");
    for (size_t s0 = 0; s0 < slots.size(); s0 += 65535) {
      const uint32_t ns = static_cast<uint32_t>(std::min<size_t>(65535, slots.size() - s0));
      kt.begin(CSM_K_AUX);
      loc_kernel<<<dim3(gx, ns), kThreads, 0, s>>>(words, d_slots + s0, cn_off, cn_len, g->hub_id, loc);
      kt.end(CSM_K_AUX, 8ull * words * ns + 32ull * cand_words * ns / slots.size());
    }
    for (uint32_t i = 0; i < nqs; ++i) out[i].loc = loc;
  }
  for (uint32_t i = 0; i < nqs; ++i) {
    out[i].cn_adj = cn_adj;
    out[i].total_lists = total;
  }
  out[0].build_bytes = 24ull * words * slots.size() + 8ull * cand_words + 16ull * entries + 4ull * total;
  return CSM_OK;
}

csm_status slot_total(csm_ctx *ctx, const CandIndexHost &ci, uint64_t base, uint64_t n, uint64_t &total) {
  total = 0;
  if (n == 0) return CSM_OK;
  std::vector<uint32_t> len(n);
  CSM_HIP_TRY(ctx, hipMemcpyAsync(len.data(), ci.cn_len + base, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (uint32_t l : len) total += l;
  return CSM_OK;
}

csm_status build_cand_index(csm_ctx *ctx, const csm_graph *g, const csm_query *q, CandIndexHost &ci,
                            WorkMem &mem, KernelTimer &kt) {
  std::vector<CandIndexHost> out;
  csm_status st = build_cand_index_batch(ctx, g, {q}, out, mem, kt);
  if (st == CSM_OK) ci = out[0];
  return st;
}

}  // namespace csm
