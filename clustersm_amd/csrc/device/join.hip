// Plain (one query per call) join on gfx950.
//
// Partial embeddings are materialised level by level, as in the reference
// (join.cu:290-435), but in bounded chunks: "BFS within a chunk, DFS across
// chunks".  Level k holds rows of k vertex ids (matching-order positions
// 0..k-1, row-major).  For a chunk of rows, every (row, neighbour of the
// row's parent) pair is one item; items are numbered through the exclusive
// prefix sum of the parents' degrees, so a workgroup tile of 256 consecutive
// items covers a contiguous run of rows (each row has at least one item).
//
//   expand<MARK>   test each item, write a 64-bit survivor mask per wave
//   scan           tile survivor counts -> output offsets
//   materialize    write surviving rows (prefix + new vertex) in item order
//   expand<COUNT>  last level: count survivors only
//
// Item order is (row, ascending neighbour), so with ascending root
// candidates every level is in lexicographic order of the matching-order
// tuple, and DFS over chunks enumerates embeddings in that order: a cap cuts
// the same embeddings as the CPU backtracking oracle.
#include <algorithm>
#include <chrono>
#include <numeric>

#include "common.hpp"

namespace csm {

csm_status run_filter(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine);

namespace {

constexpr int kTile = 256;          // items per workgroup
constexpr int kTileWaves = kTile / kWave;

struct ExpandArgs {
  const uint32_t *off;
  const uint32_t *adj;
  const uint32_t *rows;     // level k, stride k
  const uint64_t *dscan;    // level k parent-degree prefix, n_k + 1 entries
  const uint64_t *cand;     // candidate bitmap row of the new vertex
  uint64_t row_a, row_b;    // chunk rows [a, b)
  uint64_t item_base;       // dscan[a]
  uint64_t items;           // dscan[b] - dscan[a]
  uint32_t k;               // columns of the input rows
  LevelDesc lv;             // position k: parent, non-tree checks, injectivity
  uint64_t *mask;           // MARK: items / 64 words
  uint64_t *tile_cnt;       // MARK: survivors per tile
  const uint64_t *tile_off; // MATERIALIZE: exclusive prefix of tile_cnt
  uint32_t *out;            // MATERIALIZE: rows of k + 1 ids
  unsigned long long *count;  // COUNT
};

enum { kMark = 0, kCount = 1, kMaterialize = 2 };

// Largest r in [lo, hi) with dscan[r] <= x (dscan[lo] <= x < dscan[hi]).
__device__ inline uint64_t upper_row(const uint64_t *dscan, uint64_t lo, uint64_t hi, uint64_t x) {
  while (hi - lo > 1) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (dscan[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ inline bool adjacent(const uint32_t *off, const uint32_t *adj, uint32_t x, uint32_t w) {
  uint32_t lo = off[x], hi = off[x + 1];
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t y = adj[mid];
    if (y == w) return true;
    if (y < w) lo = mid + 1; else hi = mid;
  }
  return false;
}

template <int MODE>
__global__ __launch_bounds__(kTile) void expand_kernel(ExpandArgs a) {
  __shared__ uint64_t sd[kTile + 1];
  __shared__ uint64_t s_rlo;
  __shared__ uint32_t s_wcnt[kTileWaves];
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kTile;
  if (tid == 0) s_rlo = upper_row(a.dscan, a.row_a, a.row_b, a.item_base + first);
  __syncthreads();
  const uint64_t rlo = s_rlo;
  const uint64_t nwin = min(static_cast<uint64_t>(kTile + 1), a.row_b + 1 - rlo);
  for (uint64_t i = tid; i < nwin; i += kTile) sd[i] = a.dscan[rlo + i];
  __syncthreads();

  const uint64_t item = first + tid;
  const bool valid = item < a.items;
  bool ok = false;
  uint32_t w = 0;
  uint64_t r = 0;
  if (valid) {
    const uint64_t gi = a.item_base + item;
    uint64_t lo = 0, hi = nwin - 1;  // sd[lo] <= gi < sd[hi] within the window
    if (gi >= sd[nwin - 1]) {        // only with zero-degree rows: global search
      r = upper_row(a.dscan, rlo + nwin - 1, a.row_b, gi);
    } else {
      while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (sd[mid] <= gi) lo = mid; else hi = mid;
      }
      r = rlo + lo;
    }
    const uint32_t *row = a.rows + r * a.k;
    const uint64_t rs = (r - rlo < nwin) ? sd[r - rlo] : a.dscan[r];
    const uint32_t p = row[a.lv.parent];
    w = a.adj[a.off[p] + static_cast<uint32_t>(gi - rs)];
    if (MODE == kMaterialize) {
      ok = (a.mask[item >> 6] >> (item & 63)) & 1ull;
    } else {
      ok = (a.cand[w >> 6] >> (w & 63)) & 1ull;
      for (uint32_t t = 0; ok && t < a.lv.ninj; ++t) ok = row[a.lv.inj[t]] != w;
      for (uint32_t t = 0; ok && t < a.lv.nback; ++t) ok = adjacent(a.off, a.adj, row[a.lv.back[t]], w);
    }
  }
  const uint64_t m = ballot(ok);
  if (MODE == kMaterialize) {
    if (lane == 0) s_wcnt[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (ok) {
      uint64_t pos = a.tile_off[blockIdx.x] + lanes_below(m);
      for (int x = 0; x < wave; ++x) pos += s_wcnt[x];
      const uint32_t *row = a.rows + r * a.k;
      uint32_t *dst = a.out + pos * (a.k + 1);
      for (uint32_t j = 0; j < a.k; ++j) dst[j] = row[j];
      dst[a.k] = w;
    }
    return;
  }
  if (lane == 0) {
    s_wcnt[wave] = static_cast<uint32_t>(__popcll(m));
    if (MODE == kMark) a.mask[(first >> 6) + wave] = m;
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t c = 0;
    for (int x = 0; x < kTileWaves; ++x) c += s_wcnt[x];
    if (MODE == kMark) a.tile_cnt[blockIdx.x] = c;
    if (MODE == kCount && c) atomicAdd(a.count, static_cast<unsigned long long>(c));
  }
}

// dscan[i] = degree of the parent vertex of row i; dscan[n] = 0.
__global__ void parent_degree_kernel(const uint32_t *off, const uint32_t *rows, uint32_t k, uint32_t parent,
                                     uint64_t n, uint64_t *dscan) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i <= n; i += 256ull * gridDim.x) {
    if (i == n) {
      dscan[i] = 0;
    } else {
      uint32_t p = rows[i * k + parent];
      dscan[i] = off[p + 1] - off[p];
    }
  }
}

__global__ void word_popcount_kernel(const uint64_t *bits, uint32_t words, uint64_t *cnt) {
  for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b <= words; b += 256ull * gridDim.x)
    cnt[b] = b < words ? static_cast<uint64_t>(__popcll(bits[b])) : 0;
}

// Root candidates in ascending order (firstJoinKernel, join.cu:70-97).
__global__ void compact_bits_kernel(const uint64_t *bits, uint32_t words, const uint64_t *pos, uint32_t *out) {
  for (uint64_t b = blockIdx.x * 256ull + threadIdx.x; b < words; b += 256ull * gridDim.x) {
    uint64_t x = bits[b];
    uint64_t p = pos[b];
    while (x) {
      int i = __builtin_ctzll(x);
      x &= x - 1;
      out[p++] = static_cast<uint32_t>(b * 64 + i);
    }
  }
}

// Chunk [c, b): the largest b with dscan[b] - dscan[c] <= cap (at least c+1).
__global__ void choose_chunk_kernel(const uint64_t *dscan, uint64_t n, uint64_t c, uint64_t cap, uint64_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t lo = c + 1, hi = n;  // answer in [c+1, n]
  const uint64_t base = dscan[c];
  if (dscan[n] - base <= cap) {
    lo = n;
  } else {
    while (lo < hi) {
      uint64_t mid = lo + (hi - lo + 1) / 2;
      if (dscan[mid] - base <= cap) lo = mid; else hi = mid - 1;
    }
  }
  out[0] = lo;
  out[1] = dscan[lo] - base;
  out[2] = base;
}

uint32_t grid_for(uint64_t n) { return static_cast<uint32_t>(std::min<uint64_t>(65536, (n + 255) / 256 + 1)); }

// Control block slots (device and pinned host mirror)
enum : int {
  kCtlCount = 0,     // survivors counted at the last level
  kCtlTotal = 1,     // scan total (rows materialised by the last step)
  kCtlChunkA = 2,    // chunk choice for the new level (b, items, base)
  kCtlChunkB = 5,    // chunk choice for the current level (b, items, base)
  kCtlSize = 8
};

struct Workspace {
  csm_ctx *ctx = nullptr;
  std::vector<void *> blocks;
  uint32_t *rows[CSM_MAX_QV + 1] = {};
  uint64_t *dscan[CSM_MAX_QV + 1] = {};
  uint64_t *mask = nullptr, *tiles = nullptr, *partials = nullptr, *ctl = nullptr;
  uint64_t item_cap = 0;
  ~Workspace() {
    for (void *p : blocks) ctx->arena.free(p);
  }
  void *get(uint64_t bytes) {
    void *p = ctx->arena.alloc(bytes);
    if (p) blocks.push_back(p);
    return p;
  }
};

// Per-item bytes of the workspace across all levels.
uint64_t bytes_per_item(uint32_t nq, bool list) {
  uint64_t b = 0;
  for (uint32_t k = 2; k < nq; ++k) b += 4ull * k + 8;  // rows of k ids + degree prefix
  if (list) b += 4ull * nq;
  b += 1 + 8 / 256 + 1;  // mask bit, tile counter, scan partials (rounded up)
  return b;
}

}  // namespace

csm_status run_join(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts &o, csm_result &res) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const uint32_t nq = q->desc.nq;
  const uint32_t words = q->desc.words;
  const bool want_list = o.list != nullptr;
  const uint64_t cap = o.cap ? o.cap : ~0ull;
  const LevelDesc *lv = q->level;
  KernelTimer kt(ctx, res.kernels);

  const uint64_t n1 = q->counts[lv[0].u];
  res.count = 0;
  res.capped = 0;
  res.listed = 0;
  if (n1 == 0) {
    res.ms_join = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CSM_OK;
  }
  Workspace ws;
  ws.ctx = ctx;
  // chunk size: requested, else what fits in half of the largest free block
  uint64_t item_cap = o.chunk_items ? o.chunk_items : (64ull << 20);
  const uint64_t per_item = bytes_per_item(nq, want_list);
  const uint64_t fixed = 8ull * (n1 + 1) + 4ull * n1 + 8ull * (words + 1) + (1 << 20);
  const uint64_t avail = ctx->arena.largest_free();
  if (avail < fixed + per_item * g->max_degree * 2)
    return set_error(ctx, CSM_E_NOMEM, "arena too small for the join workspace");
  item_cap = std::min(item_cap, (avail - fixed) / 2 / per_item);
  item_cap = std::max<uint64_t>(item_cap, g->max_degree);
  item_cap = (item_cap + kTile - 1) / kTile * kTile;
  ws.item_cap = item_cap;
  const uint64_t max_tiles = item_cap / kTile + 2;
  const uint64_t scan_n = std::max<uint64_t>({item_cap + 1, n1 + 1, words + 1ull, max_tiles});
  ws.ctl = static_cast<uint64_t *>(ws.get(8 * kCtlSize));
  ws.mask = static_cast<uint64_t *>(ws.get(8 * (item_cap / 64 + kTileWaves)));
  ws.tiles = static_cast<uint64_t *>(ws.get(8 * max_tiles));
  ws.partials = static_cast<uint64_t *>(ws.get(8 * (scan_n / kScanTile + 2)));
  ws.rows[1] = static_cast<uint32_t *>(ws.get(4 * n1));
  ws.dscan[1] = static_cast<uint64_t *>(ws.get(8 * std::max<uint64_t>(n1 + 1, words + 1ull)));
  bool okalloc = ws.ctl && ws.mask && ws.tiles && ws.partials && ws.rows[1] && ws.dscan[1];
  for (uint32_t k = 2; k < nq && okalloc; ++k) {
    ws.rows[k] = static_cast<uint32_t *>(ws.get(4ull * k * item_cap));
    ws.dscan[k] = static_cast<uint64_t *>(ws.get(8ull * (item_cap + 1)));
    okalloc = ws.rows[k] && ws.dscan[k];
  }
  if (okalloc && want_list) {
    ws.rows[nq] = static_cast<uint32_t *>(ws.get(4ull * nq * item_cap));
    okalloc = ws.rows[nq] != nullptr;
  }
  if (!okalloc) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (join workspace)");
  CSM_HIP_TRY(ctx, hipMemsetAsync(ws.ctl, 0, 8 * kCtlSize, s));

  // level 1: root candidates, ascending
  const uint64_t *root_bits = q->bits + static_cast<uint64_t>(lv[0].u) * words;
  kt.begin(CSM_K_AUX);
  word_popcount_kernel<<<grid_for(words + 1ull), 256, 0, s>>>(root_bits, words, ws.dscan[1]);
  scan_u64_exclusive(ws.dscan[1], words + 1ull, ws.partials, nullptr, s);
  compact_bits_kernel<<<grid_for(words), 256, 0, s>>>(root_bits, words, ws.dscan[1], ws.rows[1]);
  kt.end(CSM_K_AUX, 16ull * words + 4ull * n1);

  uint64_t nrows[CSM_MAX_QV + 1] = {};
  uint64_t cursor[CSM_MAX_QV + 1] = {};
  uint64_t next_b[CSM_MAX_QV + 1] = {}, next_items[CSM_MAX_QV + 1] = {}, next_base[CSM_MAX_QV + 1] = {};
  nrows[1] = n1;
  uint64_t *pin = ctx->pinned;

  // prefix of parent degrees for level k and the first chunk choice
  auto prepare_level = [&](uint32_t k) -> csm_status {
    const uint32_t par = lv[k].parent;
    kt.begin(CSM_K_AUX);
    parent_degree_kernel<<<grid_for(nrows[k] + 1), 256, 0, s>>>(g->off, ws.rows[k], k, par, nrows[k], ws.dscan[k]);
    scan_u64_exclusive(ws.dscan[k], nrows[k] + 1, ws.partials, nullptr, s);
    choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], 0, item_cap, ws.ctl + kCtlChunkA);
    kt.end(CSM_K_AUX, nrows[k] * (4ull + 8 + 16));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    cursor[k] = 0;
    next_b[k] = pin[kCtlChunkA];
    next_items[k] = pin[kCtlChunkA + 1];
    next_base[k] = pin[kCtlChunkA + 2];
    return CSM_OK;
  };

  csm_status st = prepare_level(1);
  if (st != CSM_OK) return st;
  std::vector<uint32_t> host_rows;
  uint64_t total = 0;
  uint32_t k = 1;
  const uint32_t last = nq - 1;  // level whose extension completes an embedding
  bool stop = false;
  while (!stop) {
    if (cursor[k] >= nrows[k]) {
      if (k == 1) break;
      --k;
      continue;
    }
    const uint64_t a = cursor[k], b = next_b[k], items = next_items[k];
    ExpandArgs ea{};
    ea.off = g->off;
    ea.adj = g->adj;
    ea.rows = ws.rows[k];
    ea.dscan = ws.dscan[k];
    ea.cand = q->bits + static_cast<uint64_t>(lv[k].u) * words;
    ea.row_a = a;
    ea.row_b = b;
    ea.k = k;
    ea.lv = lv[k];
    ea.mask = ws.mask;
    ea.tile_cnt = ws.tiles;
    ea.tile_off = ws.tiles;
    ea.count = reinterpret_cast<unsigned long long *>(ws.ctl + kCtlCount);
    ea.item_base = next_base[k];
    ea.items = items;
    const uint32_t ntiles = ceil_div(items, kTile);
    const uint64_t rows_in = b - a;
    res.items += items;
    // algorithmic bytes of one extension step: rows read, parent offsets,
    // (adjacency id + bitmap word) per item, rows written
    const uint64_t step_in = rows_in * (4ull * k + 8) + items * 12ull;
    if (k == last && !want_list) {
      kt.begin(CSM_K_COUNT);
      expand_kernel<kCount><<<ntiles, kTile, 0, s>>>(ea);
      kt.end(CSM_K_COUNT, step_in);
      cursor[k] = b;
      choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], b, item_cap, ws.ctl + kCtlChunkB);
      CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      next_b[k] = pin[kCtlChunkB];
      next_items[k] = pin[kCtlChunkB + 1];
      next_base[k] = pin[kCtlChunkB + 2];
      total = pin[kCtlCount];
      if (total > cap) stop = true;
      continue;
    }
    // mark, scan, materialise into level k + 1
    kt.begin(CSM_K_EXPAND);
    expand_kernel<kMark><<<ntiles, kTile, 0, s>>>(ea);
    kt.end(CSM_K_EXPAND, step_in);
    kt.begin(CSM_K_AUX);
    scan_u64_exclusive(ws.tiles, ntiles, ws.partials, ws.ctl + kCtlTotal, s);
    kt.end(CSM_K_AUX, 16ull * ntiles);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    const uint64_t produced = pin[kCtlTotal];
    ea.out = ws.rows[k + 1];
    if (produced) {
      kt.begin(CSM_K_MATERIALIZE);
      expand_kernel<kMaterialize><<<ntiles, kTile, 0, s>>>(ea);
      kt.end(CSM_K_MATERIALIZE, rows_in * (4ull * k + 8) + items * (4ull + 1) + produced * 4ull * (k + 1));
    }
    res.partial_rows += produced;
    cursor[k] = b;
    if (k == last) {  // list mode: rows are complete embeddings in lexicographic order
      const uint64_t lim = std::min(cap, o.list_capacity);
      const uint64_t want = total < lim ? std::min(produced, lim - total) : 0;
      if (want) {
        const size_t before = host_rows.size();
        host_rows.resize(before + want * nq);
        CSM_HIP_TRY(ctx, hipMemcpyAsync(host_rows.data() + before, ws.rows[nq], 4ull * nq * want,
                                        hipMemcpyDeviceToHost, s));
      }
      total += produced;
      choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], b, item_cap, ws.ctl + kCtlChunkB);
      CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      next_b[k] = pin[kCtlChunkB];
      next_items[k] = pin[kCtlChunkB + 1];
      next_base[k] = pin[kCtlChunkB + 2];
      if (total > cap) stop = true;
      continue;
    }
    if (cursor[k] < nrows[k]) {
      choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], b, item_cap, ws.ctl + kCtlChunkB);
      CSM_HIP_TRY(ctx, hipMemcpyAsync(pin + kCtlChunkB, ws.ctl + kCtlChunkB, 24, hipMemcpyDeviceToHost, s));
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      next_b[k] = pin[kCtlChunkB];
      next_items[k] = pin[kCtlChunkB + 1];
      next_base[k] = pin[kCtlChunkB + 2];
    }
    if (produced) {
      nrows[k + 1] = produced;
      ++k;
      st = prepare_level(k);
      if (st != CSM_OK) return st;
    }
  }
  kt.flush();
  res.count = std::min(total, cap);
  res.capped = total > cap;
  if (want_list) {
    // matching-order rows -> query-id order, then sort lexicographically
    const uint64_t nl = host_rows.size() / nq;
    std::vector<uint32_t> qrows(host_rows.size());
    for (uint64_t i = 0; i < nl; ++i)
      for (uint32_t p = 0; p < nq; ++p) qrows[i * nq + lv[p].u] = host_rows[i * nq + p];
    std::vector<uint64_t> idx(nl);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](uint64_t x, uint64_t y) {
      return std::lexicographical_compare(qrows.begin() + x * nq, qrows.begin() + (x + 1) * nq,
                                          qrows.begin() + y * nq, qrows.begin() + (y + 1) * nq);
    });
    for (uint64_t i = 0; i < nl; ++i)
      std::copy(qrows.begin() + idx[i] * nq, qrows.begin() + (idx[i] + 1) * nq, o.list + i * nq);
    res.listed = nl;
  }
  res.ms_join = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return CSM_OK;
}

}  // namespace csm

using namespace csm;

extern "C" csm_status csm_match(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts *opts,
                                csm_result *out) {
  if (!ctx || !g || !q || !out || q->graph_n != g->n) return CSM_E_ARG;
  csm_match_opts o{};
  o.refine = 1;
  if (opts) o = *opts;
  if (o.list && o.list_capacity == 0) o.list = nullptr;
  std::memset(out, 0, sizeof(*out));
  hipSetDevice(ctx->device);
  auto t0 = std::chrono::steady_clock::now();
  if (!(o.reuse_filter && q->filtered)) {
    csm_status st = run_filter(ctx, g, q, o.refine);
    if (st != CSM_OK) {
      out->status = st;
      return st;
    }
  }
  for (int c = 0; c < CSM_K_NUM; ++c) out->kernels[c] = q->kstat[c];
  out->ms_filter = q->ms_filter;
  csm_status st = run_join(ctx, g, q, o, *out);
  out->status = st;
  out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return st;
}
