// Plain (one query per call) join on gfx950.
//
// Partial embeddings are materialised level by level, as in the reference
// (join.cu:290-435), but in bounded chunks: "BFS within a chunk, DFS across
// chunks".  Level k holds rows of k vertex ids (matching-order positions
// 0..k-1, row-major).
//
// Extending a row to position k needs a candidate of u_k adjacent to the
// row's vertex at every anchor of k (its tree parent and its non-tree
// neighbours).  Each row enumerates the shortest of its anchors' candidate
// lists N(x) ∩ C(u_k) (the "driver", from the candidate adjacency index) and
// probes the other anchors' adjacency (hub bitmap or binary search).  Every
// (row, driver entry) pair is one item; items are numbered through the
// exclusive prefix sum of the driver list lengths, so a workgroup tile of 256
// consecutive items covers a contiguous run of rows.
//
//   expand<MARK>        test each item, 64-bit survivor mask per wave
//   scan                tile survivor counts -> output offsets
//   expand<MATERIALIZE> write surviving rows (prefix + new vertex), item order
//   count<false>        last position: count survivors, four tiles per
//                       workgroup and four items per thread in lockstep
//   count<true>         last position without non-tree edges, counted from
//                       the position before it: each surviving item adds the
//                       length of N(x) ∩ C(u_last) for its parent vertex x
//                       minus the row's own vertices in that list
//
// Item order is (row, ascending new vertex): every level is in lexicographic
// order of the matching-order tuple and DFS over chunks enumerates in that
// order, so a cap cuts the same embeddings as the CPU backtracking oracle.
#include <algorithm>
#include <chrono>
#include <numeric>
#include <rocprim/device/device_radix_sort.hpp>

#include "../host/host_graph.hpp"
#include "cand_index.hpp"
#include "filter.hpp"

namespace csm {

namespace {

constexpr int kTile = 256;          // items per workgroup
constexpr int kTileWaves = kTile / kWave;

// Symmetry-breaking conditions at one position: the new vertex must exceed
// the row's ids at the `gt` positions and stay below those at `lt`.
struct LevelSym {
  uint32_t ngt, nlt;
  uint32_t gt[CSM_MAX_QV], lt[CSM_MAX_QV];
};

struct ExpandArgs {
  AdjProbe probe;
  CandIndex ci;
  const uint32_t *rows;     // level k, stride k
  const uint64_t *dscan;    // level k driver-length prefix, n_k + 1 entries
  const uint8_t *drv;       // driver anchor index per row
  const uint32_t *tile_row; // first row of each tile, relative to row_a (ntiles + 1 entries)
  uint64_t row_a, row_b;    // chunk rows [a, b)
  uint64_t item_base;       // dscan[a]
  uint64_t items;           // dscan[b] - dscan[a]
  uint32_t k;               // columns of the input rows = position being matched
  LevelJoin lv;             // position k
  uint64_t *mask;           // MARK: items / 64 words
  uint64_t *tile_cnt;       // MARK / COUNT / FUSED: per tile
  const uint64_t *tile_off; // MATERIALIZE: exclusive prefix of tile_cnt
  uint32_t *out;            // MATERIALIZE: rows of k + 1 ids
  unsigned long long *count;  // COUNT / FUSED total
  LevelJoin lv_last;        // FUSED: the last position (k + 1)
  // FUSED with the last position's parent in an earlier column: per row, the
  // list of that parent's id and its length minus the row's own members
  const uint32_t *fbase;
  const uint64_t *fbeg;
  const uint32_t *flen;
  uint32_t fused_rows;
  uint32_t last_inj_k;      // the last position's label equals position k's
  LevelSym sym;             // symmetry-breaking conditions at position k (MARK / COUNT)
  unsigned long long *extra;  // count kernel: bytes read beyond rows, prefix, drivers and list entries
};

enum { kMark = 0, kMaterialize = 2 };

template <typename T>
__device__ inline T block_sum(T v, T *lds) {
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) lds[wave] = v;
  __syncthreads();
  T t = 0;
  for (int x = 0; x < kTileWaves; ++x) t += lds[x];
  return t;
}

constexpr int kTileRows = kTile + 2;  // rows staged in LDS per tile

// Largest r in [lo, hi) with dscan[r] <= x (dscan[lo] <= x < dscan[hi]).
__device__ inline uint64_t upper_row(const uint64_t *dscan, uint64_t lo, uint64_t hi, uint64_t x) {
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (dscan[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Dynamic LDS: the tile's rows (kTileRows x k ids).  A tile whose items span
// more than kTileRows rows (rows with empty driver lists in between) reads
// its rows from global memory instead.
template <int MODE>
__global__ __launch_bounds__(kTile) void expand_kernel(ExpandArgs a) {
  extern __shared__ uint32_t s_rows[];
  __shared__ uint64_t sd[kTileRows + 1];
  __shared__ uint64_t s_base[kTileRows];
  __shared__ uint32_t s_wcnt[kTileWaves];
  __shared__ uint16_t s_ori[MODE == kMaterialize ? kTile : 1];  // MATERIALIZE: survivors' staged rows
  __shared__ uint32_t s_ow[MODE == kMaterialize ? kTile : 1];   // and new ids, in output order
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kTile;
  const uint64_t rlo = a.row_a + a.tile_row[blockIdx.x];
  const uint64_t rend = min(a.row_a + a.tile_row[blockIdx.x + 1] + 1, a.row_b);  // exclusive
  const uint64_t span = rend - rlo;
  const bool wide = span > static_cast<uint64_t>(kTileRows);
  const uint32_t nr = wide ? 0u : static_cast<uint32_t>(span);
  const uint32_t k = a.k;
  if (!wide) {  // stage: driver-length prefix, rows (coalesced), driver list start per row
    for (uint32_t i = tid; i <= nr; i += kTile) sd[i] = a.dscan[rlo + i];
    for (uint32_t i = tid; i < nr * k; i += kTile) s_rows[i] = a.rows[rlo * k + i];
    __syncthreads();
    for (uint32_t i = tid; i < nr; i += kTile) {
      const uint32_t d = a.drv[rlo + i];
      uint64_t beg, end;
      a.ci.list(a.lv, d, s_rows[i * k + a.lv.anchor[d]], beg, end);
      s_base[i] = beg;
    }
  }
  __syncthreads();

  const uint64_t item = first + tid;
  const bool valid = item < a.items;
  bool ok = false;
  uint32_t w = 0;
  uint32_t rs = 0;  // the item's staged row
  const uint32_t *row = nullptr;
  if (valid) {
    const uint64_t gi = a.item_base + item;
    uint32_t d;
    if (!wide) {
      uint32_t lo = 0, hi = nr;  // sd[lo] <= gi < sd[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sd[mid] <= gi) lo = mid; else hi = mid;
      }
      rs = lo;
      row = s_rows + lo * k;
      d = a.drv[rlo + lo];
      w = a.ci.cn_adj[s_base[lo] + (gi - sd[lo])];
    } else {
      const uint64_t r = upper_row(a.dscan, rlo, rend, gi);
      row = a.rows + r * k;
      d = a.drv[r];
      uint64_t beg, end;
      a.ci.list(a.lv, d, row[a.lv.anchor[d]], beg, end);
      w = a.ci.cn_adj[beg + (gi - a.dscan[r])];
    }
    if (MODE == kMaterialize) {
      ok = (a.mask[item >> 6] >> (item & 63)) & 1ull;
    } else {
      ok = true;  // w is a candidate of u_k adjacent to the driver anchor
      for (uint32_t t = 0; ok && t < a.lv.ninj; ++t) ok = row[a.lv.inj[t]] != w;
      for (uint32_t t = 0; ok && t < a.sym.ngt; ++t) ok = w > row[a.sym.gt[t]];
      for (uint32_t t = 0; ok && t < a.sym.nlt; ++t) ok = w < row[a.sym.lt[t]];
      for (uint32_t t = 0; ok && t < a.lv.nanchor; ++t)
        if (t != d) ok = a.probe.has(row[a.lv.anchor[t]], w);
    }
  }
  const uint64_t m = ballot(ok);
  if (MODE == kMaterialize) {
    if (lane == 0) s_wcnt[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int x = 0; x < kTileWaves; ++x) {
      before += x < wave ? s_wcnt[x] : 0u;
      tot += s_wcnt[x];
    }
    if (wide) {  // rows in global memory: each survivor writes its own row
      if (ok) {
        uint32_t *dst = a.out + (a.tile_off[blockIdx.x] + before + lanes_below(m)) * (k + 1);
        for (uint32_t j = 0; j < k; ++j) dst[j] = row[j];
        dst[k] = w;
      }
      return;
    }
    // survivors (staged row, new id) to LDS in output order, then the tile's
    // rows leave as one contiguous, coalesced run of words
    if (ok) {
      const uint32_t p = before + lanes_below(m);
      s_ori[p] = static_cast<uint16_t>(rs);
      s_ow[p] = w;
    }
    __syncthreads();
    const uint32_t kout = k + 1;
    const uint32_t kout_magic = static_cast<uint32_t>((0x100000000ull + kout - 1) / kout);  // i / kout for i < 2^17
    uint32_t *dst = a.out + a.tile_off[blockIdx.x] * kout;
    const uint32_t nw = tot * kout;
    for (uint32_t i = tid; i < nw; i += kTile) {
      const uint32_t r = __umulhi(i, kout_magic), c = i - r * kout;
      dst[i] = c < k ? s_rows[s_ori[r] * k + c] : s_ow[r];
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
    a.tile_cnt[blockIdx.x] = c;
  }
}

// ---------------------------------------------------------------- count kernel
// The count level's items (COUNT; FUSED: counted from the position before the
// last), four 256-item tiles per workgroup and four items per thread, checked
// in lockstep so that their dependent loads overlap.  Besides the driver
// list, each row looks up the candidate list of every other kept anchor: w is
// a candidate of u_k already, so w ~ x iff w is in N(x) ∩ C(u_k), which is
// shorter than N(x).  Lists that many of a row's items search are copied into
// LDS once per workgroup and searched there: the row's run of driver entries
// is intersected with the LDS copies of its other sorted lists.
constexpr int kCIlp = 4;
constexpr uint32_t kCMaxA = 3;          // anchors whose candidate lists are kept per row (more: adjacency probes)
constexpr uint32_t kCListWords = 1536;  // LDS words for copied lists
constexpr uint64_t kCStageMin = 16;     // items a row needs in the workgroup before its lists are copied
constexpr uint64_t kCStageRatio = 8;    // and a copied list has at most this many entries per such item
constexpr uint64_t kCTagHub = 1ull << 63;  // list slot of a hub anchor: tag | hub index
constexpr uint64_t kCTagLds = 1ull << 62;  // list slot copied into LDS: tag | list start

// Membership of v[j] in sorted lists, all searches advancing together; for j
// with in_lds[j], lo[j] is an offset into the LDS copies s, else a position
// in cn.  n[j] = 0: not found.  `loads` gains the global entries read.
template <uint32_t N>
__device__ __forceinline__ void find_mixed(const uint32_t *cn, const uint32_t *s, uint64_t (&lo)[N], uint32_t (&n)[N],
                                           const uint32_t (&v)[N], const bool (&in_lds)[N], bool (&found)[N],
                                           unsigned long long &loads) {
  bool has[N];
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) has[j] = n[j] != 0;
  for (;;) {
    bool any = false;
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) any |= n[j] > 1;
    if (!any) break;
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
      if (n[j] > 1) {
        const uint32_t half = n[j] >> 1;
        const uint32_t e = in_lds[j] ? s[static_cast<uint32_t>(lo[j]) + half] : cn[lo[j] + half];
        if (e <= v[j]) lo[j] += half;
        n[j] -= half;
        loads += in_lds[j] ? 0 : 1;
      }
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    const uint32_t e = !has[j] ? 0u : in_lds[j] ? s[static_cast<uint32_t>(lo[j])] : cn[lo[j]];
    found[j] = has[j] && e == v[j];
    loads += has[j] && !in_lds[j] ? 1 : 0;
  }
}

__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t *lds, uint32_t &total) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t x = v;
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) lds[wave] = x;
  __syncthreads();
  uint32_t off = 0, all = 0;
  for (int w = 0; w < kTileWaves; ++w) {
    if (w < wave) off += lds[w];
    all += lds[w];
  }
  __syncthreads();
  total = all;
  return off + x - v;
}

// v in the sorted list cn[beg .. beg + len); `loads` gains the entries read.
__device__ inline bool in_list_n(const uint32_t *cn, uint64_t beg, uint64_t len, uint32_t v, unsigned long long &loads) {
  while (len > 1) {
    const uint64_t half = len >> 1;
    if (cn[beg + half] <= v) beg += half;
    len -= half;
    ++loads;
  }
  loads += len == 1 ? 1 : 0;
  return len == 1 && cn[beg] == v;
}

// One item of a workgroup whose rows are not staged (they span more than
// kTileRows rows): rows, lists and adjacency from global memory.  `bytes`
// gains what it reads beyond its list entry.
template <bool FUSED>
__device__ unsigned long long count_item_wide(const ExpandArgs &a, uint64_t rlo, uint64_t rend, uint64_t gi,
                                              unsigned long long &bytes) {
  const uint32_t k = a.k;
  uint64_t lo = rlo, hi = rend;  // dscan[lo] <= gi < dscan[hi]
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (a.dscan[mid] <= gi) lo = mid; else hi = mid;
    bytes += 8;
  }
  const uint64_t r = lo;
  const uint32_t *row = a.rows + r * k;
  const uint32_t d = a.drv[r];
  uint64_t beg, end;
  a.ci.list(a.lv, d, row[a.lv.anchor[d]], beg, end);
  const uint32_t w = a.ci.cn_adj[beg + (gi - a.dscan[r])];
  bytes += 4ull * k + 8 + 1 + 28;  // row, its prefix entry, driver, list lookup
  unsigned long long loads = 0;
  bool ok = true;
  for (uint32_t t = 0; ok && t < a.lv.ninj; ++t) ok = row[a.lv.inj[t]] != w;
  for (uint32_t t = 0; ok && t < a.sym.ngt; ++t) ok = w > row[a.sym.gt[t]];
  for (uint32_t t = 0; ok && t < a.sym.nlt; ++t) ok = w < row[a.sym.lt[t]];
  for (uint32_t t = 0; ok && t < a.lv.nanchor; ++t) {
    if (t == d) continue;
    const uint32_t x = row[a.lv.anchor[t]];
    const int32_t h = a.probe.hub_id[x];
    if (h >= 0) {
      ok = (a.probe.hub_bits[static_cast<uint64_t>(h) * a.probe.words + (w >> 6)] >> (w & 63)) & 1ull;
      bytes += 12;
    } else {
      ok = in_list_n(a.probe.adj, a.probe.off[x], a.probe.off[x + 1] - a.probe.off[x], w, loads);
      bytes += 12;
    }
  }
  unsigned long long c = ok ? 1 : 0;
  if (FUSED && ok) {
    const uint32_t pk = a.lv_last.parent;
    if (a.fused_rows) {
      c = a.fbase[r];
      bytes += 4;
      if (a.last_inj_k && w != row[pk]) {
        bytes += 12;
        if (in_list_n(a.ci.cn_adj, a.fbeg[r], a.flen[r], w, loads)) --c;
      }
    } else {
      uint64_t lb, le;
      a.ci.list(a.lv_last, 0, w, lb, le);
      bytes += 28;
      c = le - lb;
      for (uint32_t t = 0; t < a.lv_last.ninj; ++t) {
        const uint32_t j = a.lv_last.inj[t];
        if (j != k && row[j] != w && in_list_n(a.ci.cn_adj, lb, le - lb, row[j], loads)) --c;
      }
    }
  }
  bytes += 4 * loads;
  return c;
}

// Grid: ceil(ntiles / kCIlp) workgroups; tile_cnt[blockIdx.x] = the
// workgroup's count; *a.extra gains the bytes of list lookups, searches,
// hub probes and LDS copies.
template <bool FUSED, bool LISTS>
__global__ __launch_bounds__(kTile) void count_kernel(ExpandArgs a, uint32_t ntiles) {
  extern __shared__ uint32_t s_rows[];              // kTileRows x k ids
  __shared__ uint64_t sd[kTileRows + 1];            // driver-length prefix of the rows
  __shared__ uint64_t s_base[kTileRows];            // driver list start
  __shared__ uint64_t s_lb[kCMaxA + 1][kTileRows];  // kept anchors' lists (slot kCMaxA: the fused list): start | tag
  __shared__ uint32_t s_ll[kCMaxA + 1][kTileRows];  // and length
  __shared__ uint16_t s_lo[kCMaxA + 1][kTileRows];  // offset of a copied list in s_list
  __shared__ uint32_t s_fbase[kTileRows];
  __shared__ uint8_t s_drv[kTileRows];
  __shared__ uint32_t s_spre[kTileRows];            // words copied before each row
  __shared__ uint32_t s_list[kCListWords];
  __shared__ uint32_t s_scan[kTileWaves];
  __shared__ uint32_t s_tot;
  __shared__ unsigned long long s_sum[kTileWaves];
  const int tid = threadIdx.x;
  const uint32_t g0 = blockIdx.x * kCIlp, g1 = min(g0 + kCIlp, ntiles);
  const uint64_t first = static_cast<uint64_t>(g0) * kTile;                  // items [first, last) of the chunk
  const uint64_t last = min(static_cast<uint64_t>(g1) * kTile, a.items);
  const uint64_t rlo = a.row_a + a.tile_row[g0];
  const uint64_t rend = min(a.row_a + a.tile_row[g1] + 1, a.row_b);  // exclusive
  const uint64_t span = rend - rlo;
  const bool wide = span > static_cast<uint64_t>(kTileRows);
  const uint32_t nr = wide ? 0u : static_cast<uint32_t>(span);
  const uint32_t k = a.k;
  const uint32_t nanchor = a.lv.nanchor, nkept = min(nanchor, kCMaxA);
  const bool frows = FUSED && a.fused_rows != 0;
  const bool flist = frows && a.last_inj_k != 0;  // items search the row's fused list
  unsigned long long contrib = 0, loads = 0, bytes = 0;
  if (wide) {
#pragma unroll
    for (int j = 0; j < kCIlp; ++j) {
      const uint64_t item = first + static_cast<uint64_t>(j) * kTile + tid;
      if (item < last) contrib += count_item_wide<FUSED>(a, rlo, rend, a.item_base + item, bytes);
    }
  } else {
    // stage: prefix, rows (coalesced), drivers, the fused count's per-row terms
    for (uint32_t i = tid; i <= nr; i += kTile) sd[i] = a.dscan[rlo + i];
    for (uint32_t i = tid; i < nr * k; i += kTile) s_rows[i] = a.rows[rlo * k + i];
    for (uint32_t i = tid; i < nr; i += kTile) {
      s_drv[i] = a.drv[rlo + i];
      if (frows) s_fbase[i] = a.fbase[rlo + i];
      if (LISTS && flist) {
        s_lb[kCMaxA][i] = a.fbeg[rlo + i];
        s_ll[kCMaxA][i] = a.flen[rlo + i];
      }
    }
    __syncthreads();
    if constexpr (LISTS) {
      // list lookups, one (row, anchor) pair per thread, all in flight at once;
      // a hub anchor other than the driver is tested on its adjacency bitmap
      for (uint32_t p = tid; p < nr * nkept; p += kTile) {
        const uint32_t r = p / nkept, t = p - r * nkept;
        const uint32_t x = s_rows[r * k + a.lv.anchor[t]];
        const int32_t h = t == s_drv[r] ? -1 : a.probe.hub_id[x];
        if (h >= 0) {
          s_lb[t][r] = kCTagHub | static_cast<uint64_t>(h);
        } else {
          uint64_t beg, end;
          a.ci.list(a.lv, t, x, beg, end);
          if (t == s_drv[r]) s_base[r] = beg;
          s_lb[t][r] = beg;
          s_ll[t][r] = static_cast<uint32_t>(end - beg);
        }
      }
      for (uint32_t r = tid; r < nr; r += kTile) {  // a driver beyond the kept anchors
        const uint32_t d = s_drv[r];
        if (d >= nkept) {
          uint64_t beg, end;
          a.ci.list(a.lv, d, s_rows[r * k + a.lv.anchor[d]], beg, end);
          s_base[r] = beg;
        }
      }
      // beyond the rows, prefix and drivers the host counts: a list lookup
      // (rank word, prefix, start, length) or hub index per kept anchor, and
      // the fused count's per-row terms
      if (tid == 0) bytes += nr * (32ull * nkept + (frows ? 4 : 0) + (flist ? 12 : 0));
      __syncthreads();
      // copies: a row with at least kCStageMin items here copies its short
      // non-hub lists (slot order) at its place in the prefix of the words the
      // rows need; rows past the buffer's end keep searching global memory
      uint32_t need = 0;
      uint64_t most = 0;
      if (static_cast<uint32_t>(tid) < nr) {
        const uint64_t lo = max(sd[tid], a.item_base + first), hi = min(sd[tid + 1], a.item_base + last);
        const uint64_t cnt = hi > lo ? hi - lo : 0;
        if (cnt >= kCStageMin) {
          most = min<uint64_t>(kCListWords, kCStageRatio * cnt);
          for (uint32_t t = 0; t < nkept; ++t)
            if (t != s_drv[tid] && !(s_lb[t][tid] & kCTagHub) && s_ll[t][tid] <= most) need += s_ll[t][tid];
          if (flist && s_ll[kCMaxA][tid] <= most) need += s_ll[kCMaxA][tid];
        }
      }
      uint32_t tot = 0;
      if (__syncthreads_or(need != 0)) {
        if (tid == 0) s_tot = 0;
        uint32_t all;
        const uint32_t off = block_excl_scan_u32(need, s_scan, all);
        if (static_cast<uint32_t>(tid) < nr) s_spre[tid] = off;
        if (need && off + need <= kCListWords) {
          uint32_t o = off;
          for (uint32_t t = 0; t < nkept; ++t) {
            if (t != s_drv[tid] && !(s_lb[t][tid] & kCTagHub) && s_ll[t][tid] <= most) {
              s_lo[t][tid] = static_cast<uint16_t>(o);
              s_lb[t][tid] |= kCTagLds;
              o += s_ll[t][tid];
            }
          }
          if (flist && s_ll[kCMaxA][tid] <= most) {
            s_lo[kCMaxA][tid] = static_cast<uint16_t>(o);
            s_lb[kCMaxA][tid] |= kCTagLds;
          }
          atomicMax(&s_tot, off + need);
        }
        __syncthreads();
        tot = s_tot;
      }
      // word i: the last row whose prefix is <= i (a row that copies nothing has
      // the next row's prefix), then the list of that row holding it
      for (uint32_t i0 = 0; i0 < tot; i0 += 4 * kTile) {
        uint32_t v[4];
  #pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          const uint32_t i = i0 + j * kTile + tid;
          v[j] = 0;
          if (i < tot) {
            uint32_t r = 0;
  #pragma unroll
            for (uint32_t step = 256; step > 0; step >>= 1)
              if (r + step < nr && s_spre[r + step] <= i) r += step;
            uint32_t il = i - s_spre[r];
            uint64_t src = ~0ull;
            for (uint32_t t = 0; t < nkept && src == ~0ull; ++t) {
              const uint64_t b = s_lb[t][r];
              if (t != s_drv[r] && (b & kCTagLds)) {
                if (il < s_ll[t][r]) src = (b & ~kCTagLds) + il;
                else il -= s_ll[t][r];
              }
            }
            if (src == ~0ull) src = (s_lb[kCMaxA][r] & ~kCTagLds) + il;  // the fused list comes last
            v[j] = a.ci.cn_adj[src];
          }
        }
  #pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          const uint32_t i = i0 + j * kTile + tid;
          if (i < tot) s_list[i] = v[j];
        }
      }
      bytes += tid == 0 ? 4ull * tot : 0;
      if (tot) __syncthreads();
    } else {  // the driver's list only; other anchors are probed on their adjacency
      for (uint32_t r = tid; r < nr; r += kTile) {
        const uint32_t d = s_drv[r];
        uint64_t beg, end;
        a.ci.list(a.lv, d, s_rows[r * k + a.lv.anchor[d]], beg, end);
        s_base[r] = beg;
      }
      if (tid == 0) bytes += nr * (28ull + (frows ? 4 : 0));
      __syncthreads();
    }

    // items: row by binary search over the prefix, then the new vertex
    bool act[kCIlp], ok[kCIlp];
    uint64_t gi[kCIlp];
    uint32_t r[kCIlp], w[kCIlp];
#pragma unroll
    for (int j = 0; j < kCIlp; ++j) {
      const uint64_t item = first + static_cast<uint64_t>(j) * kTile + tid;
      act[j] = item < last;
      gi[j] = a.item_base + item;
      r[j] = 0;
    }
#pragma unroll
    for (uint32_t step = 256; step > 0; step >>= 1) {
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) {
        const uint32_t c = r[j] + step;
        if (act[j] && c < nr && sd[c] <= gi[j]) r[j] = c;
      }
    }
#pragma unroll
    for (int j = 0; j < kCIlp; ++j) w[j] = act[j] ? a.ci.cn_adj[s_base[r[j]] + (gi[j] - sd[r[j]])] : 0u;
#pragma unroll
    for (int j = 0; j < kCIlp; ++j) ok[j] = act[j];
    for (uint32_t x = 0; x < a.lv.ninj; ++x) {
      const uint32_t col = a.lv.inj[x];
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) ok[j] = ok[j] && s_rows[r[j] * k + col] != w[j];
    }
    for (uint32_t x = 0; x < a.sym.ngt; ++x) {
      const uint32_t col = a.sym.gt[x];
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) ok[j] = ok[j] && w[j] > s_rows[r[j] * k + col];
    }
    for (uint32_t x = 0; x < a.sym.nlt; ++x) {
      const uint32_t col = a.sym.lt[x];
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) ok[j] = ok[j] && w[j] < s_rows[r[j] * k + col];
    }
    if constexpr (LISTS) {
      // the other anchors: hub bitmap word, or the anchor's list (LDS copy or
      // global), searched in lockstep
      for (uint32_t t = 0; t < nanchor; ++t) {
        if (t < nkept) {
          uint64_t lo[kCIlp];
          uint32_t n[kCIlp];
          bool in_lds[kCIlp], hub[kCIlp], hbit[kCIlp], f[kCIlp];
  #pragma unroll
          for (int j = 0; j < kCIlp; ++j) {
            const bool need_j = ok[j] && t != s_drv[r[j]];
            const uint64_t b = need_j ? s_lb[t][r[j]] : 0ull;
            hub[j] = need_j && (b & kCTagHub) != 0;
            in_lds[j] = need_j && (b & kCTagLds) != 0;
            lo[j] = in_lds[j] ? s_lo[t][r[j]] : (b & ~kCTagHub);
            n[j] = need_j && !hub[j] ? s_ll[t][r[j]] : 0u;
            hbit[j] = false;
            if (hub[j]) {
              hbit[j] = (a.probe.hub_bits[lo[j] * a.probe.words + (w[j] >> 6)] >> (w[j] & 63)) & 1ull;
              bytes += 8;
            }
          }
          find_mixed<kCIlp>(a.ci.cn_adj, s_list, lo, n, w, in_lds, f, loads);
  #pragma unroll
          for (int j = 0; j < kCIlp; ++j)
            if (ok[j] && t != s_drv[r[j]]) ok[j] = hub[j] ? hbit[j] : f[j];
        } else {
  #pragma unroll
          for (int j = 0; j < kCIlp; ++j)
            if (ok[j] && t != s_drv[r[j]]) ok[j] = a.probe.has(s_rows[r[j] * k + a.lv.anchor[t]], w[j]);
        }
      }
    } else {
      // the other anchors x: hub bitmap word, else a search of N(x); the
      // hub index and both adjacency offsets are loaded together
      for (uint32_t t = 0; t < nanchor; ++t) {
        uint64_t lo[kCIlp];
        uint32_t n[kCIlp], x[kCIlp], o1[kCIlp];
        int32_t h[kCIlp];
        bool need[kCIlp], hbit[kCIlp], f[kCIlp], in_lds[kCIlp];
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) {
          need[j] = ok[j] && t != s_drv[r[j]];
          x[j] = need[j] ? s_rows[r[j] * k + a.lv.anchor[t]] : 0u;
          h[j] = need[j] ? a.probe.hub_id[x[j]] : -1;
          lo[j] = need[j] ? a.probe.off[x[j]] : 0u;
          o1[j] = need[j] ? a.probe.off[x[j] + 1] : 0u;
          in_lds[j] = false;
        }
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) {
          hbit[j] = false;
          n[j] = need[j] && h[j] < 0 ? o1[j] - static_cast<uint32_t>(lo[j]) : 0u;
          if (need[j] && h[j] >= 0) {
            hbit[j] = (a.probe.hub_bits[static_cast<uint64_t>(h[j]) * a.probe.words + (w[j] >> 6)] >> (w[j] & 63)) & 1ull;
            bytes += 8;
          }
          bytes += need[j] ? 12 : 0;
        }
        find_mixed<kCIlp>(a.probe.adj, s_rows, lo, n, w, in_lds, f, loads);
#pragma unroll
        for (int j = 0; j < kCIlp; ++j)
          if (need[j]) ok[j] = h[j] >= 0 ? hbit[j] : f[j];
      }
    }
    if (!FUSED) {
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) contrib += ok[j] ? 1 : 0;
    } else if (frows) {
      // x = the row's parent of the last position: the row's count, minus w
      // when w is in that list
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) contrib += ok[j] ? s_fbase[r[j]] : 0u;
      if (flist) {
        const uint32_t pk = a.lv_last.parent;
        uint64_t lo[kCIlp];
        uint32_t n[kCIlp];
        bool in_lds[kCIlp], f[kCIlp];
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) {
          const bool need_j = ok[j] && w[j] != s_rows[r[j] * k + pk];
          if constexpr (LISTS) {
            const uint64_t b = need_j ? s_lb[kCMaxA][r[j]] : 0ull;
            in_lds[j] = need_j && (b & kCTagLds) != 0;
            lo[j] = in_lds[j] ? s_lo[kCMaxA][r[j]] : b;
            n[j] = need_j ? s_ll[kCMaxA][r[j]] : 0u;
          } else {
            in_lds[j] = false;
            lo[j] = need_j ? a.fbeg[rlo + r[j]] : 0ull;
            n[j] = need_j ? a.flen[rlo + r[j]] : 0u;
            bytes += need_j ? 12 : 0;
          }
        }
        find_mixed<kCIlp>(a.ci.cn_adj, LISTS ? s_list : s_rows, lo, n, w, in_lds, f, loads);
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) contrib -= f[j] ? 1u : 0u;
      }
    } else {
      // x = w: w's list of the last position, minus the row's own members
      uint64_t lb[kCIlp], le[kCIlp];
#pragma unroll
      for (int j = 0; j < kCIlp; ++j) {
        lb[j] = le[j] = 0;
        if (ok[j]) {
          a.ci.list(a.lv_last, 0, w[j], lb[j], le[j]);
          bytes += 28;
        }
        contrib += le[j] - lb[j];
      }
      for (uint32_t x = 0; x < a.lv_last.ninj; ++x) {
        const uint32_t col = a.lv_last.inj[x];
        if (col == k) continue;
        uint64_t lo[kCIlp];
        uint32_t n[kCIlp], v[kCIlp];
        bool in_lds[kCIlp], f[kCIlp];
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) {
          v[j] = s_rows[r[j] * k + col];
          lo[j] = lb[j];
          n[j] = ok[j] && v[j] != w[j] ? static_cast<uint32_t>(le[j] - lb[j]) : 0u;
          in_lds[j] = false;
        }
        find_mixed<kCIlp>(a.ci.cn_adj, s_rows, lo, n, v, in_lds, f, loads);
#pragma unroll
        for (int j = 0; j < kCIlp; ++j) contrib -= f[j] ? 1u : 0u;
      }
    }
  }
  const unsigned long long tot = block_sum(contrib, s_sum);
  __syncthreads();  // s_sum is read again below
  const unsigned long long b = block_sum(bytes + 4ull * loads, s_sum);
  if (tid == 0) {  // summed by reduce_add2_kernel (one atomic per reducing workgroup, not per workgroup here)
    a.tile_cnt[2 * blockIdx.x] = tot;
    a.tile_cnt[2 * blockIdx.x + 1] = b;
  }
}

// *out += sum of vals[0..n): a few workgroups (one atomic each, not one per tile).
__global__ __launch_bounds__(256) void reduce_add_kernel(const uint64_t *vals, uint64_t n, unsigned long long *out) {
  __shared__ unsigned long long lds[256 / kWave];
  unsigned long long acc = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) acc += vals[i];
  for (int d = kWave / 2; d > 0; d >>= 1) acc += __shfl_xor(acc, d, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 256 / kWave; ++w) t += lds[w];
    if (t) atomicAdd(out, t);
  }
}

// *out0 += sum of vals[2i], *out1 += sum of vals[2i + 1] over i < n.
__global__ __launch_bounds__(256) void reduce_add2_kernel(const uint64_t *vals, uint64_t n, unsigned long long *out0,
                                                          unsigned long long *out1) {
  __shared__ unsigned long long lds[2][256 / kWave];
  unsigned long long x = 0, y = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
    x += vals[2 * i];
    y += vals[2 * i + 1];
  }
  for (int d = kWave / 2; d > 0; d >>= 1) {
    x += __shfl_xor(x, d, kWave);
    y += __shfl_xor(y, d, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    lds[0][threadIdx.x / kWave] = x;
    lds[1][threadIdx.x / kWave] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0;
    for (int w = 0; w < 256 / kWave; ++w) {
      t0 += lds[0][w];
      t1 += lds[1][w];
    }
    if (t0) atomicAdd(out0, t0);
    if (t1) atomicAdd(out1, t1);
  }
}

uint32_t reduce_grid(uint64_t n) { return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(256, n / 2048))); }

// tile_row[t] = row (relative to a) holding item t * kTile of the chunk
// [a, b); tile_row[ntiles] = b - a.
__global__ __launch_bounds__(256) void tile_rows_kernel(const uint64_t *dscan, uint64_t a, uint64_t b, uint64_t ntiles,
                                                        uint32_t *tile_row) {
  const uint64_t base = dscan[a];
  for (uint64_t r = a + blockIdx.x * 256ull + threadIdx.x; r < b; r += 256ull * gridDim.x) {
    const uint64_t lo = dscan[r] - base, hi = dscan[r + 1] - base;  // items [lo, hi)
    for (uint64_t t = (lo + kTile - 1) / kTile; t * kTile < hi; ++t) tile_row[t] = static_cast<uint32_t>(r - a);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tile_row[ntiles] = static_cast<uint32_t>(b - a);
}

// Per-row pass of a level.
// dscan[i] = shortest candidate list among the row's anchors, drv[i] = that
// anchor's index; dscan[n] = 0.
// At the count level, when the last position's parent is an earlier column
// (fused_rows): fbase/fbeg/flen[i] = that parent's list of the last position
// (fbeg/flen only when the count kernel reads them: store_list)
// and its length minus the row's own members.  With row_count (one anchor,
// and no per-item correction), the row's count is its driver list minus the
// row's own ids, times fbase when fused: it goes to partials[block] and the
// row gets no items (dscan[i] = 0).  With deg_sums, for the reference-byte
// accounting, deg_sums[block] = the degrees of the rows' parent vertices.
struct RowPass {
  LevelJoin lv_last;
  uint32_t fused_rows, row_count;
  uint32_t store_list;  // fbeg/flen are read by the count kernel (the last label equals position k's)
  uint32_t *fbase;
  uint64_t *fbeg;
  uint32_t *flen;
  uint64_t *partials;  // row_count: one per workgroup
  const uint32_t *off;  // graph offsets, for deg_sums
  uint64_t *deg_sums;   // one per workgroup, or null
};

__global__ __launch_bounds__(256) void anchor_len_kernel(CandIndex ci, const uint32_t *rows, uint32_t k, LevelJoin lv,
                                                         uint64_t n, uint64_t *dscan, uint8_t *drv, RowPass rp) {
  __shared__ unsigned long long lds[256 / kWave];
  unsigned long long acc = 0, deg = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i <= n; i += 256ull * gridDim.x) {
    if (i == n) {
      dscan[i] = 0;
      continue;
    }
    const uint32_t *row = rows + i * k;
    // the first kRowLists anchors' lists and the fused count's list are
    // looked up together: every rank word and prefix in one round, then every
    // start and length in a second
    constexpr uint32_t kRowLists = 4;
    const uint32_t na = min(lv.nanchor, kRowLists);
    uint64_t ix[kRowLists + 1];
#pragma unroll
    for (uint32_t t = 0; t <= kRowLists; ++t) {
      ix[t] = 0;
      if (t < na) ix[t] = lv.off_base[t] + ci.rank(lv.src_u[t], row[lv.anchor[t]]);
      else if (t == kRowLists && rp.fused_rows)
        ix[t] = rp.lv_last.off_base[0] + ci.rank(rp.lv_last.src_u[0], row[rp.lv_last.parent]);
    }
    uint64_t lb[kRowLists + 1];
    uint32_t ln[kRowLists + 1];
#pragma unroll
    for (uint32_t t = 0; t <= kRowLists; ++t) {
      const bool use = t < na || (t == kRowLists && rp.fused_rows);
      lb[t] = use ? ci.cn_off[ix[t]] : 0;
      ln[t] = use ? ci.cn_len[ix[t]] : 0;
    }
    uint64_t best = ~0ull, bbeg = 0;
    uint32_t bi = 0;
#pragma unroll
    for (uint32_t t = 0; t < kRowLists; ++t) {
      if (t < na && ln[t] < best) {
        best = ln[t];
        bbeg = lb[t];
        bi = t;
      }
    }
    for (uint32_t t = kRowLists; t < lv.nanchor; ++t) {
      uint64_t beg, end;
      ci.list(lv, t, row[lv.anchor[t]], beg, end);
      if (end - beg < best) {
        best = end - beg;
        bbeg = beg;
        bi = t;
      }
    }
    unsigned long long loads = 0;  // not counted here: the host charges the row pass per row
    uint32_t fb = 1;
    if (rp.fused_rows) {
      // the list minus the row's own members of it, searched in lockstep
      const uint32_t xv = row[rp.lv_last.parent];
      const uint64_t flb = lb[kRowLists];
      const uint32_t fln = ln[kRowLists];
      uint64_t c = fln;
      for (uint32_t t0 = 0; t0 < rp.lv_last.ninj; t0 += 4) {
        uint64_t lo[4];
        uint32_t n4[4], v[4];
        bool f[4], in_lds[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          const uint32_t t = t0 + j;
          const uint32_t col = t < rp.lv_last.ninj ? rp.lv_last.inj[t] : k;
          v[j] = col < k ? row[col] : 0u;
          lo[j] = flb;
          n4[j] = col < k && v[j] != xv ? fln : 0u;
          in_lds[j] = false;
        }
        find_mixed<4>(ci.cn_adj, ci.cn_adj, lo, n4, v, in_lds, f, loads);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) c -= f[j] ? 1 : 0;
      }
      fb = static_cast<uint32_t>(c);
      rp.fbase[i] = fb;
      if (rp.store_list) {
        rp.fbeg[i] = flb;
        rp.flen[i] = fln;
      }
    }
    if (rp.row_count) {  // the driver list minus the row's own ids, searched in lockstep
      uint64_t surv = best;
      for (uint32_t t0 = 0; t0 < lv.ninj; t0 += 4) {
        uint64_t lo[4];
        uint32_t n4[4], v[4];
        bool f[4], in_lds[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          const uint32_t t = t0 + j;
          v[j] = t < lv.ninj ? row[lv.inj[t]] : 0u;
          lo[j] = bbeg;
          n4[j] = t < lv.ninj ? static_cast<uint32_t>(best) : 0u;
          in_lds[j] = false;
        }
        find_mixed<4>(ci.cn_adj, ci.cn_adj, lo, n4, v, in_lds, f, loads);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) surv -= f[j] ? 1 : 0;
      }
      acc += surv * fb;
      best = 0;
    }
    dscan[i] = best;
    drv[i] = static_cast<uint8_t>(bi);
    if (rp.deg_sums) {
      const uint32_t x = row[lv.anchor[0]];  // the parent's vertex
      deg += rp.off[x + 1] - rp.off[x];
    }
  }
  if (rp.row_count) {
    for (int d = kWave / 2; d > 0; d >>= 1) acc += __shfl_xor(acc, d, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < 256 / kWave; ++w) t += lds[w];
      rp.partials[blockIdx.x] = t;
    }
  }
  if (rp.deg_sums) {
    __syncthreads();  // lds may still be read above
    for (int d = kWave / 2; d > 0; d >>= 1) deg += __shfl_xor(deg, d, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = deg;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < 256 / kWave; ++w) t += lds[w];
      rp.deg_sums[blockIdx.x] = t;
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
  kCtlCount = 0,     // embeddings counted at the last position
  kCtlTotal = 1,     // scan total (rows materialised by the last step)
  kCtlChunkA = 2,    // chunk choice for a new level (b, items, base)
  kCtlChunkB = 5,    // chunk choice for the current level (b, items, base)
  kCtlDegMark = 8,   // parent degrees summed over the rows of marked chunks
  kCtlDegCount = 9,  // and over the rows of counted chunks
  kCtlCountBytes = 10,  // the count kernel's list lookups, searches, hub probes and LDS copies
  kCtlSize = 11
};

// List mode.  Query vertex of each matching-order position, by value.
struct ColMap {
  uint32_t u[CSM_MAX_QV];
};

// n rows of nq ids in matching order -> query-id order, written at dst.
__global__ void to_query_order_kernel(const uint32_t *src, uint64_t n, uint32_t nq, ColMap cm, uint32_t *dst) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n * nq; i += 256ull * gridDim.x) {
    const uint64_t r = i / nq;
    const uint32_t p = static_cast<uint32_t>(i - r * nq);
    dst[r * nq + cm.u[p]] = src[i];
  }
}

__global__ void iota_kernel(uint32_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) p[i] = static_cast<uint32_t>(i);
}

// keys[i] = column c of row perm[i]
__global__ void gather_col_kernel(const uint32_t *rows, uint32_t nq, uint32_t c, const uint32_t *perm, uint64_t n,
                                  uint32_t *keys) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x)
    keys[i] = rows[static_cast<uint64_t>(perm[i]) * nq + c];
}

// out row i = row perm[i]
__global__ void gather_rows_kernel(const uint32_t *rows, uint32_t nq, const uint32_t *perm, uint64_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n * nq; i += 256ull * gridDim.x) {
    const uint64_t r = i / nq;
    out[i] = rows[static_cast<uint64_t>(perm[r]) * nq + (i - r * nq)];
  }
}

// Sorts nl rows of nq ids (query-id order) lexicographically on the host.
void sort_rows_host(std::vector<uint32_t> &qrows, uint64_t nl, uint32_t nq, uint32_t *out) {
  std::vector<uint64_t> idx(nl);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](uint64_t x, uint64_t y) {
    return std::lexicographical_compare(qrows.begin() + x * nq, qrows.begin() + (x + 1) * nq,
                                        qrows.begin() + y * nq, qrows.begin() + (y + 1) * nq);
  });
  for (uint64_t i = 0; i < nl; ++i) std::copy(qrows.begin() + idx[i] * nq, qrows.begin() + (idx[i] + 1) * nq, out + i * nq);
}

struct Workspace {
  WorkMem *mem = nullptr;
  uint32_t *rows[CSM_MAX_QV + 1] = {};
  uint64_t *dscan[CSM_MAX_QV + 1] = {};
  uint8_t *drv[CSM_MAX_QV + 1] = {};
  uint64_t *mask = nullptr, *tiles = nullptr, *partials = nullptr, *ctl = nullptr;
  uint32_t *tile_row = nullptr;
  uint32_t *fbase = nullptr, *flen = nullptr;  // count level, per row
  uint64_t *fbeg = nullptr, *row_sums = nullptr;
  uint64_t *deg_sums = nullptr;  // row pass: parent degrees per workgroup
  void *get(uint64_t bytes) { return mem->get(bytes); }
};

}  // namespace

csm_status run_join(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts &o, csm_result &res,
                    WorkMem *mem) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const uint32_t nq = q->desc.nq;
  const uint32_t words = q->desc.words;
  const bool want_list = o.list != nullptr;
  const uint64_t cap = o.cap ? o.cap : ~0ull;
  const LevelDesc *lv = q->level;
  const uint32_t L = nq - 1;  // last position
  KernelTimer kt(ctx, res.kernels);
  AdjProbe probe{g->off, g->adj, g->hub_id, g->hub_bits, words};

  // Symmetry breaking (count only): conditions per position; the last
  // position is then counted per item, since its list lengths would count
  // the images the conditions exclude.
  LevelSym sym[CSM_MAX_QV] = {};
  res.automorphisms = 1;
  const bool symmetry = o.symmetry && !want_list;
  if (symmetry) {
    std::vector<std::pair<uint32_t, uint32_t>> less;
    uint64_t aut = 1;
    if (query_symmetry(q->plan, aut, less) == CSM_OK) {
      uint32_t pos[CSM_MAX_QV];
      for (uint32_t p = 0; p < nq; ++p) pos[lv[p].u] = p;
      for (const auto &c : less) {  // f(v) < f(w): checked where the later of the two is matched
        const uint32_t pv = pos[c.first], pw = pos[c.second];
        if (pw > pv) sym[pw].gt[sym[pw].ngt++] = pv;
        else sym[pv].lt[sym[pv].nlt++] = pw;
      }
      res.automorphisms = aut;
    }
  }
  // How the last position is counted.
  const bool precount = !want_list && lv[L].nback == 0 && !symmetry;
  const bool fused = precount && L >= 2;     // counted from level L-1 items
  const bool rows2 = precount && L == 1;     // two-vertex query: count level-1 rows
  const uint32_t stop = fused ? L - 1 : L;   // deepest level whose items are expanded
  // fused, last position's parent before `stop`: its list is a per-row constant
  const bool fused_rows = fused && lv[L].parent != stop;

  const uint64_t n1 = q->counts[lv[0].u];
  res.count = 0;
  res.capped = 0;
  res.listed = 0;
  auto finish_time = [&]() {
    res.ms_join = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  if (n1 == 0) {
    finish_time();
    return CSM_OK;
  }
  WorkMem own(ctx);
  WorkMem &wm = mem ? *mem : own;
  Workspace ws;
  ws.mem = &wm;
  // chunk size: requested, else what fits in half of the largest free block
  uint64_t per_item = 16;  // mask bit, tile counter, scan partials, rounded up
  for (uint32_t k = 2; k <= stop; ++k) per_item += 4ull * k + 8 + 1;
  if (want_list) per_item += 4ull * nq;
  if (fused_rows) per_item += 16;
  // candidate adjacency index first: its size is data dependent
  CandIndexHost cih;
  csm_status st = build_cand_index(ctx, g, q, cih, wm, kt);
  if (st != CSM_OK) return st;
  CandIndex ci{q->bits, cih.rank_pref, cih.cn_off, cih.cn_len, cih.cn_adj, words};
  const LevelJoin *lj = cih.level;
  // the count level's rows count outright when one anchor drives them and no
  // per-item correction applies (the new vertex cannot be in the last list)
  bool last_inj_k = false;
  for (uint32_t t = 0; t < lj[L].ninj; ++t) last_inj_k |= lj[L].inj[t] == stop;
  const bool row_count = fused_rows && !last_inj_k && lj[stop].nanchor == 1;
  const uint64_t fixed = 13ull * (n1 + 1) + 8ull * (words + 1) + (fused_rows ? 16ull * (n1 + 1) : 0) + (2 << 20);
  const uint64_t avail = wm.largest_free();
  if (avail < fixed + per_item * g->max_degree * 2)
    return set_error(ctx, CSM_E_NOMEM, "arena too small for the join workspace");
  uint64_t item_cap = o.chunk_items ? o.chunk_items : (64ull << 20);
  item_cap = std::min(item_cap, (avail - fixed) / 2 / per_item);
  item_cap = std::max<uint64_t>(item_cap, g->max_degree);
  item_cap = (item_cap + kTile - 1) / kTile * kTile;
  const uint64_t max_tiles = item_cap / kTile + 2;
  const uint64_t scan_n = std::max<uint64_t>({item_cap + 1, n1 + 1, words + 1ull, max_tiles});
  ws.ctl = static_cast<uint64_t *>(ws.get(8 * kCtlSize));
  ws.mask = static_cast<uint64_t *>(ws.get(8 * (item_cap / 64 + kTileWaves)));
  ws.tiles = static_cast<uint64_t *>(ws.get(16 * max_tiles));  // count kernel: (count, bytes) per workgroup
  ws.tile_row = static_cast<uint32_t *>(ws.get(4 * (max_tiles + 1)));
  ws.partials = static_cast<uint64_t *>(ws.get(8 * (scan_n / kScanTile + 2)));
  ws.rows[1] = static_cast<uint32_t *>(ws.get(4 * n1));
  ws.dscan[1] = static_cast<uint64_t *>(ws.get(8 * std::max<uint64_t>(n1 + 1, words + 1ull)));
  ws.drv[1] = static_cast<uint8_t *>(ws.get(n1 + 1));
  ws.deg_sums = static_cast<uint64_t *>(ws.get(8 * 65536));  // one per row-pass workgroup (grid_for's bound)
  bool okalloc = ws.ctl && ws.mask && ws.tiles && ws.tile_row && ws.partials && ws.rows[1] && ws.dscan[1] && ws.drv[1] &&
                 ws.deg_sums;
  for (uint32_t k = 2; k <= stop && okalloc; ++k) {
    ws.rows[k] = static_cast<uint32_t *>(ws.get(4ull * k * item_cap));
    ws.dscan[k] = static_cast<uint64_t *>(ws.get(8ull * (item_cap + 1)));
    ws.drv[k] = static_cast<uint8_t *>(ws.get(item_cap + 1));
    okalloc = ws.rows[k] && ws.dscan[k] && ws.drv[k];
  }
  if (okalloc && want_list) {
    ws.rows[nq] = static_cast<uint32_t *>(ws.get(4ull * nq * item_cap));
    okalloc = ws.rows[nq] != nullptr;
  }
  if (okalloc && fused_rows) {
    const uint64_t nr = (stop == 1 ? n1 : item_cap) + 1;
    ws.fbase = static_cast<uint32_t *>(ws.get(4 * nr));
    ws.flen = static_cast<uint32_t *>(ws.get(4 * nr));
    ws.fbeg = static_cast<uint64_t *>(ws.get(8 * nr));
    ws.row_sums = static_cast<uint64_t *>(ws.get(8 * 65536));
    okalloc = ws.fbase && ws.flen && ws.fbeg && ws.row_sums;
  }
  if (!okalloc) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (join workspace)");
  // list mode: the list stays on the device in query-id order and is sorted
  // there when the arena holds it; otherwise each chunk is copied out and the
  // list is sorted on the host
  const uint64_t list_lim = want_list ? std::min(cap, o.list_capacity) : 0;
  uint32_t *dlist = nullptr;
  if (list_lim && list_lim < (1ull << 32)) dlist = static_cast<uint32_t *>(ws.get(4ull * nq * list_lim));
  ColMap cmap{};
  for (uint32_t p = 0; p < nq; ++p) cmap.u[p] = lv[p].u;
  CSM_HIP_TRY(ctx, hipMemsetAsync(ws.ctl, 0, 8 * kCtlSize, s));
  uint64_t *pin = ctx->pinned;
  if (rows2) {  // two vertices: the count is the total length of the one slot's lists
    uint64_t total2 = 0;
    st = slot_total(ctx, cih, lj[1].off_base[0], n1, total2);
    if (st != CSM_OK) return st;
    kt.flush();
    res.count = std::min<uint64_t>(total2, cap);
    res.capped = total2 > cap;
    res.items = n1;
    finish_time();
    return CSM_OK;
  }

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

  std::vector<uint32_t> host_rows;
  uint64_t total = 0;
  uint64_t ref_mark = 0, ref_count = 0, ref_mat = 0;  // reference-algorithm bytes by class (host part)
  uint32_t k = 1;
  bool stop_all = false;

  // driver-degree prefix for level k and the first chunk choice
  auto prepare_level = [&](uint32_t k) -> csm_status {
    RowPass rp{};
    const bool count_level = k == stop && fused_rows;
    if (count_level) {
      rp.lv_last = lj[L];
      rp.fused_rows = 1;
      rp.row_count = row_count;
      rp.fbase = ws.fbase;
      rp.fbeg = ws.fbeg;
      rp.flen = ws.flen;
      rp.store_list = last_inj_k;
      rp.partials = ws.row_sums;
    }
    const uint32_t grid = grid_for(nrows[k] + 1);
    // per row: the row, a rank word + prefix + two offsets per list looked up
    const uint64_t per_row = 4ull * k + 24ull * (lv[k].nanchor + (count_level ? 1 : 0));
    if (count_level && row_count) {  // the whole level is counted here: no items
      kt.begin(CSM_K_AUX);
      anchor_len_kernel<<<grid, 256, 0, s>>>(ci, ws.rows[k], k, lj[k], nrows[k], ws.dscan[k], ws.drv[k], rp);
      reduce_add_kernel<<<reduce_grid(grid), 256, 0, s>>>(ws.row_sums, grid,
                                                          reinterpret_cast<unsigned long long *>(ws.ctl + kCtlCount));
      kt.end(CSM_K_AUX, nrows[k] * per_row);
      CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      res.items += nrows[k];
      total = pin[kCtlCount];
      if (total > cap) stop_all = true;
      cursor[k] = nrows[k];
      return CSM_OK;
    }
    // the reference-byte accounting's parent degrees, over every row of the
    // level: the rows of counted levels and of marked ones
    rp.off = g->off;
    rp.deg_sums = ws.deg_sums;
    const bool counted = k == stop && !want_list;
    kt.begin(CSM_K_AUX);
    anchor_len_kernel<<<grid, 256, 0, s>>>(ci, ws.rows[k], k, lj[k], nrows[k], ws.dscan[k], ws.drv[k], rp);
    reduce_add_kernel<<<reduce_grid(grid), 256, 0, s>>>(
        ws.deg_sums, grid, reinterpret_cast<unsigned long long *>(ws.ctl + (counted ? kCtlDegCount : kCtlDegMark)));
    scan_u64_exclusive(ws.dscan[k], nrows[k] + 1, ws.partials, nullptr, s);
    choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], 0, item_cap, ws.ctl + kCtlChunkA);
    kt.end(CSM_K_AUX, nrows[k] * (per_row + 9 + 8 + (count_level ? (last_inj_k ? 16 : 4) : 0)));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    cursor[k] = 0;
    next_b[k] = pin[kCtlChunkA];
    next_items[k] = pin[kCtlChunkA + 1];
    next_base[k] = pin[kCtlChunkA + 2];
    return CSM_OK;
  };
  auto next_chunk = [&](uint32_t k, uint64_t from, bool read_count) -> csm_status {
    choose_chunk_kernel<<<1, 64, 0, s>>>(ws.dscan[k], nrows[k], from, item_cap, ws.ctl + kCtlChunkB);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    next_b[k] = pin[kCtlChunkB];
    next_items[k] = pin[kCtlChunkB + 1];
    next_base[k] = pin[kCtlChunkB + 2];
    (void)read_count;
    return CSM_OK;
  };

  const auto deadline = t0 + std::chrono::milliseconds(o.time_limit_ms);
  st = prepare_level(1);
  if (st != CSM_OK) return st;
  while (!stop_all) {
    if (o.time_limit_ms && std::chrono::steady_clock::now() > deadline) {
      res.status = CSM_E_TIMEOUT;  // count so far is a lower bound
      break;
    }
    if (o.item_budget && res.items >= o.item_budget) {
      res.status = CSM_E_BUDGET;  // count so far is a lower bound
      break;
    }
    if (cursor[k] >= nrows[k]) {
      if (k == 1) break;
      --k;
      continue;
    }
    const uint64_t a = cursor[k], b = next_b[k], items = next_items[k];
    ExpandArgs ea{};
    ea.probe = probe;
    ea.ci = ci;
    ea.rows = ws.rows[k];
    ea.dscan = ws.dscan[k];
    ea.drv = ws.drv[k];
    ea.tile_row = ws.tile_row;
    ea.row_a = a;
    ea.row_b = b;
    ea.k = k;
    ea.lv = lj[k];
    ea.sym = sym[k];
    ea.mask = ws.mask;
    ea.tile_cnt = ws.tiles;
    ea.tile_off = ws.tiles;
    ea.count = reinterpret_cast<unsigned long long *>(ws.ctl + kCtlCount);
    ea.extra = reinterpret_cast<unsigned long long *>(ws.ctl + kCtlCountBytes);
    ea.item_base = next_base[k];
    ea.items = items;
    if (items == 0) {  // every row of the chunk has an empty driver list
      cursor[k] = b;
      if (cursor[k] < nrows[k]) {
        st = next_chunk(k, b, false);
        if (st != CSM_OK) return st;
      }
      continue;
    }
    const uint32_t ntiles = ceil_div(items, kTile);
    const uint64_t rows_in = b - a;
    const size_t lds_rows = 4ull * kTileRows * k;
    const bool counted = k == stop && !want_list;
    kt.begin(CSM_K_AUX);
    tile_rows_kernel<<<grid_for(rows_in), 256, 0, s>>>(ws.dscan[k], a, b, ntiles, ws.tile_row);
    kt.end(CSM_K_AUX, rows_in * 16ull + 4ull * ntiles);
    // the reference's extension pass over these rows, without the scan of the
    // parents' adjacency (added from the device sums at the end)
    (counted ? ref_count : ref_mark) += rows_in * (4ull * k + 8);
    res.items += items;
    // algorithmic bytes of one step: rows read, driver offsets, (adjacency id
    // + candidate bitmap word) per item; the rows written are added below
    const uint64_t step_in = rows_in * (4ull * k + 9) + items * 4ull;
    if (k == stop && !want_list) {
      // four tiles per workgroup; (count, bytes) per workgroup in ws.tiles
      const uint32_t ngroups = ceil_div(ntiles, static_cast<uint32_t>(kCIlp));
      if (fused) {
        ea.lv_last = lj[L];
        ea.fused_rows = fused_rows;
        ea.last_inj_k = last_inj_k;
        ea.fbase = ws.fbase;
        ea.fbeg = ws.fbeg;
        ea.flen = ws.flen;
      }
      kt.begin(CSM_K_COUNT);
      if (fused && o.lds_lists) count_kernel<true, true><<<ngroups, kTile, lds_rows, s>>>(ea, ntiles);
      else if (fused) count_kernel<true, false><<<ngroups, kTile, lds_rows, s>>>(ea, ntiles);
      else if (o.lds_lists) count_kernel<false, true><<<ngroups, kTile, lds_rows, s>>>(ea, ntiles);
      else count_kernel<false, false><<<ngroups, kTile, lds_rows, s>>>(ea, ntiles);
      kt.end(CSM_K_COUNT, step_in);
      reduce_add2_kernel<<<reduce_grid(ngroups), 256, 0, s>>>(ws.tiles, ngroups, ea.count, ea.extra);
      cursor[k] = b;
      st = next_chunk(k, b, true);
      if (st != CSM_OK) return st;
      total = pin[kCtlCount];
      if (total > cap) stop_all = true;
      continue;
    }
    // mark, scan, materialise into level k + 1
    kt.begin(CSM_K_EXPAND);
    expand_kernel<kMark><<<ntiles, kTile, lds_rows, s>>>(ea);
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
      expand_kernel<kMaterialize><<<ntiles, kTile, lds_rows, s>>>(ea);
      kt.end(CSM_K_MATERIALIZE, rows_in * (4ull * k + 9) + items * (4ull + 1) + produced * 4ull * (k + 1));
      ref_mat += produced * 4ull * (k + 1);
    }
    res.partial_rows += produced;
    cursor[k] = b;
    if (k == L) {  // list mode: rows are complete embeddings in lexicographic order
      const uint64_t lim = std::min(cap, o.list_capacity);
      const uint64_t want = total < lim ? std::min(produced, lim - total) : 0;
      if (want && dlist) {  // appended after the `total` rows listed so far
        to_query_order_kernel<<<grid_for(want * nq), 256, 0, s>>>(ws.rows[nq], want, nq, cmap, dlist + total * nq);
      } else if (want) {
        const size_t before = host_rows.size();
        host_rows.resize(before + want * nq);
        CSM_HIP_TRY(ctx, hipMemcpyAsync(host_rows.data() + before, ws.rows[nq], 4ull * nq * want,
                                        hipMemcpyDeviceToHost, s));
      }
      total += produced;
      st = next_chunk(k, b, false);
      if (st != CSM_OK) return st;
      if (total > cap) stop_all = true;
      continue;
    }
    if (cursor[k] < nrows[k]) {
      st = next_chunk(k, b, false);
      if (st != CSM_OK) return st;
    }
    if (produced) {
      nrows[k + 1] = produced;
      ++k;
      st = prepare_level(k);
      if (st != CSM_OK) return st;
    }
  }
  kt.flush();
  {  // join classes: bytes_ref = the reference's extension of the same rows
     // (join.cu:99-177): 4k + 8 per row, 12 per adjacency entry of its parent
     // vertex (id + candidate word), 4(k + 1) per row written; the work the
     // candidate index avoids (bytes: the kernels' own reads and writes,
     // without the entries read by membership searches)
    CSM_HIP_TRY(ctx, hipMemcpyAsync(pin, ws.ctl, 8 * kCtlSize, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    res.kernels[CSM_K_COUNT].bytes += pin[kCtlCountBytes];
    res.kernels[CSM_K_EXPAND].bytes_ref = ref_mark + 12ull * pin[kCtlDegMark];
    res.kernels[CSM_K_COUNT].bytes_ref = ref_count + 12ull * pin[kCtlDegCount];
    res.kernels[CSM_K_MATERIALIZE].bytes_ref = ref_mat;
  }
  res.count = std::min(total, cap);
  res.capped = total > cap;
  if (want_list && dlist) {
    // stable LSD radix sort over the columns, last to first, carrying a row
    // permutation; then one gather and one copy out
    const uint64_t nl = std::min(total, list_lim);
    res.listed = nl;
    if (nl) {
      uint32_t *k0 = static_cast<uint32_t *>(ws.get(4 * nl)), *k1 = static_cast<uint32_t *>(ws.get(4 * nl));
      uint32_t *p0 = static_cast<uint32_t *>(ws.get(4 * nl)), *p1 = static_cast<uint32_t *>(ws.get(4 * nl));
      uint32_t *out = static_cast<uint32_t *>(ws.get(4ull * nl * nq));
      size_t tmp_bytes = 0;
      void *tmp = nullptr;
      if (k0 && k1 && p0 && p1 && out) {
        rocprim::double_buffer<uint32_t> kd(k0, k1), pd(p0, p1);
        CSM_HIP_TRY(ctx, rocprim::radix_sort_pairs(nullptr, tmp_bytes, kd, pd, nl, 0, 32, s));
        tmp = ws.get(std::max<size_t>(tmp_bytes, 1));
      }
      if (tmp) {
        kt.begin(CSM_K_AUX);
        iota_kernel<<<grid_for(nl), 256, 0, s>>>(p0, nl);
        uint32_t *perm = p0;
        for (int c = static_cast<int>(nq) - 1; c >= 0; --c) {
          gather_col_kernel<<<grid_for(nl), 256, 0, s>>>(dlist, nq, static_cast<uint32_t>(c), perm, nl, k0);
          rocprim::double_buffer<uint32_t> kd(k0, k1), pd(perm, perm == p0 ? p1 : p0);
          CSM_HIP_TRY(ctx, rocprim::radix_sort_pairs(tmp, tmp_bytes, kd, pd, nl, 0, 32, s));
          perm = pd.current();
        }
        gather_rows_kernel<<<grid_for(nl * nq), 256, 0, s>>>(dlist, nq, perm, nl, out);
        kt.end(CSM_K_AUX, 24ull * nl * nq + 8ull * nl * nq);
        CSM_HIP_TRY(ctx, hipMemcpyAsync(o.list, out, 4ull * nl * nq, hipMemcpyDeviceToHost, s));
        CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      } else {  // no room for the sort: sort the copy on the host
        std::vector<uint32_t> qrows(nl * nq);
        CSM_HIP_TRY(ctx, hipMemcpyAsync(qrows.data(), dlist, 4ull * nl * nq, hipMemcpyDeviceToHost, s));
        CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
        sort_rows_host(qrows, nl, nq, o.list);
      }
    }
  } else if (want_list) {
    // matching-order rows -> query-id order, then sort lexicographically
    const uint64_t nl = host_rows.size() / nq;
    std::vector<uint32_t> qrows(host_rows.size());
    for (uint64_t i = 0; i < nl; ++i)
      for (uint32_t p = 0; p < nq; ++p) qrows[i * nq + lv[p].u] = host_rows[i * nq + p];
    sort_rows_host(qrows, nl, nq, o.list);
    res.listed = nl;
  }
  finish_time();
  return CSM_OK;
}

// Candidate-aware matching order: the root has the fewest candidates per
// query degree; each next position takes the unmatched vertex with the most
// matched neighbours, then the fewest candidates, then the highest degree,
// then the lowest id.  Its parent is its earliest matched neighbour, the
// others are its non-tree (back) positions.
csm_status plan_reorder_by_candidates(csm_query_plan &plan, const uint64_t *counts) {
  const uint32_t nq = plan.num_v;
  bool adj[CSM_MAX_QV][CSM_MAX_QV] = {};
  for (uint32_t e = 0; e < plan.num_e; ++e) {
    const uint32_t a = plan.edge_src[e], b = plan.edge_dst[e];
    if (a >= nq || b >= nq) return CSM_E_PLAN;
    if (a != b) adj[a][b] = adj[b][a] = true;
  }
  uint32_t order[CSM_MAX_QV], pos[CSM_MAX_QV];
  for (uint32_t u = 0; u < nq; ++u) pos[u] = CSM_NO_PARENT;
  uint32_t root = 0;
  for (uint32_t u = 1; u < nq; ++u)  // counts[u] / (deg[u] + 1) < counts[root] / (deg[root] + 1)
    if (static_cast<double>(counts[u]) * (plan.degree[root] + 1) <
        static_cast<double>(counts[root]) * (plan.degree[u] + 1))
      root = u;
  order[0] = root;
  pos[root] = 0;
  for (uint32_t p = 1; p < nq; ++p) {
    uint32_t best = CSM_NO_PARENT, best_nb = 0;
    for (uint32_t v = 0; v < nq; ++v) {
      if (pos[v] != CSM_NO_PARENT) continue;
      uint32_t nb = 0;
      for (uint32_t w = 0; w < nq; ++w) nb += adj[v][w] && pos[w] != CSM_NO_PARENT;
      if (nb == 0) continue;
      const bool better = best == CSM_NO_PARENT || nb > best_nb ||
                          (nb == best_nb && (counts[v] < counts[best] ||
                                             (counts[v] == counts[best] && plan.degree[v] > plan.degree[best])));
      if (better) {
        best = v;
        best_nb = nb;
      }
    }
    if (best == CSM_NO_PARENT) return CSM_E_QUERY;  // not connected
    order[p] = best;
    pos[best] = p;
  }
  for (uint32_t p = 0; p < nq; ++p) {
    const uint32_t u = order[p];
    plan.match_vertex[p] = u;
    plan.match_parent[p] = CSM_NO_PARENT;
    plan.num_back[p] = 0;
    if (p == 0) continue;
    for (uint32_t j = 0; j < p; ++j) {  // earlier positions in order
      if (!adj[u][order[j]]) continue;
      if (plan.match_parent[p] == CSM_NO_PARENT) plan.match_parent[p] = j;
      else plan.back_pos[p][plan.num_back[p]++] = j;
    }
  }
  return CSM_OK;
}

csm_status apply_reorder(csm_ctx *ctx, const csm_graph *g, csm_query *q) {
  csm_query_plan plan = q->plan;
  csm_status st = plan_reorder_by_candidates(plan, q->counts);
  if (st != CSM_OK) return set_error(ctx, st, "query is not connected");
  QueryDesc d;
  LevelDesc lv[CSM_MAX_QV];
  st = query_build_desc(plan, g->n, d, lv);
  if (st != CSM_OK) return set_error(ctx, st, "reordered plan is invalid");
  q->plan = plan;
  std::memcpy(q->level, lv, sizeof(lv));
  return CSM_OK;
}

}  // namespace csm

using namespace csm;

extern "C" csm_status csm_query_reorder(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !q || q->graph_n != g->n) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    if (!q->filtered) {
      const csm_status st = run_filter(ctx, g, q, refine);
      if (st != CSM_OK) return st;
    }
    return apply_reorder(ctx, g, q);
  });
}

extern "C" csm_status csm_match(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts *opts,
                                csm_result *out) {
  return csm::guarded(ctx, [&]() -> csm_status {
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
    if (o.reorder) {
      const csm_status st = apply_reorder(ctx, g, q);
      if (st != CSM_OK) {
        out->status = st;
        return st;
      }
    }
    for (int c = 0; c < CSM_K_NUM; ++c) out->kernels[c] = q->kstat[c];
    out->ms_filter = q->ms_filter;
    csm_status st = run_join(ctx, g, q, o, *out);
    if (st != CSM_OK || (out->status != CSM_E_TIMEOUT && out->status != CSM_E_BUDGET)) out->status = st;
    out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
  });
}

// csm_match (count only) with every workspace buffer in the library's static
// device scratch: no hipMalloc, no arena block, no allocating library call.
extern "C" csm_status csm_match_noalloc(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts *opts,
                                        csm_result *out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !q || !out || q->graph_n != g->n) return CSM_E_ARG;
    csm_match_opts o{};
    o.refine = 1;
    if (opts) o = *opts;
    std::memset(out, 0, sizeof(*out));
    if (o.list && o.list_capacity)
      return set_error(ctx, CSM_E_ARG, "csm_match_noalloc counts only: the sorted list needs arena memory");
    o.list = nullptr;
    hipSetDevice(ctx->device);
    void *region = scratch_region(ctx->device);
    if (!region) return set_error(ctx, CSM_E_DEVICE, "static scratch not resolved on this device (csm_open first)");
    WorkMem mem(ctx, region, kScratchBytes);
    const bool profiling = ctx->profiling;  // no HIP events: they may allocate
    ctx->profiling = false;
    auto t0 = std::chrono::steady_clock::now();
    csm_status st = CSM_OK;
    if (!(o.reuse_filter && q->filtered)) st = run_filter(ctx, g, q, o.refine, &mem);
    if (st == CSM_OK && o.reorder) st = apply_reorder(ctx, g, q);
    if (st == CSM_OK) {
      for (int c = 0; c < CSM_K_NUM; ++c) out->kernels[c] = q->kstat[c];
      out->ms_filter = q->ms_filter;
      mem.release();  // the filter's scratch is not needed by the join
      st = run_join(ctx, g, q, o, *out, &mem);
    }
    ctx->profiling = profiling;
    if (st == CSM_E_NOMEM)
      set_error(ctx, st, "workspace beyond the static scratch of " + std::to_string(kScratchBytes >> 20) +
                             " MiB (" + ctx->last_error + ")");
    if (st != CSM_OK || (out->status != CSM_E_TIMEOUT && out->status != CSM_E_BUDGET)) out->status = st;
    out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
  });
}
