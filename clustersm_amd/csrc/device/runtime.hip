// Context, arena, data-graph CSR build on the device, query descriptors.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "../host/host_graph.hpp"
#include "common.hpp"

namespace csm {

// ---------------------------------------------------------- static scratch
// Workspace of csm_match_noalloc: a file-scope device array, one copy per GPU
// the library runs on, held until the process ends.  Host code never uses the
// array's name as an address: its device address is resolved once per device
// (hipGetSymbolAddress, with that device current, at csm_open) and kept here.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

csm_status scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return CSM_E_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return CSM_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CSM_E_DEVICE;
  g_scratch_ptr[device] = p;
  return CSM_OK;
}
}  // namespace

void *scratch_region(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------------ arena
csm_status Arena::init(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  bytes = bytes / 4096 * 4096;
  if (bytes == 0) return CSM_E_ARG;
  if (hipMalloc(reinterpret_cast<void **>(&base_), bytes) != hipSuccess) {
    base_ = nullptr;
    return CSM_E_NOMEM;
  }
  size_ = bytes;
  free_.clear();
  live_.clear();
  free_[0] = bytes;
  used_ = peak_ = 0;
  return CSM_OK;
}

void Arena::release() {
  std::lock_guard<std::mutex> g(mu_);
  if (base_) hipFree(base_);
  base_ = nullptr;
  size_ = used_ = 0;
  free_.clear();
  live_.clear();
}

void *Arena::alloc(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  bytes = (std::max<uint64_t>(bytes, 1) + 255) / 256 * 256;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    uint64_t off = it->first, sz = it->second;
    free_.erase(it);
    if (sz > bytes) free_[off + bytes] = sz - bytes;
    live_[off] = bytes;
    used_ += bytes;
    peak_ = std::max(peak_, used_);
    return base_ + off;
  }
  return nullptr;
}

void Arena::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  uint64_t off = static_cast<uint64_t>(static_cast<char *>(p) - base_);
  auto it = live_.find(off);
  if (it == live_.end()) return;
  uint64_t sz = it->second;
  live_.erase(it);
  used_ -= sz;
  auto nx = free_.lower_bound(off);
  if (nx != free_.end() && off + sz == nx->first) {  // merge with next
    sz += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {                        // merge with previous
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      pv->second += sz;
      return;
    }
  }
  free_[off] = sz;
}

uint64_t Arena::largest_free() const {
  std::lock_guard<std::mutex> g(mu_);
  uint64_t best = 0;
  for (auto &kv : free_) best = std::max(best, kv.second);
  return best;
}

// --------------------------------------------------------------- CSR build
namespace {

__global__ void make_keys(const uint32_t *edges, uint64_t m, uint32_t shift, uint64_t *keys) {
  const uint64_t sentinel = (1ull << (2 * shift)) - 1;
  for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < m; k += 256ull * gridDim.x) {
    uint64_t s = edges[2 * k], d = edges[2 * k + 1];
    keys[2 * k] = s == d ? sentinel : (s << shift) | d;
    keys[2 * k + 1] = s == d ? sentinel : (d << shift) | s;
  }
}

__global__ void mark_unique(const uint64_t *keys, uint64_t n, uint32_t shift, uint64_t *flag) {
  const uint64_t sentinel = (1ull << (2 * shift)) - 1;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
    uint64_t k = keys[i];
    flag[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
  }
}

// Unique key i (at output position pos[i]) writes the adjacency entry; the
// first key of each source also writes the offsets of every vertex between
// the previous source and itself.
__global__ void scatter_csr(const uint64_t *keys, uint64_t n, uint32_t shift, const uint64_t *pos,
                            uint64_t total, uint32_t nv, uint32_t *off, uint32_t *adj) {
  const uint64_t sentinel = (1ull << (2 * shift)) - 1;
  const uint64_t dmask = (1ull << shift) - 1;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
    uint64_t k = keys[i];
    if (k == sentinel) continue;
    uint64_t src = k >> shift;
    if (i + 1 == n || keys[i + 1] == sentinel) {  // last valid key (possibly a duplicate)
      for (uint64_t v = src + 1; v <= nv; ++v) off[v] = static_cast<uint32_t>(total);
    }
    if (i > 0 && keys[i - 1] == k) continue;
    uint64_t p = pos[i];
    adj[p] = static_cast<uint32_t>(k & dmask);
    int64_t prev = i == 0 ? -1 : static_cast<int64_t>(keys[i - 1] >> shift);
    for (int64_t v = prev + 1; v <= static_cast<int64_t>(src); ++v) off[v] = static_cast<uint32_t>(p);
  }
}

__global__ void fill_off_empty(uint32_t *off, uint32_t nv) {  // graph without edges
  for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v <= nv; v += 256ull * gridDim.x) off[v] = 0;
}

__global__ void max_degree_kernel(const uint32_t *off, uint32_t nv, uint32_t *out) {
  uint32_t best = 0;
  for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v < nv; v += 256ull * gridDim.x)
    best = max(best, off[v + 1] - off[v]);
  for (int d = kWave / 2; d > 0; d >>= 1) best = max(best, __shfl_xor(best, d, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(out, best);
}

uint32_t grid_for(uint64_t n) { return static_cast<uint32_t>(std::min<uint64_t>(65536, (n + 255) / 256 + 1)); }

// One workgroup per hub: set the bit of every neighbour.
__global__ void hub_bits_kernel(const uint32_t *off, const uint32_t *adj, const uint32_t *hubs, uint32_t words,
                                uint64_t *bits) {
  const uint32_t h = blockIdx.x;
  const uint32_t x = hubs[h];
  uint64_t *row = bits + static_cast<uint64_t>(h) * words;
  for (uint32_t o = off[x] + threadIdx.x; o < off[x + 1]; o += blockDim.x) {
    const uint32_t w = adj[o];
    atomicOr(reinterpret_cast<unsigned long long *>(&row[w >> 6]), 1ull << (w & 63));
  }
}

}  // namespace

// Streaming copy, grid-stride, 16 bytes per lane per step.
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) dst[i] = src[i];
}

// Neighbour-label table: one wave per 64 vertices, a byte histogram per lane
// in LDS (saturating), heavy vertices counted by the whole wave.
__global__ __launch_bounds__(256) void nlc_table_kernel(const uint32_t *__restrict__ off,
                                                         const uint32_t *__restrict__ adj,
                                                         const uint32_t *__restrict__ label, uint32_t n,
                                                         uint32_t nlabels, uint8_t *__restrict__ nlc) {
  __shared__ uint32_t hist[4][kNlcStride / 4][kWave];  // [wave][label / 4][lane], 4 byte counters per word
  __shared__ uint32_t heavy[4][kNlcStride];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t words = (n + 63) / 64;
  for (uint32_t b = blockIdx.x * 4 + wave; b < words; b += gridDim.x * 4) {
    const uint32_t v = b * 64u + static_cast<uint32_t>(lane);
    uint32_t o0 = 0, dv = 0;
    if (v < n) {
      o0 = off[v];
      dv = off[v + 1] - o0;
    }
    for (uint32_t i = 0; i < kNlcStride / 4; ++i) hist[wave][i][lane] = 0;
    const bool light = dv <= 64;
    if (light) {
      for (uint32_t o = o0; o < o0 + dv; ++o) {
        const uint32_t l = label[adj[o]];
        if (l < nlabels) {
          const uint32_t sh = (l & 3) * 8, cur = (hist[wave][l >> 2][lane] >> sh) & 0xFF;
          if (cur < 255) hist[wave][l >> 2][lane] += 1u << sh;
        }
      }
    }
    uint64_t hv = ballot(!light);
    while (hv) {
      const int h = __builtin_ctzll(hv);
      hv &= hv - 1;
      const uint32_t ho = __shfl(o0, h, kWave), hd = __shfl(dv, h, kWave);
      heavy[wave][lane] = 0;
      wave_sync();
      for (uint32_t o = ho + lane; o < ho + hd; o += kWave) {
        const uint32_t l = label[adj[o]];
        if (l < nlabels) atomicAdd(&heavy[wave][l], 1u);
      }
      wave_sync();
      if (lane < static_cast<int>(kNlcStride / 4)) {
        uint32_t wv = 0;
        for (int j = 0; j < 4; ++j) wv |= min(heavy[wave][lane * 4 + j], 255u) << (8 * j);
        hist[wave][lane][h] = wv;
      }
      wave_sync();
    }
    wave_sync();
    if (v < n) {
      uint4 *row = reinterpret_cast<uint4 *>(nlc + static_cast<uint64_t>(v) * kNlcStride);
      for (uint32_t i = 0; i < kNlcStride / 16; ++i)
        row[i] = make_uint4(hist[wave][4 * i][lane], hist[wave][4 * i + 1][lane], hist[wave][4 * i + 2][lane],
                            hist[wave][4 * i + 3][lane]);
    }
    wave_sync();
  }
}

// Hubs: vertices of degree >= kHubMinDegree, highest degree first, while the
// bitmaps fit in the budget (1/16 of the arena, at most 4 GiB).
csm_status build_hubs(csm_ctx *ctx, csm_graph *g) {
  constexpr uint32_t kHubMinDegree = 128;
  hipStream_t s = ctx->stream;
  const uint32_t nv = g->n;
  const uint32_t words = ceil_div(nv, 64);
  std::vector<uint32_t> off(nv + 1ull);
  CSM_HIP_TRY(ctx, hipMemcpyAsync(off.data(), g->off, 4ull * (nv + 1ull), hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  std::vector<uint32_t> cand;
  g->num_heavy = 0;
  for (uint32_t v = 0; v < nv; ++v) {
    if (off[v + 1] - off[v] >= kHubMinDegree) cand.push_back(v);
    if (off[v + 1] - off[v] >= kHeavyDegree) ++g->num_heavy;
  }
  std::sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
    uint32_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    return da != db ? da > db : a < b;
  });
  const uint64_t budget = std::min<uint64_t>(ctx->arena.capacity() / 16, 4ull << 30);
  const uint64_t per = 8ull * words;
  size_t nh = std::min<size_t>(cand.size(), per ? budget / per : 0);
  cand.resize(nh);
  g->hub_id = static_cast<int32_t *>(ctx->arena.alloc(4ull * std::max<uint32_t>(nv, 1)));
  if (!g->hub_id) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (hub index)");
  std::vector<int32_t> id(nv, -1);
  for (size_t h = 0; h < nh; ++h) id[cand[h]] = static_cast<int32_t>(h);
  if (nv) CSM_HIP_TRY(ctx, hipMemcpyAsync(g->hub_id, id.data(), 4ull * nv, hipMemcpyHostToDevice, s));
  g->num_hubs = static_cast<uint32_t>(nh);
  g->hub_min_degree = nh ? off[cand.back() + 1] - off[cand.back()] : 0;
  if (nh) {
    g->hub_bits = static_cast<uint64_t *>(ctx->arena.alloc(per * nh));
    uint32_t *d_hubs = static_cast<uint32_t *>(ctx->arena.alloc(4ull * nh));
    if (!g->hub_bits || !d_hubs) {
      ctx->arena.free(d_hubs);
      return set_error(ctx, CSM_E_NOMEM, "arena exhausted (hub bitmaps)");
    }
    CSM_HIP_TRY(ctx, hipMemsetAsync(g->hub_bits, 0, per * nh, s));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(d_hubs, cand.data(), 4ull * nh, hipMemcpyHostToDevice, s));
    hub_bits_kernel<<<static_cast<uint32_t>(nh), 256, 0, s>>>(g->off, g->adj, d_hubs, words, g->hub_bits);
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->arena.free(d_hubs);
  }
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  return CSM_OK;
}

csm_status build_nlc(csm_ctx *ctx, csm_graph *g) {
  g->nlc_labels = std::min<uint32_t>(kNlcStride, g->max_label + 1);
  g->nlc = static_cast<uint8_t *>(ctx->arena.alloc(static_cast<uint64_t>(kNlcStride) * std::max<uint32_t>(g->n, 1)));
  if (!g->nlc) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (neighbour-label table)");
  if (g->n) {
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(65535, (ceil_div(g->n, 64) + 3) / 4));
    nlc_table_kernel<<<blocks, 256, 0, ctx->stream>>>(g->off, g->adj, g->label, g->n, g->nlc_labels, g->nlc);
  }
  CSM_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return CSM_OK;
}

csm_status upload_graph_device(csm_ctx *ctx, uint32_t nv, const uint32_t *labels, uint64_t m,
                               const uint32_t *edges, csm_graph **out) {
  hipStream_t s = ctx->stream;
  uint32_t shift = 1;
  while ((1ull << shift) <= nv) ++shift;  // 2^shift > nv: the sentinel is above every key
  // keys are (src << shift) | dst below a sentinel of 2^(2 shift) - 1
  if (2 * shift >= 64) return set_error(ctx, CSM_E_LIMIT, "vertex ids too wide (2^31 vertices or more)");
  const uint64_t nk = 2 * m;
  std::unique_ptr<csm_graph> g(new csm_graph());
  g->n = nv;
  g->label = static_cast<uint32_t *>(ctx->arena.alloc(4ull * std::max<uint32_t>(nv, 1)));
  g->off = static_cast<uint32_t *>(ctx->arena.alloc(4ull * (nv + 1ull)));
  if (!g->label || !g->off) {
    ctx->arena.free(g->label);
    ctx->arena.free(g->off);
    return set_error(ctx, CSM_E_NOMEM, "arena exhausted (graph arrays)");
  }
  // temporaries of the CSR build (freed on every exit) and the graph's own
  // arrays (freed on a failure)
  uint32_t *d_edges = nullptr;
  uint64_t *keys = nullptr, *keys2 = nullptr, *pos = nullptr, *partials = nullptr, *total_d = nullptr;
  auto free_tmp = [&]() {
    ctx->arena.free(d_edges);
    ctx->arena.free(keys);
    ctx->arena.free(keys2);
    ctx->arena.free(pos);
    ctx->arena.free(partials);
    ctx->arena.free(total_d);
    d_edges = nullptr;
    keys = keys2 = pos = partials = total_d = nullptr;
  };
  auto fail = [&](csm_status st, const std::string &msg) {
    hipStreamSynchronize(s);  // nothing in flight uses what is freed
    free_tmp();
    ctx->arena.free(g->label);
    ctx->arena.free(g->off);
    ctx->arena.free(g->adj);
    ctx->arena.free(g->hub_id);
    ctx->arena.free(g->hub_bits);
    ctx->arena.free(g->nlc);
    return set_error(ctx, st, msg);
  };
#define UPLOAD_TRY(expr)                                                             \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return fail(CSM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)
  if (nv) UPLOAD_TRY(hipMemcpyAsync(g->label, labels, 4ull * nv, hipMemcpyHostToDevice, s));
  uint32_t mx = 0;
  for (uint32_t v = 0; v < nv; ++v) mx = std::max(mx, labels[v]);
  g->max_label = mx;
  if (m == 0) {
    fill_off_empty<<<grid_for(nv + 1ull), 256, 0, s>>>(g->off, nv);
    g->adj = static_cast<uint32_t *>(ctx->arena.alloc(4));
    if (!g->adj) return fail(CSM_E_NOMEM, "arena exhausted (adjacency)");
    UPLOAD_TRY(hipStreamSynchronize(s));
    csm_status hst = build_hubs(ctx, g.get());
    if (hst == CSM_OK) hst = build_nlc(ctx, g.get());
    if (hst != CSM_OK) return fail(hst, std::string(ctx->last_error));
    *out = g.release();
    return CSM_OK;
  }
  // temporaries: edges, keys x2, flags/positions, partials, sort scratch
  d_edges = static_cast<uint32_t *>(ctx->arena.alloc(8 * m));
  keys = static_cast<uint64_t *>(ctx->arena.alloc(8 * nk));
  keys2 = static_cast<uint64_t *>(ctx->arena.alloc(8 * nk));
  pos = static_cast<uint64_t *>(ctx->arena.alloc(8 * nk));
  partials = static_cast<uint64_t *>(ctx->arena.alloc(8 * (nk / kScanTile + 2)));
  total_d = static_cast<uint64_t *>(ctx->arena.alloc(64));
  if (!d_edges || !keys || !keys2 || !pos || !partials || !total_d) {
    free_tmp();
    return fail(CSM_E_NOMEM, "arena exhausted (CSR build)");
  }
  UPLOAD_TRY(hipMemcpyAsync(d_edges, edges, 8 * m, hipMemcpyHostToDevice, s));
  make_keys<<<grid_for(m), 256, 0, s>>>(d_edges, m, shift, keys);
  rocprim::double_buffer<uint64_t> db(keys, keys2);
  size_t tmp_bytes = 0;
  rocprim::radix_sort_keys(nullptr, tmp_bytes, db, nk, 0, 2 * shift, s);
  void *tmp = ctx->arena.alloc(tmp_bytes);
  if (!tmp) {
    free_tmp();
    return fail(CSM_E_NOMEM, "arena exhausted (sort scratch)");
  }
  if (rocprim::radix_sort_keys(tmp, tmp_bytes, db, nk, 0, 2 * shift, s) != hipSuccess) {
    ctx->arena.free(tmp);
    free_tmp();
    return fail(CSM_E_DEVICE, "radix sort failed");
  }
  ctx->arena.free(tmp);
  const uint64_t *sorted = db.current();
  mark_unique<<<grid_for(nk), 256, 0, s>>>(sorted, nk, shift, pos);
  scan_u64_exclusive(pos, nk, partials, total_d, s);
  uint64_t total = 0;
  UPLOAD_TRY(hipMemcpyAsync(&total, total_d, 8, hipMemcpyDeviceToHost, s));
  UPLOAD_TRY(hipStreamSynchronize(s));
  if (total >= 0xFFFFFFFFull) {
    free_tmp();
    return fail(CSM_E_LIMIT, "2|E| does not fit the uint32 CSR");
  }
  g->adj = static_cast<uint32_t *>(ctx->arena.alloc(4 * std::max<uint64_t>(total, 1)));
  if (!g->adj) {
    free_tmp();
    return fail(CSM_E_NOMEM, "arena exhausted (adjacency)");
  }
  if (total == 0) fill_off_empty<<<grid_for(nv + 1ull), 256, 0, s>>>(g->off, nv);
  else scatter_csr<<<grid_for(nk), 256, 0, s>>>(sorted, nk, shift, pos, total, nv, g->off, g->adj);
  uint32_t *d_max = reinterpret_cast<uint32_t *>(total_d);
  UPLOAD_TRY(hipMemsetAsync(d_max, 0, 4, s));
  max_degree_kernel<<<grid_for(nv), 256, 0, s>>>(g->off, nv, d_max);
  uint32_t maxdeg = 0;
  UPLOAD_TRY(hipMemcpyAsync(&maxdeg, d_max, 4, hipMemcpyDeviceToHost, s));
  UPLOAD_TRY(hipStreamSynchronize(s));
  free_tmp();
  g->m = total / 2;
  g->max_degree = maxdeg;
  csm_status hst = build_hubs(ctx, g.get());
  if (hst == CSM_OK) hst = build_nlc(ctx, g.get());
  if (hst != CSM_OK) return fail(hst, std::string(ctx->last_error));
  *out = g.release();
  return CSM_OK;
#undef UPLOAD_TRY
}

// ------------------------------------------------------------ query desc
csm_status query_build_desc(const csm_query_plan &plan, uint32_t graph_n, QueryDesc &d, LevelDesc *lv) {
  std::memset(&d, 0, sizeof(d));
  const uint32_t nq = plan.num_v;
  if (nq < 2 || nq > CSM_MAX_QV) return CSM_E_QUERY;
  d.nq = nq;
  d.words = ceil_div(graph_n, 64);
  for (uint32_t u = 0; u < nq; ++u) {
    d.label[u] = plan.labels[u];
    d.deg[u] = plan.degree[u];
    uint32_t s = 0;
    while (s < d.num_slots && d.slot_label[s] != plan.labels[u]) ++s;
    if (s == d.num_slots) {
      d.slot_label[s] = plan.labels[u];
      d.slot_min_deg[s] = plan.degree[u];
      ++d.num_slots;
    }
    d.slot_min_deg[s] = std::min(d.slot_min_deg[s], plan.degree[u]);
    d.slot[u] = s;
  }
  // NLC signature from the full query (filter.cu:555-566), labels -> slots
  uint32_t cnt[CSM_MAX_QV][CSM_MAX_QV] = {};
  for (uint32_t k = 0; k < plan.num_e; ++k) {
    uint32_t a = plan.edge_src[k], b = plan.edge_dst[k];
    ++cnt[a][d.slot[b]];
    ++cnt[b][d.slot[a]];
  }
  uint32_t w = 0;
  for (uint32_t u = 0; u < nq; ++u) {
    d.sig_off[u] = w;
    for (uint32_t s = 0; s < d.num_slots; ++s)
      if (cnt[u][s]) {
        d.sig_slot[w] = static_cast<uint8_t>(s);
        d.sig_cnt[w] = static_cast<uint8_t>(cnt[u][s]);
        ++w;
      }
  }
  d.sig_off[nq] = w;
  for (uint32_t p = 0; p < nq; ++p) {
    LevelDesc &L = lv[p];
    std::memset(&L, 0, sizeof(L));
    L.u = plan.match_vertex[p];
    L.parent = p == 0 ? 0 : plan.match_parent[p];
    if (p > 0 && L.parent >= p) return CSM_E_PLAN;
    L.nback = plan.num_back[p];
    for (uint32_t t = 0; t < L.nback; ++t) {
      L.back[t] = plan.back_pos[p][t];
      if (L.back[t] >= p) return CSM_E_PLAN;
    }
    for (uint32_t j = 0; j < p; ++j)
      if (plan.labels[plan.match_vertex[j]] == plan.labels[L.u]) L.inj[L.ninj++] = j;
    L.nanchor = p == 0 ? 0 : 1 + L.nback;
    L.anchor[0] = L.parent;
    for (uint32_t t = 0; t < L.nback; ++t) L.anchor[1 + t] = L.back[t];
  }
  return CSM_OK;
}

}  // namespace csm

using namespace csm;

// ================================================================ C ABI
extern "C" {

csm_status csm_open(int device, uint64_t arena_bytes, csm_ctx **out) {
  return csm::guarded(nullptr, [&]() -> csm_status {
    if (!out) return CSM_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return CSM_E_NO_GPU;
    if (device < 0 || device >= count) return CSM_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return CSM_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CSM_E_DEVICE;
    std::unique_ptr<csm_ctx> ctx(new csm_ctx());
    ctx->device = device;
    ctx->num_cus = static_cast<uint32_t>(prop.multiProcessorCount);
    ctx->arch = prop.gcnArchName;
    if (ctx->arch.rfind("gfx950", 0) != 0) return CSM_E_NO_GPU;  // built for gfx950 only
    if (scratch_setup(device) != CSM_OK) return CSM_E_DEVICE;  // device current
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return CSM_E_DEVICE;
    if (hipHostMalloc(reinterpret_cast<void **>(&ctx->pinned), 4096, hipHostMallocDefault) != hipSuccess) {
      hipStreamDestroy(ctx->stream);
      return CSM_E_DEVICE;
    }
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    if (arena_bytes == 0) arena_bytes = static_cast<uint64_t>(free_b * 0.85);
    if (arena_bytes > free_b) arena_bytes = free_b - (free_b >> 6);
    csm_status st = ctx->arena.init(arena_bytes);
    if (st != CSM_OK) {
      hipHostFree(ctx->pinned);
      hipStreamDestroy(ctx->stream);
      return st;
    }
    *out = ctx.release();
    return CSM_OK;
  });
}

csm_status csm_close(csm_ctx *ctx) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx) return CSM_E_ARG;
    csm_serve_stop(ctx);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
    ctx->arena.release();
    if (ctx->pinned) hipHostFree(ctx->pinned);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return CSM_OK;
  });
}

const char *csm_last_error(const csm_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

csm_status csm_set_profiling(csm_ctx *ctx, int on) {
  if (!ctx) return CSM_E_ARG;
  ctx->profiling = on != 0;
  return CSM_OK;
}

csm_status csm_copy_probe(csm_ctx *ctx, uint64_t bytes, uint32_t iters, double *ms_per_copy) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !ms_per_copy || bytes < 16 || iters == 0) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    bytes &= ~15ull;
    void *src = ctx->arena.alloc(bytes), *dst = ctx->arena.alloc(bytes);
    if (!src || !dst) {
      ctx->arena.free(src);
      ctx->arena.free(dst);
      return set_error(ctx, CSM_E_NOMEM, "arena exhausted (copy probe)");
    }
    const uint64_t n = bytes / 16;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(ctx->num_cus * 8ull, (n + 255) / 256));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    copy_kernel<<<blocks, 256, 0, ctx->stream>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);  // warm
    hipEventRecord(e0, ctx->stream);
    for (uint32_t i = 0; i < iters; ++i)
      copy_kernel<<<blocks, 256, 0, ctx->stream>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    hipEventRecord(e1, ctx->stream);
    const hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->arena.free(src);
    ctx->arena.free(dst);
    if (err != hipSuccess) return set_error(ctx, CSM_E_DEVICE, hipGetErrorString(err));
    *ms_per_copy = ms / iters;
    return CSM_OK;
  });
}

csm_status csm_ctx_query(const csm_ctx *ctx, csm_ctx_info *out) {
  if (!ctx || !out) return CSM_E_ARG;
  std::memset(out, 0, sizeof(*out));
  out->device = ctx->device;
  out->num_cus = ctx->num_cus;
  out->arena_bytes = ctx->arena.capacity();
  out->arena_used = ctx->arena.used();
  out->arena_peak = ctx->arena.peak();
  std::strncpy(out->arch, ctx->arch.c_str(), sizeof(out->arch) - 1);
  return CSM_OK;
}

csm_status csm_graph_upload(csm_ctx *ctx, const csm_host_graph *g, csm_graph **out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !out) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    return upload_graph_device(ctx, g->n, g->labels.data(), g->m, g->edges.data(), out);
  });
}

csm_status csm_graph_upload_edges(csm_ctx *ctx, uint32_t num_v, const uint32_t *labels, uint64_t num_e,
                                  const uint32_t *edges, csm_graph **out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !out || (num_v && !labels) || (num_e && !edges)) return CSM_E_ARG;
    for (uint64_t i = 0; i < 2 * num_e; ++i)
      if (edges[i] >= num_v) return set_error(ctx, CSM_E_ARG, "edge endpoint out of range");
    hipSetDevice(ctx->device);
    return upload_graph_device(ctx, num_v, labels, num_e, edges, out);
  });
}

csm_status csm_graph_free(csm_ctx *ctx, csm_graph *g) {
  if (!ctx || !g) return CSM_E_ARG;
  hipStreamSynchronize(ctx->stream);
  ctx->arena.free(g->off);
  ctx->arena.free(g->adj);
  ctx->arena.free(g->label);
  ctx->arena.free(g->hub_id);
  ctx->arena.free(g->hub_bits);
  ctx->arena.free(g->nlc);
  delete g;
  return CSM_OK;
}

csm_status csm_graph_info(const csm_graph *g, uint32_t *num_v, uint64_t *num_e, uint32_t *max_degree) {
  if (!g) return CSM_E_ARG;
  if (num_v) *num_v = g->n;
  if (num_e) *num_e = g->m;
  if (max_degree) *max_degree = g->max_degree;
  return CSM_OK;
}

csm_status csm_graph_download(csm_ctx *ctx, const csm_graph *g, uint32_t *offsets, uint32_t *adj,
                              uint32_t *labels) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    if (offsets) CSM_HIP_TRY(ctx, hipMemcpyAsync(offsets, g->off, 4ull * (g->n + 1ull), hipMemcpyDeviceToHost, ctx->stream));
    if (adj && g->m) CSM_HIP_TRY(ctx, hipMemcpyAsync(adj, g->adj, 8ull * g->m, hipMemcpyDeviceToHost, ctx->stream));
    if (labels && g->n) CSM_HIP_TRY(ctx, hipMemcpyAsync(labels, g->label, 4ull * g->n, hipMemcpyDeviceToHost, ctx->stream));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CSM_OK;
  });
}

csm_status csm_graph_to_host(csm_ctx *ctx, const csm_graph *g, csm_host_graph **out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !out) return CSM_E_ARG;
    std::vector<uint32_t> off(g->n + 1ull), adj(2 * g->m);
    std::unique_ptr<csm_host_graph> h(new csm_host_graph());
    h->n = g->n;
    h->labels.resize(g->n);
    csm_status st = csm_graph_download(ctx, g, off.data(), adj.data(), h->labels.data());
    if (st != CSM_OK) return st;
    h->edges.reserve(2 * g->m);
    for (uint32_t v = 0; v < g->n; ++v)
      for (uint32_t o = off[v]; o < off[v + 1]; ++o)
        if (adj[o] > v) {
          h->edges.push_back(v);
          h->edges.push_back(adj[o]);
        }
    h->m = h->edges.size() / 2;
    h->off.assign(off.begin(), off.end());
    h->adj = std::move(adj);
    h->csr_ready = true;
    *out = h.release();
    return CSM_OK;
  });
}

csm_status csm_query_prepare(csm_ctx *ctx, const csm_graph *g, const csm_query_plan *plan, csm_query **out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !plan || !out) return CSM_E_ARG;
    std::unique_ptr<csm_query> q(new csm_query());
    q->plan = *plan;
    csm_status st = query_build_desc(*plan, g->n, q->desc, q->level);
    if (st != CSM_OK) return set_error(ctx, st, "invalid query plan");
    q->graph_n = g->n;
    uint64_t bytes = 8ull * q->desc.words * q->desc.nq;
    q->bits = static_cast<uint64_t *>(ctx->arena.alloc(bytes));
    q->d_desc = static_cast<QueryDesc *>(ctx->arena.alloc(sizeof(QueryDesc)));
    q->d_cluster_us = static_cast<uint32_t *>(ctx->arena.alloc(4ull * std::max<uint32_t>(plan->num_cluster_us, 1)));
    q->d_cluster_off = static_cast<uint32_t *>(ctx->arena.alloc(4ull * (plan->num_clusters + 1ull)));
    if (!q->bits || !q->d_desc || !q->d_cluster_us || !q->d_cluster_off) {
      ctx->arena.free(q->bits);
      ctx->arena.free(q->d_desc);
      ctx->arena.free(q->d_cluster_us);
      ctx->arena.free(q->d_cluster_off);
      return set_error(ctx, CSM_E_NOMEM, "arena exhausted (query)");
    }
    hipSetDevice(ctx->device);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(q->d_desc, &q->desc, sizeof(QueryDesc), hipMemcpyHostToDevice, ctx->stream));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(q->d_cluster_us, plan->cluster_us, 4ull * plan->num_cluster_us,
                                    hipMemcpyHostToDevice, ctx->stream));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(q->d_cluster_off, plan->cluster_offsets, 4ull * (plan->num_clusters + 1ull),
                                    hipMemcpyHostToDevice, ctx->stream));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *out = q.release();
    return CSM_OK;
  });
}

csm_status csm_query_free(csm_ctx *ctx, csm_query *q) {
  if (!ctx || !q) return CSM_E_ARG;
  hipStreamSynchronize(ctx->stream);
  ctx->arena.free(q->bits);
  ctx->arena.free(q->d_desc);
  ctx->arena.free(q->d_cluster_us);
  ctx->arena.free(q->d_cluster_off);
  delete q;
  return CSM_OK;
}

}  // extern "C"
