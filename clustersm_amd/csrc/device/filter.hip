// Candidate filtering on gfx950, for one query or a batch (blockIdx.y = query).
//
// nlc_filter: one wave per block of 64 consecutive data vertices (lane =
// vertex).  Low-degree vertices count their neighbour labels serially into a
// per-lane LDS histogram over the query's label slots; vertices above
// kLightDegree are counted by the whole wave with LDS atomics.  The
// predicate of filter.cu:486-507 is then evaluated per query vertex and the
// 64 answers become one bitmap word through a ballot.
//
// refine_kernel: one step of the cluster refinement, a "pull" test: for each
// set bit v of the bitmap being refined, is there a neighbour of v in every
// one of the given support bitmaps?  Early exit on the first hit.  Step s of
// every query of the batch runs in the same launch.
#include <algorithm>
#include <chrono>

#include "common.hpp"
#include "filter.hpp"

namespace csm {

namespace {

constexpr int kFilterThreads = 256;
constexpr int kFilterWaves = kFilterThreads / kWave;
constexpr uint32_t kLightDegree = 64;
constexpr uint32_t kLabelTable = 1024;
constexpr uint8_t kNoSlot = 0xFF;

struct FilterSmem {
  QueryDesc d;
  uint8_t slot_of[kLabelTable];
  uint8_t hist[kFilterWaves][CSM_MAX_QV][kWave];
  uint32_t heavy[kFilterWaves][CSM_MAX_QV];
  unsigned long long cnt[CSM_MAX_QV];
};

__device__ inline uint32_t lookup_slot(const FilterSmem &sm, uint32_t l) {
  if (l < kLabelTable) return sm.slot_of[l];
  for (uint32_t s = 0; s < sm.d.num_slots; ++s)
    if (sm.d.slot_label[s] == l) return s;
  return kNoSlot;
}

__global__ __launch_bounds__(kFilterThreads) void nlc_filter_kernel(const uint32_t *__restrict__ off,
                                                                    const uint32_t *__restrict__ adj,
                                                                    const uint32_t *__restrict__ label, uint32_t n,
                                                                    const FilterQ *__restrict__ qs) {
  __shared__ FilterSmem sm;
  const FilterQ fq = qs[blockIdx.y];
  if (fq.use_table) return;
  const int tid = threadIdx.x;
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(fq.desc);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.d);
    for (uint32_t i = tid; i < sizeof(QueryDesc) / 4; i += kFilterThreads) dst[i] = src[i];
    for (uint32_t i = tid; i < kLabelTable; i += kFilterThreads) sm.slot_of[i] = kNoSlot;
    if (tid < CSM_MAX_QV) sm.cnt[tid] = 0;
  }
  __syncthreads();
  if (tid < static_cast<int>(sm.d.num_slots) && sm.d.slot_label[tid] < kLabelTable)
    sm.slot_of[sm.d.slot_label[tid]] = static_cast<uint8_t>(tid);
  __syncthreads();

  const int wave = tid / kWave, lane = tid % kWave;
  const uint32_t words = sm.d.words, nq = sm.d.nq, nslots = sm.d.num_slots;
  uint64_t *bits = fq.bits;
  uint64_t scanned = 0;
  for (uint32_t b = blockIdx.x * kFilterWaves + wave; b < words; b += gridDim.x * kFilterWaves) {
    const uint32_t v = b * 64u + static_cast<uint32_t>(lane);
    uint32_t s = kNoSlot, dv = 0, o0 = 0;
    if (v < n) {
      s = lookup_slot(sm, label[v]);
      o0 = off[v];
      dv = off[v + 1] - o0;
    }
    const bool active = s != kNoSlot && dv >= sm.d.slot_min_deg[s];
    for (uint32_t sl = 0; sl < nslots; ++sl) sm.hist[wave][sl][lane] = 0;
    if (active) scanned += dv;
    if (active && dv <= kLightDegree) {
      for (uint32_t o = o0; o < o0 + dv; ++o) {
        const uint32_t s2 = lookup_slot(sm, label[adj[o]]);
        if (s2 != kNoSlot && sm.hist[wave][s2][lane] < 255) sm.hist[wave][s2][lane]++;
      }
    }
    uint64_t heavy = ballot(active && dv > kLightDegree);
    while (heavy) {
      const int h = __builtin_ctzll(heavy);
      heavy &= heavy - 1;
      const uint32_t ho = __shfl(o0, h, kWave), hd = __shfl(dv, h, kWave);
      if (lane < CSM_MAX_QV) sm.heavy[wave][lane] = 0;
      wave_sync();
      for (uint32_t o = ho + lane; o < ho + hd; o += kWave) {
        const uint32_t s2 = lookup_slot(sm, label[adj[o]]);
        if (s2 != kNoSlot) atomicAdd(&sm.heavy[wave][s2], 1u);
      }
      wave_sync();
      if (lane < static_cast<int>(nslots)) sm.hist[wave][lane][h] = static_cast<uint8_t>(min(sm.heavy[wave][lane], 255u));
      wave_sync();
    }
    wave_sync();
    for (uint32_t u = 0; u < nq; ++u) {
      bool ok = active && sm.d.slot[u] == s && sm.d.deg[u] <= dv;
      if (ok)
        for (uint32_t t = sm.d.sig_off[u]; t < sm.d.sig_off[u + 1]; ++t)
          if (sm.hist[wave][sm.d.sig_slot[t]][lane] < sm.d.sig_cnt[t]) {
            ok = false;
            break;
          }
      const uint64_t m = ballot(ok);
      if (lane == 0) {
        bits[static_cast<uint64_t>(u) * words + b] = m;
        if (m) atomicAdd(&sm.cnt[u], static_cast<unsigned long long>(__popcll(m)));
      }
    }
  }
  for (int d = kWave / 2; d > 0; d >>= 1) scanned += __shfl_xor(scanned, d, kWave);
  __shared__ unsigned long long s_scan;
  if (tid == 0) s_scan = 0;
  __syncthreads();
  if (lane == 0 && scanned) atomicAdd(&s_scan, static_cast<unsigned long long>(scanned));
  __syncthreads();
  if (tid == 0 && s_scan) atomicAdd(&fq.stats[kStatFilterScanned], s_scan);
  if (tid < static_cast<int>(nq) && sm.cnt[tid]) atomicAdd(&fq.stats[tid], sm.cnt[tid]);
}

// The same predicate read from the graph's neighbour-label table: one wave
// per 64 vertices, each lane reads its vertex's label and degree; a wave with
// a vertex that can match copies its 64 table rows (4 KB, contiguous) into
// LDS with four coalesced 16-byte loads per lane, and every query vertex's
// signature is then checked against LDS.  A coalesced stream instead of an
// adjacency scan with label gathers, and one round trip per 64 vertices
// instead of one per query vertex.
__global__ __launch_bounds__(kFilterThreads) void nlc_table_filter_kernel(const uint32_t *__restrict__ off,
                                                                          const uint32_t *__restrict__ label,
                                                                          const uint8_t *__restrict__ nlc, uint32_t n,
                                                                          const FilterQ *__restrict__ qs) {
  static_assert(kNlcStride == 64, "a wave's 64 table rows are 256 16-byte pieces");
  // rows at a stride of 17 words, so that the lanes' reads of one byte
  // position of their own rows fall in different banks
  constexpr uint32_t kRowWords = kNlcStride / 4 + 1;
  __shared__ FilterSmem sm;
  __shared__ uint32_t s_row[kFilterWaves][kWave * kRowWords];
  const FilterQ fq = qs[blockIdx.y];
  if (!fq.use_table) return;
  const int tid = threadIdx.x;
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(fq.desc);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.d);
    for (uint32_t i = tid; i < sizeof(QueryDesc) / 4; i += kFilterThreads) dst[i] = src[i];
    for (uint32_t i = tid; i < kLabelTable; i += kFilterThreads) sm.slot_of[i] = kNoSlot;
    if (tid < CSM_MAX_QV) sm.cnt[tid] = 0;
  }
  __syncthreads();
  if (tid < static_cast<int>(sm.d.num_slots) && sm.d.slot_label[tid] < kLabelTable)
    sm.slot_of[sm.d.slot_label[tid]] = static_cast<uint8_t>(tid);
  __syncthreads();
  const int wave = tid / kWave, lane = tid % kWave;
  const uint32_t words = sm.d.words, nq = sm.d.nq;
  uint64_t *bits = fq.bits;
  uint64_t rows = 0;  // table rows copied (lane 0)
  const uint8_t *my_row = reinterpret_cast<const uint8_t *>(s_row[wave] + lane * kRowWords);
  // the same number of iterations in every wave of the workgroup (barriers)
  for (uint32_t base = blockIdx.x * kFilterWaves; base < words; base += gridDim.x * kFilterWaves) {
    const uint32_t b = base + static_cast<uint32_t>(wave);
    const bool wv = b < words;
    const uint32_t v = b * 64u + static_cast<uint32_t>(lane);
    uint32_t s = kNoSlot, dv = 0;
    if (wv && v < n) {
      s = lookup_slot(sm, label[v]);
      dv = off[v + 1] - off[v];
    }
    const bool active = s != kNoSlot && dv >= sm.d.slot_min_deg[s];
    const uint64_t am = ballot(active);
    if (am) {  // uniform in the wave
      const uint32_t nrow = min(64u, n - b * 64u);  // rows of this word that exist
      const uint4 *src = reinterpret_cast<const uint4 *>(nlc + static_cast<uint64_t>(b) * 64u * kNlcStride);
      uint4 q[4];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t c = j * kWave + static_cast<uint32_t>(lane);  // piece c holds bytes of row c / 4
        q[j] = c / 4 < nrow ? src[c] : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t c = j * kWave + static_cast<uint32_t>(lane);
        uint32_t *d = s_row[wave] + (c / 4) * kRowWords + (c % 4) * 4;
        d[0] = q[j].x;
        d[1] = q[j].y;
        d[2] = q[j].z;
        d[3] = q[j].w;
      }
      rows += lane == 0 ? nrow : 0;
    }
    __syncthreads();
    for (uint32_t u = 0; u < nq; ++u) {
      bool ok = active && sm.d.slot[u] == s && sm.d.deg[u] <= dv;
      if (ok)
        for (uint32_t t = sm.d.sig_off[u]; t < sm.d.sig_off[u + 1]; ++t)
          if (my_row[sm.d.slot_label[sm.d.sig_slot[t]]] < sm.d.sig_cnt[t]) {
            ok = false;
            break;
          }
      const uint64_t m = ballot(ok);
      if (lane == 0 && wv) {
        bits[static_cast<uint64_t>(u) * words + b] = m;
        if (m) atomicAdd(&sm.cnt[u], static_cast<unsigned long long>(__popcll(m)));
      }
    }
    __syncthreads();  // the rows are overwritten by the next iteration
  }
  // table rows copied (kNlcStride bytes each), counted as 8-byte units in the
  // scanned statistic
  for (int d = kWave / 2; d > 0; d >>= 1) rows += __shfl_xor(rows, d, kWave);
  __shared__ unsigned long long s_rows;
  if (tid == 0) s_rows = 0;
  __syncthreads();
  if (lane == 0 && rows) atomicAdd(&s_rows, static_cast<unsigned long long>(rows) * (kNlcStride / 8));
  __syncthreads();
  if (tid == 0 && s_rows) atomicAdd(&fq.stats[kStatFilterScanned], s_rows);
  if (tid < static_cast<int>(nq) && sm.cnt[tid]) atomicAdd(&fq.stats[tid], sm.cnt[tid]);
}

// CORE: x = core of cluster `step`, supports = its members (all must hold).
// MEMBER: x = member blockIdx.z, support = the core.
template <bool CORE>
__global__ __launch_bounds__(kFilterThreads) void refine_kernel(const uint32_t *__restrict__ off,
                                                                const uint32_t *__restrict__ adj, uint32_t n,
                                                                uint32_t words, const FilterQ *__restrict__ qs,
                                                                uint32_t step) {
  const FilterQ fq = qs[blockIdx.y];
  if (step >= fq.num_clusters) return;
  const uint32_t cb = fq.cl_off[step], ce = fq.cl_off[step + 1];
  if (ce - cb < 2) return;
  if (!CORE && blockIdx.z >= ce - cb - 1) return;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint32_t *cl_us = fq.cl_us;
  uint64_t *bits = fq.bits;
  const uint32_t x = CORE ? cl_us[cb] : cl_us[cb + 1 + blockIdx.z];
  const uint32_t y0 = CORE ? cb + 1 : cb, y1 = CORE ? ce : cb + 1;
  uint64_t *xrow = bits + static_cast<uint64_t>(x) * words;
  uint64_t scanned = 0;
  for (uint32_t b = blockIdx.x * kFilterWaves + wave; b < words; b += gridDim.x * kFilterWaves) {
    const uint64_t xw = xrow[b];
    if (xw == 0) continue;
    const uint32_t v = b * 64u + static_cast<uint32_t>(lane);
    const bool act = (xw >> lane) & 1ull;
    uint32_t o0 = 0, dv = 0;
    if (act && v < n) {
      o0 = off[v];
      dv = off[v + 1] - o0;
    }
    bool ok = act;
    if (act && dv <= kLightDegree) {
      for (uint32_t yi = y0; yi < y1 && ok; ++yi) {
        const uint64_t *yrow = bits + static_cast<uint64_t>(cl_us[yi]) * words;
        bool found = false;
        for (uint32_t o = o0; o < o0 + dv && !found; ++o) {
          const uint32_t w = adj[o];
          found = (yrow[w >> 6] >> (w & 63)) & 1ull;
          scanned++;
        }
        ok = found;
      }
    }
    uint64_t heavy = ballot(act && dv > kLightDegree);
    while (heavy) {
      const int h = __builtin_ctzll(heavy);
      heavy &= heavy - 1;
      const uint32_t ho = __shfl(o0, h, kWave), hd = __shfl(dv, h, kWave);
      bool hok = true;
      for (uint32_t yi = y0; yi < y1 && hok; ++yi) {
        const uint64_t *yrow = bits + static_cast<uint64_t>(cl_us[yi]) * words;
        bool any = false;
        for (uint32_t o = ho; o < ho + hd && !any; o += kWave) {
          bool f = false;
          if (o + lane < ho + hd) {
            const uint32_t w = adj[o + lane];
            f = (yrow[w >> 6] >> (w & 63)) & 1ull;
            scanned++;
          }
          any = ballot(f) != 0;
        }
        hok = any;
      }
      if (lane == h) ok = hok;
    }
    const uint64_t nw = ballot(ok);
    if (lane == 0 && nw != xw) xrow[b] = nw;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) scanned += __shfl_xor(scanned, d, kWave);
  __shared__ unsigned long long s_scan;
  if (threadIdx.x == 0) s_scan = 0;
  __syncthreads();
  if (lane == 0 && scanned) atomicAdd(&s_scan, static_cast<unsigned long long>(scanned));
  __syncthreads();
  if (threadIdx.x == 0 && s_scan) atomicAdd(&fq.stats[kStatRefineScanned], s_scan);
}

__global__ __launch_bounds__(kFilterThreads) void count_bits_kernel(uint32_t words, const FilterQ *__restrict__ qs) {
  const FilterQ fq = qs[blockIdx.y];
  const uint32_t u = blockIdx.z;
  if (u >= fq.nq) return;
  unsigned long long c = 0;
  for (uint32_t b = blockIdx.x * kFilterThreads + threadIdx.x; b < words; b += gridDim.x * kFilterThreads)
    c += __popcll(fq.bits[static_cast<uint64_t>(u) * words + b]);
  for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(&fq.stats[u], c);
}

// Zero the per-query statistics blocks (a kernel, not a memset: launch
// sequences may be captured into a graph).
__global__ void zero_stats_kernel(const FilterQ *__restrict__ qs, uint32_t nqueries, int counts_only) {
  const uint32_t q = blockIdx.x;
  if (q >= nqueries) return;
  const uint32_t n = counts_only ? CSM_MAX_QV : kStatSize;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) qs[q].stats[i] = 0;
}

uint32_t grid_words(uint32_t words, uint32_t per_block, uint32_t cap) {
  return std::max<uint32_t>(1, std::min<uint32_t>(cap, ceil_div(words, per_block)));
}

}  // namespace

void launch_filter_batch(hipStream_t s, const csm_graph *g, const FilterQ *d_qs, uint32_t nqueries, uint32_t words,
                         const std::vector<uint32_t> &step_members, uint32_t max_nq, bool refine, KernelTimer *kt) {
  // fewer blocks per query when many queries share the launch
  const uint32_t cap = std::max<uint32_t>(16, 4096 / std::max<uint32_t>(1, nqueries));
  const uint32_t gx = grid_words(words, kFilterWaves, cap);
  zero_stats_kernel<<<nqueries, 64, 0, s>>>(d_qs, nqueries, 0);
  if (kt) kt->begin(CSM_K_FILTER);
  nlc_filter_kernel<<<dim3(gx, nqueries), kFilterThreads, 0, s>>>(g->off, g->adj, g->label, g->n, d_qs);
  if (g->nlc)
    nlc_table_filter_kernel<<<dim3(gx, nqueries), kFilterThreads, 0, s>>>(g->off, g->label, g->nlc, g->n, d_qs);
  if (kt) kt->end(CSM_K_FILTER, 0);
  if (!refine) return;
  for (uint32_t c = 0; c < step_members.size(); ++c) {
    if (step_members[c] == 0) continue;  // no cluster with members at this step
    if (kt) kt->begin(CSM_K_REFINE);
    refine_kernel<true><<<dim3(gx, nqueries), kFilterThreads, 0, s>>>(g->off, g->adj, g->n, words, d_qs, c);
    refine_kernel<false><<<dim3(gx, nqueries, step_members[c]), kFilterThreads, 0, s>>>(g->off, g->adj, g->n, words,
                                                                                       d_qs, c);
    if (kt) kt->end(CSM_K_REFINE, 0);
  }
  zero_stats_kernel<<<nqueries, 64, 0, s>>>(d_qs, nqueries, 1);
  if (kt) kt->begin(CSM_K_AUX);
  count_bits_kernel<<<dim3(grid_words(words, kFilterThreads, cap), nqueries, max_nq), kFilterThreads, 0, s>>>(words,
                                                                                                           d_qs);
  if (kt) kt->end(CSM_K_AUX, 0);
}

std::vector<uint32_t> step_members_of(const std::vector<const csm_query *> &qs) {
  std::vector<uint32_t> m;
  for (const csm_query *q : qs) {
    const csm_query_plan &p = q->plan;
    if (m.size() < p.num_clusters) m.resize(p.num_clusters, 0);
    for (uint32_t c = 0; c < p.num_clusters; ++c) {
      const uint32_t sz = p.cluster_offsets[c + 1] - p.cluster_offsets[c];
      if (sz >= 2) m[c] = std::max(m[c], sz - 1);
    }
  }
  return m;
}

FilterQ filter_entry(const csm_graph *g, csm_query *q, uint64_t *stats) {
  FilterQ f{};
  bool covered = g->nlc != nullptr;
  for (uint32_t u = 0; u < q->desc.nq; ++u) covered = covered && q->desc.label[u] < g->nlc_labels;
  f.use_table = covered ? 1u : 0u;
  f.desc = q->d_desc;
  f.bits = q->bits;
  f.cl_us = q->d_cluster_us;
  f.cl_off = q->d_cluster_off;
  f.num_clusters = q->plan.num_clusters;
  f.nq = q->desc.nq;
  f.stats = reinterpret_cast<unsigned long long *>(stats);
  return f;
}

// Algorithmic bytes of one query's filter steps, from its statistics block:
// labels + offsets of every vertex, (neighbour id + neighbour label) per
// scanned adjacency entry, bitmap words written; refinement: (adjacency id +
// bitmap word) per probe, bitmap words read; counting: bitmap words read.
void add_filter_bytes(const csm_graph *g, const csm_query *q, const uint64_t *stats, bool refine, csm_kernel_stat *ks) {
  const uint64_t words = q->desc.words, nq = q->desc.nq;
  const uint64_t fb = 8ull * g->n + 4 + 8ull * stats[kStatFilterScanned] + 8ull * words * nq;
  uint64_t clusters_bytes = 0;
  for (uint32_t c = 0; c < q->plan.num_clusters; ++c)
    clusters_bytes += 24ull * words * (q->plan.cluster_offsets[c + 1] - q->plan.cluster_offsets[c]);
  const uint64_t rb = refine ? 12ull * stats[kStatRefineScanned] + clusters_bytes : 0;
  ks[CSM_K_FILTER].bytes += fb;
  ks[CSM_K_FILTER].bytes_ref += fb;
  ks[CSM_K_REFINE].bytes += rb;
  ks[CSM_K_REFINE].bytes_ref += rb;
  ks[CSM_K_AUX].bytes += 8ull * words * nq;
  ks[CSM_K_AUX].bytes_ref += 8ull * words * nq;
}

// Filter + refinement for one query on ctx->stream; fills q->counts.
csm_status run_filter(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine, WorkMem *mem) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const uint32_t words = q->desc.words, nq = q->desc.nq;
  WorkMem own(ctx);
  WorkMem &wm = mem ? *mem : own;
  uint64_t *stats = static_cast<uint64_t *>(wm.get(8 * kStatSize));
  FilterQ *d_q = static_cast<FilterQ *>(wm.get(sizeof(FilterQ)));
  if (!stats || !d_q) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (filter)");
  FilterQ fq = filter_entry(g, q, stats);
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_q, &fq, sizeof(FilterQ), hipMemcpyHostToDevice, s));
  std::memset(q->kstat, 0, sizeof(q->kstat));
  {
    KernelTimer kt(ctx, q->kstat);
    launch_filter_batch(s, g, d_q, 1, words, refine ? step_members_of({q}) : std::vector<uint32_t>(), nq, refine != 0,
                        &kt);
    uint64_t host[kStatSize];
    CSM_HIP_TRY(ctx, hipMemcpyAsync(host, stats, sizeof(host), hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    kt.flush();
    for (uint32_t u = 0; u < nq; ++u) q->counts[u] = host[u];
    add_filter_bytes(g, q, host, refine != 0, q->kstat);
  }
  q->filtered = true;
  q->ms_filter = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return CSM_OK;
}

}  // namespace csm

using namespace csm;

extern "C" {

csm_status csm_filter(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine, uint64_t *counts) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !q || q->graph_n != g->n) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    csm_status st = run_filter(ctx, g, q, refine);
    if (st != CSM_OK) return st;
    if (counts)
      for (uint32_t u = 0; u < q->desc.nq; ++u) counts[u] = q->counts[u];
    return CSM_OK;
  });
}

csm_status csm_candidates_download(csm_ctx *ctx, const csm_query *q, uint64_t *bits) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !q || !bits) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(bits, q->bits, 8ull * q->desc.words * q->desc.nq, hipMemcpyDeviceToHost,
                                    ctx->stream));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CSM_OK;
  });
}

}  // extern "C"
