// Shared device-side definitions for the gfx950 kernels and the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "csm.h"

struct csm_host_graph;

namespace csm {

constexpr int kWave = 64;  // CDNA wavefront

// ------------------------------------------------------------------ arena
// One hipMalloc per context; first-fit sub-allocation with 256-byte
// alignment.  Replaces the reference's per-step cudaMalloc/cudaFree.
class Arena {
 public:
  csm_status init(uint64_t bytes);
  void release();
  void *alloc(uint64_t bytes);
  void free(void *p);
  uint64_t capacity() const { return size_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  uint64_t largest_free() const;

 private:
  char *base_ = nullptr;
  uint64_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<uint64_t, uint64_t> free_;   // offset -> size
  std::map<uint64_t, uint64_t> live_;   // offset -> size
  mutable std::mutex mu_;
};

class Server;

}  // namespace csm

struct csm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t num_cus = 0;
  std::string arch;
  csm::Arena arena;
  std::string last_error;
  // small pinned staging block for control reads (hipHostMalloc)
  uint64_t *pinned = nullptr;
  csm::Server *server = nullptr;  // owned; released by csm_serve_stop
  std::mutex mu;
  bool profiling = false;
  std::vector<hipEvent_t> event_pool;
};

struct csm_graph {
  uint32_t n = 0;
  uint64_t m = 0;          // undirected edges after dedup
  uint32_t max_degree = 0;
  uint32_t *off = nullptr; // n + 1
  uint32_t *adj = nullptr; // 2m
  uint32_t *label = nullptr;
  uint32_t max_label = 0;
  // Dense adjacency bitmaps of the highest-degree vertices: an edge probe
  // against a hub is one load instead of a binary search.
  int32_t *hub_id = nullptr;     // n entries, -1 for non-hubs
  uint64_t *hub_bits = nullptr;  // num_hubs x ceil(n / 64)
  uint32_t num_hubs = 0;
  uint32_t hub_min_degree = 0;
  // Neighbour-label counts per vertex, saturating at 255: row v holds
  // kNlcStride bytes, byte l = |{w in N(v) : label(w) = l}| for l < nlc_labels.
  uint8_t *nlc = nullptr;
  uint32_t nlc_labels = 0;
  uint32_t num_heavy = 0;  // vertices of degree >= csm::kHeavyDegree
};

namespace csm {

constexpr uint32_t kNlcStride = 64;  // labels covered by csm_graph::nlc
// Candidate lists of vertices at least this degree are built by a whole
// workgroup each, outside the per-word walk of the other candidates.
constexpr uint32_t kHeavyDegree = 1024;

// Per-query descriptor, passed to kernels by value.  Replaces the reference's
// __constant__ globals (globals.cu:25-38) so several queries can be resident.
struct QueryDesc {
  uint32_t nq;
  uint32_t words;                    // ceil(V / 64)
  uint32_t num_slots;                // distinct query labels
  uint32_t label[CSM_MAX_QV];
  uint32_t slot[CSM_MAX_QV];         // label slot of u
  uint32_t deg[CSM_MAX_QV];          // full-query degree
  uint32_t slot_label[CSM_MAX_QV];   // label of slot s
  uint32_t slot_min_deg[CSM_MAX_QV]; // smallest degree among query vertices of slot s
  // neighbour-label-count signature: for u, pairs (slot, count) with count > 0
  uint32_t sig_off[CSM_MAX_QV + 1];
  uint8_t sig_slot[CSM_MAX_QV * CSM_MAX_QV];
  uint8_t sig_cnt[CSM_MAX_QV * CSM_MAX_QV];
};

// Matching order as the join kernels read it (one level = one position).
// The anchors of a position are its parent followed by its non-tree
// neighbours: a new vertex must be adjacent to the vertices of every anchor.
struct LevelDesc {
  uint32_t u;            // query vertex matched at this position
  uint32_t parent;       // parent position
  uint32_t nback;        // non-tree checks against earlier positions
  uint32_t back[CSM_MAX_QV];
  uint32_t ninj;         // earlier positions with the same label (injectivity)
  uint32_t inj[CSM_MAX_QV];
  uint32_t nanchor;      // 1 + nback
  uint32_t anchor[CSM_MAX_QV + 1];
};

// Edge probe: dense bitmap for hubs, binary search in the sorted list otherwise.
struct AdjProbe {
  const uint32_t *off;
  const uint32_t *adj;
  const int32_t *hub_id;
  const uint64_t *hub_bits;
  uint32_t words;
  __device__ inline bool has(uint32_t x, uint32_t w) const {
    const int32_t h = hub_id[x];
    if (h >= 0) return (hub_bits[static_cast<uint64_t>(h) * words + (w >> 6)] >> (w & 63)) & 1ull;
    uint32_t lo = off[x], hi = off[x + 1];
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t y = adj[mid];
      if (y == w) return true;
      if (y < w) lo = mid + 1; else hi = mid;
    }
    return false;
  }
};

}  // namespace csm

struct csm_query {
  csm_query_plan plan;
  csm::QueryDesc desc;
  csm::LevelDesc level[CSM_MAX_QV];
  uint32_t graph_n = 0;
  uint64_t *bits = nullptr;          // nq x words candidate bitmaps (arena)
  csm::QueryDesc *d_desc = nullptr;  // device copy of desc (arena)
  uint32_t *d_cluster_us = nullptr;  // flattened clusters (arena)
  uint32_t *d_cluster_off = nullptr; // cluster offsets, num_clusters + 1 (arena)
  uint64_t counts[CSM_MAX_QV] = {};  // |C(u)| after the last filter
  bool filtered = false;
  double ms_filter = 0;
  csm_kernel_stat kstat[CSM_K_NUM] = {};  // filter-side kernel accounting
};

namespace csm {

inline csm_status set_error(csm_ctx *ctx, csm_status st, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return st;
}

// Body of an exported entry point: an exception thrown inside (host memory
// exhausted, a thread that cannot start) becomes a status and a message
// instead of crossing the C ABI.
template <class F>
csm_status guarded(csm_ctx *ctx, F &&body) noexcept {
  csm_status st = CSM_E_DEVICE;
  const char *what = "unknown exception";
  try {
    return body();
  } catch (const std::bad_alloc &) {
    st = CSM_E_NOMEM;
    what = "host memory exhausted";
  } catch (const std::length_error &) {
    st = CSM_E_NOMEM;
    what = "host memory exhausted (container length)";
  } catch (const std::exception &e) {
    try {  // the text lives in the exception object: copy it while it exists
      set_error(ctx, st, std::string("runtime error: ") + e.what());
    } catch (...) {
    }
    return st;
  } catch (...) {
  }
  try {
    set_error(ctx, st, what);
  } catch (...) {
  }
  return st;
}

// Device memory for one call's workspace: either sub-allocations of the
// context arena (returned to it when the WorkMem goes away), or a bump region
// over a fixed device buffer (nothing allocated, nothing returned).
class WorkMem {
 public:
  explicit WorkMem(csm_ctx *ctx) : ctx_(ctx) {}
  WorkMem(csm_ctx *ctx, void *region, uint64_t size)
      : ctx_(ctx), region_(static_cast<char *>(region)), size_(size) {}
  ~WorkMem() { release(); }
  WorkMem(const WorkMem &) = delete;
  WorkMem &operator=(const WorkMem &) = delete;
  void *get(uint64_t bytes) {
    bytes = (std::max<uint64_t>(bytes, 1) + 255) / 256 * 256;
    if (region_) {
      if (bytes > size_ - used_) return nullptr;
      void *p = region_ + used_;
      used_ += bytes;
      return p;
    }
    void *p = ctx_->arena.alloc(bytes);
    if (p) blocks_.push_back(p);
    return p;
  }
  uint64_t largest_free() const { return region_ ? size_ - used_ : ctx_->arena.largest_free(); }
  bool fixed() const { return region_ != nullptr; }
  uint64_t capacity() const { return size_; }
  void release() {
    for (void *p : blocks_) ctx_->arena.free(p);
    blocks_.clear();
    used_ = 0;
  }

 private:
  csm_ctx *ctx_;
  char *region_ = nullptr;
  uint64_t size_ = 0, used_ = 0;
  std::vector<void *> blocks_;
};

// The library's static device scratch (runtime.hip): base address on
// `device`, resolved once per device at csm_open; nullptr before that.
constexpr uint64_t kScratchBytes = CSM_NOALLOC_SCRATCH_BYTES;
void *scratch_region(int device);

#define CSM_HIP_TRY(ctx, expr)                                                             \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return ::csm::set_error((ctx), CSM_E_DEVICE,                                         \
                              std::string(#expr) + ": " + hipGetErrorString(_e));          \
  } while (0)

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// Wave-level helpers --------------------------------------------------------
__device__ inline uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane.
__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// Host helpers (implemented in runtime.hip)
csm_status query_build_desc(const csm_query_plan &plan, uint32_t graph_n, QueryDesc &d, LevelDesc *lv);
csm_status upload_graph_device(csm_ctx *ctx, uint32_t n, const uint32_t *labels, uint64_t m,
                               const uint32_t *edges, csm_graph **out);

// Per-launch accounting; device time only when ctx->profiling is set.
class KernelTimer {
 public:
  KernelTimer(csm_ctx *ctx, csm_kernel_stat *stats) : ctx_(ctx), stats_(stats) {}
  ~KernelTimer() { flush(); }
  void begin(int cls) {
    stats_[cls].launches++;
    if (!ctx_->profiling) return;
    hipEvent_t e = get();
    hipEventRecord(e, ctx_->stream);
    open_ = e;
  }
  void end(int cls, uint64_t bytes) {
    stats_[cls].bytes += bytes;
    stats_[cls].bytes_ref += bytes;
    if (!ctx_->profiling) return;
    hipEvent_t e = get();
    hipEventRecord(e, ctx_->stream);
    pending_.push_back({cls, open_, e});
  }
  void flush() {
    for (auto &p : pending_) {
      hipEventSynchronize(p.e1);
      float ms = 0;
      hipEventElapsedTime(&ms, p.e0, p.e1);
      stats_[p.cls].ns += static_cast<uint64_t>(static_cast<double>(ms) * 1e6);
      ctx_->event_pool.push_back(p.e0);
      ctx_->event_pool.push_back(p.e1);
    }
    pending_.clear();
  }

 private:
  struct P {
    int cls;
    hipEvent_t e0, e1;
  };
  hipEvent_t get() {
    if (ctx_->event_pool.empty()) {
      hipEvent_t e;
      hipEventCreate(&e);
      return e;
    }
    hipEvent_t e = ctx_->event_pool.back();
    ctx_->event_pool.pop_back();
    return e;
  }
  csm_ctx *ctx_;
  csm_kernel_stat *stats_;
  hipEvent_t open_ = nullptr;
  std::vector<P> pending_;
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Device scans (scan.hip): in-place exclusive scan of u64, returns total.
struct ScanScratch {
  uint64_t *partials = nullptr;  // >= ceil(n / kScanTile) + 1 entries
  uint64_t cap = 0;
};
constexpr uint32_t kScanTile = 4096;
void scan_u64_exclusive(uint64_t *data, uint64_t n, uint64_t *partials, uint64_t *total_out, hipStream_t s);

}  // namespace csm
