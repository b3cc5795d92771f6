// Batched matching: every query of a batch in one persistent kernel.
//
// The reference runs one process per query (runClusterSM.sh:9-13) and one
// host-synchronised launch per query edge (join.cu:348-434).  Here the
// filter, the refinement steps and the candidate indexes of all queries are
// built with one launch per step for the whole batch, then one persistent
// kernel extends the partial embeddings of all queries together:
//
//   task   = (query, level k, rows of k ids in a pool block, resume point);
//            level-1 tasks name a range of the root candidate bitmap
//   stacks = one LIFO stack of tasks per query: the deepest, lexicographically
//            first rows are extended first (depth-first per query), so a
//            capped query stops after little more than the work it needs;
//            workgroups start spread over the queries and move to another
//            query only when theirs has no task
//   pool   = fixed-size blocks of device memory for rows; a workgroup writes
//            survivors of level k into blocks of level k + 1 and pushes them
//            when the task ends or yields
//   stop   = a task at a query's count level counts (or, when the last
//            position has no non-tree edge, adds candidate-list lengths) and
//            nothing is written; a query whose count passes its cap is done
//            and its remaining tasks are dropped
//
// A task yields after filling kMaxHeld blocks, or when the pool is empty,
// and is pushed back with its resume point below the blocks it produced.  If
// the pool stays empty without progress, or a time limit passes, the kernel
// aborts and the host finishes the affected queries with the plain path;
// results are exact either way.
// Counts only: a capped list needs the plain path's ordered enumeration.
#include <algorithm>
#include <chrono>
#include <numeric>

#include "cand_index.hpp"
#include "filter.hpp"

namespace csm {

csm_status run_join(csm_ctx *ctx, const csm_graph *g, csm_query *q, const csm_match_opts &o, csm_result &res);

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr uint32_t kRowBatch = 128;   // rows staged in LDS at a time (<= kThreads)
constexpr uint32_t kMaxRootWords = 16;
constexpr uint64_t kAbortAfterFails = 1ull << 22;

struct BQuery {
  CandIndex ci;
  const LevelJoin *lv;        // nq entries
  const uint64_t *root_bits;
  const uint64_t *cand_last;  // C(u_last), for the fused correction
  uint64_t cap;
  uint32_t nq;
  uint32_t stop;              // level whose items are counted
  uint32_t fused;
  uint32_t active;            // 0: answered on the host, no tasks
};

struct BState {
  unsigned long long count;
  unsigned int done;
  unsigned int pending;  // tasks of the query pushed and not yet finished
};

// One record per pool block (a block's rows form at most one task) and one
// per root range; `next` links the records of a query's stack.
struct Task {
  uint32_t q, k, block, r0, r1, skip_rows, skip_items, next;
};

struct QStack {
  unsigned int lock;
  unsigned int top;
};

constexpr uint32_t kNil = 0xFFFFFFFFu;
constexpr uint32_t kMaxHeld = 4;  // full output blocks a task keeps before yielding

struct FreeSlot {
  uint32_t id, pad;
  unsigned long long seq;
};

struct BCtl {
  unsigned long long outstanding;
  unsigned long long alloc_fail;
  unsigned long long items;        // (row, list entry) pairs examined
  unsigned long long rows_in;      // rows read
  unsigned long long rows_out;     // rows written
  unsigned long long tasks;
  unsigned int abort;
  unsigned int pad;
};

struct BArgs {
  const uint32_t *off, *adj;
  AdjProbe probe;
  const BQuery *qs;
  BState *st;
  BCtl *ctl;
  Task *recs;                      // num_blocks + root tasks
  QStack *stacks;                  // one per query
  uint32_t nqueries;
  FreeSlot *free_ring;
  unsigned long long *free_head, *free_tail;
  uint32_t *pool;
  uint64_t num_blocks;
  uint64_t block_words;            // u32 words per block
  uint32_t root_words;             // root bitmap words per level-1 task
  uint32_t max_k;                  // widest row (columns) in LDS
  uint64_t deadline_ticks;         // s_memrealtime ticks (100 MHz) before the kernel gives up
};

__device__ inline unsigned long long ld_rlx(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rel(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

constexpr uint32_t kMaxSpin = 1u << 24;

__device__ inline uint32_t ld_rlx32(unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-query lock: held by one thread for a few loads and stores, never
// across a wait, so the spin always ends.
__device__ inline void qlock(unsigned int *l) {
  while (atomicCAS(l, 0u, 1u) != 0u) __builtin_amdgcn_s_sleep(1);
  acquire_agent();
}
__device__ inline void qunlock(unsigned int *l) {
  __hip_atomic_store(l, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Record fields must be written before the push (the unlock publishes them).
__device__ void stack_push(const BArgs &a, uint32_t q, uint32_t rec) {
  QStack *st = &a.stacks[q];
  qlock(&st->lock);
  a.recs[rec].next = ld_rlx32(&st->top);
  __hip_atomic_store(&st->top, rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  qunlock(&st->lock);
}

__device__ bool stack_pop(const BArgs &a, uint32_t q, uint32_t &rec) {
  QStack *st = &a.stacks[q];
  if (ld_rlx32(&st->top) == kNil) return false;
  qlock(&st->lock);
  const uint32_t t = ld_rlx32(&st->top);
  if (t != kNil) __hip_atomic_store(&st->top, a.recs[t].next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  qunlock(&st->lock);
  rec = t;
  return t != kNil;
}

__device__ void free_push(const BArgs &a, uint32_t id) {  // bounded MPMC ring, seq 2t free / 2t+1 full
  const unsigned long long pos = atomicAdd(a.free_tail, 1ull);
  FreeSlot *s = &a.free_ring[pos % a.num_blocks];
  for (uint32_t spin = 0; ld_rlx(&s->seq) != 2 * pos && spin < kMaxSpin; ++spin) __builtin_amdgcn_s_sleep(1);
  acquire_agent();
  s->id = id;
  st_rel(&s->seq, 2 * pos + 1);
}

__device__ bool free_pop(const BArgs &a, uint32_t &id) {
  unsigned long long pos = ld_rlx(a.free_head);
  for (;;) {
    FreeSlot *s = &a.free_ring[pos % a.num_blocks];
    const unsigned long long seq = ld_rlx(&s->seq);
    const long long diff = static_cast<long long>(seq) - static_cast<long long>(2 * pos + 1);
    if (diff == 0) {
      if (atomicCAS(a.free_head, pos, pos + 1) == pos) {
        acquire_agent();
        id = s->id;
        st_rel(&s->seq, 2 * (pos + a.num_blocks));
        return true;
      }
      pos = ld_rlx(a.free_head);
    } else if (diff < 0) {
      return false;
    } else {
      pos = ld_rlx(a.free_head);
    }
  }
}

__device__ inline void push_task(const BArgs &a, uint32_t rec, const Task &t) {
  Task &r = a.recs[rec];
  r.q = t.q;
  r.k = t.k;
  r.block = t.block;
  r.r0 = t.r0;
  r.r1 = t.r1;
  r.skip_rows = t.skip_rows;
  r.skip_items = t.skip_items;
  atomicAdd(&a.ctl->outstanding, 1ull);
  atomicAdd(&a.st[t.q].pending, 1u);
  stack_push(a, t.q, rec);
}

// Initial state: every block in the free ring, each query's root tasks
// linked in ascending order on its stack, counters zero.
__global__ void batch_init_kernel(BArgs a, uint32_t tasks_per_query, uint32_t words) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x, stride = 256ull * gridDim.x;
  const uint32_t nqs = a.nqueries;
  const uint64_t nroot = static_cast<uint64_t>(nqs) * tasks_per_query;
  for (uint64_t i = tid; i < nroot; i += stride) {
    const uint32_t q = static_cast<uint32_t>(i / tasks_per_query), c = static_cast<uint32_t>(i % tasks_per_query);
    Task &t = a.recs[a.num_blocks + i];
    t.q = q;
    t.k = 1;
    t.block = 0;
    t.r0 = c * a.root_words;
    t.r1 = min(words, (c + 1) * a.root_words);
    t.skip_rows = 0;
    t.skip_items = 0;
    t.next = c + 1 < tasks_per_query ? static_cast<uint32_t>(a.num_blocks + i + 1) : kNil;
  }
  for (uint64_t i = tid; i < a.num_blocks; i += stride) {
    a.free_ring[i].id = static_cast<uint32_t>(i);
    a.free_ring[i].seq = 2 * i + 1;
  }
  for (uint64_t i = tid; i < nqs; i += stride) {
    const bool act = a.qs[i].active != 0;
    a.st[i].count = 0;
    a.st[i].done = act ? 0u : 1u;
    a.st[i].pending = act ? tasks_per_query : 0u;
    a.stacks[i].lock = 0;
    a.stacks[i].top = act ? static_cast<uint32_t>(a.num_blocks + i * tasks_per_query) : kNil;
  }
  if (tid == 0) {
    uint64_t live = 0;
    for (uint32_t q = 0; q < nqs; ++q) live += a.qs[q].active ? tasks_per_query : 0;
    *a.free_head = 0;
    *a.free_tail = a.num_blocks;
    a.ctl->outstanding = live;
    a.ctl->alloc_fail = 0;
    a.ctl->items = a.ctl->rows_in = a.ctl->rows_out = a.ctl->tasks = 0;
    a.ctl->abort = 0;
  }
}

struct Smem {
  Task task;
  BQuery q;
  LevelJoin lv;
  LevelJoin lv_last;
  uint32_t roots[kMaxRootWords * 64];
  uint32_t nroots;
  uint64_t pre[kRowBatch + 1];
  uint64_t beg[kRowBatch];
  uint8_t drv[kRowBatch];
  uint32_t wcnt[kWaves];
  unsigned long long red[kWaves];
  uint32_t out_block;
  uint32_t out_fill;
  uint32_t held[kMaxHeld];         // full output blocks of this task, in row order
  uint32_t held_fill[kMaxHeld];
  uint32_t nheld;
  uint32_t rec;                    // record of the task being processed
  uint32_t cur_q;
  int flag;      // 0 go on, 1 yield with continuation
  int state;     // loop control: 0 work, 1 idle, 2 exit
  int stop;      // broadcast of the query's done flag
};

template <typename T>
__device__ inline T block_reduce(T v, T *lds) {
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < kWaves; ++w) t += lds[w];
  __syncthreads();
  return t;
}

__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t *lds, uint64_t &total) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint64_t x = v;
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) lds[wave] = x;
  __syncthreads();
  uint64_t off = 0, all = 0;
  for (int w = 0; w < kWaves; ++w) {
    if (w < wave) off += lds[w];
    all += lds[w];
  }
  __syncthreads();
  total = all;
  return off + x - v;
}

__device__ inline bool bit_of(const uint64_t *bits, uint32_t v) { return (bits[v >> 6] >> (v & 63)) & 1ull; }

// Process one task; returns with all threads converged.
__device__ void process_task(const BArgs &a, Smem &sm, uint32_t *s_rows) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const Task t = sm.task;
  const BQuery &Q = sm.q;
  const uint32_t k = t.k;
  const bool count_level = k == Q.stop;
  const uint32_t kout = k + 1;
  const uint32_t out_cap = static_cast<uint32_t>(a.block_words / kout);
  // rows of the task
  uint32_t nrows;
  const uint32_t *grows = nullptr;
  if (k == 1) {
    if (tid == 0) sm.nroots = 0;
    __syncthreads();
    // each wave takes words; set bits appended in order through a scan
    const uint32_t nw = t.r1 - t.r0;
    uint64_t w = 0;
    uint32_t c = 0;
    if (static_cast<uint32_t>(tid) < nw) {
      w = Q.root_bits[t.r0 + tid];
      c = __popcll(w);
    }
    uint64_t tot;
    const uint64_t off0 = block_excl_scan(c, sm.pre, tot);
    uint32_t p = static_cast<uint32_t>(off0);
    while (w) {
      const int i = __builtin_ctzll(w);
      w &= w - 1;
      sm.roots[p++] = (t.r0 + tid) * 64u + static_cast<uint32_t>(i);
    }
    nrows = static_cast<uint32_t>(tot);
    __syncthreads();
  } else {
    nrows = t.r1 - t.r0;
    grows = a.pool + static_cast<uint64_t>(t.block) * a.block_words + static_cast<uint64_t>(t.r0) * k;
  }
  unsigned long long contrib = 0;
  unsigned long long items_seen = 0, rows_seen = 0, rows_written = 0;
  uint32_t row0 = t.skip_rows;
  uint32_t skip_items = t.skip_items;
  if (tid == 0) {
    sm.flag = 0;
    sm.nheld = 0;
    sm.out_fill = 0;
    sm.out_block = kNil;
  }
  __syncthreads();
  for (; row0 < nrows; row0 += kRowBatch) {
    if (tid == 0)
      sm.stop = __hip_atomic_load(&a.st[t.q].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                __hip_atomic_load(&a.ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    __syncthreads();
    if (sm.stop) break;
    const uint32_t nb = min(kRowBatch, nrows - row0);
    // stage rows and choose each row's driver list
    for (uint32_t i = tid; i < nb * k; i += kThreads)
      s_rows[i] = k == 1 ? sm.roots[row0 + i] : grows[static_cast<uint64_t>(row0) * k + i];
    __syncthreads();
    uint64_t len = 0;
    if (static_cast<uint32_t>(tid) < nb) {
      const uint32_t *row = s_rows + tid * k;
      uint64_t best = ~0ull, bbeg = 0;
      uint32_t bi = 0;
      for (uint32_t x = 0; x < sm.lv.nanchor; ++x) {
        uint64_t b0, e0;
        Q.ci.list(sm.lv, x, row[sm.lv.anchor[x]], b0, e0);
        if (e0 - b0 < best) {
          best = e0 - b0;
          bbeg = b0;
          bi = x;
        }
      }
      if (tid == 0 && skip_items) {  // resume inside the first row
        bbeg += skip_items;
        best -= skip_items;
      }
      sm.beg[tid] = bbeg;
      sm.drv[tid] = static_cast<uint8_t>(bi);
      len = best;
    }
    const uint32_t first_skip = skip_items;  // items of row row0 already done before this batch
    skip_items = 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan(len, reinterpret_cast<uint64_t *>(sm.red), total);
    if (static_cast<uint32_t>(tid) < nb) sm.pre[tid] = ex;
    if (tid == 0) sm.pre[nb] = total;
    __syncthreads();
    rows_seen += nb;
    for (uint64_t base = 0; base < total; base += kThreads) {
      const uint64_t it = base + tid;
      bool ok = false;
      uint32_t w = 0, ri = 0;
      if (it < total) {
        uint32_t lo = 0, hi = nb;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (sm.pre[mid] <= it) lo = mid; else hi = mid;
        }
        ri = lo;
        const uint32_t *row = s_rows + ri * k;
        const uint32_t d = sm.drv[ri];
        w = Q.ci.cn_adj[sm.beg[ri] + (it - sm.pre[ri])];
        ok = true;
        for (uint32_t x = 0; ok && x < sm.lv.ninj; ++x) ok = row[sm.lv.inj[x]] != w;
        for (uint32_t x = 0; ok && x < sm.lv.nanchor; ++x)
          if (x != d) ok = a.probe.has(row[sm.lv.anchor[x]], w);
        if (count_level && ok) {
          if (Q.fused) {
            const uint32_t pk = sm.lv_last.parent;
            const uint32_t xv = pk == k ? w : row[pk];
            uint64_t lb, le;
            Q.ci.list(sm.lv_last, 0, xv, lb, le);
            long long c = static_cast<long long>(le - lb);
            for (uint32_t x = 0; x < sm.lv_last.ninj; ++x) {
              const uint32_t j = sm.lv_last.inj[x];
              const uint32_t v = j == k ? w : row[j];
              if (v != xv && bit_of(Q.cand_last, v) && a.probe.has(xv, v)) --c;
            }
            contrib += static_cast<unsigned long long>(c);
          } else {
            contrib += 1;
          }
        }
      }
      if (count_level) {
        items_seen += it < total ? 1 : 0;
        continue;
      }
      // survivors go to the current output block, in item order
      const uint64_t m = ballot(ok);
      if (lane == 0) sm.wcnt[wave] = static_cast<uint32_t>(__popcll(m));
      __syncthreads();
      uint32_t before = 0, tile = 0;
      for (int x = 0; x < kWaves; ++x) {
        if (x < wave) before += sm.wcnt[x];
        tile += sm.wcnt[x];
      }
      if (tid == 0 && tile && (sm.out_block == kNil || sm.out_fill + tile > out_cap)) {
        if (sm.out_block != kNil) {  // the current block is full: keep it
          sm.held[sm.nheld++] = sm.out_block | 0u;
          sm.held_fill[sm.nheld - 1] = sm.out_fill;
          sm.out_block = kNil;
        }
        uint32_t id;
        if (sm.nheld == kMaxHeld) {
          sm.flag = 1;  // bounded output per task: yield before this tile
        } else if (free_pop(a, id)) {
          sm.out_block = id;
          sm.out_fill = 0;
          atomicExch(&a.ctl->alloc_fail, 0ull);
        } else {
          sm.flag = 2;  // pool empty: yield before this tile, move to another query
          if (atomicAdd(&a.ctl->alloc_fail, 1ull) > kAbortAfterFails) atomicExch(&a.ctl->abort, 1u);
        }
        if (sm.flag) {  // resume point: the row holding item `base`
          uint32_t lo = 0, hi = nb;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sm.pre[mid] <= base) lo = mid; else hi = mid;
          }
          sm.task.skip_rows = row0 + lo;
          sm.task.skip_items = static_cast<uint32_t>(base - sm.pre[lo] + (lo == 0 ? first_skip : 0));
        }
      }
      __syncthreads();
      if (sm.flag) break;
      items_seen += it < total ? 1 : 0;
      if (ok) {
        uint32_t *dst = a.pool + static_cast<uint64_t>(sm.out_block) * a.block_words +
                        static_cast<uint64_t>(sm.out_fill + before + lanes_below(m)) * kout;
        const uint32_t *row = s_rows + ri * k;
        for (uint32_t j = 0; j < k; ++j) dst[j] = row[j];
        dst[k] = w;
      }
      rows_written += ok ? 1 : 0;
      __syncthreads();
      if (tid == 0) sm.out_fill += tile;
      __syncthreads();
    }
    if (sm.flag) break;
    __syncthreads();
  }
  // finish: publish rows, push continuation below the new blocks, count, free input
  __threadfence();
  const unsigned long long tot_contrib = block_reduce(contrib, sm.red);
  const unsigned long long tot_items = block_reduce(items_seen, sm.red);
  const unsigned long long tot_written = block_reduce(rows_written, sm.red);
  if (tid == 0) {
    __threadfence();
    const bool done = __hip_atomic_load(&a.st[t.q].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    const bool cont = sm.flag != 0 && !done;
    if (cont) {
      push_task(a, sm.rec, sm.task);  // same rows, resume point updated
    } else if (k >= 2) {
      free_push(a, t.block);
    }
    // lexicographically first rows on top: newest (last) block first
    if (sm.out_block != kNil) {
      if (sm.out_fill) push_task(a, sm.out_block, Task{t.q, kout, sm.out_block, 0, sm.out_fill, 0, 0, kNil});
      else free_push(a, sm.out_block);
    }
    for (int i = static_cast<int>(sm.nheld) - 1; i >= 0; --i)
      push_task(a, sm.held[i], Task{t.q, kout, sm.held[i], 0, sm.held_fill[i], 0, 0, kNil});
    if (count_level && tot_contrib) {
      const unsigned long long c = atomicAdd(&a.st[t.q].count, tot_contrib) + tot_contrib;
      if (c > Q.cap) atomicExch(&a.st[t.q].done, 1u);
    }
    atomicAdd(&a.ctl->items, tot_items);
    atomicAdd(&a.ctl->rows_in, rows_seen);
    atomicAdd(&a.ctl->rows_out, tot_written);
    atomicAdd(&a.ctl->tasks, 1ull);
    if (sm.flag == 2) sm.cur_q = (sm.cur_q + 1) % a.nqueries;  // let this query's pool users finish
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void batch_join_kernel(BArgs a) {
  extern __shared__ uint32_t s_rows[];  // kRowBatch x max_k
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
  if (tid == 0) sm.cur_q = blockIdx.x % a.nqueries;
  for (;;) {
    if (tid == 0) {
      if (__builtin_amdgcn_s_memrealtime() - t_start > a.deadline_ticks) atomicExch(&a.ctl->abort, 1u);
      sm.state = 1;
      for (uint32_t i = 0; i < a.nqueries; ++i) {
        const uint32_t q = (sm.cur_q + i) % a.nqueries;
        uint32_t rec;
        if (stack_pop(a, q, rec)) {
          const Task &r = a.recs[rec];
          sm.task.q = r.q;
          sm.task.k = r.k;
          sm.task.block = r.block;
          sm.task.r0 = r.r0;
          sm.task.r1 = r.r1;
          sm.task.skip_rows = r.skip_rows;
          sm.task.skip_items = r.skip_items;
          sm.rec = rec;
          sm.cur_q = q;
          sm.state = 0;
          break;
        }
      }
      if (sm.state == 1 &&
          (ld_rlx(&a.ctl->outstanding) == 0 || __hip_atomic_load(&a.ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        sm.state = 2;
      if (sm.state == 0)
        sm.stop = __hip_atomic_load(&a.st[sm.task.q].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                  __hip_atomic_load(&a.ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    }
    __syncthreads();
    const int state = sm.state;
    if (state == 2) break;
    if (state == 1) {
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    __threadfence();  // acquire the task's rows written by another workgroup
    const uint32_t q = sm.task.q;
    {  // stage the query header and the level descriptors
      const uint32_t *src = reinterpret_cast<const uint32_t *>(&a.qs[q]);
      uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.q);
      for (uint32_t i = tid; i < sizeof(BQuery) / 4; i += kThreads) dst[i] = src[i];
    }
    __syncthreads();
    {
      const uint32_t k = sm.task.k;
      const uint32_t *src = reinterpret_cast<const uint32_t *>(&sm.q.lv[k]);
      uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.lv);
      for (uint32_t i = tid; i < sizeof(LevelJoin) / 4; i += kThreads) dst[i] = src[i];
      if (sm.q.fused) {
        const uint32_t *s2 = reinterpret_cast<const uint32_t *>(&sm.q.lv[sm.q.nq - 1]);
        uint32_t *d2 = reinterpret_cast<uint32_t *>(&sm.lv_last);
        for (uint32_t i = tid; i < sizeof(LevelJoin) / 4; i += kThreads) d2[i] = s2[i];
      }
    }
    __syncthreads();
    if (!sm.stop) {
      process_task(a, sm, s_rows);
    } else if (tid == 0 && sm.task.k >= 2) {
      free_push(a, sm.task.block);
    }
    __syncthreads();
    if (tid == 0) {  // this task is finished
      atomicSub(&a.st[q].pending, 1u);
      atomicAdd(&a.ctl->outstanding, ~0ull);
    }
  }
}

uint32_t grid_for(uint64_t n) { return static_cast<uint32_t>(std::min<uint64_t>(65536, (n + 255) / 256 + 1)); }

struct TuneParams {
  uint32_t wgs_per_cu, root_words, block_kb;
};

}  // namespace

// Runs the batch; results[i] for queries[i].  Queries must be prepared.
csm_status run_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<csm_query *> &queries,
                     const csm_batch_opts &o, csm_result *results) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const uint32_t nqs = static_cast<uint32_t>(queries.size());
  if (nqs == 0) return CSM_OK;
  const uint32_t words = queries[0]->desc.words;
  const uint64_t cap = o.cap ? o.cap : ~0ull;
  std::vector<void *> blocks;
  struct Guard {
    csm_ctx *ctx;
    std::vector<void *> &b;
    ~Guard() {
      hipStreamSynchronize(ctx->stream);
      for (void *p : b) ctx->arena.free(p);
    }
  } guard{ctx, blocks};
  auto get = [&](uint64_t bytes) {
    void *p = ctx->arena.alloc(bytes);
    if (p) blocks.push_back(p);
    return p;
  };
  csm_kernel_stat kstat[CSM_K_NUM] = {};
  KernelTimer kt(ctx, kstat);
  uint32_t max_nq = 0;
  for (auto *q : queries) max_nq = std::max(max_nq, q->desc.nq);

  // 1. filter + refinement + counts for every query (one launch per step)
  uint64_t *stats = static_cast<uint64_t *>(get(8ull * kStatSize * nqs));
  FilterQ *d_fq = static_cast<FilterQ *>(get(sizeof(FilterQ) * nqs));
  if (!stats || !d_fq) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch filter)");
  {
    std::vector<FilterQ> fq(nqs);
    std::vector<const csm_query *> cq(queries.begin(), queries.end());
    for (uint32_t i = 0; i < nqs; ++i) fq[i] = filter_entry(queries[i], stats + static_cast<uint64_t>(i) * kStatSize);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(d_fq, fq.data(), sizeof(FilterQ) * nqs, hipMemcpyHostToDevice, s));
    const std::vector<uint32_t> steps = o.refine ? step_members_of(cq) : std::vector<uint32_t>();
    if (o.use_graph) {
      hipGraph_t graph;
      hipGraphExec_t exec;
      CSM_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch_filter_batch(s, g, d_fq, nqs, words, steps, max_nq, o.refine != 0, nullptr);
      CSM_HIP_TRY(ctx, hipStreamEndCapture(s, &graph));
      CSM_HIP_TRY(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      kt.begin(CSM_K_FILTER);
      CSM_HIP_TRY(ctx, hipGraphLaunch(exec, s));
      kt.end(CSM_K_FILTER, 0);
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      hipGraphExecDestroy(exec);
      hipGraphDestroy(graph);
    } else {
      launch_filter_batch(s, g, d_fq, nqs, words, steps, max_nq, o.refine != 0, &kt);
    }
    std::vector<uint64_t> host(static_cast<size_t>(kStatSize) * nqs);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(host.data(), stats, 8ull * host.size(), hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    for (uint32_t i = 0; i < nqs; ++i) {
      for (uint32_t u = 0; u < queries[i]->desc.nq; ++u) queries[i]->counts[u] = host[i * kStatSize + u];
      queries[i]->filtered = true;
    }
  }
  const double ms_filter = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

  // 2. candidate indexes of every query
  std::vector<CandIndexHost> cih;
  {
    std::vector<const csm_query *> cq(queries.begin(), queries.end());
    csm_status st = build_cand_index_batch(ctx, g, cq, cih, blocks, kt);
    if (st != CSM_OK) return st;
  }

  // 3. per-query descriptors; queries answered without the join
  std::vector<BQuery> bq(nqs);
  std::vector<uint64_t> host_count(nqs, 0);
  LevelJoin *d_lv = static_cast<LevelJoin *>(get(sizeof(LevelJoin) * CSM_MAX_QV * nqs));
  BQuery *d_bq = static_cast<BQuery *>(get(sizeof(BQuery) * nqs));
  BState *d_st = static_cast<BState *>(get(sizeof(BState) * nqs));
  if (!d_lv || !d_bq || !d_st) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch descriptors)");
  std::vector<LevelJoin> lvs(static_cast<size_t>(CSM_MAX_QV) * nqs);
  std::vector<std::pair<uint64_t, uint64_t>> two_vertex;  // (query, cn_off index pair)
  for (uint32_t i = 0; i < nqs; ++i) {
    const csm_query *q = queries[i];
    const uint32_t nq = q->desc.nq, L = nq - 1;
    for (uint32_t k = 0; k < nq; ++k) lvs[static_cast<size_t>(i) * CSM_MAX_QV + k] = cih[i].level[k];
    BQuery &b = bq[i];
    std::memset(&b, 0, sizeof(b));
    b.ci = CandIndex{q->bits, cih[i].rank_pref, cih[i].cn_off, cih[i].cn_adj, words};
    b.lv = d_lv + static_cast<size_t>(i) * CSM_MAX_QV;
    b.root_bits = q->bits + static_cast<uint64_t>(q->level[0].u) * words;
    b.cand_last = q->bits + static_cast<uint64_t>(q->level[L].u) * words;
    b.cap = cap;
    b.nq = nq;
    b.fused = q->level[L].nback == 0 && L >= 2;
    b.stop = b.fused ? L - 1 : L;
    b.active = q->counts[q->level[0].u] > 0;
    if (L == 1 && b.active) {  // two vertices: total length of the one slot
      b.active = 0;
      two_vertex.push_back({i, cih[i].level[1].off_base[0]});
    }
  }
  for (auto &tv : two_vertex) {
    const uint64_t n1 = queries[tv.first]->counts[queries[tv.first]->level[0].u];
    CSM_HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, cih[tv.first].cn_off + tv.second, 8, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned + 1, cih[tv.first].cn_off + tv.second + n1, 8, hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    host_count[tv.first] = ctx->pinned[1] - ctx->pinned[0];
  }
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_lv, lvs.data(), sizeof(LevelJoin) * lvs.size(), hipMemcpyHostToDevice, s));
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_bq, bq.data(), sizeof(BQuery) * nqs, hipMemcpyHostToDevice, s));

  // 4. queues and pool
  TuneParams tp{o.wgs_per_cu ? o.wgs_per_cu : 4u, 16u, 128u};
  if (o.chunk_rows) tp.root_words = std::min<uint32_t>(kMaxRootWords, std::max<uint32_t>(1, o.chunk_rows / 64));
  const uint64_t block_words = static_cast<uint64_t>(tp.block_kb) * 1024 / 4;
  if (block_words < static_cast<uint64_t>(kThreads) * CSM_MAX_QV) return set_error(ctx, CSM_E_ARG, "block too small");
  uint64_t pool_bytes = o.pool_bytes;
  const uint64_t avail = ctx->arena.largest_free();
  if (!pool_bytes) pool_bytes = std::min<uint64_t>(avail / 2, 32ull << 30);
  pool_bytes = std::min<uint64_t>(pool_bytes, avail * 3 / 4);
  const uint64_t num_blocks = std::max<uint64_t>(64, pool_bytes / (4 * block_words));
  const uint32_t tasks_per_query = ceil_div(words, tp.root_words);
  const uint64_t nroot = static_cast<uint64_t>(nqs) * tasks_per_query;
  if (num_blocks + nroot >= kNil) return set_error(ctx, CSM_E_LIMIT, "too many pool blocks and root tasks");
  uint32_t *pool = static_cast<uint32_t *>(get(4 * block_words * num_blocks));
  Task *recs = static_cast<Task *>(get(sizeof(Task) * (num_blocks + nroot)));
  QStack *stacks = static_cast<QStack *>(get(sizeof(QStack) * nqs));
  FreeSlot *free_ring = static_cast<FreeSlot *>(get(sizeof(FreeSlot) * num_blocks));
  unsigned long long *ctr = static_cast<unsigned long long *>(get(8ull * 2));
  BCtl *d_ctl = static_cast<BCtl *>(get(sizeof(BCtl)));
  if (!pool || !recs || !stacks || !free_ring || !ctr || !d_ctl)
    return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch pool)");
  BArgs a{};
  a.off = g->off;
  a.adj = g->adj;
  a.probe = AdjProbe{g->off, g->adj, g->hub_id, g->hub_bits, words};
  a.qs = d_bq;
  a.st = d_st;
  a.ctl = d_ctl;
  a.recs = recs;
  a.stacks = stacks;
  a.nqueries = nqs;
  a.free_head = ctr;
  a.free_tail = ctr + 1;
  a.free_ring = free_ring;
  a.pool = pool;
  a.num_blocks = num_blocks;
  a.block_words = block_words;
  a.root_words = tp.root_words;
  a.max_k = max_nq;
  a.deadline_ticks = 100ull * 1000 * 1000 * 120;  // 120 s at the 100 MHz real-time clock
  const size_t lds_rows = 4ull * kRowBatch * std::max<uint32_t>(1, max_nq - 1);
  const uint32_t grid = ctx->num_cus * tp.wgs_per_cu;
  auto launch_join = [&]() {
    batch_init_kernel<<<grid_for(std::max<uint64_t>(num_blocks, nroot)), 256, 0, s>>>(a, tasks_per_query, words);
    batch_join_kernel<<<grid, kThreads, lds_rows, s>>>(a);
  };
  if (o.use_graph) {
    hipGraph_t graph;
    hipGraphExec_t exec;
    CSM_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_join();
    CSM_HIP_TRY(ctx, hipStreamEndCapture(s, &graph));
    CSM_HIP_TRY(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    kt.begin(CSM_K_BATCH);
    CSM_HIP_TRY(ctx, hipGraphLaunch(exec, s));
    kt.end(CSM_K_BATCH, 0);
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
  } else {
    kt.begin(CSM_K_BATCH);
    launch_join();
    kt.end(CSM_K_BATCH, 0);
  }
  std::vector<BState> hst(nqs);
  BCtl hctl{};
  CSM_HIP_TRY(ctx, hipMemcpyAsync(hst.data(), d_st, sizeof(BState) * nqs, hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipMemcpyAsync(&hctl, d_ctl, sizeof(BCtl), hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  kt.flush();
  // algorithmic bytes of the persistent join: rows read, per-row anchor list
  // lookups, list entries, rows written
  const uint64_t join_bytes = hctl.rows_in * (4ull * 8 + 32) + hctl.items * 4 + hctl.rows_out * 4ull * 9;
  kstat[CSM_K_BATCH].bytes = join_bytes;
  const double ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (uint32_t i = 0; i < nqs; ++i) {
    csm_result &r = results[i];
    std::memset(&r, 0, sizeof(r));
    const uint64_t total = bq[i].active ? hst[i].count : host_count[i];
    r.count = std::min(total, cap);
    r.capped = total > cap;
    r.status = CSM_OK;
    r.ms_filter = ms_filter;
    r.ms_total = ms_total;
    r.ms_join = ms_total - ms_filter;
    if (i == 0) {
      for (int c = 0; c < CSM_K_NUM; ++c) r.kernels[c] = kstat[c];
      r.items = hctl.items;
      r.partial_rows = hctl.rows_out;
    }
  }
  if (hctl.abort) {  // pool exhausted without progress: finish unfinished queries one by one
    for (uint32_t i = 0; i < nqs; ++i) {
      if (!bq[i].active || hst[i].done || hst[i].pending == 0) continue;
      csm_match_opts mo{};
      mo.cap = o.cap;
      mo.refine = o.refine;
      mo.reuse_filter = 1;
      csm_result r{};
      csm_status st = run_join(ctx, g, queries[i], mo, r);
      if (st != CSM_OK) return st;
      results[i].count = r.count;
      results[i].capped = r.capped;
    }
  }
  return CSM_OK;
}

}  // namespace csm

using namespace csm;

extern "C" csm_status csm_batch_match(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                                      const csm_batch_opts *opts, csm_result *out) {
  if (!ctx || !g || (n && (!qs || !out))) return CSM_E_ARG;
  csm_batch_opts o{};
  o.refine = 1;
  o.use_graph = 1;
  if (opts) o = *opts;
  std::vector<csm_query *> v(qs, qs + n);
  for (auto *q : v)
    if (!q || q->graph_n != g->n) return CSM_E_ARG;
  hipSetDevice(ctx->device);
  return run_batch(ctx, g, v, o, out);
}
