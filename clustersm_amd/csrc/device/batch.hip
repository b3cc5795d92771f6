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
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <numeric>

#include "batch.hpp"
#include "filter.hpp"

namespace csm {


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
  unsigned int done;         // kRunning, kCapped or kAbandoned
  unsigned int lost;         // a task of the query was dropped or cut short by an early stop
  unsigned long long pending;  // tasks of the query queued or running
  unsigned long long items;  // (row, list entry) pairs examined for the query
  unsigned long long rows;   // partial rows written for the query
  unsigned int gen;          // resident mode: admission generation of the slot
  unsigned int pad;
  unsigned long long adm;    // resident mode: admission index of the slot's query
};

enum : unsigned int { kRunning = 0, kCapped = 1, kAbandoned = 2 };
enum : unsigned int { kAbortPool = 1, kAbortTime = 2, kAbortBadTask = 3 };

// A task record: one per pool block (a block's rows form at most one task)
// and one per root range.
// `seq` tags each write of a record; the deque entry that names the record
// carries the same tag, so a reader can tell the record it expects from an
// older one at the same index.
struct alignas(16) Task {
  uint32_t q, k, block, r0, r1, skip_rows, skip_items, seq;
};

constexpr uint32_t kNil = 0xFFFFFFFFu;

// Resident (serving) mode.  The host publishes prepared queries through a
// block of pinned, coherent host memory that the running kernel polls:
// admission a names a slot and the slot's generation; `doorbell` counts the
// admissions written.  The kernel initialises each admitted slot's state
// itself (device atomics) before any of its tasks run, and writes a
// completion record per slot, generation last, when the query's last task
// ends.  Admissions in flight never exceed the slots, so kAdmRing > slots.
constexpr uint32_t kAdmRing = 1024;
constexpr uint32_t kMaxSlots = 256;
struct SlotAdmit {
  uint32_t slot, gen;
};
struct SlotDone {
  unsigned long long count, items, rows;
  unsigned int lost;
  unsigned int gen;  // written last, with a release: the record is complete
};
struct ServeShared {
  unsigned long long doorbell;  // admissions published by the host
  unsigned long long stop;      // 1: the kernel exits once idle
  SlotAdmit adm[kAdmRing];
  SlotDone done[kMaxSlots];
};
constexpr unsigned long long kNilEntry = ~0ull;  // deque entries: record id | seq << 32
constexpr uint32_t kDequeCap = 1024;  // > levels x (held blocks + 2): a deque never fills

// One task deque per workgroup, split in two: entries [top, pub) are
// published and may be stolen (oldest first, at `top`); entries [pub,
// bottom) are private to the owner, which pushes and pops them at `bottom`
// (depth-first, lexicographically first rows on top) with no atomics and
// no fences.  The owner publishes, with one release, only while some
// workgroup is idle.  A short lock orders thieves with the owner's pops
// from the published part.  `bottom` lives in the owner's LDS.
struct Deque {
  unsigned int lock;
  unsigned int pad;
  unsigned long long top;
  unsigned long long pub;
  unsigned long long pad2;
};

struct FreeSlot {
  uint32_t id, pad;
  unsigned long long seq;
};

struct BCtl {
  unsigned long long root_next;    // next root task to hand out
  unsigned long long idle;         // workgroups without a task
  unsigned long long started;      // workgroups that have begun (all of them unless the GPU is shared)
  unsigned long long items;        // (row, list entry) pairs examined
  unsigned long long rows_in;      // rows read
  unsigned long long rows_out;     // rows written
  unsigned long long bytes_own;    // this kernel's own logical bytes
  unsigned long long tasks;
  unsigned long long steals;
  unsigned long long fails;        // failed block allocations
  unsigned long long busy_ticks;   // real-time ticks inside tasks, summed over workgroups
  unsigned long long idle_ticks;   // real-time ticks without a task
  unsigned long long sync_ticks;   // real-time ticks in deque and free-ring operations
  unsigned long long phase_ticks[7];  // rows staged + drivers, count tiles, extension tiles, task ends;
                                      // staging's parts: rows, list lookups
  unsigned long long alloc_fail;   // consecutive failed allocations from the free ring
  unsigned int abort;              // kAbortPool, kAbortTime or kAbortBadTask once the kernel stops early
  unsigned int overflow;           // a deque was full (never expected)
  unsigned long long late_recs;    // records whose tag matched only after a re-read (never expected)
  unsigned long long bad_recs;     // records never matching their entry, or out of range (never expected)
  unsigned int bad_why;            // first bad record: 1 earlier epoch, 2 id range, 3 tag never matched, 4 bounds
  unsigned int bad_wg;             // and the workgroup that read it
  unsigned long long bad_ent;      // its entry
  Task bad_task;                   // and the record as read
  unsigned int bad_src, bad_victim;  // where the entry came from (0 own private, 1 own locked, 2 stolen)
  unsigned long long bad_t, bad_b;   // slot index and bound it was taken at
  unsigned int late_src, late_victim;  // the same for the first late record
  unsigned long long late_t, late_b;
  // resident mode: admissions whose slots are initialised, the host doorbell
  // and stop flag as last read (polled by one workgroup at a time, at most
  // every few microseconds), and the slot of each admission
  unsigned long long adm_seen;
  unsigned long long db_mirror;
  unsigned long long stop_mirror;
  unsigned long long last_poll;
  unsigned int adm_lock;
  unsigned int exiting;            // resident mode: set once, under adm_lock; no admission follows
  unsigned int adm_slot[kAdmRing];
};

struct BArgs {
  AdjProbe probe;
  const BQuery *qs;
  BState *st;
  BCtl *ctl;
  Task *recs;                      // num_blocks + root tasks
  Deque *dq;                       // one per workgroup
  unsigned long long *dq_buf;      // kDequeCap entries per workgroup
  uint32_t epoch;                  // differs between launches: part of every record tag
  uint32_t nqueries;
  uint32_t tasks_per_query;        // root tasks per query
  uint32_t words;                  // bitmap words
  uint32_t root_words;             // root bitmap words per root task
  FreeSlot *free_ring;
  unsigned long long *free_head, *free_tail;
  uint32_t *pool;
  uint64_t num_blocks;
  uint64_t block_words;            // u32 words per block
  uint32_t ilp;                    // items per thread per tile (1..kIlp)
  uint32_t max_out;                // full output blocks a task holds before it yields
  uint64_t budget;                 // items per query before it stops with a lower bound (the caller's budget)
  uint32_t cache_cap;              // free blocks a workgroup may keep (<= kCache)
  uint64_t deadline_ticks;         // s_memrealtime ticks (100 MHz) before the kernel gives up
  uint32_t serve;                  // resident mode: queries arrive through `sh` while the kernel runs
  uint32_t pad;
  ServeShared *sh;                 // pinned host block (resident mode)
  uint64_t idle_exit_ticks;        // resident mode: exit after this long with nothing to do anywhere
};

__device__ inline unsigned long long ld_rlx(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rlx(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rel(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ld_done(const BArgs &a, uint32_t q) {
  return __hip_atomic_load(&a.st[q].done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool aborted(const BArgs &a) {
  return __hip_atomic_load(&a.ctl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ inline void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

constexpr uint32_t kMaxSpin = 1u << 24;
// An idle workgroup's scheduling round reads the deques and counters of the
// others with device-scope atomics; between rounds it sleeps
// min(rounds idle, kMaxIdleRounds) x 16 x 64 cycles (up to about 60 us).
// With a cap of 16 (about 8 us) the polling of several hundred idle
// workgroups slowed the few busy ones' memory accesses by an order of
// magnitude: one deep, narrow query ran 6x slower in the resident kernel.
constexpr uint32_t kMaxIdleRounds = 128;

// Deque lock: held by one thread for a few loads and stores, never across a
// wait, and contended only by the owner and an occasional thief.
__device__ inline void dlock(unsigned int *l) {
  while (atomicCAS(l, 0u, 1u) != 0u) __builtin_amdgcn_s_sleep(2);
  acquire_agent();
}
__device__ inline void dunlock(unsigned int *l) {
  __hip_atomic_store(l, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Thief: oldest entry of workgroup v, or kNilEntry.  A thief counted idle
// leaves the idle count while it holds v's lock: v can only declare itself
// idle after the unlock, so the count never reaches every workgroup while a
// task is in hand.
__device__ unsigned long long deque_steal(const BArgs &a, uint32_t v, bool counted_idle, unsigned long long &at,
                                          unsigned long long &bound) {
  Deque *d = &a.dq[v];
  if (ld_rlx(&d->pub) <= ld_rlx(&d->top)) return kNilEntry;
  unsigned long long *buf = a.dq_buf + static_cast<uint64_t>(v) * kDequeCap;
  dlock(&d->lock);
  // acquire: the entries below pub, their records and rows were written
  // before the owner's release of pub
  const unsigned long long b = __hip_atomic_load(&d->pub, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long t = ld_rlx(&d->top);
  unsigned long long r = kNilEntry;
  if (b > t) {
    r = __hip_atomic_load(&buf[t % kDequeCap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    at = t;
    bound = b;
    st_rlx(&d->top, t + 1);
    if (counted_idle) atomicAdd(&a.ctl->idle, ~0ull);
  }
  dunlock(&d->lock);
  return r;
}

__device__ void free_push(const BArgs &a, uint32_t id) {  // bounded MPMC ring, seq 2t free / 2t+1 full
  const unsigned long long pos = atomicAdd(a.free_tail, 1ull);
  FreeSlot *s = &a.free_ring[pos % a.num_blocks];
  for (uint32_t spin = 0; ld_rlx(&s->seq) != 2 * pos && spin < kMaxSpin; ++spin) __builtin_amdgcn_s_sleep(1);
  acquire_agent();
  __hip_atomic_store(&s->id, id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        id = __hip_atomic_load(&s->id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Initial state: every block in the free ring, deques empty, counters zero.
// Every field the join kernel synchronises on is stored with an agent-scope
// atomic, so the join kernel's atomics see this launch's state and never a
// value an earlier launch left at the same address (the arena hands the same
// blocks to consecutive calls).
template <typename T>
__device__ inline void st_init(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__global__ void batch_init_kernel(BArgs a, uint32_t nwgs) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x, stride = 256ull * gridDim.x;
  for (uint64_t i = tid; i < a.num_blocks; i += stride) {
    st_init(&a.free_ring[i].id, static_cast<uint32_t>(i));
    st_init(&a.free_ring[i].seq, static_cast<unsigned long long>(2 * i + 1));
  }
  for (uint64_t i = tid; i < nwgs; i += stride) {
    st_init(&a.dq[i].lock, 0u);
    st_init(&a.dq[i].top, 0ull);
    st_init(&a.dq[i].pub, 0ull);
  }
  // resident mode: slot states are set when each query is admitted
  for (uint64_t i = tid; i < (a.serve ? 0u : a.nqueries); i += stride) {
    st_init(&a.st[i].count, 0ull);
    st_init(&a.st[i].done, a.qs[i].active ? static_cast<unsigned int>(kRunning) : static_cast<unsigned int>(kCapped));
    st_init(&a.st[i].lost, 0u);
    st_init(&a.st[i].pending, a.qs[i].active ? static_cast<unsigned long long>(a.tasks_per_query) : 0ull);
    st_init(&a.st[i].items, 0ull);
    st_init(&a.st[i].rows, 0ull);
  }
  if (tid == 0) {
    st_init(a.free_head, 0ull);
    st_init(a.free_tail, static_cast<unsigned long long>(a.num_blocks));
    BCtl &c = *a.ctl;
    if (!a.serve) {  // resident mode: root and admission progress carry over to a relaunch
      st_init(&c.root_next, 0ull);
      st_init(&c.adm_seen, 0ull);
      st_init(&c.db_mirror, 0ull);
    }
    st_init(&c.stop_mirror, 0ull);
    st_init(&c.last_poll, 0ull);
    st_init(&c.adm_lock, 0u);
    st_init(&c.exiting, 0u);
    unsigned long long *c64[] = {&c.idle, &c.started, &c.items, &c.rows_in, &c.rows_out,
                                 &c.bytes_own, &c.tasks,
                                 &c.steals, &c.fails, &c.busy_ticks, &c.idle_ticks, &c.sync_ticks, &c.phase_ticks[0],
                                 &c.phase_ticks[1], &c.phase_ticks[2], &c.phase_ticks[3], &c.phase_ticks[4],
                                 &c.phase_ticks[5], &c.phase_ticks[6], &c.alloc_fail,
                                 &c.late_recs, &c.bad_recs};
    for (unsigned long long *f : c64) st_init(f, 0ull);
    st_init(&c.abort, 0u);
    st_init(&c.overflow, 0u);
  }
  __threadfence();
}

constexpr uint32_t kIlp = 4;     // items per thread per tile, checked in lockstep
constexpr uint64_t kHubTag = 1ull << 63;  // abeg entry of a hub anchor: tag | hub index
constexpr uint32_t kRowLoads = 8;         // row words each thread has in flight while staging rows
constexpr uint32_t kMaxA = 4;    // anchors whose candidate lists are kept per row
constexpr uint32_t kMaxOut = 8;  // most full output blocks a task holds before it yields
constexpr uint32_t kCache = 16;  // free blocks a workgroup keeps for itself
constexpr uint32_t kDefaultBlocksPerWg = 16;  // pool blocks per workgroup the default block size keeps

static_assert(kRowBatch <= 256 && kRowBatch <= kThreads, "row index in a batch is a byte");
// dynamic LDS words at least: a row batch of one column, then a root task's ids
constexpr uint32_t kRootLdsWords = kRowBatch + kMaxRootWords * 64;

struct Smem {
  Task task;
  BQuery q;
  LevelJoin lv;
  LevelJoin lv_last;
  uint64_t pre[kRowBatch + 1];
  // candidate list of each kept anchor, anchor-major so that lanes reading
  // one anchor of different rows hit consecutive banks: begin
  uint64_t abeg[kMaxA][kRowBatch];
  uint32_t alen[kMaxA][kRowBatch];  // and length
  int32_t ahub[kMaxA][kRowBatch];   // and the anchor vertex's hub index (-1: none)
  uint64_t dbeg[kRowBatch];         // driver list start (after a resume skip)
  // fused count, when the last position's parent is an earlier column: per
  // row, |N(x) ∩ C(u_last)| minus the row's own members of that list, and the list
  uint32_t fbase[kRowBatch];
  uint32_t flen[kRowBatch];
  uint64_t fbeg[kRowBatch];
  uint8_t drv[kRowBatch];
  uint8_t out_ri[kThreads * kIlp];  // a tile's survivors in output order: row in the batch
  uint32_t out_w[kThreads * kIlp];  // and the new id
  unsigned long long red[2 * kWaves];
  uint32_t tile_n[3];               // rotating survivor counters, one per tile in flight
  uint32_t out_block;
  uint32_t rec;                     // record of the task being processed
  uint32_t staged_q, staged_k;      // query header and level descriptor held in LDS
  uint32_t staged_gen, task_gen;    // resident mode: slot generation staged, and of the task in hand
  unsigned long long task_adm;      // resident mode: admission index of the task's query
  unsigned long long acq_seen;      // resident mode: admissions published before this workgroup's last invalidation
  unsigned long long idle_since;    // resident mode: when this workgroup last found nothing to do
  uint32_t nheld;                   // full output blocks of the current task, in row order
  uint32_t held[kMaxOut];
  uint32_t held_fill[kMaxOut];
  unsigned long long push[kMaxOut + 2];  // entries pushed when the task ends
  uint32_t seq_ctr;                 // record writes by this workgroup
  uint32_t ent_src;                 // where the last entry came from: 0 own private, 1 own locked, 2 stolen
  uint32_t task_stolen;             // the task in hand came from another workgroup's deque
  uint32_t ent_victim;
  unsigned long long ent_t, ent_b;  // the indices it was taken at (slot, bound)
  uint32_t cache[kCache];           // free blocks owned by this workgroup
  uint32_t ncache;
  unsigned long long dq_bottom;     // owner's end of its deque
  unsigned long long dq_pub;        // published bound (only the owner changes it)
  // counts and statistics not yet added to the query's totals
  uint32_t acc_q;
  unsigned long long acc_count, acc_items, acc_rows;
  unsigned long long sync_ticks;
  unsigned long long phase[7];  // staging, count tiles, extension tiles, task ends; staging's rows, lists
  int flag;                         // 0 go on, 1 output bound reached, 2 pool empty: yield with continuation
  int state;                        // loop control: 0 work, 1 idle, 2 exit
  int stop;                         // broadcast: the task's query is finished or the kernel aborts
};

// Thread 0.  Blocks come from the workgroup's cache first.
__device__ bool blk_alloc(const BArgs &a, Smem &sm, uint32_t &id) {
  if (sm.ncache) {
    id = sm.cache[--sm.ncache];
    return true;
  }
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  const bool ok = free_pop(a, id);
  if (ok) {
    if (ld_rlx(&a.ctl->alloc_fail)) atomicExch(&a.ctl->alloc_fail, 0ull);
  } else if (atomicAdd(&a.ctl->alloc_fail, 1ull) > kAbortAfterFails) {
    atomicCAS(&a.ctl->abort, 0u, kAbortPool);  // no block freed for a long time: the host finishes the batch
  }
  sm.sync_ticks += __builtin_amdgcn_s_memrealtime() - t0;
  return ok;
}
__device__ void blk_free(const BArgs &a, Smem &sm, uint32_t id) {
  if (sm.ncache < a.cache_cap) {
    sm.cache[sm.ncache++] = id;
    return;
  }
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  free_push(a, id);
  sm.sync_ticks += __builtin_amdgcn_s_memrealtime() - t0;
}

// Owner (thread 0): entries [pub, upto) become stealable.  Everything this
// workgroup wrote before (entries, records, rows, after the barrier that
// ended their task) is released with them.
__device__ void deque_publish(const BArgs &a, Smem &sm, uint32_t wg, unsigned long long upto) {
  // every store of this wave complete before the release begins its L2
  // write-back (the entries and records are agent-scope stores as well)
  __builtin_amdgcn_s_waitcnt(0);
  __hip_atomic_store(&a.dq[wg].pub, upto, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  sm.dq_pub = upto;
}

// Owner: push n entries in order (the last one ends on top), privately;
// published at once when a workgroup is idle.
__device__ void deque_push(const BArgs &a, Smem &sm, uint32_t wg, const unsigned long long *recs, uint32_t n) {
  if (n == 0) return;
  unsigned long long *buf = a.dq_buf + static_cast<uint64_t>(wg) * kDequeCap;
  const unsigned long long b = sm.dq_bottom, t = ld_rlx(&a.dq[wg].top);
  if (b - t + n > kDequeCap) {
    atomicExch(&a.ctl->overflow, 1u);
    atomicCAS(&a.ctl->abort, 0u, kAbortPool);
    return;
  }
  for (uint32_t i = 0; i < n; ++i)
    __hip_atomic_store(&buf[(b + i) % kDequeCap], recs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  sm.dq_bottom = b + n;
  if (ld_rlx(&a.ctl->idle) > 0) deque_publish(a, sm, wg, sm.dq_bottom);
}

// Owner: newest entry, or kNilEntry.  Private entries need no
// synchronisation; while others are idle, all but the one taken are
// published first.
__device__ unsigned long long deque_pop(const BArgs &a, Smem &sm, uint32_t wg) {
  unsigned long long *buf = a.dq_buf + static_cast<uint64_t>(wg) * kDequeCap;
  if (sm.dq_bottom > sm.dq_pub) {
    if (sm.dq_bottom - sm.dq_pub > 1 && ld_rlx(&a.ctl->idle) > 0) deque_publish(a, sm, wg, sm.dq_bottom - 1);
    sm.ent_src = 0;
    sm.ent_t = sm.dq_bottom - 1;
    sm.ent_b = sm.dq_pub;
    return buf[(--sm.dq_bottom) % kDequeCap];
  }
  Deque *d = &a.dq[wg];
  if (sm.dq_pub <= ld_rlx(&d->top)) return kNilEntry;
  dlock(&d->lock);
  const unsigned long long t = ld_rlx(&d->top), p = sm.dq_pub;
  unsigned long long r = kNilEntry;
  if (p > t) {
    r = __hip_atomic_load(&buf[(p - 1) % kDequeCap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sm.ent_src = 1;
    sm.ent_t = p - 1;
    sm.ent_b = t;
    st_rlx(&d->pub, p - 1);
    sm.dq_pub = sm.dq_bottom = p - 1;
  }
  dunlock(&d->lock);
  return r;
}

// Thread 0: adds the workgroup's pending count and statistics to query
// acc_q and marks the query capped or over its budget.
__device__ void flush_acc(const BArgs &a, Smem &sm) {
  const uint32_t q = sm.acc_q;
  if (q == kNil) return;
  if (sm.acc_count) {
    const unsigned long long c = atomicAdd(&a.st[q].count, sm.acc_count) + sm.acc_count;
    if (c > a.qs[q].cap) atomicExch(&a.st[q].done, kCapped);
  }
  if (sm.acc_items) {
    const unsigned long long it = atomicAdd(&a.st[q].items, sm.acc_items) + sm.acc_items;
    if (it > a.budget) atomicCAS(&a.st[q].done, kRunning, kAbandoned);
  }
  if (sm.acc_rows) atomicAdd(&a.st[q].rows, sm.acc_rows);
  sm.acc_count = sm.acc_items = sm.acc_rows = 0;
}

// Two sums at once: one LDS round and two barriers for both (lds holds
// 2 * kWaves values).
__device__ inline void block_reduce2(unsigned long long &x, unsigned long long &y, unsigned long long *lds) {
  for (int d = kWave / 2; d > 0; d >>= 1) {
    x += __shfl_xor(x, d, kWave);
    y += __shfl_xor(y, d, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    lds[threadIdx.x / kWave] = x;
    lds[kWaves + threadIdx.x / kWave] = y;
  }
  __syncthreads();
  x = y = 0;
  for (int w = 0; w < kWaves; ++w) {
    x += lds[w];
    y += lds[kWaves + w];
  }
  __syncthreads();
}

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

__device__ inline uint32_t row_of(const Smem &sm, uint32_t nb, uint64_t it) {
  uint32_t lo = 0, hi = nb;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sm.pre[mid] <= it) lo = mid; else hi = mid;
  }
  return lo;
}

template <uint32_t N>
__device__ inline void rows_of(const Smem &sm, uint32_t nb, const uint64_t (&it)[N], const bool (&act)[N],
                               uint32_t (&r)[N]) {
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) r[j] = 0;
#pragma unroll
  for (uint32_t step = kRowBatch / 2; step > 0; step >>= 1) {
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
      const uint32_t c = r[j] + step;
      if (act[j] && c < nb && sm.pre[c] <= it[j]) r[j] = c;
    }
  }
}

// Membership of v[j] in the sorted lists cn[lo[j] .. lo[j] + n[j]), all
// searches advancing together so their loads overlap.  n[j] = 0: not found.
// `loads` gains the number of list entries read.
template <uint32_t N>
__device__ inline void lockstep_find(const uint32_t *cn, uint64_t (&lo)[N], uint32_t (&n)[N], const uint32_t (&v)[N],
                                     bool (&found)[N], unsigned long long &loads) {
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
        if (cn[lo[j] + half] <= v[j]) lo[j] += half;
        n[j] -= half;
        ++loads;
      }
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < N; ++j) {
    found[j] = has[j] && cn[lo[j]] == v[j];
    loads += has[j] ? 1 : 0;
  }
}

// Thread 0: writes task t into record `rec` under a new tag and returns the
// deque entry that names it.
__device__ inline unsigned long long write_rec(const BArgs &a, Smem &sm, uint32_t wg, uint32_t rec, const Task &t) {
  const uint32_t seq = ((a.epoch & 0xFFu) << 24) | ((wg & 0x7FFu) << 13) | (sm.seq_ctr++ & 0x1FFFu);
  // agent-scope stores, read back with agent-scope loads (read_rec)
  unsigned long long *w = reinterpret_cast<unsigned long long *>(&a.recs[rec]);
  const unsigned long long v[4] = {t.q | (static_cast<unsigned long long>(t.k) << 32),
                                   t.block | (static_cast<unsigned long long>(t.r0) << 32),
                                   t.r1 | (static_cast<unsigned long long>(t.skip_rows) << 32),
                                   t.skip_items | (static_cast<unsigned long long>(seq) << 32)};
#pragma unroll
  for (int i = 0; i < 4; ++i) __hip_atomic_store(w + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return (static_cast<unsigned long long>(seq) << 32) | rec;
}

// Thread 0: the record an entry names, once its tag matches (re-read after
// an acquire while it does not, a bounded number of times), checked against
// the bounds every later access relies on.  0: usable; else why not (the
// kernel must stop).
__device__ uint32_t read_rec(const BArgs &a, unsigned long long ent, Task &t, bool &late) {
  late = false;
  const uint32_t rec = static_cast<uint32_t>(ent), seq = static_cast<uint32_t>(ent >> 32);
  t = Task{};
  // an entry written by an earlier launch at the same address
  if ((seq >> 24) != (a.epoch & 0xFFu)) return 1;
  if (rec >= a.num_blocks + static_cast<uint64_t>(a.nqueries) * a.tasks_per_query) return 2;
  unsigned long long *w = reinterpret_cast<unsigned long long *>(&a.recs[rec]);
  for (uint32_t spin = 0;; ++spin) {
    unsigned long long v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = __hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t.q = static_cast<uint32_t>(v[0]);
    t.k = static_cast<uint32_t>(v[0] >> 32);
    t.block = static_cast<uint32_t>(v[1]);
    t.r0 = static_cast<uint32_t>(v[1] >> 32);
    t.r1 = static_cast<uint32_t>(v[2]);
    t.skip_rows = static_cast<uint32_t>(v[2] >> 32);
    t.skip_items = static_cast<uint32_t>(v[3]);
    t.seq = static_cast<uint32_t>(v[3] >> 32);
    if (t.seq == seq) {
      late = spin != 0;
      break;
    }
    if (spin >= (1u << 16)) return 3;
    acquire_agent();
    __builtin_amdgcn_s_sleep(2);
  }
  if (t.q >= a.nqueries || t.k == 0 || t.k >= a.qs[t.q].nq) return 4;
  if (t.k == 1)
    return t.r0 < t.r1 && t.r1 <= a.words && t.r1 - t.r0 <= a.root_words && t.skip_rows <= 64u * (t.r1 - t.r0) ? 0
                                                                                                             : 4;
  return t.block < a.num_blocks && t.r0 <= t.r1 && static_cast<uint64_t>(t.r1) * t.k <= a.block_words &&
                 t.skip_rows <= t.r1 - t.r0
             ? 0
             : 4;
}

// Process one task; returns with all threads converged.
__device__ void process_task(const BArgs &a, Smem &sm, uint32_t *s_rows, uint32_t wg, unsigned long long &n_items,
                             unsigned long long &n_rows_in, unsigned long long &n_rows_out,
                             unsigned long long &b_own) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const Task t = sm.task;
  const BQuery &Q = sm.q;
  const uint32_t k = t.k;
  const bool count_level = k == Q.stop;
  const uint32_t kout = k + 1;
  const uint32_t out_cap = static_cast<uint32_t>(a.block_words / kout);
  const uint32_t ilp = count_level ? kIlp : min(a.ilp, out_cap / kThreads);
  // i / kout == __umulhi(i, kout_magic) for every word index of a tile (< 2^17)
  const uint32_t kout_magic = static_cast<uint32_t>((0x100000000ull + kout - 1) / kout);
  const uint32_t nanchor = sm.lv.nanchor, nkept = min(nanchor, kMaxA);
  const bool fused_rows = count_level && Q.fused && sm.lv_last.parent != k;
  bool last_inj_k = false;  // the last position's label equals this one's (lv_last is staged only when fused)
  if (Q.fused)
    for (uint32_t x = 0; x < sm.lv_last.ninj; ++x) last_inj_k |= sm.lv_last.inj[x] == k;
  // count-level rows whose count needs no per-item test
  const bool row_counted = count_level && nanchor == 1 && (!Q.fused || (fused_rows && !last_inj_k));
  // rows of the task; a stolen task's rows were written on another workgroup
  // (possibly another XCD): staged with agent-scope loads
  const bool stolen = sm.task_stolen != 0;
  // a root task's ids, after the first row batch's single column in the
  // dynamic LDS (which holds at least kRootLdsWords words)
  uint32_t *s_roots = s_rows + kRowBatch;
  uint32_t nrows;
  uint32_t *grows = nullptr;
  if (k == 1) {
    // set bits of the root words, appended in order through a scan
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
      s_roots[p++] = (t.r0 + tid) * 64u + static_cast<uint32_t>(i);
    }
    nrows = static_cast<uint32_t>(tot);
    __syncthreads();
  } else {
    nrows = t.r1 - t.r0;
    grows = a.pool + static_cast<uint64_t>(t.block) * a.block_words + static_cast<uint64_t>(t.r0) * k;
  }
  unsigned long long contrib = 0, items_seen = 0, rows_written = 0;
  unsigned long long probe_loads = 0;  // candidate-list entries read by membership searches
  unsigned long long hub_probes = 0;   // adjacency-bitmap words read
  uint32_t row0 = t.skip_rows;
  uint32_t skip_items = t.skip_items;
  uint32_t fill = 0;  // rows in the current output block (same value in every thread)
  uint32_t tiles = 0;
  if (tid == 0) {
    sm.flag = 0;
    sm.nheld = 0;
    sm.out_block = kNil;
    sm.tile_n[0] = sm.tile_n[1] = sm.tile_n[2] = 0;
    if (sm.acc_q != t.q) {  // counts held back belong to one query at a time
      flush_acc(a, sm);
      sm.acc_q = t.q;
    }
  }
  __syncthreads();
  uint64_t tp0 = __builtin_amdgcn_s_memrealtime();
  for (; row0 < nrows; row0 += kRowBatch) {
    // the early-stop check and the batch's rows in one round trip (rows of a
    // batch that turns out stopped are read in bounds and dropped)
    if (tid == 0) {  // both flags read at once
      const uint32_t dn = ld_done(a, t.q);
      const bool ab = aborted(a);
      sm.stop = dn != kRunning || ab;
      if (ab && dn == kRunning) __hip_atomic_store(&a.st[t.q].lost, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // cut short by an early stop
    }
    const uint32_t nb = min(kRowBatch, nrows - row0);
    const uint32_t nwords = nb * k;
    if (k == 1) {
      for (uint32_t i = tid; i < nwords; i += kThreads) s_rows[i] = s_roots[row0 + i];
    } else {
      // kRowLoads loads per thread issued before any is waited for
      const uint32_t *src = grows + static_cast<uint64_t>(row0) * k;
      for (uint32_t i0 = 0; i0 < nwords; i0 += kRowLoads * kThreads) {
        uint32_t v[kRowLoads];
#pragma unroll
        for (uint32_t j = 0; j < kRowLoads; ++j) {
          const uint32_t i = i0 + j * kThreads + tid;
          v[j] = i >= nwords ? 0u
                 : stolen   ? __hip_atomic_load(&src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : src[i];
        }
#pragma unroll
        for (uint32_t j = 0; j < kRowLoads; ++j) {
          const uint32_t i = i0 + j * kThreads + tid;
          if (i < nwords) s_rows[i] = v[j];
        }
      }
    }
    __syncthreads();
    if (sm.stop) break;
    uint64_t tq = 0;  // staging sub-phases (thread 0): rows, list lookups, drivers and scan
    if (tid == 0) {
      tq = __builtin_amdgcn_s_memrealtime();
      sm.phase[4] += tq - tp0;
    }
    // candidate lists of the kept anchors, one (row, anchor) pair per thread,
    // with each anchor vertex's hub index, and for a fused count the last
    // position's list as one more pair per row: every load here depends on
    // the row alone, so they are all in flight at once
    const uint32_t nslot = nkept + (fused_rows ? 1u : 0u);
    for (uint32_t p = tid; p < nb * nslot; p += kThreads) {
      const uint32_t r = p / nslot, x = p - r * nslot;
      const uint32_t *row = s_rows + r * k;
      if (x < nkept) {
        const CnLoc L = Q.ci.locate(sm.lv, x, row[sm.lv.anchor[x]]);
        sm.abeg[x][r] = L.start;
        sm.alen[x][r] = L.len;
        sm.ahub[x][r] = L.hub;
      } else {
        const CnLoc L = Q.ci.locate(sm.lv_last, 0, row[sm.lv_last.parent]);
        sm.fbeg[r] = L.start;
        sm.flen[r] = L.len;
      }
    }
    __syncthreads();
    if (tid == 0) {
      const uint64_t t = __builtin_amdgcn_s_memrealtime();
      sm.phase[5] += t - tq;
    }
    // each row: the shortest list drives
    uint64_t len = 0;
    if (static_cast<uint32_t>(tid) < nb) {
      const uint32_t *row = s_rows + tid * k;
      uint64_t best = ~0ull, bbeg = 0;
      uint32_t bi = 0;
      for (uint32_t x = 0; x < nkept; ++x) {
        if (sm.alen[x][tid] < best) {
          best = sm.alen[x][tid];
          bbeg = sm.abeg[x][tid];
          bi = x;
        }
      }
      for (uint32_t x = kMaxA; x < nanchor; ++x) {
        const CnLoc L = Q.ci.locate(sm.lv, x, row[sm.lv.anchor[x]]);
        if (L.len < best) {
          best = L.len;
          bbeg = L.start;
          bi = x;
        }
      }
      if (tid == 0 && skip_items) {  // resume inside the first row
        bbeg += skip_items;
        best -= skip_items;
      }
      sm.dbeg[tid] = bbeg;
      sm.drv[tid] = static_cast<uint8_t>(bi);
      len = best;
      // kept anchors other than the driver whose vertex is a hub are tested
      // against its adjacency bitmap (the driver's start is taken above)
      for (uint32_t x = 0; x < nkept; ++x) {
        if (x == bi) continue;
        const int32_t h = sm.ahub[x][tid];
        if (h >= 0) sm.abeg[x][tid] = kHubTag | static_cast<uint64_t>(h);
      }
      // this kernel reads the row and two 16-byte locators per list it looks up
      b_own += 4ull * k + 32ull * (nanchor + (fused_rows ? 1 : 0));
      if (fused_rows) {  // the row's part of the last position's count: the
        // list minus the row's own members of it, searched in lockstep
        const uint32_t xv = row[sm.lv_last.parent];
        const uint64_t lb = sm.fbeg[tid];
        const uint32_t ln = sm.flen[tid];
        uint32_t c = ln;
        for (uint32_t x0 = 0; x0 < sm.lv_last.ninj; x0 += kIlp) {
          uint64_t lo[kIlp];
          uint32_t n[kIlp], v[kIlp];
          bool f[kIlp];
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            const uint32_t x = x0 + j;
            const uint32_t col = x < sm.lv_last.ninj ? sm.lv_last.inj[x] : k;
            v[j] = col < k ? row[col] : 0u;
            lo[j] = lb;
            n[j] = col < k && v[j] != xv ? ln : 0u;  // column k is this level's new id: per item
          }
          lockstep_find<kIlp>(Q.ci.cn_adj, lo, n, v, f, probe_loads);
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) c -= f[j] ? 1u : 0u;
        }
        sm.fbase[tid] = c;
      }
      if (row_counted) {  // one anchor: the row's survivors are its list minus the row's own ids
        uint64_t surv = len;
        for (uint32_t x0 = 0; x0 < sm.lv.ninj; x0 += kIlp) {
          uint64_t lo[kIlp];
          uint32_t n[kIlp], v[kIlp];
          bool f[kIlp];
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            const uint32_t x = x0 + j;
            v[j] = x < sm.lv.ninj ? row[sm.lv.inj[x]] : 0u;
            lo[j] = bbeg;
            n[j] = x < sm.lv.ninj ? static_cast<uint32_t>(len) : 0u;
          }
          lockstep_find<kIlp>(Q.ci.cn_adj, lo, n, v, f, probe_loads);
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) surv -= f[j] ? 1u : 0u;
        }
        contrib += surv * (fused_rows ? sm.fbase[tid] : 1u);
        items_seen += len;
        len = 0;
      }
    }
    const uint32_t first_skip = skip_items;  // items of row row0 already done before this batch
    skip_items = 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan(len, reinterpret_cast<uint64_t *>(sm.red), total);
    if (static_cast<uint32_t>(tid) < nb) sm.pre[tid] = ex;
    if (tid == 0) sm.pre[nb] = total;
    __syncthreads();
    n_rows_in += nb;
    if (tid == 0) {
      const uint64_t tn = __builtin_amdgcn_s_memrealtime();
      sm.phase[0] += tn - tp0;
      tp0 = tn;
    }
    const uint64_t step = static_cast<uint64_t>(kThreads) * ilp;
    for (uint64_t base = 0; base < total; base += step) {
      uint32_t w[kIlp], ri[kIlp];
      uint64_t its[kIlp];
      bool ok[kIlp];
#pragma unroll
      for (uint32_t j = 0; j < kIlp; ++j) {
        its[j] = base + j * kThreads + tid;
        ok[j] = j < ilp && its[j] < total;
      }
      rows_of<kIlp>(sm, nb, its, ok, ri);
#pragma unroll
      for (uint32_t j = 0; j < kIlp; ++j) w[j] = ok[j] ? Q.ci.cn_adj[sm.dbeg[ri[j]] + (its[j] - sm.pre[ri[j]])] : 0u;
#pragma unroll
      for (uint32_t j = 0; j < kIlp; ++j) items_seen += ok[j] ? 1 : 0;
      // Two anchors (uniform): the other anchor's list is searched with its
      // first two steps' entries (the middle, and the middles of both halves)
      // loaded beside the item, before the item is compared with anything, so
      // they arrive with it and two dependent rounds disappear.
      uint64_t plo[kIlp];
      uint32_t pn[kIlp], p1[kIlp], p2a[kIlp], p2b[kIlp];
      bool phub[kIlp];
      if (nanchor == 2) {
#pragma unroll
        for (uint32_t j = 0; j < kIlp; ++j) {
          const uint32_t xo = 1u - sm.drv[ri[j]];
          const uint64_t b = sm.abeg[xo][ri[j]];
          phub[j] = ok[j] && (b & kHubTag) != 0;
          plo[j] = phub[j] ? (b & ~kHubTag) : b;  // hub index, or list start
          pn[j] = ok[j] && !phub[j] ? sm.alen[xo][ri[j]] : 0u;
          p1[j] = p2a[j] = p2b[j] = 0;
          if (pn[j] > 1) {
            const uint32_t h = pn[j] >> 1, h2 = (pn[j] - h) >> 1;
            p1[j] = Q.ci.cn_adj[plo[j] + h];
            probe_loads += 1;
            if (h2) {
              p2a[j] = Q.ci.cn_adj[plo[j] + h2];
              p2b[j] = Q.ci.cn_adj[plo[j] + h + h2];
              probe_loads += 2;
            }
          }
        }
      }
      // injectivity against rows in LDS
      for (uint32_t x = 0; x < sm.lv.ninj; ++x) {
        const uint32_t col = sm.lv.inj[x];
#pragma unroll
        for (uint32_t j = 0; j < kIlp; ++j)
          if (ok[j]) ok[j] = s_rows[ri[j] * k + col] != w[j];
      }
      if (nanchor == 2) {
        bool f[kIlp];
#pragma unroll
        for (uint32_t j = 0; j < kIlp; ++j) {
          f[j] = false;
          if (phub[j]) {
            f[j] = (a.probe.hub_bits[plo[j] * a.words + (w[j] >> 6)] >> (w[j] & 63)) & 1ull;
            hub_probes += 1;
            pn[j] = 0;
          } else if (pn[j] > 1) {  // the two prefetched steps
            const uint32_t h = pn[j] >> 1;
            const bool right = p1[j] <= w[j];
            plo[j] += right ? h : 0;
            pn[j] -= h;
            if (pn[j] > 1) {
              const uint32_t h2 = pn[j] >> 1;
              if ((right ? p2b[j] : p2a[j]) <= w[j]) plo[j] += h2;
              pn[j] -= h2;
            }
          }
        }
        bool g[kIlp];
        lockstep_find<kIlp>(Q.ci.cn_adj, plo, pn, w, g, probe_loads);
#pragma unroll
        for (uint32_t j = 0; j < kIlp; ++j)
          if (ok[j]) ok[j] = phub[j] ? f[j] : g[j];
      }
      // adjacency to the other anchors: w is in C(u_k), so w ~ r[a] iff w is
      // in the anchor's candidate list; binary searches run in lockstep
      for (uint32_t x = 0; x < (nanchor == 2 ? 0u : nanchor); ++x) {
        if (x < kMaxA) {
          // an anchor with a dense adjacency bitmap (abeg holds kHubTag | its
          // hub index): one word tells w ~ r[x], since w is a candidate of
          // u_k already; the others search the anchor's candidate list
          uint64_t lo[kIlp];
          uint32_t n[kIlp];
          bool f[kIlp], hub[kIlp], hbit[kIlp];
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            const bool act = ok[j] && sm.drv[ri[j]] != x;
            const uint64_t b = sm.abeg[x][ri[j]];
            const bool tagged = (b & kHubTag) != 0;
            hub[j] = act && tagged;
            lo[j] = tagged ? 0 : b;  // a tagged entry is no list position
            hbit[j] = false;
            if (hub[j])
              hbit[j] = (a.probe.hub_bits[(b & ~kHubTag) * a.words + (w[j] >> 6)] >> (w[j] & 63)) & 1ull;
            n[j] = act && !tagged ? sm.alen[x][ri[j]] : 0u;
            hub_probes += hub[j] ? 1 : 0;
          }
          lockstep_find<kIlp>(Q.ci.cn_adj, lo, n, w, f, probe_loads);
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j)
            if (ok[j] && sm.drv[ri[j]] != x) ok[j] = hub[j] ? hbit[j] : f[j];
        } else {
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j)
            if (ok[j] && sm.drv[ri[j]] != x) ok[j] = a.probe.has(s_rows[ri[j] * k + sm.lv.anchor[x]], w[j]);
        }
      }
      if (count_level) {
        // the last position: |N(x) ∩ C(u_last)| for its parent's id x, minus
        // the list members already in the row (same label, so injectivity)
        if (fused_rows) {
          uint64_t lo[kIlp];
          uint32_t n[kIlp];
          bool f[kIlp];
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            contrib += ok[j] ? sm.fbase[ri[j]] : 0u;
            const bool act = ok[j] && last_inj_k && w[j] != s_rows[ri[j] * k + sm.lv_last.parent];
            lo[j] = sm.fbeg[ri[j]];
            n[j] = act ? sm.flen[ri[j]] : 0u;
          }
          if (last_inj_k) {
            lockstep_find<kIlp>(Q.ci.cn_adj, lo, n, w, f, probe_loads);
#pragma unroll
            for (uint32_t j = 0; j < kIlp; ++j) contrib -= f[j] ? 1u : 0u;
          }
        } else if (Q.fused) {  // the parent is this position: the list of w
          uint64_t lb[kIlp], le[kIlp];
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            lb[j] = le[j] = 0;
            if (ok[j]) {
              const CnLoc L = Q.ci.locate(sm.lv_last, 0, w[j]);
              lb[j] = L.start;
              le[j] = L.start + L.len;
              probe_loads += 8;  // two 16-byte locators, in 4-byte units
            }
            contrib += le[j] - lb[j];
          }
          for (uint32_t x = 0; x < sm.lv_last.ninj; ++x) {
            const uint32_t col = sm.lv_last.inj[x];
            if (col == k) continue;
            uint64_t lo[kIlp];
            uint32_t n[kIlp], v[kIlp];
            bool f[kIlp];
#pragma unroll
            for (uint32_t j = 0; j < kIlp; ++j) {
              v[j] = s_rows[ri[j] * k + col];
              lo[j] = lb[j];
              n[j] = ok[j] && v[j] != w[j] ? static_cast<uint32_t>(le[j] - lb[j]) : 0u;
            }
            lockstep_find<kIlp>(Q.ci.cn_adj, lo, n, v, f, probe_loads);
#pragma unroll
            for (uint32_t j = 0; j < kIlp; ++j) contrib -= f[j] ? 1u : 0u;
          }
        } else {
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) contrib += ok[j] ? 1 : 0;
        }
        continue;
      }
      // survivors: a position per wave from an LDS counter, then per lane
      uint64_t m[kIlp];
      uint32_t wcnt = 0;
#pragma unroll
      for (uint32_t j = 0; j < kIlp; ++j) {
        m[j] = ballot(ok[j]);
        wcnt += static_cast<uint32_t>(__popcll(m[j]));
      }
      const uint32_t cslot = tiles % 3;
      uint32_t wpos = 0;
      if (lane == 0 && wcnt) wpos = atomicAdd(&sm.tile_n[cslot], wcnt);
      wpos = __shfl(wpos, 0, kWave);
      if (tid == 0) sm.tile_n[(tiles + 1) % 3] = 0;  // next tile's counter, last read two tiles ago
      __syncthreads();
      const uint32_t tile = sm.tile_n[cslot];
      ++tiles;
      if (tile && (sm.out_block == kNil || fill + tile > out_cap)) {  // uniform
        __syncthreads();
        if (tid == 0) {
          if (sm.out_block != kNil) {  // keep the full block; it is pushed when the task ends or yields
            sm.held[sm.nheld] = sm.out_block;
            sm.held_fill[sm.nheld] = fill;
            sm.out_block = kNil;
            if (++sm.nheld >= a.max_out) sm.flag = 1;  // bounded output: deeper work goes first
          }
          uint32_t id;
          if (sm.flag) {
          } else if (blk_alloc(a, sm, id)) {
            sm.out_block = id;
          } else {
            sm.flag = 2;  // pool empty: yield before this tile
            atomicAdd(&a.ctl->fails, 1ull);
          }
          if (sm.flag) {  // resume at the row holding item `base`
            const uint32_t lo = row_of(sm, nb, base);
            sm.task.skip_rows = row0 + lo;
            sm.task.skip_items = static_cast<uint32_t>(base - sm.pre[lo] + (lo == 0 ? first_skip : 0));
          }
        }
        __syncthreads();
        fill = 0;
        if (sm.flag) break;
      }
      if (tile) {
        // survivors (row, new id) to LDS in output order, then the tile's
        // rows leave as one contiguous, coalesced run of words
        if (wcnt) {
          uint32_t pos = wpos;
#pragma unroll
          for (uint32_t j = 0; j < kIlp; ++j) {
            if (ok[j]) {
              const uint32_t p = pos + lanes_below(m[j]);
              sm.out_ri[p] = static_cast<uint8_t>(ri[j]);
              sm.out_w[p] = w[j];
            }
            pos += static_cast<uint32_t>(__popcll(m[j]));
          }
        }
        __syncthreads();
        uint32_t *dst = a.pool + static_cast<uint64_t>(sm.out_block) * a.block_words + static_cast<uint64_t>(fill) * kout;
        const uint32_t nw = tile * kout;
        for (uint32_t i = tid; i < nw; i += kThreads) {
          const uint32_t r = __umulhi(i, kout_magic), c = i - r * kout;
          dst[i] = c < k ? s_rows[sm.out_ri[r] * k + c] : sm.out_w[r];
        }
      }
      rows_written += tid == 0 ? tile : 0;
      fill += tile;
    }
    if (count_level) {
      // the row batch's count is added at once, so the query's total passes
      // its cap within one row batch of the embedding that crosses it and
      // every workgroup stops its tasks at their next row batch
      b_own += 4ull * items_seen;
      unsigned long long part = contrib, its = items_seen;
      block_reduce2(part, its, sm.red);
      if (tid == 0) {
        sm.acc_count += part;
        sm.acc_items += its;
        n_items += its;
        if (sm.acc_count >= Q.cap / (4ull * gridDim.x) + 1 || sm.acc_items >= (1ull << 22)) flush_acc(a, sm);
      }
      contrib = 0;
      items_seen = 0;
    }
    if (tid == 0) {
      const uint64_t tn = __builtin_amdgcn_s_memrealtime();
      sm.phase[count_level ? 1 : 2] += tn - tp0;
      tp0 = tn;
    }
    if (sm.flag) break;
    __syncthreads();
  }
  // finish: push continuation and blocks, free the input.  Every lane's row
  // stores are acknowledged by the L2 (vmcnt counts stores on gfx9) before
  // the barriers below, so they sit in this XCD's L2 when thread 0 pushes.
  // No fence here: an agent-scope fence writes back and invalidates the
  // whole L2, and at one per task it dominated the kernel (1.4x slower).
  // Rows reach other XCDs only through the release stores that make them
  // visible: publishing deque entries (deque_publish) and returning blocks
  // to the shared free ring (free_push), each an L2 write-back after these
  // stores; the taker acquires (steal, free_pop) before reading.
  __builtin_amdgcn_s_waitcnt(0);
  b_own += 4ull * items_seen + 4ull * probe_loads + 8ull * hub_probes;  // list entries, searches, bitmap words
  unsigned long long tot_contrib = contrib, tot_items = items_seen;
  block_reduce2(tot_contrib, tot_items, sm.red);
  b_own += 4ull * kout * rows_written;
  if (tid == 0) {
    const bool live = ld_done(a, t.q) == kRunning && !aborted(a);
    if (!live && ld_done(a, t.q) == kRunning) __hip_atomic_store(&a.st[t.q].lost, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // its outputs are dropped by an early stop
    uint32_t np = 0;
    if (sm.flag != 0 && live) {  // same rows, resume point updated
      sm.push[np++] = write_rec(a, sm, wg, sm.rec, sm.task);
    } else if (k >= 2) {
      blk_free(a, sm, t.block);
    }
    // lexicographically first rows on top: the last block is pushed first
    if (sm.out_block != kNil) {
      if (fill && live) {
        sm.push[np++] = write_rec(a, sm, wg, sm.out_block, Task{t.q, kout, sm.out_block, 0, fill, 0, 0, 0});
      } else {
        blk_free(a, sm, sm.out_block);
      }
    }
    for (int i = static_cast<int>(sm.nheld) - 1; i >= 0; --i) {
      if (live) {
        sm.push[np++] = write_rec(a, sm, wg, sm.held[i], Task{t.q, kout, sm.held[i], 0, sm.held_fill[i], 0, 0, 0});
      } else {
        blk_free(a, sm, sm.held[i]);
      }
    }
    if (np) atomicAdd(&a.st[t.q].pending, static_cast<unsigned long long>(np));
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    deque_push(a, sm, wg, sm.push, np);
    sm.sync_ticks += __builtin_amdgcn_s_memrealtime() - t0;
    if (sm.acc_q != t.q) {
      flush_acc(a, sm);
      sm.acc_q = t.q;
    }
    sm.acc_count += count_level ? tot_contrib : 0;
    sm.acc_items += tot_items;
    sm.acc_rows += rows_written;
    // held-back counts over all workgroups stay below a quarter of the cap
    if (sm.acc_count >= Q.cap / (4ull * gridDim.x) + 1 || sm.acc_items >= (1ull << 22)) flush_acc(a, sm);
    n_items += tot_items;
    n_rows_out += rows_written;
    sm.phase[3] += __builtin_amdgcn_s_memrealtime() - tp0;
  }
  __syncthreads();
}

// ---------------------------------------------------------------- resident mode
constexpr uint64_t kPollTicks = 200;  // host doorbell read at most every 2 us (100 MHz clock)

// Thread 0: reads the host doorbell (one workgroup at a time, rate-limited)
// and initialises the slots of new admissions: state first, then the slot of
// each admission, then adm_seen with a release, so a root task is only ever
// taken for an initialised slot.
__device__ void serve_admit(const BArgs &a) {
  const unsigned long long now = __builtin_amdgcn_s_memrealtime();
  const unsigned long long last = ld_rlx(&a.ctl->last_poll);
  if (now - last < kPollTicks || atomicCAS(&a.ctl->last_poll, last, now) != last) return;
  // relaxed reads while nothing is new: an acquire invalidates caches that
  // the busy workgroups of this XCD are reading through
  const unsigned long long db = __hip_atomic_load(&a.sh->doorbell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  st_rlx(&a.ctl->stop_mirror, __hip_atomic_load(&a.sh->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
  st_rlx(&a.ctl->db_mirror, db);
  if (db <= ld_rlx(&a.ctl->adm_seen) || atomicCAS(&a.ctl->adm_lock, 0u, 1u) != 0u) return;
  // the host's ring entries were written before its release of the doorbell
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
  if (__hip_atomic_load(&a.ctl->exiting, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {  // the next launch admits
    __hip_atomic_store(&a.ctl->adm_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const unsigned long long s0 = ld_rlx(&a.ctl->adm_seen);
  unsigned long long s = s0;
  for (; s < db && s < s0 + 64; ++s) {
    const SlotAdmit *e = &a.sh->adm[s % kAdmRing];
    const uint32_t slot = __hip_atomic_load(&e->slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint32_t gen = __hip_atomic_load(&e->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (slot >= a.nqueries) {  // never expected: a host bug; stop, the host recounts
      atomicCAS(&a.ctl->abort, 0u, kAbortBadTask);
      break;
    }
    BState &q = a.st[slot];
    st_init(&q.count, 0ull);
    st_init(&q.done, static_cast<unsigned int>(kRunning));
    st_init(&q.lost, 0u);
    st_init(&q.pending, static_cast<unsigned long long>(a.tasks_per_query));
    st_init(&q.items, 0ull);
    st_init(&q.rows, 0ull);
    st_init(&q.gen, gen);
    st_init(&q.adm, s);
    __hip_atomic_store(&a.ctl->adm_slot[s % kAdmRing], slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __hip_atomic_store(&a.ctl->adm_seen, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&a.ctl->adm_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0: n tasks of query q are finished; the query's last publishes its
// completion record (the counts of finished tasks are in its state already).
__device__ void serve_tasks_done(const BArgs &a, uint32_t q, unsigned long long n) {
  // release: this workgroup's counts precede its decrement; only the last
  // decrement acquires (every other workgroup's counts, released with their
  // decrements), so the other task ends invalidate no cache
  const unsigned long long prev =
      __hip_atomic_fetch_add(&a.st[q].pending, 0ull - n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  if (prev != n) return;
  acquire_agent();
  BState &st = a.st[q];
  SlotDone *d = &a.sh->done[q];
  const unsigned int gen = __hip_atomic_load(&st.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&d->count, ld_rlx(&st.count), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&d->items, ld_rlx(&st.items), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&d->rows, ld_rlx(&st.rows), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&d->lost, __hip_atomic_load(&st.lost, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&d->gen, gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Thread 0: the next root task of an admitted query, if any: root i is range
// i % T of admission i / T (T root tasks per query).  The remaining roots of
// a query that is already finished (capped) are taken together and counted
// as finished tasks, which keeps its task accounting exact.
__device__ bool serve_take_root(const BArgs &a, Smem &sm, uint32_t &rec) {
  const unsigned long long T = a.tasks_per_query;
  // no root left: decided with relaxed loads (an idle workgroup asks often,
  // and an acquire would invalidate caches the busy ones read through)
  if (ld_rlx(&a.ctl->root_next) >= ld_rlx(&a.ctl->adm_seen) * T) return false;
  for (;;) {
    const unsigned long long seen = __hip_atomic_load(&a.ctl->adm_seen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long i = ld_rlx(&a.ctl->root_next);
    if (i >= seen * T) return false;
    const uint32_t slot =
        __hip_atomic_load(&a.ctl->adm_slot[(i / T) % kAdmRing], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ld_done(a, slot) != kRunning) {
      const unsigned long long end = (i / T + 1) * T;
      if (atomicCAS(&a.ctl->root_next, i, end) == i) serve_tasks_done(a, slot, end - i);
      continue;
    }
    if (atomicCAS(&a.ctl->root_next, i, i + 1) != i) continue;
    const uint32_t c = static_cast<uint32_t>(i % T);
    rec = static_cast<uint32_t>(a.num_blocks + static_cast<uint64_t>(slot) * T + c);
    sm.task = Task{slot, 1, 0, c * a.root_words, min(a.words, (c + 1) * a.root_words), 0, 0, 0};
    sm.task_stolen = 0;
    return true;
  }
}

// Thread 0, at the end of every task of query q in resident mode: the
// workgroup's counts of q reach the query's state before the task leaves the
// pending count.
__device__ void serve_task_end(const BArgs &a, Smem &sm, uint32_t q) {
  flush_acc(a, sm);
  serve_tasks_done(a, q, 1);
}

__device__ inline uint32_t xorshift(uint32_t &s) {
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}

// MINB: workgroups per CU the register allocation must allow (2 keeps every
// value in registers; 4 caps them at 128 with a few spills).
template <int MINB>
__global__ __launch_bounds__(kThreads, MINB) void batch_join_kernel(BArgs a) {
  extern __shared__ uint32_t s_rows[];  // kRowBatch x widest row
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const uint32_t wg = blockIdx.x, nwgs = gridDim.x;
  const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
  const uint64_t nroot = static_cast<uint64_t>(a.nqueries) * a.tasks_per_query;
  uint64_t busy = 0;
  unsigned long long n_items = 0, n_rows_in = 0, n_rows_out = 0, n_tasks = 0, n_steals = 0;
  unsigned long long b_own = 0;  // per thread, summed at exit
  uint32_t rng = 0x9E3779B9u ^ (wg * 0x85EBCA6Bu + 1u);
  bool idle = false;  // counted in ctl->idle
  uint32_t idle_rounds = 0;
  {  // no stale descriptor is ever read as a loop bound
    uint32_t *z = reinterpret_cast<uint32_t *>(&sm.lv_last);
    for (uint32_t i = tid; i < sizeof(LevelJoin) / 4; i += kThreads) z[i] = 0;
  }
  if (tid == 0) {
    sm.staged_q = kNil;
    sm.staged_k = kNil;
    sm.staged_gen = sm.task_gen = 0;
    sm.task_adm = sm.acq_seen = 0;
    sm.idle_since = 0;
    sm.ncache = 0;
    sm.seq_ctr = 0;
    sm.dq_bottom = sm.dq_pub = 0;
    sm.acc_q = kNil;
    sm.acc_count = sm.acc_items = sm.acc_rows = 0;
    sm.sync_ticks = 0;
    atomicAdd(&a.ctl->started, 1ull);
    for (int i = 0; i < 7; ++i) sm.phase[i] = 0;
  }
  for (;;) {
    if (tid == 0) {
      if (__builtin_amdgcn_s_memrealtime() - t_start > a.deadline_ticks) atomicCAS(&a.ctl->abort, 0u, kAbortTime);
      sm.state = 1;
      const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
      // own deque, then a new root task, then the oldest task of another
      // workgroup.  Resident mode keeps this order too: taking new queries'
      // roots before the own deque's deeper tasks was tried, and it starved
      // those tasks (a long query's bottom entries waited behind every later
      // arrival) without lowering the median.
      uint32_t rec = kNil;
      unsigned long long ent = idle ? kNilEntry : deque_pop(a, sm, wg);  // an idle deque is empty
      if (rec != kNil) {
      } else if (ent == kNilEntry && a.serve) {
        if (serve_take_root(a, sm, rec) && idle) {
          atomicAdd(&a.ctl->idle, ~0ull);
          idle = false;
        }
      } else if (ent == kNilEntry && ld_rlx(&a.ctl->root_next) < nroot) {
        if (idle) {
          atomicAdd(&a.ctl->idle, ~0ull);
          idle = false;
        }
        for (;;) {
          const unsigned long long i = atomicAdd(&a.ctl->root_next, 1ull);
          if (i >= nroot) break;
          // root tasks interleave the queries; each query's ranges ascend
          const uint32_t q = static_cast<uint32_t>(i % a.nqueries), c = static_cast<uint32_t>(i / a.nqueries);
          if (ld_done(a, q) != kRunning) continue;
          rec = static_cast<uint32_t>(a.num_blocks + static_cast<uint64_t>(q) * a.tasks_per_query + c);
          sm.task = Task{q, 1, 0, c * a.root_words, min(a.words, (c + 1) * a.root_words), 0, 0, 0};
          sm.task_stolen = 0;
          break;
        }
      }
      if (rec == kNil && ent == kNilEntry) {
        for (uint32_t i = 0; i < 8 && ent == kNilEntry; ++i) {
          const uint32_t v = xorshift(rng) % nwgs;
          if (v != wg) {
            ent = deque_steal(a, v, idle, sm.ent_t, sm.ent_b);
            sm.ent_src = 2;
            sm.ent_victim = v;
          }
        }
        if (ent != kNilEntry) {
          ++n_steals;
          idle = false;
        }
      }
      if (ent != kNilEntry) {  // a record from a deque: its tag and bounds first
        Task t;
        bool late;
        const uint32_t why = read_rec(a, ent, t, late);
        if (late && atomicAdd(&a.ctl->late_recs, 1ull) == 0) {
          a.ctl->late_src = sm.ent_src;
          a.ctl->late_victim = sm.ent_victim;
          a.ctl->late_t = sm.ent_t;
          a.ctl->late_b = sm.ent_b;
        }
        if (why == 0) {
          sm.task = t;
          sm.task_stolen = sm.ent_src == 2;
          rec = static_cast<uint32_t>(ent);
        } else {  // never expected: stop, and the host recounts unfinished queries exactly
          if (atomicAdd(&a.ctl->bad_recs, 1ull) == 0) {
            a.ctl->bad_why = why;
            a.ctl->bad_wg = wg;
            a.ctl->bad_ent = ent;
            a.ctl->bad_task = t;
            a.ctl->bad_src = sm.ent_src;
            a.ctl->bad_victim = sm.ent_victim;
            a.ctl->bad_t = sm.ent_t;
            a.ctl->bad_b = sm.ent_b;
          }
          atomicCAS(&a.ctl->abort, 0u, kAbortBadTask);
        }
      }
      if (rec != kNil) {
        sm.rec = rec;
        sm.state = 0;
        sm.stop = ld_done(a, sm.task.q) != kRunning || aborted(a);
        sm.task_gen = a.serve ? __hip_atomic_load(&a.st[sm.task.q].gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        sm.task_adm = a.serve ? ld_rlx(&a.st[sm.task.q].adm) : 0ull;
        if (a.serve && !sm.stop && (sm.staged_q != sm.task.q || sm.staged_gen != sm.task_gen) &&
            sm.task_adm >= sm.acq_seen) {
          // the query's descriptor, lists and bitmaps were written after this
          // workgroup's last invalidation, over memory whose older contents
          // this CU and XCD may still cache: invalidate once before anyone
          // reads them (the barrier below orders the other threads' loads);
          // admissions published later are read after adm_seen is read here
          sm.acq_seen = ld_rlx(&a.ctl->adm_seen);
          acquire_agent();
        }
      } else {
        const unsigned long long now = __builtin_amdgcn_s_memrealtime();
        if (!idle) {
          atomicAdd(&a.ctl->idle, 1ull);
          idle = true;
          sm.idle_since = now;
        }
        // done when every workgroup that has begun is idle (idle read first: it
        // never exceeds `started`, so equality held at the moment it was read)
        if (aborted(a)) {
          sm.state = 2;
        } else if (a.serve) {
          // new admissions first; the whole kernel exits once everything is
          // idle and either the host asked to stop or nothing arrived for a
          // while (the host relaunches for later arrivals).  The decision is
          // taken under the admission lock and stops admissions, so no work
          // can appear after it: every workgroup exits at its next idle turn.
          if (__hip_atomic_load(&a.ctl->exiting, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            sm.state = 2;
          } else {
            serve_admit(a);
            const unsigned long long seen = ld_rlx(&a.ctl->adm_seen);
            if (ld_rlx(&a.ctl->root_next) < seen * a.tasks_per_query) {
              sm.idle_since = now;
            } else if (now - sm.idle_since > a.idle_exit_ticks || ld_rlx(&a.ctl->stop_mirror)) {
              if (atomicCAS(&a.ctl->adm_lock, 0u, 1u) == 0u) {
                acquire_agent();
                const unsigned long long seen2 = ld_rlx(&a.ctl->adm_seen);
                if (ld_rlx(&a.ctl->idle) == ld_rlx(&a.ctl->started) && ld_rlx(&a.ctl->db_mirror) == seen2 &&
                    ld_rlx(&a.ctl->root_next) >= seen2 * a.tasks_per_query) {
                  __hip_atomic_store(&a.ctl->exiting, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  sm.state = 2;
                }
                __hip_atomic_store(&a.ctl->adm_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
              }
            }
          }
        } else if (ld_rlx(&a.ctl->root_next) >= nroot) {
          const unsigned long long nidle = ld_rlx(&a.ctl->idle);
          if (nidle == ld_rlx(&a.ctl->started)) sm.state = 2;
        }
      }
      sm.sync_ticks += __builtin_amdgcn_s_memrealtime() - t0;
    }
    __syncthreads();
    const int state = sm.state;
    if (state == 2) break;
    if (state == 1) {  // back off while the other workgroups produce work
      if (tid == 0) flush_acc(a, sm);
      idle_rounds = min(idle_rounds + 1, kMaxIdleRounds);
      for (uint32_t i = 0; i < idle_rounds; ++i) __builtin_amdgcn_s_sleep(16);
      continue;
    }
    idle_rounds = 0;
    const uint64_t t_task = __builtin_amdgcn_s_memrealtime();
    // a stolen task's rows were published by the victim's unlock and acquired
    // by the lock of the steal (thread 0); the barrier orders the other reads
    const uint32_t q = sm.task.q, k = sm.task.k;
    // a dropped task (its query finished, or the kernel stops) needs nothing staged
    const bool new_q = !sm.stop && (sm.staged_q != q || sm.staged_gen != sm.task_gen);
    if (new_q) {  // stage the query header (uniform branch)
      const uint32_t *src = reinterpret_cast<const uint32_t *>(&a.qs[q]);
      uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.q);
      for (uint32_t i = tid; i < sizeof(BQuery) / 4; i += kThreads) dst[i] = src[i];
      __syncthreads();
      if (sm.q.fused) {
        const uint32_t *s2 = reinterpret_cast<const uint32_t *>(&sm.q.lv[sm.q.nq - 1]);
        uint32_t *d2 = reinterpret_cast<uint32_t *>(&sm.lv_last);
        for (uint32_t i = tid; i < sizeof(LevelJoin) / 4; i += kThreads) d2[i] = s2[i];
      }
    }
    if (!sm.stop && (new_q || sm.staged_k != k)) {  // and the level descriptor
      const uint32_t *src = reinterpret_cast<const uint32_t *>(&sm.q.lv[k]);
      uint32_t *dst = reinterpret_cast<uint32_t *>(&sm.lv);
      for (uint32_t i = tid; i < sizeof(LevelJoin) / 4; i += kThreads) dst[i] = src[i];
    }
    __syncthreads();
    if (tid == 0 && !sm.stop) {  // what LDS holds now
      sm.staged_q = q;
      sm.staged_k = k;
      sm.staged_gen = sm.task_gen;
    }
    if (!sm.stop) {
      process_task(a, sm, s_rows, wg, n_items, n_rows_in, n_rows_out, b_own);
      if (sm.flag == 2)  // pool empty: give other workgroups time to free blocks
        for (int i = 0; i < 16; ++i) __builtin_amdgcn_s_sleep(32);
    } else if (tid == 0) {
      if (ld_done(a, q) == kRunning) __hip_atomic_store(&a.st[q].lost, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // dropped by an early stop
      if (k >= 2) blk_free(a, sm, sm.task.block);
    }
    __syncthreads();
    if (tid == 0) {  // this task is finished
      if (a.serve) serve_task_end(a, sm, q);
      else atomicAdd(&a.st[q].pending, ~0ull);
    }
    ++n_tasks;
    busy += __builtin_amdgcn_s_memrealtime() - t_task;
  }
  if (a.serve && tid == 0)  // other workgroups may still run: give back the cached blocks
    while (sm.ncache) free_push(a, sm.cache[--sm.ncache]);
  const unsigned long long wg_own = block_reduce(b_own, sm.red);
  if (tid == 0) {
    flush_acc(a, sm);
    atomicAdd(&a.ctl->bytes_own, wg_own);
    const uint64_t idle_t = __builtin_amdgcn_s_memrealtime() - t_start - busy;
    atomicAdd(&a.ctl->busy_ticks, static_cast<unsigned long long>(busy));
    atomicAdd(&a.ctl->idle_ticks, static_cast<unsigned long long>(idle_t));
    atomicAdd(&a.ctl->sync_ticks, sm.sync_ticks);
    for (int i = 0; i < 7; ++i) atomicAdd(&a.ctl->phase_ticks[i], sm.phase[i]);
    atomicAdd(&a.ctl->items, n_items);
    atomicAdd(&a.ctl->rows_in, n_rows_in);
    atomicAdd(&a.ctl->rows_out, n_rows_out);
    atomicAdd(&a.ctl->tasks, n_tasks);
    atomicAdd(&a.ctl->steals, n_steals);
  }
}

uint32_t grid_for(uint64_t n) { return static_cast<uint32_t>(std::min<uint64_t>(65536, (n + 255) / 256 + 1)); }

std::atomic<uint32_t> g_launches{0};  // join launches of this process: part of every task record's tag

struct TuneParams {
  uint32_t wgs_per_cu, root_words, block_kb, ilp, max_out;
};

}  // namespace

// Makes ctx->stream = s for the helpers that launch on the context stream
// (and time with its events) while one stage runs.
struct StreamSwap {
  csm_ctx *ctx;
  hipStream_t old;
  StreamSwap(csm_ctx *c, hipStream_t s) : ctx(c), old(c->stream) { c->stream = s; }
  ~StreamSwap() { ctx->stream = old; }
};

BatchRun::BatchRun(csm_ctx *ctx, const csm_graph *g, std::vector<csm_query *> queries, const csm_batch_opts &o)
    : ctx_(ctx), g_(g), queries_(std::move(queries)), o_(o), mem_(ctx), kstat_(CSM_K_NUM) {
  t0_ = std::chrono::steady_clock::now();
  words_ = queries_.empty() ? 0 : queries_[0]->desc.words;
  for (auto *q : queries_) max_nq_ = std::max(max_nq_, q->desc.nq);
}

void BatchRun::release_graph() {
  if (graph_exec_) hipGraphExecDestroy(graph_exec_);
  if (graph_) hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
}

BatchRun::~BatchRun() {
  if (join_stream_) hipStreamSynchronize(join_stream_);
  release_graph();
  join_kt_.reset();
  hipStreamSynchronize(ctx_->stream);
  mem_.release();
}

void *BatchRun::get(uint64_t bytes) { return mem_.get(bytes); }

csm_status BatchRun::prepare(hipStream_t s) {
  StreamSwap swap(ctx_, s);
  csm_ctx *ctx = ctx_;
  const csm_graph *g = g_;
  const uint32_t nqs = static_cast<uint32_t>(queries_.size()), words = words_;
  const uint64_t cap = o_.cap ? o_.cap : ~0ull;
  KernelTimer kt(ctx, kstat_.data());
  // 1. filter (+ refinement) and candidate counts for every query, one launch per step
  uint64_t *stats = static_cast<uint64_t *>(get(8ull * kStatSize * std::max<uint32_t>(nqs, 1)));
  FilterQ *d_fq = static_cast<FilterQ *>(get(sizeof(FilterQ) * std::max<uint32_t>(nqs, 1)));
  if (!stats || !d_fq) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch filter)");
  {
    std::vector<FilterQ> fq(nqs);
    std::vector<const csm_query *> cq(queries_.begin(), queries_.end());
    for (uint32_t i = 0; i < nqs; ++i) fq[i] = filter_entry(g, queries_[i], stats + static_cast<uint64_t>(i) * kStatSize);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(d_fq, fq.data(), sizeof(FilterQ) * nqs, hipMemcpyHostToDevice, s));
    const std::vector<uint32_t> steps = o_.refine ? step_members_of(cq) : std::vector<uint32_t>();
    if (o_.use_graph) {
      hipGraph_t graph;
      hipGraphExec_t exec;
      CSM_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      launch_filter_batch(s, g, d_fq, nqs, words, steps, max_nq_, o_.refine != 0, nullptr);
      CSM_HIP_TRY(ctx, hipStreamEndCapture(s, &graph));
      CSM_HIP_TRY(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      kt.begin(CSM_K_FILTER);
      CSM_HIP_TRY(ctx, hipGraphLaunch(exec, s));
      kt.end(CSM_K_FILTER, 0);
      CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
      hipGraphExecDestroy(exec);
      hipGraphDestroy(graph);
    } else {
      launch_filter_batch(s, g, d_fq, nqs, words, steps, max_nq_, o_.refine != 0, &kt);
    }
    std::vector<uint64_t> host(static_cast<size_t>(kStatSize) * nqs);
    CSM_HIP_TRY(ctx, hipMemcpyAsync(host.data(), stats, 8ull * host.size(), hipMemcpyDeviceToHost, s));
    CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
    for (uint32_t i = 0; i < nqs; ++i) {
      for (uint32_t u = 0; u < queries_[i]->desc.nq; ++u) queries_[i]->counts[u] = host[i * kStatSize + u];
      queries_[i]->filtered = true;
      add_filter_bytes(g, queries_[i], host.data() + static_cast<size_t>(i) * kStatSize, o_.refine != 0,
                       kstat_.data());
    }
  }
  if (o_.reorder)  // candidate-aware order from these counts: host work only
    for (csm_query *q : queries_) {
      const csm_status st = apply_reorder(ctx, g, q);
      if (st != CSM_OK) return st;
    }
  ms_filter_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();

  // 2. candidate indexes of every query
  {
    std::vector<const csm_query *> cq(queries_.begin(), queries_.end());
    csm_status st = build_cand_index_batch(ctx, g, cq, cih_, mem_, kt, true);
    if (st != CSM_OK) return st;
  }

  // 3. per-query descriptors; queries answered without the join
  std::vector<BQuery> bq(nqs);
  host_count_.assign(nqs, 0);
  active_.assign(nqs, 0);
  LevelJoin *d_lv = static_cast<LevelJoin *>(get(sizeof(LevelJoin) * CSM_MAX_QV * std::max<uint32_t>(nqs, 1)));
  d_bq_ = get(sizeof(BQuery) * std::max<uint32_t>(nqs, 1));
  d_st_ = get(sizeof(BState) * std::max<uint32_t>(nqs, 1));
  if (!d_lv || !d_bq_ || !d_st_) return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch descriptors)");
  std::vector<LevelJoin> lvs(static_cast<size_t>(CSM_MAX_QV) * nqs);
  std::vector<std::pair<uint64_t, uint64_t>> two_vertex;  // (query, cn_off index of its one slot)
  for (uint32_t i = 0; i < nqs; ++i) {
    const csm_query *q = queries_[i];
    const uint32_t nq = q->desc.nq, L = nq - 1;
    for (uint32_t k = 0; k < nq; ++k) lvs[static_cast<size_t>(i) * CSM_MAX_QV + k] = cih_[i].level[k];
    BQuery &b = bq[i];
    std::memset(&b, 0, sizeof(b));
    b.ci = CandIndex{q->bits, cih_[i].rank_pref, cih_[i].cn_off, cih_[i].cn_len, cih_[i].cn_adj, words, cih_[i].rloc, cih_[i].loc};
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
      two_vertex.push_back({i, cih_[i].level[1].off_base[0]});
    }
    active_[i] = static_cast<uint8_t>(b.active);
  }
  for (auto &tv : two_vertex) {
    const uint64_t n1 = queries_[tv.first]->counts[queries_[tv.first]->level[0].u];
    csm_status st = slot_total(ctx, cih_[tv.first], tv.second, n1, host_count_[tv.first]);
    if (st != CSM_OK) return st;
  }
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_lv, lvs.data(), sizeof(LevelJoin) * lvs.size(), hipMemcpyHostToDevice, s));
  CSM_HIP_TRY(ctx, hipMemcpyAsync(d_bq_, bq.data(), sizeof(BQuery) * nqs, hipMemcpyHostToDevice, s));
  bq_host_.resize(sizeof(BQuery) * nqs);
  std::memcpy(bq_host_.data(), bq.data(), bq_host_.size());
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  prepared_ = true;
  return CSM_OK;
}

csm_status BatchRun::launch(hipStream_t s) {
  if (!prepared_) return set_error(ctx_, CSM_E_ARG, "batch launched before it was prepared");
  StreamSwap swap(ctx_, s);
  csm_ctx *ctx = ctx_;
  const csm_graph *g = g_;
  const csm_batch_opts &o = o_;
  const uint32_t nqs = static_cast<uint32_t>(queries_.size()), words = words_;
  TuneParams tp{o.wgs_per_cu ? o.wgs_per_cu : 4u, 16u, o.block_kb,
                o.ilp ? std::min(o.ilp, kIlp) : kIlp, o.max_out ? std::min(o.max_out, kMaxOut) : 1u};
  // the caller's item budget stops a query with a lower bound; without one a
  // query runs in the kernel until it is done (or the kernel's time limit)
  const uint64_t budget = o.item_budget ? o.item_budget : ~0ull;
  if (o.chunk_rows) tp.root_words = std::min<uint32_t>(kMaxRootWords, std::max<uint32_t>(1, o.chunk_rows / 64));
  uint64_t pool_bytes = o.pool_bytes;
  const uint64_t avail = ctx->arena.largest_free();
  if (!pool_bytes) pool_bytes = std::min<uint64_t>(avail / 2, 32ull << 30);
  pool_bytes = std::min<uint64_t>(pool_bytes, avail * 3 / 4);
  if (!tp.block_kb) {
    // large blocks mean few tasks per row (each task costs a record, deque
    // operations and a descriptor staging): 512 KiB while the pool still
    // holds kDefaultBlocksPerWg of them per workgroup, else smaller, to 128
    tp.block_kb = 512;
    const uint64_t want = static_cast<uint64_t>(kDefaultBlocksPerWg) * ctx->num_cus * tp.wgs_per_cu;
    while (tp.block_kb > 128 && pool_bytes / (1024ull * tp.block_kb) < want) tp.block_kb /= 2;
  }
  if (tp.wgs_per_cu > 8 || tp.block_kb < 32 || tp.block_kb > 4096) return set_error(ctx, CSM_E_ARG, "batch tuning out of range");
  const uint64_t block_words = static_cast<uint64_t>(tp.block_kb) * 1024 / 4;
  if (block_words < static_cast<uint64_t>(kThreads) * CSM_MAX_QV) return set_error(ctx, CSM_E_ARG, "block too small");
  const uint64_t num_blocks = std::max<uint64_t>(64, pool_bytes / (4 * block_words));
  const uint32_t tasks_per_query = ceil_div(std::max<uint32_t>(words, 1), tp.root_words);
  const uint64_t nroot = static_cast<uint64_t>(nqs) * tasks_per_query;
  if (num_blocks + nroot >= kNil) return set_error(ctx, CSM_E_LIMIT, "too many pool blocks and root tasks");
  const uint32_t grid = ctx->num_cus * tp.wgs_per_cu;
  uint32_t *pool = static_cast<uint32_t *>(get(4 * block_words * num_blocks));
  Task *recs = static_cast<Task *>(get(sizeof(Task) * (num_blocks + nroot)));
  Deque *dq = static_cast<Deque *>(get(sizeof(Deque) * grid));
  unsigned long long *dq_buf = static_cast<unsigned long long *>(get(8ull * kDequeCap * grid));
  FreeSlot *free_ring = static_cast<FreeSlot *>(get(sizeof(FreeSlot) * num_blocks));
  unsigned long long *ctr = static_cast<unsigned long long *>(get(8ull * 2));
  d_ctl_ = get(sizeof(BCtl));
  if (!pool || !recs || !dq || !dq_buf || !free_ring || !ctr || !d_ctl_)
    return set_error(ctx, CSM_E_NOMEM, "arena exhausted (batch pool)");
  BArgs a{};
  a.probe = AdjProbe{g->off, g->adj, g->hub_id, g->hub_bits, words};
  a.qs = static_cast<const BQuery *>(d_bq_);
  a.st = static_cast<BState *>(d_st_);
  a.ctl = static_cast<BCtl *>(d_ctl_);
  a.recs = recs;
  a.dq = dq;
  a.dq_buf = dq_buf;
  a.epoch = g_launches.fetch_add(1) + 1;
  launch_epoch_ = a.epoch;
  tasks_per_query_ = tasks_per_query;
  a.nqueries = nqs;
  a.tasks_per_query = tasks_per_query;
  a.words = words;
  a.budget = budget;
  // workgroup caches hold at most a quarter of the pool
  a.cache_cap = static_cast<uint32_t>(std::min<uint64_t>(kCache, num_blocks / (4ull * grid)));
  a.free_head = ctr;
  a.free_tail = ctr + 1;
  a.free_ring = free_ring;
  a.pool = pool;
  a.num_blocks = num_blocks;
  a.block_words = block_words;
  a.root_words = tp.root_words;
  a.ilp = tp.ilp;
  a.max_out = std::min(tp.max_out, kMaxOut);
  const double limit_s = o.time_limit_ms ? o.time_limit_ms * 1e-3 : 120.0;
  a.deadline_ticks = static_cast<uint64_t>(limit_s * 1e8);  // 100 MHz real-time clock
  grid_ = grid;
  num_blocks_ = num_blocks;
  const size_t lds_rows = 4ull * std::max<uint32_t>(kRowBatch * std::max<uint32_t>(1, max_nq_ - 1), kRootLdsWords);
  auto launch_join = [&]() {
    // register budgets so that wgs_per_cu workgroups of 4 waves fit a CU's
    // 4 SIMDs (512 VGPRs per lane each): 128 (4 per CU), 168 (3), 256 (2).
    // More than 4 per CU needs at most 96 VGPRs: the spills made 5 per CU
    // 1.2x slower on the headline and R-MAT 20 batches, so 4 is the most.
    // Measured: 3 per CU (168 VGPRs, 4 spills) 1.09x slower than 4.
    if (tp.wgs_per_cu >= 4) batch_join_kernel<4><<<grid, kThreads, lds_rows, s>>>(a);
    else if (tp.wgs_per_cu == 3) batch_join_kernel<3><<<grid, kThreads, lds_rows, s>>>(a);
    else batch_join_kernel<2><<<grid, kThreads, lds_rows, s>>>(a);
  };
  join_kt_.reset(new KernelTimer(ctx, kstat_.data()));
  KernelTimer &kt = *join_kt_;
  // the state reset is an ordinary launch before the join (and before its
  // graph): a kernel boundary on the stream, not a node boundary in a graph
  batch_init_kernel<<<grid_for(std::max<uint64_t>(num_blocks, grid)), 256, 0, s>>>(a, grid);
  if (o.use_graph) {  // the graph lives until finish() has waited for its launch
    CSM_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_join();
    CSM_HIP_TRY(ctx, hipStreamEndCapture(s, &graph_));
    CSM_HIP_TRY(ctx, hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    kt.begin(CSM_K_BATCH);
    CSM_HIP_TRY(ctx, hipGraphLaunch(graph_exec_, s));
    kt.end(CSM_K_BATCH, 0);
  } else {
    kt.begin(CSM_K_BATCH);
    launch_join();
    kt.end(CSM_K_BATCH, 0);
  }
  join_stream_ = s;
  launched_ = true;
  return CSM_OK;
}

csm_status BatchRun::finish(csm_result *results) {
  csm_ctx *ctx = ctx_;
  const uint32_t nqs = static_cast<uint32_t>(queries_.size());
  const uint64_t cap = o_.cap ? o_.cap : ~0ull;
  if (!launched_) return set_error(ctx, CSM_E_ARG, "batch finished before it was launched");
  hipStream_t s = join_stream_;
  std::vector<BState> hst(nqs);
  BCtl hctl{};
  CSM_HIP_TRY(ctx, hipMemcpyAsync(hst.data(), d_st_, sizeof(BState) * nqs, hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipMemcpyAsync(&hctl, d_ctl_, sizeof(BCtl), hipMemcpyDeviceToHost, s));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  join_kt_->flush();
  release_graph();
  // algorithmic bytes of the persistent join: its logical reads and writes,
  // per input row of level k the row (4k) and two 16-byte locators per
  // candidate list looked up (32), 4 per list entry examined and per entry
  // read by a membership search, 8 per hub-bitmap word, 4(k + 1) per row
  // written.
  kstat_[CSM_K_BATCH].bytes = hctl.bytes_own;
  const char *tr = std::getenv("CSM_TRACE");
  const bool trace = tr && *tr == '1';
  const uint32_t grid = grid_;
  if (trace) {
    std::fprintf(stderr,
                 "csm batch: queries %u grid %u blocks %llu tasks %llu items %llu rows_in %llu rows_out %llu "
                 "steals %llu abort %u overflow %u fails %llu; per workgroup: busy %.1f ms idle %.1f ms sync %.1f ms\n",
                 nqs, grid, static_cast<unsigned long long>(num_blocks_), hctl.tasks, hctl.items, hctl.rows_in,
                 hctl.rows_out, hctl.steals, hctl.abort, hctl.overflow, hctl.fails, hctl.busy_ticks / 1e5 / grid,
                 hctl.idle_ticks / 1e5 / grid, hctl.sync_ticks / 1e5 / grid);
    std::fprintf(stderr,
                 "csm batch: per workgroup: staging %.1f ms (rows %.1f, list lookups %.1f), count tiles %.1f ms, "
                 "extension tiles %.1f ms, task ends %.1f ms\n",
                 hctl.phase_ticks[0] / 1e5 / grid, hctl.phase_ticks[4] / 1e5 / grid, hctl.phase_ticks[5] / 1e5 / grid,
                 hctl.phase_ticks[1] / 1e5 / grid, hctl.phase_ticks[2] / 1e5 / grid, hctl.phase_ticks[3] / 1e5 / grid);
  }
  if (hctl.late_recs || hctl.bad_recs || hctl.abort == kAbortBadTask)
    std::fprintf(stderr,
                 "csm batch: %llu task records seen late, %llu never matched or out of range (abort %u); first: "
                 "reason %u, workgroup %u of %u, entry rec %u tag %08x (epoch %u), read q %u k %u block %u rows %u..%u "
                 "skip %u/%u tag %08x; queries %u blocks %llu roots/query %u\n",
                 hctl.late_recs, hctl.bad_recs, hctl.abort, hctl.bad_why, hctl.bad_wg, grid_,
                 static_cast<uint32_t>(hctl.bad_ent), static_cast<uint32_t>(hctl.bad_ent >> 32),
                 static_cast<uint32_t>(launch_epoch_ & 0xFF), hctl.bad_task.q, hctl.bad_task.k, hctl.bad_task.block,
                 hctl.bad_task.r0, hctl.bad_task.r1, hctl.bad_task.skip_rows, hctl.bad_task.skip_items, hctl.bad_task.seq,
                 nqs, static_cast<unsigned long long>(num_blocks_), tasks_per_query_);
  if (hctl.late_recs || hctl.bad_recs)
    std::fprintf(stderr,
                 "csm batch: first bad entry from source %u (0 own private, 1 own locked, 2 stolen) victim %u "
                 "at %llu bound %llu; first late record from source %u victim %u at %llu bound %llu\n",
                 hctl.bad_src, hctl.bad_victim, hctl.bad_t, hctl.bad_b, hctl.late_src, hctl.late_victim, hctl.late_t,
                 hctl.late_b);
  for (uint32_t i = 0; i < nqs; ++i) {
    csm_result &r = results[i];
    std::memset(&r, 0, sizeof(r));
    const uint64_t total = active_[i] ? hst[i].count : host_count_[i];
    r.count = std::min(total, cap);
    r.capped = total > cap;
    r.status = CSM_OK;
    r.ms_filter = ms_filter_;
    r.items = hst[i].items;
    r.partial_rows = hst[i].rows;
    r.path = active_[i] ? CSM_PATH_BATCH : CSM_PATH_HOST;
    r.anomalies = static_cast<uint32_t>(std::min<unsigned long long>(hctl.late_recs + hctl.bad_recs, 0xFFFFFFFFull));
    if (i == 0) {
      for (int c = 0; c < CSM_K_NUM; ++c) r.kernels[c] = kstat_[c];
      r.kernels[CSM_K_BATCH].launches = 1;
    }
  }
  // Every unfinished query after the kernel stopped early is finished by the
  // plain path (one at a time, exact), except that with a caller's time limit
  // a query still unfinished when the limit passed keeps its count so far
  // with status CSM_E_TIMEOUT, and with a caller's item budget a query over it
  // keeps its count so far with status CSM_E_BUDGET.
  // A running query whose task accounting did not close (tasks pending, or
  // one dropped) is finished by the plain path too: its count is not trusted.
  uint32_t fallbacks = 0, unclosed = 0;
  const double ms_kernel_end = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
  for (uint32_t i = 0; i < nqs; ++i) {
    if (!active_[i] || hst[i].done == kCapped) continue;
    const bool finished = hst[i].pending == 0 && !hst[i].lost;
    if (hst[i].done == kRunning && finished) continue;  // complete before any early stop: exact
    if (o_.time_limit_ms && hctl.abort == kAbortTime && hst[i].done == kRunning) {
      results[i].status = CSM_E_TIMEOUT;
      continue;
    }
    if (o_.item_budget && hst[i].done == kAbandoned) {  // the caller's bound: a lower bound, as in csm_match
      results[i].status = CSM_E_BUDGET;
      continue;
    }
    {
      if (hst[i].done == kRunning && !hctl.abort) ++unclosed;
      csm_match_opts mo{};
      mo.cap = o_.cap;
      mo.refine = o_.refine;
      mo.reuse_filter = 1;
      if (o_.time_limit_ms) {  // what is left of the caller's limit
        const double left = o_.time_limit_ms - ms_kernel_end;
        if (left <= 1.0) {
          results[i].status = CSM_E_TIMEOUT;
          continue;
        }
        mo.time_limit_ms = static_cast<uint32_t>(left);
      }
      csm_result r{};
      csm_status st = run_join(ctx, g_, queries_[i], mo, r);
      if (st != CSM_OK) return st;
      results[i].count = r.count;
      results[i].capped = r.capped;
      results[i].status = r.status;
      results[i].items += r.items;
      results[i].partial_rows += r.partial_rows;
      results[i].path = CSM_PATH_BATCH_ABORT;
      ++fallbacks;
    }
  }
  const double ms_all = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
  for (uint32_t i = 0; i < nqs; ++i) {
    results[i].ms_total = ms_all;
    results[i].ms_join = ms_all - ms_filter_;
  }
  if (trace || unclosed)
    std::fprintf(stderr, "csm batch: %u queries finished by the plain path (%u with unclosed task accounting), %.1f ms in all\n",
                 fallbacks, unclosed, ms_all);
  return CSM_OK;
}

csm_status BatchRun::copy_descriptor(uint32_t i, void *dst, hipStream_t s) const {
  if (!prepared_ || i >= queries_.size()) return set_error(ctx_, CSM_E_ARG, "descriptor of an unprepared query");
  CSM_HIP_TRY(ctx_, hipMemcpyAsync(dst, bq_host_.data() + sizeof(BQuery) * i, sizeof(BQuery), hipMemcpyHostToDevice, s));
  return CSM_OK;
}

// ------------------------------------------------------------ resident join
ResidentJoin::ResidentJoin(csm_ctx *ctx, const csm_graph *g, uint32_t slots, uint64_t pool_bytes,
                           uint32_t chunk_rows, uint32_t wgs_per_cu)
    : ctx_(ctx), g_(g), slots_(std::min(std::max(slots, 1u), kMaxSlots)),
      root_words_(chunk_rows ? std::min<uint32_t>(kMaxRootWords, std::max<uint32_t>(1, chunk_rows / 64)) : 16u),
      wgs_per_cu_(wgs_per_cu ? std::min(wgs_per_cu, 4u) : 2u), pool_bytes_(pool_bytes), mem_(ctx) {}

ResidentJoin::~ResidentJoin() {
  shutdown();
  if (ev_) hipEventDestroy(ev_);
  delete static_cast<BArgs *>(args_);
  if (sh_) hipHostFree(sh_);
  mem_.release();
}

csm_status ResidentJoin::init(hipStream_t s) {
  s_ = s;
  csm_ctx *ctx = ctx_;
  const uint32_t words = ceil_div(std::max<uint32_t>(g_->n, 1), 64);
  const uint32_t T = ceil_div(words, root_words_);
  const uint64_t block_words = 128ull * 1024 / 4;
  const uint64_t num_blocks = std::max<uint64_t>(64, pool_bytes_ / (4 * block_words));
  if (num_blocks + static_cast<uint64_t>(slots_) * T >= kNil) return set_error(ctx, CSM_E_LIMIT, "too many blocks");
  const uint32_t grid = ctx->num_cus * wgs_per_cu_;
  BArgs *a = new BArgs{};
  args_ = a;
  a->qs = static_cast<const BQuery *>(mem_.get(sizeof(BQuery) * slots_));
  a->st = static_cast<BState *>(mem_.get(sizeof(BState) * slots_));
  a->ctl = static_cast<BCtl *>(mem_.get(sizeof(BCtl)));
  a->pool = static_cast<uint32_t *>(mem_.get(4 * block_words * num_blocks));
  a->recs = static_cast<Task *>(mem_.get(sizeof(Task) * (num_blocks + static_cast<uint64_t>(slots_) * T)));
  a->dq = static_cast<Deque *>(mem_.get(sizeof(Deque) * grid));
  a->dq_buf = static_cast<unsigned long long *>(mem_.get(8ull * kDequeCap * grid));
  a->free_ring = static_cast<FreeSlot *>(mem_.get(sizeof(FreeSlot) * num_blocks));
  unsigned long long *ctr = static_cast<unsigned long long *>(mem_.get(16));
  if (!a->qs || !a->st || !a->ctl || !a->pool || !a->recs || !a->dq || !a->dq_buf || !a->free_ring || !ctr)
    return set_error(ctx, CSM_E_NOMEM, "arena exhausted (resident join)");
  CSM_HIP_TRY(ctx, hipHostMalloc(&sh_, sizeof(ServeShared), hipHostMallocCoherent));
  std::memset(sh_, 0, sizeof(ServeShared));
  CSM_HIP_TRY(ctx, hipMemsetAsync(a->ctl, 0, sizeof(BCtl), s));
  CSM_HIP_TRY(ctx, hipMemsetAsync(const_cast<BQuery *>(a->qs), 0, sizeof(BQuery) * slots_, s));
  CSM_HIP_TRY(ctx, hipMemsetAsync(a->st, 0, sizeof(BState) * slots_, s));
  CSM_HIP_TRY(ctx, hipStreamSynchronize(s));
  CSM_HIP_TRY(ctx, hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  a->probe = AdjProbe{g_->off, g_->adj, g_->hub_id, g_->hub_bits, words};
  a->nqueries = slots_;
  a->tasks_per_query = T;
  a->words = words;
  a->root_words = root_words_;
  a->free_head = ctr;
  a->free_tail = ctr + 1;
  a->num_blocks = num_blocks;
  a->block_words = block_words;
  a->ilp = kIlp;
  a->max_out = 1;
  a->budget = ~0ull;
  a->cache_cap = static_cast<uint32_t>(std::min<uint64_t>(kCache, num_blocks / (4ull * grid)));
  // a resident kernel stops after 120 s (its unfinished queries are then
  // recounted on the plain path) and is relaunched for the rest
  a->deadline_ticks = static_cast<uint64_t>(120.0 * 1e8);
  a->serve = 1;
  a->sh = static_cast<ServeShared *>(sh_);
  a->idle_exit_ticks = 200000;  // 2 ms (100 MHz clock)
  free_.clear();
  for (uint32_t i = slots_; i-- > 0;) free_.push_back(i);
  gen_.assign(slots_, 0);
  busy_.assign(slots_, 0);
  return CSM_OK;
}

csm_status ResidentJoin::launch() {
  BArgs &a = *static_cast<BArgs *>(args_);
  a.epoch = g_launches.fetch_add(1) + 1;
  const uint32_t grid = ctx_->num_cus * wgs_per_cu_;
  const size_t lds_rows = 4ull * std::max<uint32_t>(kRowBatch * (CSM_MAX_QV - 1), kRootLdsWords);
  batch_init_kernel<<<grid_for(std::max<uint64_t>(a.num_blocks, grid)), 256, 0, s_>>>(a, grid);
  batch_join_kernel<4><<<grid, kThreads, lds_rows, s_>>>(a);
  CSM_HIP_TRY(ctx_, hipGetLastError());
  CSM_HIP_TRY(ctx_, hipEventRecord(ev_, s_));
  running_ = true;
  return CSM_OK;
}

bool ResidentJoin::reserve(uint32_t &slot) {
  if (free_.empty()) return false;
  slot = free_.back();
  free_.pop_back();
  return true;
}

csm_status ResidentJoin::fill(const BatchRun &b, uint32_t i, hipStream_t s, uint32_t slot) {
  BArgs &a = *static_cast<BArgs *>(args_);
  csm_status st = b.copy_descriptor(i, const_cast<BQuery *>(a.qs) + slot, s);
  if (st != CSM_OK) return st;
  CSM_HIP_TRY(ctx_, hipStreamSynchronize(s));
  return CSM_OK;
}

void ResidentJoin::ring(uint32_t slot) {
  busy_[slot] = 1;
  const uint32_t gen = ++gen_[slot];
  ServeShared *sh = static_cast<ServeShared *>(sh_);
  SlotAdmit &e = sh->adm[doorbell_ % kAdmRing];
  __atomic_store_n(&e.slot, slot, __ATOMIC_RELAXED);
  __atomic_store_n(&e.gen, gen, __ATOMIC_RELAXED);
  __atomic_store_n(&sh->doorbell, ++doorbell_, __ATOMIC_RELEASE);
}

csm_status ResidentJoin::poll(std::vector<Done> &done) {
  ServeShared *sh = static_cast<ServeShared *>(sh_);
  for (uint32_t slot = 0; slot < slots_; ++slot) {
    if (!busy_[slot] || __atomic_load_n(&sh->done[slot].gen, __ATOMIC_ACQUIRE) != gen_[slot]) continue;
    const SlotDone &d = sh->done[slot];
    done.push_back(Done{slot, __atomic_load_n(&d.count, __ATOMIC_RELAXED), __atomic_load_n(&d.items, __ATOMIC_RELAXED),
                        __atomic_load_n(&d.rows, __ATOMIC_RELAXED), __atomic_load_n(&d.lost, __ATOMIC_RELAXED) != 0});
    busy_[slot] = 0;
    free_.push_back(slot);
  }
  if (running_ && hipEventQuery(ev_) == hipSuccess) {
    running_ = false;
    BArgs &a = *static_cast<BArgs *>(args_);
    BCtl hctl{};
    CSM_HIP_TRY(ctx_, hipMemcpyAsync(&hctl, a.ctl, sizeof(BCtl), hipMemcpyDeviceToHost, s_));
    CSM_HIP_TRY(ctx_, hipStreamSynchronize(s_));
    const char *tr = std::getenv("CSM_TRACE");
    if (tr && *tr == '1') {
      const double g = static_cast<double>(ctx_->num_cus * wgs_per_cu_) * 1e5;
      std::fprintf(stderr,
                   "csm resident: launch ended: tasks %llu items %llu rows_in %llu rows_out %llu steals %llu "
                   "pool misses %llu; per workgroup: busy %.1f idle %.1f sync %.1f staging %.1f count %.1f "
                   "extension %.1f task ends %.1f ms\n",
                   hctl.tasks, hctl.items, hctl.rows_in, hctl.rows_out, hctl.steals, hctl.alloc_fail,
                   hctl.busy_ticks / g, hctl.idle_ticks / g, hctl.sync_ticks / g, hctl.phase_ticks[0] / g,
                   hctl.phase_ticks[1] / g, hctl.phase_ticks[2] / g, hctl.phase_ticks[3] / g);
    }
    if (hctl.abort) {
      // stopped early: every admitted, unfinished query is recounted by the
      // caller, and the next launch starts after the last admission
      std::fprintf(stderr, "csm serve: resident join stopped early (abort %u, %llu late, %llu bad records)\n",
                   hctl.abort, hctl.late_recs, hctl.bad_recs);
      for (uint32_t slot = 0; slot < slots_; ++slot) {
        if (!busy_[slot]) continue;
        done.push_back(Done{slot, 0, 0, 0, true});
        busy_[slot] = 0;
        free_.push_back(slot);
      }
      hctl.root_next = doorbell_ * a.tasks_per_query;
      hctl.adm_seen = hctl.db_mirror = doorbell_;
      CSM_HIP_TRY(ctx_, hipMemcpyAsync(&a.ctl->root_next, &hctl.root_next, 8, hipMemcpyHostToDevice, s_));
      CSM_HIP_TRY(ctx_, hipMemcpyAsync(&a.ctl->adm_seen, &hctl.adm_seen, 8, hipMemcpyHostToDevice, s_));
      CSM_HIP_TRY(ctx_, hipMemcpyAsync(&a.ctl->db_mirror, &hctl.db_mirror, 8, hipMemcpyHostToDevice, s_));
      CSM_HIP_TRY(ctx_, hipStreamSynchronize(s_));
    }
  }
  bool waiting = false;
  for (uint32_t slot = 0; slot < slots_ && !waiting; ++slot) waiting = busy_[slot] != 0;
  if (!running_ && waiting) return launch();
  return CSM_OK;
}

void ResidentJoin::shutdown() {
  if (sh_) __atomic_store_n(&static_cast<ServeShared *>(sh_)->stop, 1ull, __ATOMIC_RELEASE);
  if (running_) {
    hipEventSynchronize(ev_);
    running_ = false;
  }
}

csm_status run_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<csm_query *> &queries,
                     const csm_batch_opts &o, csm_result *results) {
  if (queries.empty()) return CSM_OK;
  BatchRun run(ctx, g, queries, o);
  csm_status st = run.prepare(ctx->stream);
  if (st == CSM_OK) st = run.launch(ctx->stream);
  if (st == CSM_OK) st = run.finish(results);
  return st;
}

}  // namespace csm

using namespace csm;

extern "C" csm_status csm_batch_match(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                                      const csm_batch_opts *opts, csm_result *out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || (n && (!qs || !out))) return CSM_E_ARG;
    csm_batch_opts o{};
    o.use_graph = 1;  // refinement off by default: the candidate lists already prune dead ends
    if (opts) o = *opts;
    std::vector<csm_query *> v(qs, qs + n);
    for (auto *q : v)
      if (!q || q->graph_n != g->n) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    return run_batch(ctx, g, v, o, out);
  });
}

extern "C" csm_status csm_batch_autotune(csm_ctx *ctx, const csm_graph *g, csm_query *const *qs, uint32_t n,
                                         const csm_batch_opts *base, uint32_t sample, csm_batch_opts *out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g || !qs || !out || n == 0) return CSM_E_ARG;
    csm_batch_opts b{};
    b.use_graph = 1;
    if (base) b = *base;
    for (uint32_t i = 0; i < n; ++i)
      if (!qs[i] || qs[i]->graph_n != g->n) return CSM_E_ARG;
    hipSetDevice(ctx->device);
    const uint32_t m = std::max<uint32_t>(1, std::min(sample ? sample : 32u, n));
    std::vector<csm_query *> v;
    for (uint32_t i = 0; i < m; ++i) v.push_back(qs[static_cast<uint64_t>(i) * n / m]);
    std::vector<csm_result> res(m);
    // (workgroups per CU, output blocks per task, block KiB, items per thread)
    static const uint32_t grid[][4] = {{4, 1, 512, 4}, {3, 1, 512, 4}, {2, 1, 512, 4}, {4, 2, 512, 4},
                                       {4, 1, 128, 4}, {4, 1, 256, 4}, {4, 1, 1024, 4}, {4, 1, 512, 2}};
    csm_status st = run_batch(ctx, g, v, b, res.data());  // warm-up: code objects, arena layout
    if (st != CSM_OK) return st;
    double best = 1e300;
    csm_batch_opts pick = b;
    for (const auto &c : grid) {
      csm_batch_opts t = b;
      t.wgs_per_cu = c[0];
      t.max_out = c[1];
      t.block_kb = c[2];
      t.ilp = c[3];
      const auto t0 = std::chrono::steady_clock::now();
      st = run_batch(ctx, g, v, t, res.data());
      if (st != CSM_OK) return st;
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (ms < best) {
        best = ms;
        pick = t;
      }
    }
    *out = pick;
    return CSM_OK;
  });
}
