// Fiber scheduler and HIP runtime subset for the CPU build of the device code
// (tests/test_emulated.py).  Kernels run one at a time, as on one stream;
// persistent workers each own a workgroup slot of up to 1024 lane fibers.
#include <sanitizer/common_interface_defs.h>
#include <setjmp.h>
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

#if defined(__SANITIZE_ADDRESS__)
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

namespace {

constexpr size_t kStack = 128 << 10;

enum State { kRun, kBlock, kWave, kDone };

struct Lane {
  ucontext_t ctx;
  jmp_buf jb;
  bool started = false, made = false;
  char *stack = nullptr;
  State st = kRun;
  int site = 0;
  uint64_t val = 0;
  EmuIdx tidx{0, 0, 0};
};

struct Block {
  std::vector<Lane> lanes;
  uint32_t n = 0;
  EmuIdx bidx{0, 0, 0}, bdim{0, 0, 0}, gdim{0, 0, 0};
  std::function<void()> *body = nullptr;
  ucontext_t sched;
  jmp_buf sched_jb;
  const void *sched_bottom = nullptr;
  size_t sched_size = 0;
  uint32_t cur = 0;
  std::vector<uint64_t> wave_vals;  // per wave: 64 values snapshot
  std::vector<uint64_t> wave_act;
  std::vector<char> lds;
  const char *kname = "";
  uint64_t bar_or = 0;  // OR of the values the lanes brought to the last block barrier
};

thread_local Block *t_blk = nullptr;

void to_sched(Block *b) {
  Lane &l = b->lanes[b->cur];
#if EMU_ASAN
  __sanitizer_start_switch_fiber(nullptr, b->sched_bottom, b->sched_size);
#endif
  if (!_setjmp(l.jb)) _longjmp(b->sched_jb, 1);
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &b->sched_bottom, &b->sched_size);
#endif
}

// A lane runs the body once per workgroup the worker takes, parking after
// each run; the worker resumes it for the next workgroup.
void lane_entry(int lo, int hi) {
  Block *b = t_blk;
  (void)lo;
  (void)hi;
#if EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &b->sched_bottom, &b->sched_size);
#endif
  for (;;) {
    (*b->body)();
    b->lanes[b->cur].st = kDone;
    to_sched(b);
  }
}

[[noreturn]] void die(Block *b, const char *what) {
  std::fprintf(stderr, "EMU: %s in kernel %s block (%u,%u,%u)\n", what, b->kname, b->bidx.x, b->bidx.y, b->bidx.z);
  for (uint32_t i = 0; i < b->n; ++i) {
    const Lane &l = b->lanes[i];
    if (l.st != kDone) std::fprintf(stderr, "  lane %u state %d site %d\n", i, l.st, l.site);
  }
  std::abort();
}

void run_block(Block *b) {
  const uint32_t n = b->n, nw = (n + 63) / 64;
  for (uint32_t i = 0; i < n; ++i) {
    Lane &l = b->lanes[i];
    l.st = kRun;
    l.tidx = EmuIdx{i % b->bdim.x, (i / b->bdim.x) % b->bdim.y, i / (b->bdim.x * b->bdim.y)};
    if (!l.started && !l.made) {
      getcontext(&l.ctx);
      l.ctx.uc_stack.ss_sp = l.stack;
      l.ctx.uc_stack.ss_size = kStack;
      l.ctx.uc_link = nullptr;
      makecontext(&l.ctx, reinterpret_cast<void (*)()>(lane_entry), 2, 0, 0);
      l.made = true;
    }
  }
  for (;;) {
    bool ran = false;
    for (uint32_t i = 0; i < n; ++i) {
      Lane &l = b->lanes[i];
      if (l.st != kRun) continue;
      ran = true;
      b->cur = i;
#if EMU_ASAN
      __sanitizer_start_switch_fiber(nullptr, l.stack, kStack);
#endif
      if (!_setjmp(b->sched_jb)) {
        if (l.started) {
          _longjmp(l.jb, 1);
        } else {
          l.started = true;
          ucontext_t tmp;
          swapcontext(&tmp, &l.ctx);
        }
      }
#if EMU_ASAN
      __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
    }
    if (ran) continue;
    // resolve wave meetings: every live lane of a wave waits at one site
    bool progress = false;
    for (uint32_t w = 0; w < nw; ++w) {
      int site = -1;
      bool any_wave = false, mixed = false;
      for (uint32_t i = w * 64; i < std::min(n, (w + 1) * 64); ++i) {
        const Lane &l = b->lanes[i];
        if (l.st == kWave) {
          if (site < 0) site = l.site;
          else if (site != l.site) mixed = true;
          any_wave = true;
        }
      }
      if (!any_wave) continue;
      if (mixed) die(b, "lanes of one wave wait at different wave operations");
      uint64_t act = 0;
      uint64_t *vals = &b->wave_vals[w * 64];
      for (uint32_t i = w * 64; i < std::min(n, (w + 1) * 64); ++i) {
        Lane &l = b->lanes[i];
        vals[i - w * 64] = 0;
        if (l.st == kWave) {
          vals[i - w * 64] = l.val;
          act |= 1ull << (i - w * 64);
        }
      }
      b->wave_act[w] = act;
      for (uint32_t i = w * 64; i < std::min(n, (w + 1) * 64); ++i)
        if (b->lanes[i].st == kWave) b->lanes[i].st = kRun;
      progress = true;
    }
    if (progress) continue;
    uint32_t live = 0, at_bar = 0;
    int bsite = -1;
    bool bmixed = false;
    for (uint32_t i = 0; i < n; ++i) {
      const Lane &l = b->lanes[i];
      if (l.st == kDone) continue;
      ++live;
      if (l.st == kBlock) {
        ++at_bar;
        if (bsite < 0) bsite = l.site;
        else if (bsite != l.site) bmixed = true;
      }
    }
    if (live == 0) return;
    if (at_bar == live) {
      if (bmixed) die(b, "lanes wait at different __syncthreads");
      b->bar_or = 0;
      for (uint32_t i = 0; i < n; ++i)
        if (b->lanes[i].st == kBlock) {
          b->bar_or |= b->lanes[i].val;
          b->lanes[i].st = kRun;
        }
      continue;
    }
    die(b, "deadlock");
  }
}

std::atomic<uint64_t> g_launches{0};

}  // namespace

const EmuIdx &emu_thread_idx() { return t_blk->lanes[t_blk->cur].tidx; }
const EmuIdx &emu_block_idx() { return t_blk->bidx; }
const EmuIdx &emu_block_dim() { return t_blk->bdim; }
const EmuIdx &emu_grid_dim() { return t_blk->gdim; }
uint32_t emu_lane() { return t_blk->cur; }
void *emu_dyn_lds() { return t_blk->lds.data(); }

void emu_syncthreads(int site) {
  Block *b = t_blk;
  Lane &l = b->lanes[b->cur];
  l.st = kBlock;
  l.site = site;
  l.val = 0;
  to_sched(b);
}

int emu_syncthreads_or(int site, int p) {
  Block *b = t_blk;
  Lane &l = b->lanes[b->cur];
  l.st = kBlock;
  l.site = site;
  l.val = p != 0;
  to_sched(b);
  return b->bar_or != 0;
}

void emu_yield() {
  Block *b = t_blk;
  to_sched(b);  // stays runnable
  std::this_thread::yield();
}

const uint64_t *emu_wave_meet(int site, uint64_t v, uint64_t *active) {
  Block *b = t_blk;
  const uint32_t me = b->cur;
  Lane &l = b->lanes[me];
  l.st = kWave;
  l.site = site;
  l.val = v;
  to_sched(b);
  const uint32_t w = b->cur / 64;
  if (active) *active = b->wave_act[w];
  return &b->wave_vals[w * 64];
}

uint64_t emu_memrealtime() {  // 100 MHz, slowed by EMU_CLOCK_DIV (the emulation runs far slower than a GPU)
  static const uint64_t div = std::getenv("EMU_CLOCK_DIV") ? std::strtoull(std::getenv("EMU_CLOCK_DIV"), nullptr, 10) : 1;
  return static_cast<uint64_t>(
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count() /
      (10 * div));
}

namespace {
// persistent workers: each owns one workgroup slot (1024 lane fibers)
struct Job {
  const char *name;
  dim3 grid, block;
  size_t lds;
  std::function<void()> *body;
  uint64_t nblocks;
  std::atomic<uint64_t> next{0};
};
// never destroyed: detached workers wait on them until the process ends
std::mutex &g_pool_mu = *new std::mutex;
std::condition_variable &g_pool_cv = *new std::condition_variable;
std::condition_variable &g_done_cv = *new std::condition_variable;
Job *g_job = nullptr;
uint64_t g_gen = 0;
uint32_t g_busy = 0;
std::vector<std::thread> g_workers;

void worker_main() {
  Block b;
  b.lanes.resize(1024);
  for (auto &l : b.lanes) l.stack = static_cast<char *>(std::malloc(kStack));
  b.wave_vals.assign(1024, 0);
  b.wave_act.assign(16, 0);
  t_blk = &b;
  uint64_t seen = 0;
  for (;;) {
    Job *j;
    {
      std::unique_lock<std::mutex> lk(g_pool_mu);
      g_pool_cv.wait(lk, [&] { return g_gen != seen; });
      seen = g_gen;
      j = g_job;
    }
    const uint32_t nthr = j->block.x * j->block.y * j->block.z;
    b.n = nthr;
    b.bdim = EmuIdx{j->block.x, j->block.y, j->block.z};
    b.gdim = EmuIdx{j->grid.x, j->grid.y, j->grid.z};
    b.body = j->body;
    b.kname = j->name;
    b.lds.assign(j->lds + 16, 0x5a);
    for (;;) {
      const uint64_t i = j->next.fetch_add(1);
      if (i >= j->nblocks) break;
      b.bidx = EmuIdx{static_cast<uint32_t>(i % j->grid.x), static_cast<uint32_t>((i / j->grid.x) % j->grid.y),
                      static_cast<uint32_t>(i / (static_cast<uint64_t>(j->grid.x) * j->grid.y))};
      run_block(&b);
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (--g_busy == 0) g_done_cv.notify_all();
  }
}
}  // namespace

void emu_launch(const char *name, dim3 grid, dim3 block, size_t lds, std::function<void()> body) {
  ++g_launches;
  const uint64_t nblocks = static_cast<uint64_t>(grid.x) * grid.y * grid.z;
  const uint32_t nthr = block.x * block.y * block.z;
  if (nblocks == 0 || nthr == 0 || nthr > 1024) {
    std::fprintf(stderr, "EMU: bad launch of %s grid %u,%u,%u block %u\n", name, grid.x, grid.y, grid.z, nthr);
    std::abort();
  }
  if (lds > 64 * 1024) {
    std::fprintf(stderr, "EMU: kernel %s asks %zu bytes of dynamic LDS\n", name, lds);
    std::abort();
  }
  static std::mutex launch_mu;  // one kernel at a time, as on one stream
  std::lock_guard<std::mutex> one(launch_mu);
  Job j;
  j.name = name;
  j.grid = grid;
  j.block = block;
  j.lds = lds;
  j.body = &body;
  j.nblocks = nblocks;
  std::unique_lock<std::mutex> lk(g_pool_mu);
  if (g_workers.empty()) {
    const char *env = std::getenv("EMU_THREADS");
    const uint32_t n = env ? static_cast<uint32_t>(std::atoi(env)) : 8u;
    for (uint32_t t = 0; t < std::max(1u, n); ++t) g_workers.emplace_back(worker_main);
    for (auto &t : g_workers) t.detach();
  }
  g_job = &j;
  g_busy = static_cast<uint32_t>(g_workers.size());
  ++g_gen;
  g_pool_cv.notify_all();
  g_done_cv.wait(lk, [&] { return g_busy == 0; });
  g_job = nullptr;
}

// ---------------------------------------------------------------- runtime API
struct EmuEvent {
  double ms = 0;
};

hipError_t hipMalloc(void **p, size_t n) {
  *p = std::aligned_alloc(256, (std::max<size_t>(n, 1) + 255) / 256 * 256);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
  if (n) std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) { return hipMemsetAsync(d, v, n, nullptr); }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  *s = reinterpret_cast<hipStream_t>(new char[8]);
  return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) {
  delete[] reinterpret_cast<char *>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new EmuEvent();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = static_cast<float>(b->ms - a->ms);
  return hipSuccess;
}
struct EmuGraph {
  int x;
};
struct EmuGraphExec {
  int x;
};
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  *g = new EmuGraph();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t, hipGraphNode_t *, char *, size_t) {
  *e = new EmuGraphExec();
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->name, "emu");
  std::strcpy(p->gcnArchName, "gfx950:emu");
  const char *c = std::getenv("EMU_CUS");
  p->multiProcessorCount = c ? std::atoi(c) : 2;
  p->totalGlobalMem = 8ull << 30;
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t *f, size_t *t) {
  *f = 2ull << 30;
  *t = 8ull << 30;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emu error"; }

// SIGUSR1: print the receiving thread's stack (for a stalled run)
#include <execinfo.h>
#include <signal.h>
#include <unistd.h>
namespace {
void emu_bt(int) {
  void *fr[64];
  const int n = backtrace(fr, 64);
  backtrace_symbols_fd(fr, n, 2);
}
struct EmuBtInit {
  EmuBtInit() { signal(SIGUSR1, emu_bt); }
} g_emu_bt_init;
}  // namespace
