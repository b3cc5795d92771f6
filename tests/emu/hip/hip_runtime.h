// CPU emulation of the HIP subset the library uses.  Workgroups run on OS
// threads (so __shared__ = static thread_local is per workgroup), lanes are
// ucontext fibers scheduled cooperatively; block barriers and wave
// operations are rendezvous points.  Only for hunting out-of-bounds accesses
// on the CPU (build with -fsanitize=address).
#pragma once
#include <ucontext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&x)

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
#define hipHostMallocCoherent 0x40000000
#define hipEventDisableTiming 0x2
typedef struct EmuStream *hipStream_t;
typedef struct EmuEvent *hipEvent_t;
typedef struct EmuGraph *hipGraph_t;
typedef struct EmuGraphExec *hipGraphExec_t;
typedef struct EmuGraphNode *hipGraphNode_t;

struct dim3 {
  uint32_t x, y, z;
  dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  size_t totalGlobalMem;
};

// ---------------------------------------------------------------- runtime API
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags = 0);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
// a device global's address: on the CPU the object itself
template <typename T>
inline hipError_t hipGetSymbolAddress(void **p, T *sym) {
  *p = static_cast<void *>(sym);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned f);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t *e);
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }  // launches complete before they return
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *n, char *log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipMemGetInfo(size_t *f, size_t *t);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);

// ---------------------------------------------------------------- execution model
struct EmuIdx {
  uint32_t x, y, z;
};
const EmuIdx &emu_thread_idx();
const EmuIdx &emu_block_idx();
const EmuIdx &emu_block_dim();
const EmuIdx &emu_grid_dim();
#define threadIdx (emu_thread_idx())
#define blockIdx (emu_block_idx())
#define blockDim (emu_block_dim())
#define gridDim (emu_grid_dim())

void *emu_dyn_lds();
void emu_syncthreads(int site);
void emu_yield();
// all live lanes of the wave meet; returns the wave's 64 values (inactive lanes 0)
// and the mask of lanes that took part
const uint64_t *emu_wave_meet(int site, uint64_t v, uint64_t *active);
uint32_t emu_lane();
uint64_t emu_memrealtime();

#define __syncthreads() emu_syncthreads(__LINE__)
int emu_syncthreads_or(int site, int p);
#define __syncthreads_or(p) emu_syncthreads_or(__LINE__, (p))
#define __threadfence() std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_fence(o, s) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() (void)emu_wave_meet(__LINE__, 0, nullptr)
#define __builtin_amdgcn_s_sleep(n) emu_yield()
#define __builtin_amdgcn_s_waitcnt(n) ((void)0)
#define __builtin_amdgcn_s_memrealtime() emu_memrealtime()
#define __lane_id() (emu_lane() & 63u)

inline uint32_t emu_mbcnt_lo(uint32_t m, uint32_t acc) {
  const uint32_t l = emu_lane() & 63u;
  const uint32_t below = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
  return acc + static_cast<uint32_t>(__builtin_popcount(m & below));
}
inline uint32_t emu_mbcnt_hi(uint32_t m, uint32_t acc) {
  const uint32_t l = emu_lane() & 63u;
  const uint32_t below = l < 32 ? 0u : ((1u << (l - 32)) - 1u);
  return acc + static_cast<uint32_t>(__builtin_popcount(m & below));
}
#define __builtin_amdgcn_mbcnt_lo(m, a) emu_mbcnt_lo((m), (a))
#define __builtin_amdgcn_mbcnt_hi(m, a) emu_mbcnt_hi((m), (a))

#define __popcll(x) __builtin_popcountll(x)
#define __popc(x) __builtin_popcount(x)
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32); }

template <class T>
inline uint64_t emu_bits(T v) {
  static_assert(sizeof(T) <= 8, "wave op on a wide type");
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T emu_from(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}
template <class T>
inline T emu_shfl_xor(int site, T v, int mask, int width = 64) {
  const uint64_t *w = emu_wave_meet(site, emu_bits(v), nullptr);
  const uint32_t l = emu_lane() & 63u;
  const uint32_t base = l & ~static_cast<uint32_t>(width - 1);
  return emu_from<T>(w[base + (((l - base) ^ static_cast<uint32_t>(mask)) & (width - 1))]);
}
template <class T>
inline T emu_shfl(int site, T v, int src, int width = 64) {
  const uint64_t *w = emu_wave_meet(site, emu_bits(v), nullptr);
  const uint32_t l = emu_lane() & 63u;
  const uint32_t base = l & ~static_cast<uint32_t>(width - 1);
  return emu_from<T>(w[base + (static_cast<uint32_t>(src) & (width - 1))]);
}
template <class T>
inline T emu_shfl_up(int site, T v, unsigned d, int width = 64) {
  const uint64_t *w = emu_wave_meet(site, emu_bits(v), nullptr);
  const uint32_t l = emu_lane() & 63u;
  const uint32_t base = l & ~static_cast<uint32_t>(width - 1);
  return (l - base) >= d ? emu_from<T>(w[l - d]) : v;
}
inline uint64_t emu_ballot(int site, bool p) {
  uint64_t act = 0;
  const uint64_t *w = emu_wave_meet(site, p ? 1 : 0, &act);
  uint64_t m = 0;
  for (int i = 0; i < 64; ++i)
    if (((act >> i) & 1) && w[i]) m |= 1ull << i;
  return m;
}
#define __shfl_xor(...) emu_shfl_xor(__LINE__, __VA_ARGS__)
#define __shfl(...) emu_shfl(__LINE__, __VA_ARGS__)
#define __shfl_up(...) emu_shfl_up(__LINE__, __VA_ARGS__)
#define __ballot(p) emu_ballot(__LINE__, (p))

// atomics: sequentially consistent host atomics
template <class T, class U>
inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, static_cast<T>(v), __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicSub(T *p, U v) { return __atomic_fetch_sub(p, static_cast<T>(v), __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, static_cast<T>(v), __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicExch(T *p, U v) { return __atomic_exchange_n(p, static_cast<T>(v), __ATOMIC_SEQ_CST); }
template <class T, class U, class V>
inline T atomicCAS(T *p, U cmp, V v) {
  T c = static_cast<T>(cmp);
  __atomic_compare_exchange_n(p, &c, static_cast<T>(v), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return c;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < static_cast<T>(v) &&
         !__atomic_compare_exchange_n(p, &cur, static_cast<T>(v), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
#define __HIP_MEMORY_SCOPE_WAVEFRONT 1
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3
#define __HIP_MEMORY_SCOPE_SYSTEM 4
#define __hip_atomic_load(p, o, s) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, o, s) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, o, s) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
  return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
  return a < b ? b : a;
}

// kernel launch: every workgroup of the grid runs `body` in each lane
void emu_launch(const char *name, dim3 grid, dim3 block, size_t lds, std::function<void()> body);
template <class F, class... A>
void emu_launch_fn(const char *name, dim3 g, dim3 b, size_t l, F *f, A &&...a) {
  auto tup = std::make_tuple(typename std::decay<A>::type(a)...);
  emu_launch(name, g, b, l, [f, tup]() { std::apply(f, tup); });
}
#define EMU_LAUNCH(kern, g, b, l, s, ...) emu_launch_fn(#kern, dim3(g), dim3(b), (l), kern, __VA_ARGS__)

struct uint2 { uint32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
struct ulonglong2 { unsigned long long x, y; };
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
