// Device-wide exclusive scan of uint64 (reduce-then-scan, three launches).
// Used for level degree prefixes, tile offsets and compaction.
#include "common.hpp"

namespace csm {

namespace {

constexpr int kScanThreads = 256;
constexpr int kPerThread = kScanTile / kScanThreads;  // 16

// Exclusive scan of one value per thread over a 256-thread block; returns the
// block total through *total.
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *lds, uint64_t *total) {
  const int tid = threadIdx.x;
  // wave-inclusive scan with DPP-free shuffles
  uint64_t x = v;
  const int lane = tid & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  const int wave = tid / kWave;
  if (lane == kWave - 1) lds[wave] = x;
  __syncthreads();
  uint64_t wave_off = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    uint64_t t = lds[w];
    if (w < wave) wave_off += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return wave_off + x - v;
}

__global__ __launch_bounds__(kScanThreads) void tile_reduce(const uint64_t *data, uint64_t n, uint64_t *partials) {
  __shared__ uint64_t lds[kScanThreads / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kScanThreads + threadIdx.x;
    if (idx < n) s += data[idx];
  }
  uint64_t total;
  block_exclusive_scan(s, lds, &total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void partials_scan(uint64_t *partials, uint64_t np, uint64_t *total_out) {
  __shared__ uint64_t lds[kScanThreads / kWave];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < np; b += kScanThreads) {
    uint64_t idx = b + threadIdx.x;
    uint64_t v = idx < np ? partials[idx] : 0;
    uint64_t total;
    uint64_t ex = block_exclusive_scan(v, lds, &total);
    if (idx < np) partials[idx] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Each thread owns kPerThread consecutive elements (blocked layout), so the
// scan order is the element order.  The staging array holds one pad element
// after every kPerThread: a thread's block then starts kPerThread + 1
// elements after its neighbour's, so the blocked accesses of a wave spread
// over the LDS banks instead of all meeting in the same ones.
__device__ inline uint32_t padded(uint32_t e) { return e + e / kPerThread; }

__global__ __launch_bounds__(kScanThreads) void tile_scan(uint64_t *data, uint64_t n, const uint64_t *partials) {
  __shared__ uint64_t lds[kScanThreads / kWave];
  __shared__ uint64_t stage[kScanTile + kScanTile / kPerThread];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kScanThreads + threadIdx.x;
    stage[padded(i * kScanThreads + threadIdx.x)] = idx < n ? data[idx] : 0;
  }
  __syncthreads();
  uint64_t v[kPerThread];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    v[i] = stage[padded(threadIdx.x * kPerThread + i)];
    s += v[i];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, lds, &total) + partials[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    stage[padded(threadIdx.x * kPerThread + i)] = run;
    run += v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kScanThreads + threadIdx.x;
    if (idx < n) data[idx] = stage[padded(i * kScanThreads + threadIdx.x)];
  }
}

}  // namespace

void scan_u64_exclusive(uint64_t *data, uint64_t n, uint64_t *partials, uint64_t *total_out, hipStream_t s) {
  if (n == 0) {
    if (total_out) hipMemsetAsync(total_out, 0, sizeof(uint64_t), s);
    return;
  }
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  tile_reduce<<<static_cast<uint32_t>(tiles), kScanThreads, 0, s>>>(data, n, partials);
  partials_scan<<<1, kScanThreads, 0, s>>>(partials, tiles, total_out);
  tile_scan<<<static_cast<uint32_t>(tiles), kScanThreads, 0, s>>>(data, n, partials);
}

}  // namespace csm
