// Exclusive prefix sum of per-query result counts (uint32 -> uint64), in
// three launches: per-tile sums, a single-workgroup scan of the tile sums
// (carried across chunks, so any number of tiles works), and the per-tile
// scan that adds the tile's offset. Also the queue reservation: one thread
// adds the grand total to the result queue's counter.
#include "rts_internal.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kTileElems = kThreads * kPerThread;  // 1024 counts per tile

__device__ __forceinline__ uint64_t block_exclusive_u64(uint64_t v, uint64_t* sh_wave,
                                                         uint64_t* total) {
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += o;
  }
  if (lane == kWave - 1) sh_wave[w] = inc;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int i = 0; i < kThreads / kWave; i++) {
    if (i < (int)w) before += sh_wave[i];
    all += sh_wave[i];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ void __launch_bounds__(kThreads) k_tile_sums(const uint32_t* __restrict__ counts, uint64_t n,
                                                        uint64_t* __restrict__ tile_sums) {
  __shared__ uint64_t sh[kThreads / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kTileElems;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
    if (i < n) s += counts[i];
  }
  s = wave_sum_u64(s);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kThreads / kWave; w++) t += sh[w];
    tile_sums[blockIdx.x] = t;
  }
}

// One workgroup: exclusive scan of the tile sums in place; total -> *total.
__global__ void __launch_bounds__(kThreads) k_scan_tiles(uint64_t* tile_sums, uint64_t tiles,
                                                         uint64_t* __restrict__ total) {
  __shared__ uint64_t sh[kThreads / kWave];
  uint64_t carry = 0;
  for (uint64_t c = 0; c < tiles; c += kThreads) {
    uint64_t i = c + threadIdx.x;
    uint64_t v = i < tiles ? tile_sums[i] : 0;
    uint64_t all;
    uint64_t ex = block_exclusive_u64(v, sh, &all);
    if (i < tiles) tile_sums[i] = carry + ex;
    carry += all;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kThreads) k_scan_apply(const uint32_t* __restrict__ counts, uint64_t n,
                                                         const uint64_t* __restrict__ tile_offsets,
                                                         uint64_t* __restrict__ offsets) {
  __shared__ uint64_t sh[kThreads / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kTileElems;
  // Thread t owns elements base + t*4 .. +3 (blocked) so that the scan order
  // is the element order.
  uint32_t c[kPerThread];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    uint64_t i = base + (uint64_t)threadIdx.x * kPerThread + k;
    c[k] = i < n ? counts[i] : 0u;
    s += c[k];
  }
  uint64_t all;
  uint64_t ex = block_exclusive_u64(s, sh, &all) + tile_offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    uint64_t i = base + (uint64_t)threadIdx.x * kPerThread + k;
    if (i < n) offsets[i] = ex;
    ex += c[k];
  }
}

// base[0] = atomicAdd(queue counter, total): the slots this launch owns.
__global__ void k_reserve(const rts_dev_queue* __restrict__ dq, const uint64_t* __restrict__ total,
                          uint64_t* __restrict__ base) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    base[0] = atomicAdd(dq->counter, (unsigned long long)total[0]);
}

}  // namespace

size_t scan_scratch_bytes(uint64_t n) {
  uint64_t tiles = (n + kTileElems - 1) / kTileElems;
  return (tiles + 2) * sizeof(uint64_t);
}

hipError_t launch_exclusive_scan(const uint32_t* counts, uint64_t n, uint64_t* offsets,
                                 uint64_t* total, void* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t tiles = (n + kTileElems - 1) / kTileElems;
  uint64_t* tile_sums = static_cast<uint64_t*>(scratch);
  hipLaunchKernelGGL(k_tile_sums, dim3((uint32_t)tiles), dim3(kThreads), 0, stream, counts, n,
                     tile_sums);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, stream, tile_sums, tiles, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)tiles), dim3(kThreads), 0, stream, counts, n,
                     tile_sums, offsets);
  return hipGetLastError();
}

hipError_t launch_reserve(const void* dev_queue, const uint64_t* total, uint64_t* base,
                          hipStream_t stream) {
  hipLaunchKernelGGL(k_reserve, dim3(1), dim3(64), 0, stream,
                     static_cast<const rts_dev_queue*>(dev_queue), total, base);
  return hipGetLastError();
}

}  // namespace rts
