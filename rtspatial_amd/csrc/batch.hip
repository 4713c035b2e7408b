// Persistent work-stealing kernels that answer many uneven query requests in
// one launch.
//
// A batch is a list of requests (queries pointer, count, point/envelope,
// predicate, tile mode) laid out in one query space: request r owns global
// queries [qbase_r, qbase_r + n_r) and tiles [tbase_r, tbase_r + tiles_r).
// A tile is either 64 queries of one request answered one per lane (lane
// tiles: points and data-contains-query, few results each) or one query
// answered by a whole wavefront (wave tiles: intersects / within). Requests
// are listed class by class (batch.h), and one persistent kernel per pass
// takes the tiles of every class: per class, each XCD first drains its own
// contiguous eighth of the class's tiles, which keeps neighbouring queries on
// one L2, and then steals from the other XCDs.
//
// Two passes: the count pass writes a bound on each query's results (exact
// for lane tiles; a wave tile's count walk counts the leaf ranges it would
// have to test as a whole and logs them), a scan turns the bounds into
// exclusive offsets into the slot's result arena, and the write pass tests
// the logged ranges, writes each query's pairs at the start of its slots
// and records its exact count. Per request, the batch reports its first
// slot, its slots and its results. Every launch reads the batch's sizes from
// device memory.
#include <algorithm>

#include "batch.h"
#include "checksum.h"
#include "morton.h"
#include "range_log.h"
#include "rtspatial/device/traverse.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;

__device__ __forceinline__ uint32_t find_request(const BatchRequest* reqs, uint32_t nreq,
                                                 uint32_t tile) {
  uint32_t lo = 0, hi = nreq - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (reqs[mid].tbase <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Tile slices: class c's tiles [lo_c, hi_c) are cut into kXcds contiguous
// slices, one counter each. Wave w on XCD x first drains its own XCD's slices
// class by class (heaviest first: neighbouring Morton-ordered queries then
// share that XCD's L2), then steals, heaviest class first, from the other
// XCDs. Counters only grow, so a slice found drained stays drained: `pos`
// (wave-uniform) walks the 5 x 8 slice positions of this wave's order once.
constexpr uint32_t kSlicePositions = kClasses * kXcds;

__device__ __forceinline__ uint32_t grab_tile(uint32_t* counters, const uint32_t* class_end,
                                              uint32_t xcd, uint32_t& pos) {
  uint32_t got = 0xFFFFFFFFu, at = kSlicePositions;
  if (lane_id() == 0) {
    for (uint32_t p = pos; p < kSlicePositions; p++) {
      uint32_t c, x;
      if (p < kClasses) {
        c = p;
        x = xcd;
      } else {
        c = (p - kClasses) / (kXcds - 1);
        x = (xcd + 1 + (p - kClasses) % (kXcds - 1)) % kXcds;
      }
      const uint32_t lo = c == 0 ? 0u : class_end[c - 1];
      const uint32_t n = class_end[c] - lo;
      const uint32_t per = (n + kXcds - 1) / kXcds;
      const uint32_t begin = x * per;
      const uint32_t end = begin + per < n ? begin + per : n;
      if (begin >= end) continue;
      const uint32_t v = atomicAdd(&counters[(c * kXcds + x) * kCounterStride], 1u);
      if (v < end - begin) {
        got = lo + begin + v;
        at = p;
        break;
      }
    }
  }
  pos = readlane_u32(at, 0);
  return readlane_u32(got, 0);
}

// One tile of the batch: a wave tile's one query (by the whole wave) or a
// lane tile's 64 queries. With skip_giants a wave tile whose query is on a
// giant list is left to that list.
template <bool WRITE>
__device__ __forceinline__ void run_tile(const TreeView& t0, TreeView& t, const BatchArgs& a,
                                         const BatchHeader& h, uint32_t tile, bool skip_giants,
                                         LogCursor& lc, WaveLds& lds) {
  const uint32_t ri = find_request(a.requests, h.n_requests, tile);
  const BatchRequest r = a.requests[ri];
  const uint32_t local = tile - r.tbase;
  const bool wave = r.flags & kReqWaveTiles;
  if (WRITE && skip_giants && wave && local < r.n &&
      a.counts[r.qbase + local] >= giant_min(kGiantBuckets - 1))
    return;
  // A request whose slots run past the arena is launched again: its pairs
  // would be dropped, so its walks are skipped (a one-query request still
  // runs, for its exact count).
  if (WRITE && r.n > 1 && a.ranges[3 * ri] + a.ranges[3 * ri + 1] > a.arena_cap) {
    if (wave) {
      if (local < r.n && lane_id() == 0) a.exact[r.qbase + local] = 0;
    } else {
      const uint32_t q = local * kWave + lane_id();
      if (q < r.n) a.exact[r.qbase + q] = 0;
    }
    return;
  }
  if (t.n == 0) {
    if (wave) {
      if (local < r.n && lane_id() == 0) a.counts[r.qbase + local] = a.exact[r.qbase + local] = 0;
    } else {
      const uint32_t q = local * kWave + lane_id();
      if (q < r.n) a.counts[r.qbase + q] = a.exact[r.qbase + q] = 0;
    }
    return;
  }
  const uint32_t lb = (r.flags >> kReqLeafShift) & 7u;
  t.leaf_size = lb ? 1u << (lb - 1) : t0.leaf_size;
  const bool points = r.flags & kReqPoints;
  const int pred = points ? (int)RTS_CONTAINS : (int)((r.flags >> kReqPredShift) & 3u);
  if (wave) {
    if (local >= r.n) return;
    const QueryRT qb{load_query(r.queries, points, local), pred};
    const uint64_t g = r.qbase + local;
    if (WRITE) {
      RangeLog lg = log_for_write(a.log, g);
      const uint64_t e = wave_query<0, false, true, true>(
          t, qb, local + r.qid_base, QueueEmit{a.arena, a.arena_cap, a.offsets[g]}, lds, &lg);
      if (lane_id() == 0) a.exact[g] = (uint32_t)e;
    } else {
      // The bound: the write pass tests the logged ranges' leaves.
      RangeLog lg = log_for_count(lc);
      const uint64_t c =
          wave_query<0, false, false, true, true>(t, qb, local, NoEmit{}, lds, &lg);
      log_commit(lc, a.log, g, lg);
      if (lane_id() == 0) {
        a.counts[g] = (uint32_t)c;
        // Giants are listed as they are counted (rare: one atomic each).
        if (a.giants && c >= giant_min(kGiantBuckets - 1)) {
          int b = kGiantBuckets - 1;
          while (b > 0 && c >= giant_min(b - 1)) b--;
          const uint32_t at =
              atomicAdd(&a.tile_counters[kGiantCounterOff + 2 * b * kCounterStride], 1u);
          if (at < a.giant_cap) a.giants[(uint64_t)b * a.giant_cap + at] = tile;
        }
      }
    }
  } else {
    const uint32_t q = local * kWave + lane_id();
    if (q >= r.n) return;
    const QueryRT qb{load_query(r.queries, points, q), pred};
    const uint64_t g = r.qbase + q;
    if (WRITE) {
      lane_query<0, false, true>(t, qb, q + r.qid_base,
                                 QueueEmit{a.arena, a.arena_cap, a.offsets[g]});
    } else {
      a.counts[g] = a.exact[g] = (uint32_t)lane_query<0, false, false>(t, qb, q, NoEmit{});
    }
  }
}

// One kernel per pass answers every tile of the batch. A request's predicate
// is a run-time value of the walks (QueryRT), so the kernel holds one wave
// walk and one lane walk. The tree view is read from device memory so that
// the launch does not depend on the index's state. One wavefront per
// workgroup: a workgroup's slot is freed as soon as its one wave runs out of
// tiles (four-wave workgroups held theirs until the slowest wave ended). The
// write pass first takes the giant tiles, largest bucket first.
// The write pass is pinned at 6 waves per SIMD (80 VGPRs), which the range
// logs would otherwise push just past (81); the count pass, which tests no
// leaves of wave tiles, needs 58 VGPRs and runs 8 (at least 7).
template <bool WRITE>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WRITE ? 6 : 7)))
k_batch(const TreeView* tv, const BatchArgs a) {
  __shared__ WaveLds s_lds;
  const TreeView t0 = *tv;
  TreeView t = t0;  // leaf_size may differ per request
  const BatchHeader h = *a.header;
  if (h.n_tiles == 0 || h.n_requests == 0) return;
  // An empty index (no boxes inserted, or cleared) has no tree: the count
  // pass still writes every query's count (zero), the write pass has nothing
  // to write.
  if (WRITE && t.n == 0) return;
  uint32_t* counters = a.tile_counters + (WRITE ? kClasses * kXcds * kCounterStride : 0u);
  const uint32_t xcd = blockIdx.x % kXcds;
  const uint32_t wave_end = h.class_end[kClassWavePoints];
  // The giant lists are used only when every one of them held its tiles
  // (a batch never has more wave tiles than the lists' capacity).
  uint32_t* gc = a.tile_counters + kGiantCounterOff;
  bool giants = false;
  if (WRITE && a.giants) {
    giants = true;
    for (int b = 0; b < kGiantBuckets; b++) giants &= gc[2 * b * kCounterStride] <= a.giant_cap;
  }
  int bucket = giants ? 0 : kGiantBuckets;  // wave-uniform: list being drained
  LogCursor lc = WRITE ? LogCursor{} : log_cursor(a.log);
  uint32_t pos = 0;
  // Every iteration takes a tile, from a list or the queue, or finds none
  // left; listed tiles are wave tiles, so 2 * n_tiles + 1 iterations bound
  // the loop.
  for (uint64_t it = 0; it <= 2ull * h.n_tiles; it++) {
    uint32_t tile = 0xFFFFFFFFu;
    for (; bucket < kGiantBuckets; bucket++) {
      uint32_t k = 0;
      if (lane_id() == 0) k = atomicAdd(&gc[(2 * bucket + 1) * kCounterStride], 1u);
      k = readlane_u32(k, 0);
      if (k < gc[2 * bucket * kCounterStride]) {
        tile = a.giants[(uint64_t)bucket * a.giant_cap + k];
        break;
      }
    }
    const bool listed = tile != 0xFFFFFFFFu;
    if (!listed) {
      // class_end from device memory: indexing the local copy by a run-time
      // class would put that copy in private memory.
      tile = grab_tile(counters, a.header->class_end, xcd, pos);
      if (tile >= h.n_tiles) break;
      // Wave tiles are Morton-ordered within their class (the sort keys
      // lead with the class).
      if (a.order && tile < wave_end) tile = a.order[tile];
    }
    run_tile<WRITE>(t0, t, a, h, tile, giants && !listed, lc, s_lds);
  }
}

// ---- scan of the per-query counts over the batch's query space, with a
// fixed grid: block b owns the contiguous segment [b*S, (b+1)*S).
constexpr uint32_t kScanBlocks = 1024;

__device__ __forceinline__ uint64_t block_scan_excl(uint64_t v, uint64_t* sh, uint64_t* total) {
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += o;
  }
  if (lane == kWave - 1) sh[w] = inc;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int i = 0; i < kWaves; i++) {
    if (i < (int)w) before += sh[i];
    all += sh[i];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ void __launch_bounds__(kThreads) k_seg_sums(const BatchArgs a) {
  __shared__ uint64_t sh[kWaves];
  const uint64_t n = a.header->n_queries;
  const uint64_t seg = (n + kScanBlocks - 1) / kScanBlocks;
  const uint64_t b0 = blockIdx.x * seg, b1 = b0 + seg < n ? b0 + seg : n;
  uint64_t s = 0;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += kThreads) s += a.counts[i];
  uint64_t all;
  block_scan_excl(s, sh, &all);
  if (threadIdx.x == 0) a.seg_sums[blockIdx.x] = all;
}

__global__ void __launch_bounds__(kThreads) k_seg_scan(const BatchArgs a) {
  __shared__ uint64_t sh[kWaves];
  uint64_t carry = 0;
  for (uint32_t c = 0; c < kScanBlocks; c += kThreads) {
    const uint32_t i = c + threadIdx.x;
    const uint64_t v = a.seg_sums[i];
    uint64_t all;
    const uint64_t ex = block_scan_excl(v, sh, &all);
    a.seg_sums[i] = carry + ex;
    carry += all;
  }
  if (threadIdx.x == 0) a.total[0] = carry;
}

__global__ void __launch_bounds__(kThreads) k_seg_apply(const BatchArgs a) {
  __shared__ uint64_t sh[kWaves];
  const uint64_t n = a.header->n_queries;
  const uint64_t seg = (n + kScanBlocks - 1) / kScanBlocks;
  const uint64_t b0 = blockIdx.x * seg, b1 = b0 + seg < n ? b0 + seg : n;
  uint64_t carry = a.seg_sums[blockIdx.x];
  for (uint64_t c = b0; c < b1; c += kThreads) {
    const uint64_t i = c + threadIdx.x;
    const uint64_t v = i < b1 ? a.counts[i] : 0;
    uint64_t all;
    const uint64_t ex = block_scan_excl(v, sh, &all);
    if (i < b1) a.offsets[i] = carry + ex;
    carry += all;
  }
}

// Per request: first arena slot and slots (the sum of its bounds).
__global__ void __launch_bounds__(kThreads) k_request_ranges(const BatchArgs a) {
  const BatchHeader h = *a.header;
  for (uint32_t r = blockIdx.x * kThreads + threadIdx.x; r < h.n_requests;
       r += gridDim.x * kThreads) {
    const BatchRequest q = a.requests[r];
    const uint64_t start = q.n ? a.offsets[q.qbase] : 0;
    const uint64_t end = (q.qbase + q.n < h.n_queries) ? a.offsets[q.qbase + q.n] : a.total[0];
    a.ranges[3 * r] = start;
    a.ranges[3 * r + 1] = q.n ? end - start : 0;
    a.ranges[3 * r + 2] = 0;  // results: summed by k_request_exact
  }
}

// Request of global query q: the last request whose qbase <= q (requests are
// laid out in query order).
__device__ __forceinline__ uint32_t request_of_query(const BatchRequest* reqs, uint32_t nreq,
                                                     uint64_t q) {
  uint32_t lo = 0, hi = nreq - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (reqs[mid].qbase <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// After the write pass: each request's results, the sum of its queries'
// exact counts, added into ranges[3r + 2]. A wavefront sums a block of
// kExactBlock counts: with one atomic when the block lies in one request
// (requests are contiguous in query space), else lane by lane, each lane
// over its own run of the block, one atomic per request it meets. (One
// atomic per query from the write pass, on one address per request, made
// that pass twice as slow.)
constexpr uint32_t kExactPerLane = 16;
constexpr uint32_t kExactBlock = kWave * kExactPerLane;

__global__ void __launch_bounds__(kThreads) k_request_exact(const BatchArgs a) {
  const BatchHeader h = *a.header;
  if (h.n_requests == 0) return;
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * kWaves;
  for (uint64_t w = (uint64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
       w * kExactBlock < h.n_queries; w += waves) {
    const uint64_t b0 = w * kExactBlock;
    const uint64_t b1 = b0 + kExactBlock < h.n_queries ? b0 + kExactBlock : h.n_queries;
    const uint32_t r0 = request_of_query(a.requests, h.n_requests, b0);
    const uint32_t r1 = request_of_query(a.requests, h.n_requests, b1 - 1);
    if (r0 == r1) {
      uint64_t s = 0;
      for (uint64_t i = b0 + lane; i < b1; i += kWave) s += a.exact[i];
      s = wave_sum_u64(s);
      if (lane == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(&a.ranges[3 * r0 + 2]), s);
    } else {
      uint64_t i = b0 + (uint64_t)lane * kExactPerLane;
      const uint64_t e = i + kExactPerLane < b1 ? i + kExactPerLane : b1;
      if (i < e) {
        uint32_t r = request_of_query(a.requests, h.n_requests, i);
        uint64_t r_end = r + 1 < h.n_requests ? a.requests[r + 1].qbase : h.n_queries;
        uint64_t s = 0;
        for (; i < e; i++) {
          while (i >= r_end) {
            if (s) atomicAdd(reinterpret_cast<unsigned long long*>(&a.ranges[3 * r + 2]), s);
            s = 0;
            r++;
            r_end = r + 1 < h.n_requests ? a.requests[r + 1].qbase : h.n_queries;
          }
          s += a.exact[i];
        }
        if (s) atomicAdd(reinterpret_cast<unsigned long long*>(&a.ranges[3 * r + 2]), s);
      }
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_copy_u32(const uint32_t* __restrict__ src,
                                                       uint32_t* __restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    dst[i] = src[i];
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

// Resident workgroups of a persistent kernel over the device (occupancy).
template <typename K>
uint32_t persistent_blocks(K kernel, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
  if (cus <= 0) cus = 256;
  if (per_cu <= 0) per_cu = 1;
  const uint32_t blocks = (uint32_t)(cus * per_cu);
  return (blocks + kXcds - 1) / kXcds * kXcds;
}

template <bool WRITE>
void launch_batch_kernel(const TreeView* t, const BatchArgs& a_in, hipStream_t stream) {
  static const uint32_t blocks = persistent_blocks(k_batch<WRITE>, kWave);
  BatchArgs a = a_in;
  a.log.per_wave = (uint32_t)(a.log.n_entries / blocks);
  hipLaunchKernelGGL((k_batch<WRITE>), dim3(blocks), dim3(kWave), 0, stream, t, a);
}

// Sort keys of the wave tiles [0, n): class in the top 2 bits (the wave
// classes are 0..3), then the 30-bit Morton code of the query's centre.
__global__ void __launch_bounds__(kThreads) k_batch_keys(const BatchArgs a, const float4* scene,
                                                         uint32_t n, uint32_t* keys) {
  const BatchHeader h = *a.header;
  const float4 sc = scene[0];
  for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
    const uint32_t ri = find_request(a.requests, h.n_requests, t);
    const BatchRequest r = a.requests[ri];
    const uint32_t local = t - r.tbase;
    const bool points = r.flags & kReqPoints;
    const float4 q = local < r.n ? load_query(r.queries, points, local) : make_float4(0, 0, 0, 0);
    const uint32_t cls = (uint32_t)tile_class(r.flags) & 3u;
    keys[t] = (cls << kMortonBits) | morton_of(0.5f * (q.x + q.z), 0.5f * (q.y + q.w), sc);
  }
}

// Scan of `counts` into `offsets`, total in total[0] (fixed grid).
void launch_seg_scan(const BatchArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(k_seg_sums, dim3(kScanBlocks), dim3(kThreads), 0, stream, a);
  hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(kThreads), 0, stream, a);
  hipLaunchKernelGGL(k_seg_apply, dim3(kScanBlocks), dim3(kThreads), 0, stream, a);
}

hipError_t launch_batch(const TreeView* t, const BatchArgs& a_in, const BatchSort& sort,
                        const BatchExt* ext, size_t n_ext, const BatchExtArgs& ea, bool collect,
                        hipStream_t stream) {
  BatchArgs a = a_in;
  a.order = nullptr;
  if (sort.n > 1) {
    const uint32_t blocks = (sort.n + kThreads - 1) / kThreads < 4096 ? (sort.n + kThreads - 1) / kThreads
                                                                       : 4096;
    hipLaunchKernelGGL(k_batch_keys, dim3(blocks), dim3(kThreads), 0, stream, a, sort.scene, sort.n,
                       sort.keys_a);
    bool in_b = false;
    const hipError_t e = radix_sort_pairs(sort.keys_a, sort.vals_a, sort.keys_b, sort.vals_b, sort.n,
                                          32, true, sort.ss, stream, &in_b);
    if (e != hipSuccess) return e;
    a.order = in_b ? sort.vals_b : sort.vals_a;
  }
  // Phase events are only recorded outside stream capture (profiling off
  // or graphs off); a captured sequence is timed as a whole by its caller.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(stream, &cap);
  const bool timed = cap == hipStreamCaptureStatusNone;
  PhaseToken ph = timed ? phase_begin(RTS_PHASE_BATCH_COUNT, stream) : PhaseToken{-1, nullptr};
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, a.tile_counters,
                     kTileCounterWords);
  launch_batch_kernel<false>(t, a, stream);
  // Requests the batch kernel does not walk: one count launch each, into the
  // same counts (so the scan below covers them). Their counts are exact.
  for (size_t i = 0; i < n_ext; i++) {
    const BatchExt& x = ext[i];
    hipError_t e;
    if (x.none) {
      e = launch_fill_u32(a.counts + x.qbase, 0u, x.ql.n, stream);
    } else {
      const WaveArgs wa{a.counts + x.qbase, a.offsets + x.qbase, ea.zero, ea.tickets, {}};
      e = launch_query_wave(x.ql, wa, kPhaseCount, stream);
    }
    if (e != hipSuccess) return e;
    if (x.ql.n) {
      const uint32_t blocks = (uint32_t)std::min<uint64_t>((x.ql.n + kThreads - 1) / kThreads, 1024);
      hipLaunchKernelGGL(k_copy_u32, dim3(blocks), dim3(kThreads), 0, stream, a.counts + x.qbase,
                         a.exact + x.qbase, (uint64_t)x.ql.n);
    }
  }
  launch_seg_scan(a, stream);
  hipLaunchKernelGGL(k_request_ranges, dim3(16), dim3(kThreads), 0, stream, a);
  phase_end(ph, stream);
  ph = timed ? phase_begin(RTS_PHASE_BATCH_WRITE, stream) : PhaseToken{-1, nullptr};
  launch_batch_kernel<true>(t, a, stream);
  for (size_t i = 0; i < n_ext; i++) {
    const BatchExt& x = ext[i];
    if (x.none) continue;
    QueryLaunch ql = x.ql;
    ql.sink.mode = RTS_HANDLER_COLLECT;
    ql.sink.arg = ea.arena_queue;
    const WaveArgs wa{a.counts + x.qbase, a.offsets + x.qbase, ea.zero, ea.tickets, {}};
    const hipError_t e = launch_query_wave(ql, wa, kPhaseWrite, stream);
    if (e != hipSuccess) return e;
  }
  phase_end(ph, stream);
  // Per request results, and the compacted offsets for requests that copy
  // their pairs out.
  hipLaunchKernelGGL(k_request_exact, dim3(1024), dim3(kThreads), 0, stream, a);
  if (collect) {
    BatchArgs x = a;
    x.counts = a.exact;
    x.offsets = a.exact_offsets;
    x.total = a.total + 2;
    launch_seg_scan(x, stream);
  }
  return hipGetLastError();
}

size_t batch_seg_sums_words() { return kScanBlocks; }

namespace {

// hashes[r] += pair_hash over request r's pairs. The queries' slot ranges
// [offsets[q], + counts[q]) tile the arena's [0, total), and query q's pairs
// are the first exact[q] slots of its range. A wavefront takes blocks of
// kHashBlock consecutive slots (lanes interleaved, so loads coalesce); each
// lane finds the query of its first slot, then follows the ranges forward,
// keeping a running sum that it adds to hashes[] when its request changes.
// Where the whole wave ends a block in one request, one wave-reduced atomic
// suffices.
constexpr uint32_t kHashBlock = kWave * 256;

__device__ __forceinline__ uint64_t query_of_slot(const uint64_t* offsets, uint64_t nq,
                                                  uint64_t i) {
  uint64_t lo = 0, hi = nq - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (offsets[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kThreads) k_hash_queries(const BatchArgs a,
                                                           unsigned long long* __restrict__ hashes) {
  const BatchHeader h = *a.header;
  const uint64_t nq = h.n_queries;
  const uint64_t hi = a.total[0] < a.arena_cap ? a.total[0] : a.arena_cap;
  if (h.n_requests == 0 || nq == 0 || hi == 0) return;
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * kWaves;
  const uint64_t w0 = (uint64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  for (uint64_t b0 = w0 * kHashBlock; b0 < hi; b0 += waves * kHashBlock) {
    const uint64_t b1 = b0 + kHashBlock < hi ? b0 + kHashBlock : hi;
    uint64_t i = b0 + lane;
    uint64_t q = query_of_slot(a.offsets, nq, i < b1 ? i : b0);
    uint64_t q_end = a.offsets[q] + a.counts[q];
    uint64_t v_end = a.offsets[q] + a.exact[q];
    uint32_t r = request_of_query(a.requests, h.n_requests, q);
    uint64_t r_end = r + 1 < h.n_requests ? a.requests[r + 1].qbase : nq;
    uint64_t hs = 0;
    for (; i < b1; i += kWave) {
      while (i >= q_end && q + 1 < nq) {
        q++;
        q_end = a.offsets[q] + a.counts[q];
        v_end = a.offsets[q] + a.exact[q];
        while (q >= r_end) {
          if (hs) atomicAdd(&hashes[r], (unsigned long long)hs);
          hs = 0;
          r++;
          r_end = r + 1 < h.n_requests ? a.requests[r + 1].qbase : nq;
        }
      }
      if (i < v_end && i >= a.offsets[q]) {
        const rts_pair p = a.arena[i];
        hs += pair_hash(p.query_id, p.geom_id);
      }
    }
    const uint32_t r0 = readlane_u32(r, 0);
    if (__all(r == r0)) {
      hs = wave_sum_u64(hs);
      if (lane == 0 && hs) atomicAdd(&hashes[r0], (unsigned long long)hs);
    } else if (hs) {
      atomicAdd(&hashes[r], (unsigned long long)hs);
    }
  }
}

// Pairs of queries [q0, q0 + n) to out[0, take), in query order: one
// wavefront per query.
__global__ void __launch_bounds__(kThreads) k_copy_queries(const BatchArgs a, uint64_t q0,
                                                           uint64_t n, rts_pair* __restrict__ out,
                                                           uint64_t take) {
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * kWaves;
  const uint64_t base = a.exact_offsets[q0];
  for (uint64_t j = (uint64_t)blockIdx.x * kWaves + threadIdx.x / kWave; j < n; j += waves) {
    const uint64_t q = q0 + j;
    const uint64_t d = a.exact_offsets[q] - base;
    if (d >= take) continue;
    const uint64_t e = a.exact[q];
    const uint64_t m = d + e < take ? e : take - d;
    const uint64_t o = a.offsets[q];
    for (uint64_t i = lane; i < m && o + i < a.arena_cap; i += kWave) out[d + i] = a.arena[o + i];
  }
}

// Pairs in arena[span[0], span[1]) get query_id += add; hash[0] += their hash.
__global__ void __launch_bounds__(kThreads) k_rebase_hash_span(rts_pair* __restrict__ arena,
                                                               uint64_t cap,
                                                               const uint64_t* __restrict__ span,
                                                               uint32_t add, int do_hash,
                                                               unsigned long long* __restrict__ hash) {
  const uint64_t lo = span[0], hi = span[1] < cap ? span[1] : cap;
  uint64_t h = 0;
  for (uint64_t i = lo + blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < hi;
       i += (uint64_t)gridDim.x * kThreads) {
    rts_pair p = arena[i];
    p.query_id += add;
    if (add) arena[i] = p;
    if (do_hash) h += pair_hash(p.query_id, p.geom_id);
  }
  if (do_hash) {
    h = wave_sum_u64(h);
    if (lane_id() == 0 && h) atomicAdd(hash, (unsigned long long)h);
  }
}

}  // namespace

hipError_t launch_hash_queries(const BatchArgs& a, unsigned long long* hashes,
                               hipStream_t stream) {
  hipLaunchKernelGGL(k_hash_queries, dim3(4096), dim3(kThreads), 0, stream, a, hashes);
  return hipGetLastError();
}

hipError_t launch_copy_queries(const BatchArgs& a, uint64_t q0, uint64_t n, rts_pair* out,
                               uint64_t take, hipStream_t stream) {
  if (n == 0 || take == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + kWaves - 1) / kWaves, 4096);
  hipLaunchKernelGGL(k_copy_queries, dim3((uint32_t)blocks), dim3(kThreads), 0, stream, a, q0, n,
                     out, take);
  return hipGetLastError();
}

hipError_t launch_rebase_hash_span(rts_pair* arena, uint64_t cap, const uint64_t* span,
                                   uint32_t add, int do_hash, unsigned long long* hash,
                                   hipStream_t stream) {
  hipLaunchKernelGGL(k_rebase_hash_span, dim3(1024), dim3(kThreads), 0, stream, arena, cap, span,
                     add, do_hash, hash);
  return hipGetLastError();
}

}  // namespace rts
