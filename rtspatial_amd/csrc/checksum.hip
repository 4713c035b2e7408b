// rts_pairs_checksum: per-query count and ghash sum of a result-pair array
// (the factored checksum of checksum.h), for checking result sets too large
// to compare pair by pair.
#include "checksum.h"
#include "rts_internal.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
// Pairs per wave-block: lanes take slots b0 + k * 64 + lane (coalesced).
constexpr uint64_t kBlock = 64 * 64;

// Adds (c, g) to query q's totals. When every lane of the wave holds the same
// query (a query's pairs are usually contiguous and many), one lane adds the
// wave's sums; otherwise each lane adds its own.
__device__ __forceinline__ void flush(uint32_t q, uint64_t c, uint64_t g, uint64_t nq,
                                      unsigned long long* counts, unsigned long long* gsums) {
  const uint32_t q0 = readlane_u32(q, 0);
  if (__all(q == q0)) {
    c = wave_sum_u64(c);
    g = wave_sum_u64(g);
    if (lane_id() == 0 && c && q0 < nq) {
      atomicAdd(&counts[q0], (unsigned long long)c);
      atomicAdd(&gsums[q0], (unsigned long long)g);
    }
  } else if (c && q < nq) {
    atomicAdd(&counts[q], (unsigned long long)c);
    atomicAdd(&gsums[q], (unsigned long long)g);
  }
}

__global__ void __launch_bounds__(kThreads) k_pair_checksums(const rts_pair* __restrict__ pairs,
                                                             uint64_t n, uint64_t nq,
                                                             unsigned long long* counts,
                                                             unsigned long long* gsums) {
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * kWaves;
  const uint64_t w0 = (uint64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  for (uint64_t b0 = w0 * kBlock; b0 < n; b0 += waves * kBlock) {
    const uint64_t b1 = b0 + kBlock < n ? b0 + kBlock : n;
    uint32_t q = 0xFFFFFFFFu;
    uint64_t c = 0, g = 0;
    for (uint64_t i = b0 + lane; i < b1; i += kWave) {
      const rts_pair p = pairs[i];
      if (p.query_id != q) {
        if (c && q < nq) {
          atomicAdd(&counts[q], (unsigned long long)c);
          atomicAdd(&gsums[q], (unsigned long long)g);
        }
        q = p.query_id;
        c = 0;
        g = 0;
      }
      c++;
      g += ghash(p.geom_id);
    }
    flush(q, c, g, nq, counts, gsums);
  }
}

}  // namespace
}  // namespace rts

using namespace rts;

extern "C" {

int rts_pairs_checksum(const rts_pair* pairs, uint64_t n, uint64_t n_queries, uint64_t* counts,
                       uint64_t* gsums, rts_stream_t stream) {
  if ((n && !pairs) || (n_queries && (!counts || !gsums))) {
    set_error("rts_pairs_checksum: null argument");
    return RTS_ERR_INVALID;
  }
  if (n_queries == 0) return RTS_OK;
  hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(counts), 0u, 2 * n_queries, stream);
  if (e == hipSuccess)
    e = launch_fill_u32(reinterpret_cast<uint32_t*>(gsums), 0u, 2 * n_queries, stream);
  if (e == hipSuccess && n) {
    const uint64_t blocks64 = (n + kBlock * kWaves - 1) / (kBlock * kWaves);
    const uint32_t blocks = (uint32_t)(blocks64 < 8192 ? blocks64 : 8192);
    hipLaunchKernelGGL(k_pair_checksums, dim3(blocks), dim3(kThreads), 0, stream, pairs, n,
                       n_queries, reinterpret_cast<unsigned long long*>(counts),
                       reinterpret_cast<unsigned long long*>(gsums));
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    set_error(std::string("rts_pairs_checksum: ") + hipGetErrorString(e));
    return RTS_ERR_HIP;
  }
  return RTS_OK;
}

}  // extern "C"
