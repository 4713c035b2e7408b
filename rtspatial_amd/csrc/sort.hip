// LDS-staged onesweep LSD radix sort of (uint32 key, uint32 value) pairs.
//
// Method: Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back" / Onesweep (2022). One histogram kernel computes every digit
// histogram in a single read of the keys; then each 8-bit digit pass is one
// kernel in which a workgroup ranks its tile in LDS, publishes its per-digit
// counts, looks back over earlier tiles for their running totals, and writes
// its keys out through LDS so that stores to each digit bucket are contiguous.
//
// CDNA specifics: wave64 multisplit ranking (eight ballots give each lane the
// mask of lanes that share its digit), tiles handed out by an atomic ticket so
// a workgroup only ever waits on tiles that are already resident, and a bounded
// wait: a look-back that spins too long sets an error word and stops, so the
// kernel always finishes.
#include "rts_internal.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kSortThreads = 512;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kKeysPerThread = 16;
// 8192 keys per workgroup: on 10M keys a sort takes 0.35 ms (0.40 ms with
// 4096-key tiles), with fewer tiles to look back over and 32 keys per digit
// and tile on average for the scatter.
constexpr int kTile = kSortThreads * kKeysPerThread;
constexpr int kRadix = 256;
constexpr int kMaxPasses = 4;
constexpr uint32_t kFlagAgg = 1u << 30;   // tile count published
constexpr uint32_t kFlagInc = 2u << 30;   // inclusive running total published
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 24;

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// All digit histograms of the keys in one pass.
__global__ void __launch_bounds__(256) k_histogram(const uint32_t* __restrict__ keys, uint32_t n,
                                                   int passes, uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[kMaxPasses][kRadix];
  for (int i = threadIdx.x; i < kMaxPasses * kRadix; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = keys[i];
    for (int p = 0; p < passes; p++) atomicAdd(&sh[p][(k >> (8 * p)) & 0xFF], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; p++) {
    uint32_t c = sh[p][threadIdx.x];
    if (c) atomicAdd(&hist[p * kRadix + threadIdx.x], c);
  }
}

// Exclusive scan of each pass's histogram: block p handles pass p.
__global__ void __launch_bounds__(256) k_scan_hist(const uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ offsets) {
  __shared__ uint32_t wave_tot[kSortWaves];
  const uint32_t t = threadIdx.x, lane = lane_id(), w = t / kWave;
  uint32_t v = hist[blockIdx.x * kRadix + t];
  uint32_t inc = wave_inclusive_sum(v);
  if (lane == kWave - 1) wave_tot[w] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t i = 0; i < w; i++) before += wave_tot[i];
  offsets[blockIdx.x * kRadix + t] = before + inc - v;
}

__global__ void __launch_bounds__(kSortThreads)
    k_onesweep(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
               uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
               int shift, const uint32_t* __restrict__ digit_offsets, uint32_t* status,
               uint32_t* ticket, uint32_t* error) {
  __shared__ uint32_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t s_wave_hist[kSortWaves][kRadix];
  __shared__ uint32_t s_tile_start[kRadix];
  __shared__ uint32_t s_global_base[kRadix];
  __shared__ uint32_t s_wave_tot[kSortWaves];
  __shared__ uint32_t s_part;

  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  if (tid == 0) s_part = atomicAdd(ticket, 1u);
  for (int i = tid; i < kSortWaves * kRadix; i += kSortThreads) (&s_wave_hist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t part = s_part;
  const uint64_t tile_base = (uint64_t)part * kTile;

  // Load: wave w owns keys [w*1024, (w+1)*1024) of the tile, iteration j
  // covers 64 consecutive keys, so (wave, j, lane) is the key order.
  uint32_t k[kKeysPerThread], v[kKeysPerThread], rank[kKeysPerThread];
#pragma unroll
  for (int j = 0; j < kKeysPerThread; j++) {
    uint64_t idx = tile_base + (uint64_t)wave * (kKeysPerThread * kWave) + j * kWave + lane;
    bool valid = idx < n;
    k[j] = valid ? keys_in[idx] : 0xFFFFFFFFu;
    v[j] = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
  }

  // Stable rank within the wave: lanes sharing a digit are found with eight
  // ballots; the last lane of each group bumps the wave's running count.
#pragma unroll
  for (int j = 0; j < kKeysPerThread; j++) {
    uint64_t idx = tile_base + (uint64_t)wave * (kKeysPerThread * kWave) + j * kWave + lane;
    bool valid = idx < n;
    uint32_t d = (k[j] >> shift) & 0xFF;
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1;
      uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t below = mask_rank(peers);
    uint32_t cnt = __popcll(peers);
    uint32_t prev = valid ? s_wave_hist[wave][d] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (valid && below == cnt - 1) s_wave_hist[wave][d] = prev + cnt;
    __builtin_amdgcn_wave_barrier();
    rank[j] = prev + below;
  }
  __syncthreads();

  // Thread d < kRadix (whole waves): tile count of digit d and the per-wave
  // exclusive offsets.
  const uint32_t d = tid;
  const bool digit_thread = tid < (uint32_t)kRadix;
  uint32_t count_d = 0, inc = 0;
  if (digit_thread) {
#pragma unroll
  for (int w = 0; w < kSortWaves; w++) {
    uint32_t c = s_wave_hist[w][d];
    s_wave_hist[w][d] = count_d;
    count_d += c;
  }
  uint32_t* st = status + (uint64_t)part * kRadix + d;
  store_relaxed_agent(st, (part == 0 ? kFlagInc : kFlagAgg) | count_d);

  // Tile-local exclusive scan over digits.
  inc = wave_inclusive_sum(count_d);
  if (lane == kWave - 1) s_wave_tot[wave] = inc;

  // Decoupled look-back for the running total of digit d before this tile.
  uint32_t excl = 0;
  if (part > 0) {
    int64_t p = (int64_t)part - 1;
    uint32_t spins = 0;
    while (p >= 0) {
      uint32_t s = load_relaxed_agent(status + (uint64_t)p * kRadix + d);
      if ((s & ~kValueMask) == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += s & kValueMask;
      if (s & kFlagInc) break;
      p--;
    }
    store_relaxed_agent(st, kFlagInc | ((excl + count_d) & kValueMask));
  }
  s_global_base[d] = digit_offsets[d] + excl;
  }
  __syncthreads();
  if (digit_thread) {
    uint32_t before = 0;
    for (uint32_t i = 0; i < wave; i++) before += s_wave_tot[i];
    s_tile_start[d] = before + inc - count_d;
  }
  __syncthreads();

  // Scatter into LDS in digit order.
#pragma unroll
  for (int j = 0; j < kKeysPerThread; j++) {
    uint64_t idx = tile_base + (uint64_t)wave * (kKeysPerThread * kWave) + j * kWave + lane;
    if (idx < n) {
      uint32_t dd = (k[j] >> shift) & 0xFF;
      uint32_t pos = s_tile_start[dd] + s_wave_hist[wave][dd] + rank[j];
      s_keys[pos] = k[j];
      s_vals[pos] = v[j];
    }
  }
  __syncthreads();

  // Write out: consecutive LDS entries of one digit go to consecutive slots.
  const uint64_t rem = (uint64_t)n - tile_base;
  const uint32_t tile_n = rem < (uint64_t)kTile ? (uint32_t)rem : (uint32_t)kTile;
  for (uint32_t i = tid; i < tile_n; i += kSortThreads) {
    uint32_t key = s_keys[i];
    uint32_t dd = (key >> shift) & 0xFF;
    uint32_t dst = s_global_base[dd] + (i - s_tile_start[dd]);
    if (dst < n) {  // holds whenever the look-back completed
      keys_out[dst] = key;
      vals_out[dst] = s_vals[i];
    }
  }
}

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, stream, p, value, n);
  return hipGetLastError();
}

static uint32_t tiles_for(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }

// Layout: [hist 4x256 | tickets 4 | error 1 | status 4 x tiles x 256] zeroed
// together, then offsets 4x256.
size_t sort_scratch_bytes(uint64_t n) {
  uint64_t tiles = tiles_for(n ? n : 1);
  size_t zeroed = (kMaxPasses * kRadix + kMaxPasses + 1 + kMaxPasses * tiles * kRadix) * 4;
  return align256(zeroed) + align256(kMaxPasses * kRadix * 4);
}

void sort_scratch_carve(void* base, uint64_t n, SortScratch* s) {
  uint64_t tiles = tiles_for(n ? n : 1);
  uint32_t* p = static_cast<uint32_t*>(base);
  s->hist = p;
  s->tickets = p + kMaxPasses * kRadix;
  s->error = s->tickets + kMaxPasses;
  s->status = s->error + 1;
  size_t zeroed = (kMaxPasses * kRadix + kMaxPasses + 1 + kMaxPasses * tiles * kRadix) * 4;
  s->offsets = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + align256(zeroed));
  s->tiles_cap = (uint32_t)tiles;
}

hipError_t radix_sort_pairs(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b, uint32_t* vals_b,
                            uint32_t n, int key_bits, bool identity_vals, const SortScratch& s,
                            hipStream_t stream, bool* in_b) {
  *in_b = false;
  if (n == 0) return hipSuccess;
  const int passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  if (passes > kMaxPasses) return hipErrorInvalidValue;
  const uint32_t tiles = tiles_for(n);
  if (tiles > s.tiles_cap) return hipErrorInvalidValue;
  // Zero histograms, tickets, error word and every pass's status array.
  const uint64_t zero_words = kMaxPasses * kRadix + kMaxPasses + 1 + (uint64_t)passes * tiles * kRadix;
  hipError_t e = launch_fill_u32(s.hist, 0u, zero_words, stream);
  if (e != hipSuccess) return e;
  uint32_t hist_blocks = tiles < 1024u ? tiles : 1024u;
  hipLaunchKernelGGL(k_histogram, dim3(hist_blocks), dim3(256), 0, stream, keys_a, n, passes,
                     s.hist);
  hipLaunchKernelGGL(k_scan_hist, dim3(passes), dim3(256), 0, stream, s.hist, s.offsets);
  uint32_t *ki = keys_a, *vi = vals_a, *ko = keys_b, *vo = vals_b;
  for (int p = 0; p < passes; p++) {
    // Pass 0 of an identity sort generates the values instead of reading them.
    const uint32_t* vin = (p == 0 && identity_vals) ? nullptr : vi;
    hipLaunchKernelGGL(k_onesweep, dim3(tiles), dim3(kSortThreads), 0, stream, ki, vin, ko, vo, n,
                       8 * p, s.offsets + p * kRadix, s.status + (uint64_t)p * tiles * kRadix,
                       s.tickets + p, s.error);
    uint32_t* tk = ki;
    uint32_t* tv = vi;
    ki = ko;
    vi = vo;
    ko = tk;
    vo = tv;
  }
  *in_b = (passes % 2) == 1;
  return hipGetLastError();
}

}  // namespace rts
