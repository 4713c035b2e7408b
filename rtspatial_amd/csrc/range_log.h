// Range logs of two-pass wave walks (traverse.h RangeLog), laid out per
// wavefront: each persistent wave of a count pass appends the entries of the
// queries it walks to its own region of the log (per_wave entries, one wave
// per workgroup), and each query records where its entries
// start and how many there are. The write pass replays a query's entries
// instead of walking the tree. A query whose entries do not fit -- more than
// kLogQueryMax, or past the end of its wave's region -- is walked again.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rts_internal.h"
#include "rtspatial/device/traverse.h"

namespace rts {

constexpr uint32_t kLogQueryMax = 4096;

// A count-pass wave's place in its region.
struct LogCursor {
  uint2* region = nullptr;
  uint32_t base = 0;  // index of region[0] in the log
  uint32_t room = 0;  // entries in the region
  uint32_t fill = 0;  // entries used (wave-uniform)
};

__device__ __forceinline__ LogCursor log_cursor(const RangeLogArgs& la) {
  LogCursor c;
  if (la.entries) {
    c.base = blockIdx.x * la.per_wave;
    c.region = la.entries + c.base;
    c.room = la.per_wave;
  }
  return c;
}

// The log of the next counting walk.
__device__ __forceinline__ RangeLog log_for_count(const LogCursor& c) {
  RangeLog lg;
  if (c.region) {
    lg.entries = c.region + c.fill;
    const uint32_t left = c.room - c.fill;
    lg.cap = left < kLogQueryMax ? left : kLogQueryMax;
  }
  return lg;
}

// After the counting walk of query g: its header, and the cursor advanced
// past its entries when they fit.
__device__ __forceinline__ void log_commit(LogCursor& c, const RangeLogArgs& la, uint64_t g,
                                           const RangeLog& lg) {
  if (!c.region) return;
  const bool fit = lg.logged <= lg.cap;
  if (lane_id() == 0) {
    la.n[g] = fit ? lg.logged : kLogOverflow;
    la.off[g] = c.base + c.fill;
  }
  if (fit) c.fill += lg.logged;
}

// The log a writing walk of query g replays (n_in = kLogOverflow: none).
__device__ __forceinline__ RangeLog log_for_write(const RangeLogArgs& la, uint64_t g) {
  RangeLog lg;
  if (la.entries) {
    // Wave-uniform values, kept in scalar registers.
    lg.n_in = readlane_u32(la.n[g], 0);
    lg.cap = kLogQueryMax;
    lg.entries = la.entries + readlane_u32(la.off[g], 0);
  }
  return lg;
}

}  // namespace rts
