// Result handlers written the way a caller of the library writes them: this
// file includes only the public headers (include/), is compiled as HIP for
// gfx950 and links librtspatial_amd. It builds librts_user_handlers.so, whose
// C functions the tests check against the oracle and bench.py times.
//
//   Summary      per query: result count, sum of splitmix64((query << 32) |
//                geom) and smallest geom id -- one call per result pair
//   CountDone    per query: result count, from done() alone (kPairs = false)
//   the reference's counting handler functions (default_handlers/
//                counting_handlers.h:7-25): one atomicAdd per result into a
//                single counter, reached through SpatialIndex-style Query with
//                RTSPATIAL_USER_HANDLERS
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

// The reference's handler functions, in the form a reference user's handler
// header defines them; they must precede the library's headers.
extern "C" __device__ __forceinline__ void rtspatial_handle_point_contains(uint32_t, uint32_t,
                                                                          void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
extern "C" __device__ __forceinline__ void rtspatial_handle_envelope_contains(uint32_t, uint32_t,
                                                                             void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
extern "C" __device__ __forceinline__ void rtspatial_handle_envelope_intersects(uint32_t, uint32_t,
                                                                               void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
#define RTSPATIAL_USER_HANDLERS 1

#include "rtspatial/device/query.h"

namespace {

thread_local std::string g_error;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Summary {
  unsigned long long* count;
  unsigned long long* hash;
  unsigned int* min_id;
  __device__ void operator()(uint32_t geom_id, uint32_t query_id) const {
    atomicAdd(count + query_id, 1ull);
    atomicAdd(hash + query_id, (unsigned long long)splitmix64(((uint64_t)query_id << 32) | geom_id));
    atomicMin(min_id + query_id, geom_id);
  }
};

struct CountDone {
  static constexpr bool kPairs = false;
  unsigned long long* count;
  __device__ void done(uint32_t query_id, uint64_t c) const { count[query_id] = c; }
};

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

const char* rtsx_last_error(void) { return g_error.c_str(); }

// count[q] += results, hash[q] += splitmix64((q << 32) | g), min_id[q] =
// min(min_id[q], g) over query q's results. The arrays are the caller's.
int rtsx_summary(rts_index* ix, int pred, const void* queries, uint64_t n, int points,
                 int doubles, uint32_t traversal, unsigned long long* count,
                 unsigned long long* hash, unsigned int* min_id, hipStream_t stream) {
  return guarded([&] {
    rtspatial::QueryWith(ix, static_cast<rtspatial::Predicate>(pred), queries, n, points != 0,
                         doubles != 0, Summary{count, hash, min_id}, stream, traversal);
  });
}

// count[q] = query q's number of results.
int rtsx_count(rts_index* ix, int pred, const void* queries, uint64_t n, int points, int doubles,
               uint32_t traversal, unsigned long long* count, hipStream_t stream) {
  return guarded([&] {
    rtspatial::QueryWith(ix, static_cast<rtspatial::Predicate>(pred), queries, n, points != 0,
                         doubles != 0, CountDone{count}, stream, traversal);
  });
}

// *counter += all results, through the reference's handler functions.
int rtsx_reference_count(rts_index* ix, int pred, const void* queries, uint64_t n, int points,
                         int doubles, unsigned long long* counter, hipStream_t stream) {
  return guarded([&] {
    rtspatial::QueryWithReferenceHandlers(ix, static_cast<rtspatial::Predicate>(pred), queries, n,
                                          points != 0, doubles != 0, counter, stream);
  });
}

}  // extern "C"
