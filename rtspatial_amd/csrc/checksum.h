// Factored checksum of (query, geom) result pairs, for verifying result sets
// at sizes where the pairs cannot be compared one by one.
//
// A pair (query q, box b) hashes to qmult(q) * ghash(b) mod 2^64, with
//   ghash(b) = splitmix64(b ^ 0xD1B54A32D192ED03)   (splitmix64: Steele, Lea
//   qmult(q) = splitmix64(q) | 1                      and Flood, OOPSLA 2014)
// A query's result set is summarised by its count and the sum of ghash(b)
// over it; a request's by the sum over its queries of qmult(q) times that.
// The hash factors so that a checker can take a subtree that matches whole
// from a stored sum. splitmix64 is a bijection, so a missing, extra or
// substituted pair changes the sums (qmult is odd, hence invertible mod 2^64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rts {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__host__ __device__ __forceinline__ uint64_t ghash(uint64_t geom_id) {
  return splitmix64(geom_id ^ 0xD1B54A32D192ED03ull);
}

__host__ __device__ __forceinline__ uint64_t qmult(uint64_t query_id) {
  return splitmix64(query_id) | 1ull;
}

__host__ __device__ __forceinline__ uint64_t pair_hash(uint64_t query_id, uint64_t geom_id) {
  return qmult(query_id) * ghash(geom_id);
}

}  // namespace rts
