// Batch descriptors shared by the batched kernels (batch.hip) and the host
// engine that builds and launches batches (server.hip).
#pragma once

#include "rts_internal.h"

namespace rts {

// BatchRequest.flags
constexpr uint32_t kReqPoints = 1u;       // queries are points (float2)
constexpr uint32_t kReqPredShift = 1;     // bits 1-2: rts_predicate
constexpr uint32_t kReqWaveTiles = 8u;    // one query per wavefront (else 64 per tile)
constexpr uint32_t kReqLeafShift = 4;     // bits 4-6: log2(leaf_size) + 1, 0 = the tree's
constexpr uint32_t kReqExt = 128u;        // walked by per-request launches (BatchExt), no tiles

struct BatchRequest {
  const void* queries;  // device (or pinned host) queries of the request
  uint32_t n;           // queries
  uint32_t flags;
  uint64_t qbase;       // first global query index of the request in the batch
  uint32_t tbase;       // first tile of the request
  uint32_t qid_base;    // added to query indices when reporting pairs
};
static_assert(sizeof(BatchRequest) == 32, "32-byte request descriptor");

// Tile classes. A batch lists its requests class by class, so each class
// owns a contiguous range of tiles: the heavy wave classes first, the light
// ones (points, lane tiles) last, where they fill the tail of the pass. One
// persistent kernel per pass takes tiles of every class (the predicate is a
// run-time, wave-uniform value of the walk), from per-class per-XCD counters.
constexpr int kClassWaveIntersects = 0;
constexpr int kClassWaveWithin = 1;
constexpr int kClassWaveContains = 2;
constexpr int kClassWavePoints = 3;
constexpr int kClassLane = 4;  // 64 queries per tile, one per lane, any predicate
constexpr int kClasses = 5;

// Requests the batch kernels do not walk (segments, double queries) come
// after every tile class; they own no tiles.
constexpr int kClassExt = kClasses;

__host__ __device__ inline int tile_class(uint32_t flags) {
  if (flags & kReqExt) return kClassExt;
  if (!(flags & kReqWaveTiles)) return kClassLane;
  if (flags & kReqPoints) return kClassWavePoints;
  switch ((flags >> kReqPredShift) & 3) {
    case RTS_INTERSECTS: return kClassWaveIntersects;
    case RTS_WITHIN: return kClassWaveWithin;
    default: return kClassWaveContains;
  }
}

// Longest first: the count pass lists each wave tile whose query has at
// least giant_min(b) results in bucket b (bucket 0 the largest), and the
// write pass takes the listed tiles, bucket by bucket, before every other
// tile, so that a query of millions of results does not start late and hold
// one wave past the end of the pass (one wave writes a few hundred million
// pairs per second).
constexpr int kGiantBuckets = 3;
__host__ __device__ constexpr uint32_t giant_min(int b) { return 1u << (20 - 2 * b); }

// Tile counters of a batch, each on its own 64-byte line, zeroed by the
// batch's first launch: for each pass (count, write), class and XCD one tile
// counter, then per giant bucket an append and a take counter.
constexpr uint32_t kCounterStride = 16;
constexpr uint32_t kTileCounterWords = (2 * kClasses * 8 + 2 * kGiantBuckets) * kCounterStride;
constexpr uint32_t kGiantCounterOff = 2 * kClasses * 8 * kCounterStride;

struct BatchHeader {
  uint32_t n_requests;
  uint32_t n_tiles;
  uint64_t n_queries;
  uint32_t class_end[kClasses];  // class c owns tiles [class_end[c-1], class_end[c])
  uint32_t pad;
};

// Counts of a batch. The count pass gives each query a bound on its results:
// exact for lane tiles and extension requests, an upper bound for wave tiles
// (their count walks do not test the leaves of partly covered ranges). The
// scan of the bounds places each query's pairs in the arena, and the write
// pass, which tests those leaves, records each query's exact count. Query q's
// pairs are then arena[offsets[q], offsets[q] + exact[q]).
struct BatchArgs {
  const BatchHeader* header;     // device
  const BatchRequest* requests;  // device, sorted by tbase
  uint32_t* tile_counters;       // kTileCounterWords: per-pass, per-class, per-XCD counters
  uint32_t* counts;              // per global query: result bound
  uint32_t* exact;               // per global query: results
  uint64_t* offsets;             // per global query: exclusive scan of counts
  uint64_t* exact_offsets;       // per global query: exclusive scan of exact (with collect)
  uint64_t* seg_sums;            // scan scratch (batch_seg_sums_words())
  uint64_t* total;               // [0] sum of the bounds, [2] of the exact counts
  uint64_t* ranges;              // per request: {first arena slot, slots (bounds), results}
  rts_pair* arena;               // results
  uint64_t arena_cap;            // pairs the arena holds
  const uint32_t* order;         // wave tiles in Morton order (tile -> tile), or null
  uint32_t* giants;              // kGiantBuckets lists of giant_cap tiles each
  uint32_t giant_cap;
  RangeLogArgs log;  // range logs of the wave tiles' count walks (range_log.h)
};

// Morton ordering of a batch's wave tiles: keys (class << 30 | Morton code of
// the query's centre) sorted with the tile index, so that neighbouring tiles
// of one class run together on one XCD. Buffers hold max_tiles entries.
struct BatchSort {
  uint32_t* keys_a;
  uint32_t* vals_a;
  uint32_t* keys_b;
  uint32_t* vals_b;
  SortScratch ss;
  const float4* scene;  // the index's scene bounds (device)
  uint32_t n;           // wave tiles of this batch (0: no ordering)
};

// Count pass, scan, per-request ranges and write pass of one batch, on
// `stream`. With sort.n = 0 every launch has a fixed grid (capturable in a
// graph and replayable for any batch whose sizes fit the buffers); otherwise
// the wave tiles are Morton-ordered first. `t` is a device copy of the tree
// view.
// A request of the batch that the batch kernels do not walk: segment
// queries, or any query on a double index. Its queries [qbase, qbase + n) of
// the batch are counted and written by one wave launch each (query.hip, the
// plain path's kernels) into the batch's counts, offsets and arena.
struct BatchExt {
  QueryLaunch ql;  // tree, queries, n, qid_base, pred, point_queries, f64, leaf size
  uint64_t qbase;
  bool none;       // a predicate that matches nothing (segments under RTS_CONTAINS)
};
struct BatchExtArgs {
  uint32_t* tickets;     // kWaveTicketWords words
  const uint64_t* zero;  // a device 0 (queue base of the arena)
  const void* arena_queue;  // device rts_dev_queue over the slot's arena
};
// With `collect`, exact_offsets is filled too (the pairs of requests that
// copy them out are gathered by launch_copy_queries).
hipError_t launch_batch(const TreeView* t, const BatchArgs& a, const BatchSort& sort,
                        const BatchExt* ext, size_t n_ext, const BatchExtArgs& ea, bool collect,
                        hipStream_t stream);
// hashes[r] = sum of pair_hash(query_id, geom_id) (checksum.h) over the pairs
// of request r of the batch (each query's arena[offsets[q], + exact[q])).
// hashes must be zero.
hipError_t launch_hash_queries(const BatchArgs& a, unsigned long long* hashes,
                               hipStream_t stream);
// Copies the pairs of the batch's queries [q0, q0 + n), in query order, to
// out[0, take): query q's pairs go to out + exact_offsets[q] - exact_offsets[q0].
hipError_t launch_copy_queries(const BatchArgs& a, uint64_t q0, uint64_t n, rts_pair* out,
                               uint64_t take, hipStream_t stream);
// Adds `add` to the query ids of the pairs in arena[span[0], span[1]) and, if
// do_hash, adds their hash to *hash.
hipError_t launch_rebase_hash_span(rts_pair* arena, uint64_t cap, const uint64_t* span,
                                   uint32_t add, int do_hash, unsigned long long* hash,
                                   hipStream_t stream);
size_t batch_seg_sums_words();

}  // namespace rts
