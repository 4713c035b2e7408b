// Internal declarations shared by the HIP translation units of
// librtspatial_amd.so. Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "rts_c.h"
#include "rtspatial/device/tree.h"

namespace rts {

constexpr uint64_t kMaxBoxes = (1ull << 30) - 1;  // ids and sorted positions fit in 30 bits
constexpr int kMortonBits = 30;                   // 15 bits per axis

// Result sink resolved from (handler, arg).
struct SinkArgs {
  int mode;          // rts_handler
  const void* arg;   // counter or rts_dev_queue device pointer
};

// ----------------------------------------------------- host-side kernels ----
// Thin launch wrappers (defined in the .hip files). All are asynchronous on
// `stream` and return the hipError_t of the launch.

// sort.hip
struct SortScratch {
  uint32_t* hist;      // 4 x 256
  uint32_t* offsets;   // 4 x 256
  uint32_t* status;    // 4 x tiles x 256
  uint32_t* tickets;   // 4
  uint32_t* error;     // 1
  uint32_t tiles_cap;  // tiles the status array holds per pass
};
size_t sort_scratch_bytes(uint64_t n);
void sort_scratch_carve(void* base, uint64_t n, SortScratch* s);
// Sorts (keys, vals) on bits [0, key_bits). keys_a/vals_a hold the input
// (with identity_vals the values 0..n-1 are generated and vals_a is only a
// buffer); keys_b/vals_b are ping-pong buffers. `*in_b` tells whether the
// result ended in the b buffers.
hipError_t radix_sort_pairs(uint32_t* keys_a, uint32_t* vals_a, uint32_t* keys_b,
                            uint32_t* vals_b, uint32_t n, int key_bits, bool identity_vals,
                            const SortScratch& s, hipStream_t stream, bool* in_b);

// build.hip
hipError_t launch_scene_bounds(const float4* boxes, uint32_t n, float4* partials,
                               uint32_t n_partials, float4* scene, hipStream_t stream);
uint32_t scene_bounds_partials(uint32_t n);
hipError_t launch_morton_boxes(const float4* boxes, uint32_t n, const float4* scene,
                               uint32_t* keys, hipStream_t stream);
hipError_t launch_morton_points(const float2* pts, uint32_t n, const float4* scene,
                                uint32_t* keys, hipStream_t stream);
// Every stride-th element of src (elem_bytes 8, 16 or 32) into dst[0, m).
hipError_t launch_gather_strided(const void* src, uint32_t elem_bytes, uint64_t stride, void* dst,
                                 uint32_t m, hipStream_t stream);
// Karras hierarchy over n sorted keys: the children's parent codes and each
// internal node's far end node_end[i] (the other end of its leaf range; n - 1
// words, kept for every refit, which writes the node records); only nodes
// spanning refit chunks get their record's meta here.
hipError_t launch_karras(const uint32_t* sorted_keys, uint32_t n, NodeRec* nodes, uint32_t* parent,
                         uint32_t* node_end, hipStream_t stream);
struct RefitArgs {
  float4* leaf_boxes;
  uint32_t n;
  const uint32_t* parent;
  NodeRec* nodes;
  float4* root_box;
  uint32_t* root_live;
  uint32_t* flags;
  uint32_t* error;
  uint32_t* boundary;  // refit_boundary_words(n): per-chunk boundary slots, then the top list
  const uint32_t* node_end;  // the hierarchy kernel's far ends (records' meta)
  NodeRec* top = nullptr;    // main tree of kTopMinLeaves or more: its top, rebuilt after the refit
  // Build refits also gather the leaves: leaf k's box is boxes[id] with
  // id = ids[k] + off (ids counted from off, rewritten to the ids themselves
  // when off > 0), stored to leaf_boxes[k], and leaf_of_id[id] = off + k. off
  // > 0 places a delta tree over ids [off, off + n) at positions [off, off +
  // n): leaf_boxes and ids are then passed advanced by off.
  bool gather = false;
  const float4* boxes = nullptr;
  uint32_t* ids = nullptr;
  uint32_t off = 0;
  uint32_t* leaf_of_id = nullptr;
};
// Refit chunks: C = 2^refit_chunk_log2(n) leaves per level-1 workgroup, whose
// LDS holds 24 bytes per node of the chunk (48 KB at 2048 leaves).
__host__ __device__ constexpr uint32_t refit_chunk_log2(uint64_t) { return 11u; }
inline size_t refit_chunk_lds_bytes(uint32_t lg) {
  return (size_t)(1u << lg) * (sizeof(float4) + 2 * sizeof(uint32_t));
}
// Parent codes (node << 1 | side) carry kSpanBit when the parent's leaf range
// crosses a refit chunk boundary.
constexpr uint32_t kSpanBit = 0x80000000u;
// A chunk's boundary codes name the spanning nodes that receive a child from
// the chunk; those nodes cross the chunk's first or last leaf boundary, so
// they lie on two chains of ancestors of at most 64 nodes each (tree depth).
constexpr uint32_t kRefitChunkSlots = 128;
constexpr uint32_t kRefitChunkStride = kRefitChunkSlots + 1;  // [0] count, then codes
// The top list: a node reaching beyond its level-2 group contains a group
// boundary, and a boundary lies in at most 64 nodes.
__host__ __device__ inline uint64_t refit_top_words(uint64_t n) {
  return 64 * (n / 2048 + 2) + 2;
}
__host__ __device__ inline uint64_t refit_boundary_words(uint64_t n) {
  return (n / 2048 + 1) * kRefitChunkStride + refit_top_words(n);
}
hipError_t launch_refit(const RefitArgs& a, hipStream_t stream);
// Trees of at least this many leaves get a top (TreeView::top) for the wave walks.
constexpr uint32_t kTopMinLeaves = 4096;
hipError_t launch_tree_cost(const NodeRec* nodes, const float4* root_box, uint32_t n,
                            double* out, hipStream_t stream);
hipError_t launch_delete(const uint64_t* ids, uint64_t n_ids, float4* boxes, uint64_t n_boxes,
                         float4* leaf_boxes, const uint32_t* leaf_of_id, uint64_t n_tree,
                         Box64* boxes64, hipStream_t stream);
hipError_t launch_update(const rts_update* ups, uint64_t n_ups, float4* boxes, uint64_t n_boxes,
                         float4* leaf_boxes, const uint32_t* leaf_of_id, uint64_t n_tree,
                         hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t stream);

// coord64.hip: double coordinates.
enum { kRoundOutward = 0, kRoundNearest = 1 };
// dst[i] = src[i] rounded to float boxes: outward (min down, max up) or to nearest.
hipError_t launch_round_boxes64(const Box64* src, uint64_t n, float4* dst, int rounding,
                                hipStream_t stream);
// Queries rounded to nearest floats (double2 -> float2 or Box64 -> float4).
hipError_t launch_round_queries64(const void* src, bool points, uint64_t n, void* dst,
                                  hipStream_t stream);
// boxes64[id] = box, and the rounded float box by id and in sorted position.
hipError_t launch_update64(const rts_update_d* ups, uint64_t n_ups, Box64* boxes64, float4* boxes,
                           uint64_t n_boxes, float4* leaf_boxes, const uint32_t* leaf_of_id,
                           uint64_t n_tree, int rounding, hipStream_t stream);
// Morton keys of double points / box centres (rounded to floats first).
hipError_t launch_morton64(const void* queries, bool points, uint32_t n, const float4* scene,
                           uint32_t* keys, hipStream_t stream);

// query.hip
struct QueryLaunch {
  TreeView tree;
  const void* queries;   // float2 points, float4 envelopes or segments (double when f64)
  uint32_t n;
  uint32_t qid_base;     // added to the query index when reporting
  const uint32_t* order; // optional permutation: lane i works on queries[order[i]]
  int pred;              // rts_predicate, or kPredLine for segment queries
  int point_queries;     // 1 = points
  int f64;               // 1 = double queries (double2 / Box64), exact double predicates
  SinkArgs sink;
  int traversal;         // rts_traversal (resolved: THREAD or WAVE)
};
hipError_t launch_query(const QueryLaunch& q, hipStream_t stream);
// Range logs of two-pass wave walks (range_log.h).
struct RangeLogArgs {
  uint2* entries = nullptr;  // null: no logs (both passes walk the tree)
  uint64_t n_entries = 0;    // below 2^32
  uint32_t per_wave = 0;     // entries of one wave's region: set by the launch (n_entries / grid)
  uint32_t* off = nullptr;   // per query: first entry
  uint32_t* n = nullptr;     // per query: entries, or kLogOverflow
};
// Range-log entries (8 bytes each) of a two-pass query call: 512 per query,
// at most 2^29 (4 GiB).
constexpr uint64_t kLogPerQuery = 512;
constexpr uint64_t kLogMaxEntries = 1ull << 29;

struct WaveArgs {
  uint32_t* counts;
  const uint64_t* offsets;
  const uint64_t* base;
  uint32_t* tickets;  // kWaveTicketWords words, cleared by the launch
  RangeLogArgs log;   // two-pass calls: the count walks' range logs
};
constexpr uint32_t kWaveTicketWords = 8 * 16;
// Wave-per-query traversal. phase 0: per-query counts -> a.counts;
// 1: counts added to the counting handler's counter; 2: pairs written at
// queue slot a.base[0] + a.offsets[query]; 3: pairs collected in one pass
// (staged per wavefront, reserved in the queue by atomics).
enum { kPhaseCount = 0, kPhaseCountSink = 1, kPhaseWrite = 2, kPhaseWriteStaged = 3 };
hipError_t launch_query_wave(const QueryLaunch& q, const WaveArgs& a, int phase,
                             hipStream_t stream);

// scan.hip
size_t scan_scratch_bytes(uint64_t n);
// offsets[i] = counts[0] + ... + counts[i-1]; *total = the sum of all counts.
hipError_t launch_exclusive_scan(const uint32_t* counts, uint64_t n, uint64_t* offsets,
                                 uint64_t* total, void* scratch, hipStream_t stream);
// *base = atomicAdd(queue counter, *total).
hipError_t launch_reserve(const void* dev_queue, const uint64_t* total, uint64_t* base,
                          hipStream_t stream);

// index.hip
void set_error(const std::string& msg);

// profile.hip
struct PhaseToken {
  int phase;
  hipEvent_t begin;
};
PhaseToken phase_begin(int phase, hipStream_t stream);
void phase_end(const PhaseToken& t, hipStream_t stream);
// While one is alive, phase_begin on this thread records nothing (a call
// that must not create events, e.g. one that allocates nothing).
struct PhasesOff {
  PhasesOff();
  ~PhasesOff();
  PhasesOff(const PhasesOff&) = delete;
  PhasesOff& operator=(const PhasesOff&) = delete;
};

}  // namespace rts
