// Device-side layout of the LBVH: the node record, the tree view the
// traversal kernels take, and the exact FP32 predicates. Shared by the
// library's kernels and by kernels a caller compiles around its own result
// handler (rtspatial/device/query.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rts_c.h"

namespace rts {

// ------------------------------------------------------------ constants ----
constexpr int kWave = 64;                 // CDNA wavefront width
constexpr uint32_t kLeafFlag = 0x80000000u;   // child ref: leaf (sorted position)
constexpr uint32_t kRangeFlag = 0x40000000u;  // stack entry: scan the range of an internal node
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;
constexpr int kStackDepth = 72;           // >= max Karras depth (64) + 2

// --------------------------------------------------------------- layout ----
// Karras LBVH over n sorted leaves: internal nodes 0..n-2 (root 0), one
// 64-byte record each, so expanding a node is one cache-line fetch:
//   box[0], box[1]  bounds of the left / right child
//   meta            {left ref, right ref, first leaf, last leaf}
//                   ref = internal index, or sorted leaf position | kLeafFlag
//   live            {live leaves below the left child, ... the right child,
//                    split (the left child's last leaf), 0}
// parent[j] = (parent << 1) | side, for internal j in [0, n-1) and leaf k at
// [n-1+k]. Leaves are the boxes in Morton order: leaf_boxes[k], id leaf_ids[k].
struct __align__(64) NodeRec {
  float4 box[2];
  uint4 meta;
  uint4 live;
};
static_assert(sizeof(NodeRec) == 64, "node record is one 64-byte line");

// A box with double coordinates (rts_envelope_d): x0, y0, x1, y1.
struct Box64 {
  double x0, y0, x1, y1;
};
static_assert(sizeof(Box64) == sizeof(rts_envelope_d), "Box64 is rts_envelope_d");

// The tree's top: a cut of the tree into 2 * kTopPairs disjoint subtrees
// (or leaves), in leaf order, two to a record in the node record's format
// (two adjacent leaf ranges: meta {ref 0, ref 1, first leaf of 0, last leaf
// of 1}, live {live 0, live 1, last leaf of 0, 0}), so that a wavefront's
// first expansion step tests all of them at once instead of descending the
// top levels one dependent step at a time. Unused entries hold the empty box
// (no test passes).
constexpr int kTopPairs = 64;

struct TreeView {
  const float4* leaf_boxes;
  const uint32_t* leaf_ids;
  const NodeRec* nodes;
  const float4* root_box;     // device scalar
  const uint32_t* root_live;  // device scalar: live leaves in the tree
  uint32_t n;                 // leaves
  uint32_t leaf_size;         // ranges with <= leaf_size leaves are scanned directly
  uint32_t* error;            // device word: bit 1 refit walk, bit 2 traversal bound
  const Box64* boxes64;       // double index: the boxes by id (else null)
  const NodeRec* top = nullptr;  // kTopPairs records of the tree's top, or null
};

// Internal predicate id of line-segment queries (closed segment meets closed
// box); not an rts_predicate value.
constexpr int kPredLine = 3;

// Device error bits (the index's error word).
constexpr uint32_t kErrSortWait = 1u;
constexpr uint32_t kErrRefitWalk = 2u;
constexpr uint32_t kErrTraversal = 4u;

// Pushes v onto a lane's private traversal stack. kStackDepth exceeds the
// deepest tree the builder emits (ties in the Morton keys are split by leaf
// index, so depth <= 64); should a push still find the stack full, the walk
// sets kErrTraversal in the index's error word (the call then fails) rather
// than dropping the subtree silently.
__device__ __forceinline__ void lane_push(uint32_t* stack, int& sp, uint32_t v, uint32_t* error) {
  if (sp < kStackDepth) {
    stack[sp++] = v;
  } else {
    atomicOr(error, kErrTraversal);
  }
}

// ----------------------------------------------------------- predicates ----
// Exact closed-interval tests in FP32 (reference geom/envelope.cuh:29-69).
// Boxes are float4 (x0, y0, x1, y1). Every comparison with a NaN is false, so
// a deleted box (NaN coordinates) never matches.
__device__ __forceinline__ bool contains_point(const float4 b, float px, float py) {
  return px >= b.x && px <= b.z && py >= b.y && py <= b.w;
}
__device__ __forceinline__ bool contains_box(const float4 a, const float4 o) {
  return o.x >= a.x && o.z <= a.z && o.y >= a.y && o.w <= a.w;
}
__device__ __forceinline__ bool intersects(const float4 a, const float4 o) {
  return o.x <= a.z && o.z >= a.x && o.y <= a.w && o.w >= a.y;
}

}  // namespace rts
