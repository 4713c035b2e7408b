// The state behind an rts_index handle (opaque to C ABI callers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "rts_internal.h"

struct rts_index {
  rts_config cfg{};
  int device = 0;
  std::mutex mu;  // serializes host-side state changes of one index

  // Boxes by id (the reference's envelopes_, spatial_index.cuh:919) and the
  // sorted position of each id in the current tree.
  uint64_t n = 0;    // ids handed out
  uint64_t cap = 0;  // capacity of the by-id arrays
  float4* boxes = nullptr;      // a double index: the boxes rounded to floats
  rts::Box64* boxes64 = nullptr;  // a double index: the boxes as given
  uint32_t* leaf_of_id = nullptr;

  // Trees (valid for n_tree leaves unless stale). The main tree holds ids
  // [0, n_main) at sorted positions [0, n_main); a small delta tree of boxes
  // inserted since the last full build holds ids [n_main, n_tree) at
  // positions [n_main, n_tree) of the same leaf arrays, its nodes at
  // nodes[n_main ...] and its parents at parent[2 n_main ...]. Ids in
  // [n_tree, n) are inserted but in no tree yet.
  uint64_t tcap = 0;
  uint64_t n_tree = 0;
  uint64_t n_main = 0;
  uint64_t n_delta = 0;
  bool stale = true;
  bool refit_needed = false;  // boxes changed in place since the last refit/build
  float4* leaf_boxes = nullptr;
  uint32_t* leaf_ids = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  rts::NodeRec* nodes = nullptr;
  uint32_t* parent = nullptr;
  uint32_t* node_end = nullptr;  // far end of each internal node's leaf range (refit meta)
  rts::NodeRec* top = nullptr;   // the main tree's top (kTopPairs records)
  uint32_t* flags = nullptr;
  char* sort_scratch = nullptr;
  rts::SortScratch ss{};
  float4* partials = nullptr;
  uint32_t n_partials_cap = 0;

  // Device scalars: scene bounds, root box, tree cost (sum, ratio).
  char* small = nullptr;
  float4* scene = nullptr;
  float4* root_box = nullptr;
  uint32_t* root_live = nullptr;
  float4* delta_scene = nullptr;
  float4* delta_root_box = nullptr;
  uint32_t* delta_root_live = nullptr;
  double* cost = nullptr;
  uint32_t* error = nullptr;  // device error bits (kErr*)

  // Tree-cost bookkeeping for the SAH rebuild policy.
  double* h_cost = nullptr;  // pinned: [0..1] after build, [2..3] after refit
  hipEvent_t ev_cost_build = nullptr;
  hipEvent_t ev_cost_refit = nullptr;
  bool cost_build_pending = false;
  bool cost_refit_pending = false;
  uint64_t build_epoch = 0;
  uint64_t refit_epoch = 0;
  double sah_at_build = 0;
  double sah_current = 0;

  uint64_t builds = 0;
  uint64_t refits = 0;
  uint64_t updated_since_build = 0;  // boxes given new envelopes since the last full build
  uint64_t delta_builds = 0;

  // Plans stored by rts_index_autotune, per [predicate][points]; kept until
  // the next tuning of that kind (Clear keeps them: they describe the
  // workload, not the stored boxes).
  struct TunedPlan {
    bool valid = false;
    uint32_t traversal = 0;
    bool sort = false;
    bool sort_tuned = false;  // ordering was timed on the whole query set
    uint64_t sort_min_n = 0;  // with sort_tuned: calls of at least this many queries are ordered
    uint32_t leaf_size = 0;
  };
  TunedPlan tuned[3][2];

  bool f64() const { return cfg.coord != RTS_COORD_F32; }
};

namespace rts {
// The main tree, and the delta tree (n = 0 when there is none).
TreeView tree_view(const rts_index* ix);
TreeView delta_view(const rts_index* ix);
// Brings the trees up to date on `stream`: a pending full build, refit or
// delta build runs. With `single`, a delta tree is merged by a full build,
// for callers that walk one tree only.
int ensure_built(rts_index* ix, hipStream_t stream, bool single = false);
// Reads the device error word (synchronous): RTS_ERR_DEVICE with a message if
// a kernel of this index gave up a bounded wait or walk.
int device_error_status(rts_index* ix);
}  // namespace rts
