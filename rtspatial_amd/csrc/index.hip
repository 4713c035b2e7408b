// Host side of the C ABI: index state, LBVH build orchestration, mutation and
// query entry points. Mirrors rtspatial::SpatialIndex<float, 2>
// (reference include/rtspatial/spatial_index.cuh:62-931) with the OptiX
// GAS/IAS machinery replaced by one LBVH over every inserted box.
//
// Build policy: Insert appends boxes by id; the next query (or
// rts_index_build) brings them into a tree once, however many Insert calls
// came before: while they are few next to the main tree (insert_delta_pct),
// a small delta tree is built over them alone and queries walk both trees;
// otherwise the whole tree is rebuilt. Update and Delete write the boxes in
// place and mark the trees for a refit, which runs once at the next query
// (or rts_index_build); under
// RTS_REBUILD_SAH a refit whose tree cost exceeds sah_rebuild_ratio times the
// cost measured right after the last build marks the tree stale instead of
// letting it degrade, and updates covering 1/200 of the boxes since the last
// build skip the refit for a rebuild (the reference never rebuilds,
// spatial_index.cuh:208-336).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "rts_internal.h"
#include "index_state.h"

namespace rts {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace rts

using namespace rts;

#define RTS_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                    \
      return RTS_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define RTS_CHECK_ARG(cond, msg)      \
  do {                                \
    if (!(cond)) {                    \
      set_error(msg);                 \
      return RTS_ERR_INVALID;         \
    }                                 \
  } while (0)

namespace {

template <typename T>
hipError_t dev_alloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int ensure_id_capacity(rts_index* ix, uint64_t need, hipStream_t stream) {
  if (need <= ix->cap) return RTS_OK;
  uint64_t cap = std::max<uint64_t>(need, ix->cap * 2);
  cap = std::max<uint64_t>(cap, 1024);
  float4* nb = nullptr;
  uint32_t* nl = nullptr;
  Box64* nd = nullptr;
  RTS_HIP(dev_alloc(&nb, cap));
  RTS_HIP(dev_alloc(&nl, cap));
  if (ix->f64()) RTS_HIP(dev_alloc(&nd, cap));
  if (ix->n) {
    RTS_HIP(hipMemcpyAsync(nb, ix->boxes, ix->n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    RTS_HIP(hipMemcpyAsync(nl, ix->leaf_of_id, ix->n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                           stream));
    if (nd)
      RTS_HIP(hipMemcpyAsync(nd, ix->boxes64, ix->n * sizeof(Box64), hipMemcpyDeviceToDevice,
                             stream));
  }
  // The old buffers may still be read by work queued on any stream.
  RTS_HIP(hipDeviceSynchronize());
  dev_free(ix->boxes);
  dev_free(ix->leaf_of_id);
  dev_free(ix->boxes64);
  ix->boxes = nb;
  ix->leaf_of_id = nl;
  ix->boxes64 = nd;
  ix->cap = cap;
  return RTS_OK;
}

int ensure_tree_capacity(rts_index* ix, uint64_t n) {
  if (n <= ix->tcap) return RTS_OK;
  RTS_HIP(hipDeviceSynchronize());
  // Room for a delta tree of up to insert_delta_pct percent on top of n.
  uint64_t cap = std::max<uint64_t>(n + n * ix->cfg.insert_delta_pct / 100, ix->tcap + ix->tcap / 2);
  cap = std::max<uint64_t>(cap, 1024);
  dev_free(ix->leaf_boxes);
  dev_free(ix->leaf_ids);
  dev_free(ix->keys);
  dev_free(ix->keys_alt);
  dev_free(ix->vals_alt);
  dev_free(ix->nodes);
  dev_free(ix->parent);
  dev_free(ix->node_end);
  dev_free(ix->top);
  dev_free(ix->flags);
  dev_free(ix->sort_scratch);
  dev_free(ix->partials);
  RTS_HIP(dev_alloc(&ix->leaf_boxes, cap));
  RTS_HIP(dev_alloc(&ix->leaf_ids, cap));
  RTS_HIP(dev_alloc(&ix->keys, cap));
  RTS_HIP(dev_alloc(&ix->keys_alt, cap));
  RTS_HIP(dev_alloc(&ix->vals_alt, cap));
  RTS_HIP(dev_alloc(&ix->nodes, cap));
  RTS_HIP(dev_alloc(&ix->parent, 2 * cap));
  RTS_HIP(dev_alloc(&ix->node_end, cap));
  RTS_HIP(dev_alloc(&ix->top, kTopPairs));
  RTS_HIP(dev_alloc(&ix->flags, cap + refit_boundary_words(cap)));
  RTS_HIP(dev_alloc(&ix->sort_scratch, sort_scratch_bytes(cap)));
  sort_scratch_carve(ix->sort_scratch, cap, &ix->ss);
  ix->n_partials_cap = scene_bounds_partials((uint32_t)cap);
  RTS_HIP(dev_alloc(&ix->partials, ix->n_partials_cap));
  ix->tcap = cap;
  return RTS_OK;
}

// Reads back finished tree-cost measurements without blocking.
void poll_cost(rts_index* ix) {
  if (ix->cost_build_pending && hipEventQuery(ix->ev_cost_build) == hipSuccess) {
    ix->sah_at_build = ix->h_cost[1];
    ix->sah_current = ix->h_cost[1];
    ix->cost_build_pending = false;
  }
  if (ix->cost_refit_pending && hipEventQuery(ix->ev_cost_refit) == hipSuccess) {
    ix->sah_current = ix->h_cost[3];
    ix->cost_refit_pending = false;
    if (ix->cfg.rebuild_policy == RTS_REBUILD_SAH && ix->sah_at_build > 0 &&
        ix->sah_current > ix->cfg.sah_rebuild_ratio * ix->sah_at_build && ix->refit_epoch == ix->build_epoch) {
      ix->stale = true;
    }
  }
}

int measure_cost(rts_index* ix, bool after_build, hipStream_t stream) {
  if (ix->n_main == 0) return RTS_OK;
  if (after_build) {
    if (ix->cost_build_pending) RTS_HIP(hipEventSynchronize(ix->ev_cost_build));
  } else if (ix->cost_refit_pending || ix->cost_build_pending) {
    return RTS_OK;  // a measurement is still in flight; skip this one
  }
  RTS_HIP(launch_tree_cost(ix->nodes, ix->root_box, (uint32_t)ix->n_main, ix->cost, stream));
  double* dst = after_build ? ix->h_cost : ix->h_cost + 2;
  RTS_HIP(hipMemcpyAsync(dst, ix->cost, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (after_build) {
    RTS_HIP(hipEventRecord(ix->ev_cost_build, stream));
    ix->cost_build_pending = true;
  } else {
    RTS_HIP(hipEventRecord(ix->ev_cost_refit, stream));
    ix->cost_refit_pending = true;
    ix->refit_epoch = ix->build_epoch;
  }
  return RTS_OK;
}

// Refit of the main tree, or (delta) of the delta tree, whose leaves, nodes
// and parents start at n_main, n_main and 2 n_main. Both share the flag and
// boundary words; refits run one after the other on a stream.
RefitArgs refit_args(rts_index* ix, bool delta) {
  const uint64_t off = delta ? ix->n_main : 0;
  RefitArgs a;
  a.leaf_boxes = ix->leaf_boxes + off;
  a.n = (uint32_t)(delta ? ix->n_delta : ix->n_main);
  a.parent = ix->parent + 2 * off;
  a.node_end = ix->node_end + off;
  a.top = delta ? nullptr : ix->top;
  a.nodes = ix->nodes + off;
  a.root_box = delta ? ix->delta_root_box : ix->root_box;
  a.root_live = delta ? ix->delta_root_live : ix->root_live;
  a.flags = ix->flags;
  a.error = ix->error;
  a.boundary = ix->flags + ix->tcap;  // after the n - 1 node flags
  return a;
}

int build_tree(rts_index* ix, hipStream_t stream) {
  const uint64_t N = ix->n;
  ix->n_tree = 0;
  if (N == 0) {
    ix->stale = false;
    return RTS_OK;
  }
  int rc = ensure_tree_capacity(ix, N);
  if (rc) return rc;
  const uint32_t n = (uint32_t)N;
  const uint32_t parts = scene_bounds_partials(n);
  PhaseToken ph = phase_begin(RTS_PHASE_BUILD_MORTON, stream);
  RTS_HIP(launch_scene_bounds(ix->boxes, n, ix->partials, parts, ix->scene, stream));
  RTS_HIP(launch_morton_boxes(ix->boxes, n, ix->scene, ix->keys, stream));
  phase_end(ph, stream);
  bool in_b = false;
  ph = phase_begin(RTS_PHASE_BUILD_SORT, stream);
  RTS_HIP(radix_sort_pairs(ix->keys, ix->leaf_ids, ix->keys_alt, ix->vals_alt, n, kMortonBits, true,
                           ix->ss, stream, &in_b));
  phase_end(ph, stream);
  if (in_b) {
    std::swap(ix->keys, ix->keys_alt);
    std::swap(ix->leaf_ids, ix->vals_alt);
  }
  ph = phase_begin(RTS_PHASE_BUILD_TREE, stream);
  RTS_HIP(launch_karras(ix->keys, n, ix->nodes, ix->parent, ix->node_end, stream));
  ix->n_main = N;
  ix->n_delta = 0;
  RefitArgs ra = refit_args(ix, false);  // the refit also gathers the leaves
  ra.gather = true;
  ra.boxes = ix->boxes;
  ra.ids = ix->leaf_ids;
  ra.off = 0;
  ra.leaf_of_id = ix->leaf_of_id;
  RTS_HIP(launch_refit(ra, stream));
  phase_end(ph, stream);
  ix->n_tree = N;
  ix->stale = false;
  ix->refit_needed = false;
  ix->updated_since_build = 0;
  ix->builds++;
  ix->build_epoch++;
  if (ix->cfg.rebuild_policy == RTS_REBUILD_SAH) return measure_cost(ix, true, stream);
  return RTS_OK;
}

// Builds the delta tree over ids [n_main, n) alone: its own scene bounds,
// Morton sort and hierarchy, in the tails of the shared arrays. The main tree
// is left as it is (refitted first if a mutation left it pending).
int build_delta(rts_index* ix, hipStream_t stream) {
  const uint64_t off = ix->n_main;
  const uint32_t nd = (uint32_t)(ix->n - off);
  const uint32_t parts = scene_bounds_partials(nd);
  PhaseToken ph = phase_begin(RTS_PHASE_BUILD_MORTON, stream);
  RTS_HIP(launch_scene_bounds(ix->boxes + off, nd, ix->partials, parts, ix->delta_scene, stream));
  RTS_HIP(launch_morton_boxes(ix->boxes + off, nd, ix->delta_scene, ix->keys + off, stream));
  phase_end(ph, stream);
  bool in_b = false;
  ph = phase_begin(RTS_PHASE_BUILD_SORT, stream);
  RTS_HIP(radix_sort_pairs(ix->keys + off, ix->leaf_ids + off, ix->keys_alt + off,
                           ix->vals_alt + off, nd, kMortonBits, true, ix->ss, stream, &in_b));
  phase_end(ph, stream);
  if (in_b) {  // the main tree's arrays stay where they are: copy the tail back
    RTS_HIP(hipMemcpyAsync(ix->keys + off, ix->keys_alt + off, nd * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, stream));
    RTS_HIP(hipMemcpyAsync(ix->leaf_ids + off, ix->vals_alt + off, nd * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, stream));
  }
  ph = phase_begin(RTS_PHASE_BUILD_TREE, stream);
  RTS_HIP(launch_karras(ix->keys + off, nd, ix->nodes + off, ix->parent + 2 * off,
                        ix->node_end + off, stream));
  ix->n_delta = nd;
  RefitArgs ra = refit_args(ix, true);  // the refit also gathers the leaves
  ra.gather = true;
  ra.boxes = ix->boxes;
  ra.ids = ix->leaf_ids + off;
  ra.off = (uint32_t)off;
  ra.leaf_of_id = ix->leaf_of_id;
  RTS_HIP(launch_refit(ra, stream));
  phase_end(ph, stream);
  ix->n_tree = ix->n;
  ix->delta_builds++;
  return RTS_OK;
}

// Whether the ids inserted since the last full build go to the delta tree:
// few enough of them, a tree to add them to, and room in the arrays (growing
// the arrays rebuilds anyway).
bool delta_ok(const rts_index* ix) {
  const uint64_t extra = ix->n - ix->n_main;
  return ix->cfg.insert_delta_pct > 0 && ix->cfg.rebuild_policy != RTS_REBUILD_ALWAYS &&
         ix->n_main > 0 && ix->n <= ix->tcap &&
         extra * 100 <= ix->n_main * (uint64_t)ix->cfg.insert_delta_pct;
}

// Under RTS_REBUILD_SAH, updates that give new envelopes to at least 1/200
// of the tree's boxes since its last build mark it stale: a full build costs
// ~1.5x a refit here (10M boxes: 2.8 vs 1.7 ms), while a refit over boxes
// moved beyond their old extents leaves the tree slower to query. On the
// served stream's 100k-box update batches (1% of 10M boxes, each moved by up
// to 10x its size) the tree cost grew ~2% per batch and never reached the
// 1.5x ratio; rebuilding instead raised saturated throughput from 27.8 to
// 30.1 Mq/s.
constexpr uint64_t kRebuildUpdateShare = 200;

// Update and Delete only mark the tree: the refit (or rebuild) runs once, at
// the next query or rts_index_build, however many mutations came before it,
// and not at all when an Insert in between makes the tree stale anyway.
// `updated`: boxes given new envelopes (0 for deletes, which only shrink
// node bounds).
int after_mutation(rts_index* ix, uint64_t updated) {
  if (ix->stale) return RTS_OK;  // the next query rebuilds anyway
  if (ix->cfg.rebuild_policy == RTS_REBUILD_ALWAYS) {
    ix->stale = true;
    return RTS_OK;
  }
  if (ix->cfg.rebuild_policy == RTS_REBUILD_SAH) {
    ix->updated_since_build += updated;
    if (ix->updated_since_build * kRebuildUpdateShare >= ix->n_tree) {
      ix->stale = true;
      return RTS_OK;
    }
  }
  ix->refit_needed = true;
  return RTS_OK;
}

int refit_tree(rts_index* ix, hipStream_t stream) {
  ix->refit_needed = false;
  PhaseToken ph = phase_begin(RTS_PHASE_REFIT, stream);
  RTS_HIP(launch_refit(refit_args(ix, false), stream));
  if (ix->n_delta) RTS_HIP(launch_refit(refit_args(ix, true), stream));
  phase_end(ph, stream);
  ix->refits++;
  if (ix->cfg.rebuild_policy == RTS_REBUILD_SAH) return measure_cost(ix, false, stream);
  return RTS_OK;
}

int check_device_error(rts_index* ix) {
  uint32_t err = 0, sort_err = 0;
  RTS_HIP(hipMemcpy(&err, ix->error, sizeof(err), hipMemcpyDeviceToHost));
  if (ix->ss.error)
    RTS_HIP(hipMemcpy(&sort_err, ix->ss.error, sizeof(sort_err), hipMemcpyDeviceToHost));
  if (sort_err) err |= kErrSortWait;
  if (err) {
    std::string m = "device error:";
    if (err & kErrSortWait) m += " radix-sort look-back exceeded its wait bound;";
    if (err & kErrRefitWalk) m += " refit walk left the tree;";
    if (err & kErrTraversal) m += " traversal exceeded its step bound;";
    set_error(m);
    return RTS_ERR_DEVICE;
  }
  return RTS_OK;
}

}  // namespace

namespace rts {

TreeView tree_view(const rts_index* ix) {
  TreeView t;
  t.leaf_boxes = ix->leaf_boxes;
  t.leaf_ids = ix->leaf_ids;
  t.nodes = ix->nodes;
  t.root_box = ix->root_box;
  t.root_live = ix->root_live;
  t.n = (uint32_t)ix->n_main;
  t.leaf_size = ix->cfg.leaf_size;
  t.error = ix->error;
  t.boxes64 = ix->cfg.coord == RTS_COORD_F64 ? ix->boxes64 : nullptr;
  // The top is rebuilt by every refit of the main tree (builds included).
  t.top = ix->n_main >= kTopMinLeaves ? ix->top : nullptr;
  return t;
}

TreeView delta_view(const rts_index* ix) {
  TreeView t = tree_view(ix);
  const uint64_t off = ix->n_main;
  t.leaf_boxes = ix->leaf_boxes + off;
  t.leaf_ids = ix->leaf_ids + off;
  t.nodes = ix->nodes + off;
  t.root_box = ix->delta_root_box;
  t.root_live = ix->delta_root_live;
  t.n = (uint32_t)ix->n_delta;
  t.top = nullptr;  // the delta tree is walked from its root
  return t;
}

int device_error_status(rts_index* ix) { return check_device_error(ix); }

int ensure_built(rts_index* ix, hipStream_t stream, bool single) {
  poll_cost(ix);
  if (ix->stale) return build_tree(ix, stream);
  const bool inserted = ix->n > ix->n_tree;
  if ((inserted && !delta_ok(ix)) || (single && (ix->n_delta || inserted)))
    return build_tree(ix, stream);  // a refit pending now is subsumed
  if (ix->refit_needed) {
    int rc = refit_tree(ix, stream);
    if (rc) return rc;
  }
  if (inserted) return build_delta(ix, stream);
  return RTS_OK;
}

}  // namespace rts

extern "C" {

void rts_config_default(rts_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->max_geometries = 1000ull * 1000 * 10;
  c->max_queries = 1000ull * 1000;
  c->preallocate = 0;
  c->geom_sample_rate = 0.2f;
  c->query_sample_rate = 0.2f;
  c->intersect_cost_weight = 0.95f;
  c->max_geom_samples = 10000;
  c->max_query_samples = 1000;
  c->max_parallelism = 512;
  c->leaf_size = 8;
  c->rebuild_policy = RTS_REBUILD_SAH;
  c->sah_rebuild_ratio = 1.5f;
  c->traversal = RTS_TRAVERSAL_AUTO;
  c->sort_queries = 0;
  c->coord = RTS_COORD_F32;
  c->insert_delta_pct = 2;
}

const char* rts_last_error(void) { return g_last_error.c_str(); }

const char* rts_version(void) { return "rtspatial_amd 0.1 (gfx950)"; }

int rts_index_create(const rts_config* cfg, rts_index** out) {
  RTS_CHECK_ARG(out != nullptr, "rts_index_create: out is null");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("rts_index_create: no HIP device");
    return RTS_ERR_MISSING;
  }
  rts_index* ix = new rts_index();
  if (cfg) ix->cfg = *cfg; else rts_config_default(&ix->cfg);
  if (ix->cfg.leaf_size == 0) ix->cfg.leaf_size = 8;
  if (ix->cfg.sah_rebuild_ratio <= 1.0f) ix->cfg.sah_rebuild_ratio = 1.5f;
  if (ix->cfg.coord > RTS_COORD_F64_AS_F32) {
    delete ix;
    set_error("rts_index_create: unknown coord");
    return RTS_ERR_INVALID;
  }
  RTS_HIP(hipGetDevice(&ix->device));
  RTS_HIP(hipMalloc(&ix->small, 256));
  ix->scene = reinterpret_cast<float4*>(ix->small);
  ix->root_box = reinterpret_cast<float4*>(ix->small + 16);
  ix->cost = reinterpret_cast<double*>(ix->small + 32);
  ix->error = reinterpret_cast<uint32_t*>(ix->small + 64);
  ix->root_live = reinterpret_cast<uint32_t*>(ix->small + 80);
  ix->delta_scene = reinterpret_cast<float4*>(ix->small + 96);
  ix->delta_root_box = reinterpret_cast<float4*>(ix->small + 112);
  ix->delta_root_live = reinterpret_cast<uint32_t*>(ix->small + 128);
  RTS_HIP(hipMemset(ix->small, 0, 256));
  // Complete before any stream (including non-blocking ones) uses the index.
  RTS_HIP(hipDeviceSynchronize());
  RTS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ix->h_cost), 4 * sizeof(double)));
  RTS_HIP(hipEventCreateWithFlags(&ix->ev_cost_build, hipEventDisableTiming));
  RTS_HIP(hipEventCreateWithFlags(&ix->ev_cost_refit, hipEventDisableTiming));
  if (ix->cfg.preallocate && ix->cfg.max_geometries) {
    int rc = ensure_id_capacity(ix, ix->cfg.max_geometries, nullptr);
    if (!rc) rc = ensure_tree_capacity(ix, ix->cfg.max_geometries);
    if (rc) {
      rts_index_destroy(ix);
      return rc;
    }
  }
  *out = ix;
  return RTS_OK;
}

int rts_index_destroy(rts_index* ix) {
  if (!ix) return RTS_OK;
  (void)hipDeviceSynchronize();
  dev_free(ix->boxes);
  dev_free(ix->boxes64);
  dev_free(ix->leaf_of_id);
  dev_free(ix->leaf_boxes);
  dev_free(ix->leaf_ids);
  dev_free(ix->keys);
  dev_free(ix->keys_alt);
  dev_free(ix->vals_alt);
  dev_free(ix->nodes);
  dev_free(ix->parent);
  dev_free(ix->node_end);
  dev_free(ix->top);
  dev_free(ix->flags);
  dev_free(ix->sort_scratch);
  dev_free(ix->partials);
  dev_free(ix->small);
  if (ix->h_cost) (void)hipHostFree(ix->h_cost);
  if (ix->ev_cost_build) (void)hipEventDestroy(ix->ev_cost_build);
  if (ix->ev_cost_refit) (void)hipEventDestroy(ix->ev_cost_refit);
  delete ix;
  return RTS_OK;
}

int rts_index_clear(rts_index* ix) {
  RTS_CHECK_ARG(ix, "rts_index_clear: null index");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  RTS_HIP(hipDeviceSynchronize());
  ix->n = 0;
  ix->n_tree = 0;
  ix->n_main = 0;
  ix->n_delta = 0;
  ix->stale = true;
  ix->refit_needed = false;
  ix->cost_build_pending = ix->cost_refit_pending = false;
  ix->sah_at_build = ix->sah_current = 0;
  return RTS_OK;
}

int rts_index_insert(rts_index* ix, const rts_envelope* boxes, uint64_t n, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_insert: null index");
  RTS_CHECK_ARG(!ix->f64(), "rts_index_insert: a double index takes rts_index_insert_d");
  if (n == 0) return RTS_OK;  // spatial_index.cuh:144-146
  RTS_CHECK_ARG(boxes, "rts_index_insert: null boxes");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_CHECK_ARG(ix->n + n <= kMaxBoxes, "rts_index_insert: more than 2^30-1 boxes");
  RTS_HIP(hipSetDevice(ix->device));
  int rc = ensure_id_capacity(ix, ix->n + n, stream);
  if (rc) return rc;
  RTS_HIP(hipMemcpyAsync(ix->boxes + ix->n, boxes, n * sizeof(float4), hipMemcpyDefault, stream));
  ix->n += n;  // ids [n_tree, n) wait for the next build (delta or full)
  return RTS_OK;
}

int rts_index_insert_d(rts_index* ix, const rts_envelope_d* boxes, uint64_t n,
                       rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_insert_d: null index");
  RTS_CHECK_ARG(ix->f64(), "rts_index_insert_d: a float index takes rts_index_insert");
  if (n == 0) return RTS_OK;
  RTS_CHECK_ARG(boxes, "rts_index_insert_d: null boxes");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_CHECK_ARG(ix->n + n <= kMaxBoxes, "rts_index_insert_d: more than 2^30-1 boxes");
  RTS_HIP(hipSetDevice(ix->device));
  int rc = ensure_id_capacity(ix, ix->n + n, stream);
  if (rc) return rc;
  RTS_HIP(hipMemcpyAsync(ix->boxes64 + ix->n, boxes, n * sizeof(Box64), hipMemcpyDefault, stream));
  RTS_HIP(launch_round_boxes64(ix->boxes64 + ix->n, n, ix->boxes + ix->n,
                               ix->cfg.coord == RTS_COORD_F64 ? kRoundOutward : kRoundNearest,
                               stream));
  ix->n += n;  // ids [n_tree, n) wait for the next build (delta or full)
  return RTS_OK;
}

int rts_index_delete(rts_index* ix, const uint64_t* ids, uint64_t n, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_delete: null index");
  if (n == 0) return RTS_OK;
  RTS_CHECK_ARG(ids, "rts_index_delete: null ids");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  poll_cost(ix);
  if (ix->n == 0) return RTS_OK;
  PhaseToken ph = phase_begin(RTS_PHASE_MUTATE, stream);
  RTS_HIP(launch_delete(ids, n, ix->boxes, ix->n, ix->leaf_boxes, ix->leaf_of_id,
                        ix->stale ? 0 : ix->n_tree, ix->boxes64, stream));
  phase_end(ph, stream);
  return after_mutation(ix, 0);
}

int rts_index_update(rts_index* ix, const rts_update* ups, uint64_t n, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_update: null index");
  RTS_CHECK_ARG(!ix->f64(), "rts_index_update: a double index takes rts_index_update_d");
  if (n == 0) return RTS_OK;
  RTS_CHECK_ARG(ups, "rts_index_update: null updates");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  poll_cost(ix);
  if (ix->n == 0) return RTS_OK;
  PhaseToken ph = phase_begin(RTS_PHASE_MUTATE, stream);
  RTS_HIP(launch_update(ups, n, ix->boxes, ix->n, ix->leaf_boxes, ix->leaf_of_id,
                        ix->stale ? 0 : ix->n_tree, stream));
  phase_end(ph, stream);
  return after_mutation(ix, n);
}

int rts_index_update_d(rts_index* ix, const rts_update_d* ups, uint64_t n, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_update_d: null index");
  RTS_CHECK_ARG(ix->f64(), "rts_index_update_d: a float index takes rts_index_update");
  if (n == 0) return RTS_OK;
  RTS_CHECK_ARG(ups, "rts_index_update_d: null updates");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  poll_cost(ix);
  if (ix->n == 0) return RTS_OK;
  PhaseToken ph = phase_begin(RTS_PHASE_MUTATE, stream);
  RTS_HIP(launch_update64(ups, n, ix->boxes64, ix->boxes, ix->n, ix->leaf_boxes, ix->leaf_of_id,
                          ix->stale ? 0 : ix->n_tree,
                          ix->cfg.coord == RTS_COORD_F64 ? kRoundOutward : kRoundNearest, stream));
  phase_end(ph, stream);
  return after_mutation(ix, n);
}

}  // extern "C"

namespace {

// Per-call device scratch, carved from one stream-ordered allocation.
// Stream-ordered scratch of query calls (range logs take gigabytes) comes
// from one memory pool per device that keeps its memory: the default pool
// hands freed memory back at every synchronisation, so each call paid for
// the driver's allocation again (the per-request path's p99 rose from 272
// ms to 1.69 s with range logs). The pools live until the process ends.
hipError_t scratch_alloc(void** p, size_t bytes, hipStream_t stream) {
  static std::mutex mu;
  static hipMemPool_t pools[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  hipMemPool_t pool;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!pools[dev]) {
      hipMemPoolProps props = {};
      props.allocType = hipMemAllocationTypePinned;
      props.handleTypes = hipMemHandleTypeNone;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      e = hipMemPoolCreate(&pools[dev], &props);
      if (e != hipSuccess) {
        pools[dev] = nullptr;
        return e;
      }
      uint64_t keep = UINT64_MAX;
      e = hipMemPoolSetAttribute(pools[dev], hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) return e;
    }
    pool = pools[dev];
  }
  return hipMallocFromPoolAsync(p, bytes, pool, stream);
}

struct QueryScratch {
  char* block = nullptr;
  const uint32_t* order = nullptr;  // Morton permutation (optional)
  uint32_t* counts = nullptr;       // wave two-pass: per-query counts
  uint64_t* offsets = nullptr;      //                exclusive scan of counts
  uint64_t* total = nullptr;        //                [0] total, [1] queue base
  void* scan_tmp = nullptr;
  uint32_t* tickets = nullptr;      // wave traversal: per-XCD query tickets
  RangeLogArgs log;                 // wave two-pass: the count walks' range logs
};

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Carves the scratch of a query call over nq queries. With `ext` (the
// allocation-free entry point) the block is ext's first ext_bytes bytes, and
// a call needing more is an error; otherwise it is allocated on `stream`
// (and qs->block is freed by the caller).
int alloc_scratch(uint64_t nq, bool sort, bool wave, bool two_pass, hipStream_t stream,
                  QueryScratch* qs, size_t* order_off, size_t* sort_off, char* ext = nullptr,
                  size_t ext_bytes = 0, char** base_out = nullptr) {
  size_t off = 0;
  size_t o_keys = 0, o_sort = 0, o_counts = 0, o_offsets = 0, o_total = 0, o_scan = 0, o_tk = 0;
  size_t o_log = 0, o_log_off = 0, o_log_n = 0;
  // Range logs: not in a caller's static scratch.
  const bool logs = two_pass && !ext;
  const uint64_t log_entries = std::min<uint64_t>(nq * kLogPerQuery, kLogMaxEntries);
  if (wave) {
    o_tk = off;
    off = align_up(off + kWaveTicketWords * sizeof(uint32_t));
  }
  if (sort) {
    o_keys = off;
    off = align_up(off + 4 * nq * sizeof(uint32_t));
    o_sort = off;
    off = align_up(off + sort_scratch_bytes(nq));
  }
  if (two_pass) {
    o_counts = off;
    off = align_up(off + nq * sizeof(uint32_t));
    o_offsets = off;
    off = align_up(off + nq * sizeof(uint64_t));
    o_total = off;
    off = align_up(off + 2 * sizeof(uint64_t));
    o_scan = off;
    off = align_up(off + scan_scratch_bytes(nq));
  }
  if (logs) {
    o_log = off;
    off = align_up(off + log_entries * sizeof(uint2));
    o_log_off = off;
    off = align_up(off + nq * sizeof(uint32_t));
    o_log_n = off;
    off = align_up(off + nq * sizeof(uint32_t));
  }
  char* b = nullptr;
  if (ext) {
    RTS_CHECK_ARG(off <= ext_bytes, "query: the call's scratch exceeds the static scratch array");
    b = ext;
  } else {
    if (off == 0) return RTS_OK;
    RTS_HIP(scratch_alloc(reinterpret_cast<void**>(&qs->block), off, stream));
    b = qs->block;
  }
  if (base_out) *base_out = b;
  if (wave) qs->tickets = reinterpret_cast<uint32_t*>(b + o_tk);
  if (two_pass) {
    qs->counts = reinterpret_cast<uint32_t*>(b + o_counts);
    qs->offsets = reinterpret_cast<uint64_t*>(b + o_offsets);
    qs->total = reinterpret_cast<uint64_t*>(b + o_total);
    qs->scan_tmp = b + o_scan;
  }
  if (logs) {
    qs->log.entries = reinterpret_cast<uint2*>(b + o_log);
    qs->log.n_entries = log_entries;
    qs->log.off = reinterpret_cast<uint32_t*>(b + o_log_off);
    qs->log.n = reinterpret_cast<uint32_t*>(b + o_log_n);
  }
  *order_off = o_keys;
  *sort_off = o_sort;
  return RTS_OK;
}

// Morton-orders the queries for traversal coherence; the permutation lands in
// the scratch block.
int order_queries(rts_index* ix, const void* queries, uint64_t nq, bool points, bool f64,
                  char* keys_base, char* sort_base, hipStream_t stream,
                  const uint32_t** order_out) {
  const size_t words = (size_t)nq;
  uint32_t* keys = reinterpret_cast<uint32_t*>(keys_base);
  uint32_t* vals = keys + words;
  uint32_t* keys_b = vals + words;
  uint32_t* vals_b = keys_b + words;
  SortScratch ss;
  sort_scratch_carve(sort_base, nq, &ss);
  ss.error = ix->error;  // a look-back that gives up is reported by the index's error word
  if (f64) {
    RTS_HIP(launch_morton64(queries, points, (uint32_t)nq, ix->scene, keys, stream));
  } else if (points) {
    RTS_HIP(launch_morton_points(static_cast<const float2*>(queries), (uint32_t)nq, ix->scene, keys,
                                 stream));
  } else {
    RTS_HIP(launch_morton_boxes(static_cast<const float4*>(queries), (uint32_t)nq, ix->scene, keys,
                                stream));
  }
  bool in_b = false;
  RTS_HIP(radix_sort_pairs(keys, vals, keys_b, vals_b, (uint32_t)nq, kMortonBits, true, ss, stream,
                           &in_b));
  *order_out = in_b ? vals_b : vals;
  return RTS_OK;
}

// How one query call runs: traversal mode, query ordering, leaf-range size.
struct Plan {
  int traversal;
  bool sort;
  uint32_t leaf_size;
};

// The plan of a query call. Explicit config settings win; a plan stored by
// rts_index_autotune fills the settings left on auto; otherwise: one query
// per wavefront where queries have many results (envelope intersects /
// within), one per lane otherwise (points, envelope contains), and Morton
// ordering of batches of 64K queries or more (neighbouring lanes share their
// descent; neighbouring waves, which take consecutive queries on one XCD,
// share that XCD's L2).
// Calls of at least this many queries are Morton-ordered unless a tuning
// moved the threshold (TunedPlan::sort_min_n).
constexpr uint64_t kSortMinQueries = 65536;
// The autotuner times query ordering on the whole query set up to this size.
constexpr uint64_t kOrderTuneQueries = 1ull << 22;

Plan plan_query(const rts_index* ix, int pred, bool points, uint64_t nq) {
  // Segment queries (kPredLine) are not tuned: they take the config or the defaults.
  static const rts_index::TunedPlan kUntuned{};
  const rts_index::TunedPlan& t =
      pred == kPredLine ? kUntuned : ix->tuned[pred][points ? 1 : 0];
  Plan p;
  if (ix->cfg.traversal == RTS_TRAVERSAL_THREAD || ix->cfg.traversal == RTS_TRAVERSAL_WAVE)
    p.traversal = (int)ix->cfg.traversal;
  else if (t.valid)
    p.traversal = (int)t.traversal;
  else
    p.traversal = (points || pred == RTS_CONTAINS) ? RTS_TRAVERSAL_THREAD : RTS_TRAVERSAL_WAVE;
  if (ix->cfg.sort_queries == 1 || ix->cfg.sort_queries == 2)
    p.sort = ix->cfg.sort_queries == 1;
  else if (t.valid && t.sort_tuned)
    p.sort = nq >= t.sort_min_n;
  else
    p.sort = nq >= kSortMinQueries;
  p.leaf_size = t.valid ? t.leaf_size : ix->cfg.leaf_size;
  return p;
}

// Bytes of one query: float or double point, or envelope / segment.
size_t query_bytes(bool points, bool f64) { return (points ? 8 : 16) * (f64 ? 2 : 1); }

// Runs one query call under `plan`; the caller holds the index lock and has
// brought the tree up to date. f64: double queries on an exact double index.
// ext / ext_bytes: device scratch supplied by the caller (nothing is then
// allocated; the call must fit it), else scratch is allocated per call.
int run_plan(rts_index* ix, const Plan& plan, int pred, const void* queries, uint64_t nq,
             bool points, bool f64, rts_handler handler, void* arg, hipStream_t stream,
             char* ext = nullptr, size_t ext_bytes = 0) {
  const bool sort = plan.sort;
  const int trav = plan.traversal;
  // Collecting wave walks count, scan and then write each query's pairs
  // contiguously; segment walks, whose count pass would cost as much as the
  // write (no subtree lies inside a segment), collect in one staged pass.
  const bool staged =
      trav == RTS_TRAVERSAL_WAVE && handler == RTS_HANDLER_COLLECT && pred == kPredLine;
  const bool two_pass = trav == RTS_TRAVERSAL_WAVE && handler == RTS_HANDLER_COLLECT && !staged;
  // At most 2^26 queries per launch: one wavefront per query keeps
  // grid x workgroup size below 2^32 work-items.
  const uint64_t kChunk = 1ull << 26;
  for (uint64_t off = 0; off < nq; off += kChunk) {
    const uint64_t m = std::min(kChunk, nq - off);
    const void* qp = static_cast<const char*>(queries) + off * query_bytes(points, f64);
    QueryScratch qs;
    size_t o_keys = 0, o_sort = 0;
    char* base = nullptr;
    int rc = alloc_scratch(m, sort, trav == RTS_TRAVERSAL_WAVE, two_pass, stream, &qs, &o_keys,
                           &o_sort, ext, ext_bytes, &base);
    if (rc) return rc;
    if (sort) {
      PhaseToken ph = phase_begin(RTS_PHASE_QUERY_ORDER, stream);
      rc = order_queries(ix, qp, m, points, f64, base + o_keys, base + o_sort, stream,
                         &qs.order);
      phase_end(ph, stream);
      if (rc) return rc;
    }
    QueryLaunch ql;
    ql.queries = qp;
    ql.n = (uint32_t)m;
    ql.qid_base = (uint32_t)off;
    ql.order = qs.order;
    ql.pred = pred;
    ql.point_queries = points ? 1 : 0;
    ql.f64 = f64 ? 1 : 0;
    ql.sink.mode = handler;
    ql.sink.arg = arg;
    ql.traversal = trav;
    // The main tree, then the delta tree of recent inserts when there is one;
    // the second walk reuses the scratch once the first is done (one stream).
    const TreeView trees[2] = {tree_view(ix), delta_view(ix)};
    for (const TreeView& tv : trees) {
      if (tv.n == 0) continue;
      ql.tree = tv;
      ql.tree.leaf_size = plan.leaf_size;
      if (trav == RTS_TRAVERSAL_THREAD) {
        PhaseToken ph = phase_begin(RTS_PHASE_TRAVERSE, stream);
        RTS_HIP(launch_query(ql, stream));
        phase_end(ph, stream);
      } else if (!two_pass) {
        WaveArgs wa{nullptr, nullptr, nullptr, qs.tickets};
        PhaseToken ph = phase_begin(RTS_PHASE_TRAVERSE, stream);
        RTS_HIP(launch_query_wave(ql, wa, staged ? kPhaseWriteStaged : kPhaseCountSink, stream));
        phase_end(ph, stream);
      } else {
        WaveArgs wa{qs.counts, qs.offsets, qs.total + 1, qs.tickets, qs.log};
        PhaseToken ph = phase_begin(RTS_PHASE_QUERY_COUNT, stream);
        RTS_HIP(launch_query_wave(ql, wa, kPhaseCount, stream));
        RTS_HIP(launch_exclusive_scan(qs.counts, m, qs.offsets, qs.total, qs.scan_tmp, stream));
        RTS_HIP(launch_reserve(arg, qs.total, qs.total + 1, stream));
        phase_end(ph, stream);
        ph = phase_begin(RTS_PHASE_TRAVERSE, stream);
        RTS_HIP(launch_query_wave(ql, wa, kPhaseWrite, stream));
        phase_end(ph, stream);
      }
    }
    if (qs.block) RTS_HIP(hipFreeAsync(qs.block, stream));
  }
  return RTS_OK;
}

int check_query_args(int pred, const void* queries, uint64_t nq, bool points, rts_handler handler,
                     const void* arg) {
  RTS_CHECK_ARG(handler == RTS_HANDLER_COUNT || handler == RTS_HANDLER_COLLECT,
                "query: unknown handler");
  RTS_CHECK_ARG(arg, "query: null handler argument");
  RTS_CHECK_ARG((pred >= RTS_INTERSECTS && pred <= RTS_WITHIN) || pred == kPredLine,
                "query: unknown predicate");
  RTS_CHECK_ARG(!points || pred == RTS_CONTAINS,
                "query: point queries support only RTS_CONTAINS (spatial_index.cuh:385-389)");
  RTS_CHECK_ARG(nq == 0 || queries, "query: null queries");
  return RTS_OK;
}

// Queries of a double index with coord RTS_COORD_F64_AS_F32, rounded to
// nearest floats in a stream-ordered buffer that the destructor releases.
struct RoundedQueries {
  void* p = nullptr;
  hipStream_t stream = nullptr;
  ~RoundedQueries() {
    if (p) (void)hipFreeAsync(p, stream);
  }
};

// The queries as the kernels take them: unchanged, or, on an index that
// rounds doubles to floats, rounded into `rq`. *f64 tells which.
int kernel_queries(rts_index* ix, const void* queries, uint64_t nq, bool points,
                   hipStream_t stream, RoundedQueries* rq, const void** out, bool* f64) {
  *out = queries;
  *f64 = ix->cfg.coord == RTS_COORD_F64;
  if (ix->cfg.coord != RTS_COORD_F64_AS_F32) return RTS_OK;
  rq->stream = stream;
  RTS_HIP(scratch_alloc(&rq->p, nq * query_bytes(points, false), stream));
  RTS_HIP(launch_round_queries64(queries, points, nq, rq->p, stream));
  *out = rq->p;
  return RTS_OK;
}

int run_query(rts_index* ix, int pred, const void* queries, uint64_t nq, bool points,
              bool double_queries, rts_handler handler, void* arg, hipStream_t stream) {
  int rc = check_query_args(pred, queries, nq, points, handler, arg);
  if (rc) return rc;
  RTS_CHECK_ARG(double_queries == ix->f64(),
                double_queries ? "query: a float index takes rts_index_query_points/envelopes"
                               : "query: a double index takes rts_index_query_points_d/envelopes_d");
  if (nq == 0) return RTS_OK;
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  if (ix->n == 0) return RTS_OK;  // spatial_index.cuh:354-356
  rc = ensure_built(ix, stream);
  if (rc) return rc;
  RoundedQueries rq;
  const void* kq = nullptr;
  bool f64 = false;
  rc = kernel_queries(ix, queries, nq, points, stream, &rq, &kq, &f64);
  if (rc) return rc;
  return run_plan(ix, plan_query(ix, pred, points, nq), pred, kq, nq, points, f64, handler, arg,
                  stream);
}

// Segment queries (spatial_index.cuh:517-547): RTS_INTERSECTS runs the
// segment walk; RTS_CONTAINS reports nothing, as the reference's (its
// kContains case does no work, :524-526).
int run_lines(rts_index* ix, int pred, const void* queries, uint64_t nq, bool doubles,
              rts_handler handler, void* arg, hipStream_t stream) {
  RTS_CHECK_ARG(pred == RTS_INTERSECTS || pred == RTS_CONTAINS,
                "query_lines: segment queries take RTS_INTERSECTS or RTS_CONTAINS "
                "(spatial_index.cuh:523-545)");
  if (pred == RTS_INTERSECTS)
    return run_query(ix, kPredLine, queries, nq, false, doubles, handler, arg, stream);
  int rc = check_query_args(RTS_CONTAINS, queries, nq, false, handler, arg);
  if (rc) return rc;
  RTS_CHECK_ARG(doubles == ix->f64(),
                doubles ? "query: a float index takes rts_index_query_lines"
                        : "query: a double index takes rts_index_query_lines_d");
  return RTS_OK;
}

// Device buffers of one tuning run, released on every exit path.
struct TuneBuffers {
  void* sample = nullptr;
  rts_pair* pairs = nullptr;
  rts_pair* all_pairs = nullptr;  // results of the whole query set (ordering runs)
  char* small = nullptr;  // [0, 24): rts_dev_queue, [32, 40): counter, [40, 64): a second queue
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~TuneBuffers() {
    if (sample) (void)hipFree(sample);
    if (pairs) (void)hipFree(pairs);
    if (all_pairs) (void)hipFree(all_pairs);
    if (small) (void)hipFree(small);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace

extern "C" {

int rts_index_query_points(rts_index* ix, rts_predicate pred, const rts_point* q, uint64_t n,
                           rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_points: null index");
  return run_query(ix, pred, q, n, true, false, handler, arg, stream);
}

int rts_index_query_envelopes(rts_index* ix, rts_predicate pred, const rts_envelope* q, uint64_t n,
                              rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_envelopes: null index");
  return run_query(ix, pred, q, n, false, false, handler, arg, stream);
}

int rts_index_query_points_d(rts_index* ix, rts_predicate pred, const rts_point_d* q, uint64_t n,
                             rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_points_d: null index");
  return run_query(ix, pred, q, n, true, true, handler, arg, stream);
}

int rts_index_query_envelopes_d(rts_index* ix, rts_predicate pred, const rts_envelope_d* q,
                                uint64_t n, rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_envelopes_d: null index");
  return run_query(ix, pred, q, n, false, true, handler, arg, stream);
}

int rts_index_query_lines(rts_index* ix, rts_predicate pred, const rts_line* q, uint64_t n,
                          rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_lines: null index");
  return run_lines(ix, pred, q, n, false, handler, arg, stream);
}

int rts_index_query_lines_d(rts_index* ix, rts_predicate pred, const rts_line_d* q, uint64_t n,
                            rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_lines_d: null index");
  return run_lines(ix, pred, q, n, true, handler, arg, stream);
}

}  // extern "C"

namespace {

// Scratch of rts_index_query_static: one array per GPU in the library's device
// image, held until the process ends. kStaticMaxQueries queries under any plan
// need at most ~29.2 MiB of it (Morton keys and values 16 MiB, radix-sort
// status 1 MiB, counts 4 MiB, offsets 8 MiB, scan tiles and tickets); a call
// that would need more is refused before anything is launched. In host code
// `scratch` names a host-side stand-in, so kernels are only ever given the
// device address hipGetSymbolAddress returns for the current device.
constexpr uint64_t kStaticMaxQueries = 1ull << 20;
constexpr size_t kStaticScratchBytes = 32ull << 20;
__device__ __align__(256) char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once.
int static_scratch(int device, char** out) {
  RTS_CHECK_ARG(device >= 0 && device < kMaxDevices, "rts_index_query_static: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    RTS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RTS_CHECK_ARG(p, "rts_index_query_static: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return RTS_OK;
}

}  // namespace

extern "C" {

int rts_index_query_static(rts_index* ix, rts_predicate pred, int kind, const void* q, uint64_t n,
                           rts_handler handler, void* arg, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_query_static: null index");
  RTS_CHECK_ARG(kind >= RTS_QUERY_ENVELOPES && kind <= RTS_QUERY_LINES,
                "rts_index_query_static: unknown query kind");
  RTS_CHECK_ARG(n <= kStaticMaxQueries,
                "rts_index_query_static: at most 1048576 queries per call (the static scratch)");
  RTS_CHECK_ARG(ix->cfg.coord != RTS_COORD_F64_AS_F32,
                "rts_index_query_static: an index with coord RTS_COORD_F64_AS_F32 rounds its "
                "queries into an allocated buffer; use the allocating query calls");
  const bool points = kind == RTS_QUERY_POINTS;
  int pred_k = pred;
  if (kind == RTS_QUERY_LINES) {
    RTS_CHECK_ARG(pred == RTS_INTERSECTS || pred == RTS_CONTAINS,
                  "rts_index_query_static: segment queries take RTS_INTERSECTS or RTS_CONTAINS");
    int rc = check_query_args(RTS_CONTAINS, q, n, false, handler, arg);
    if (rc) return rc;
    if (pred == RTS_CONTAINS) return RTS_OK;  // no results, as rts_index_query_lines
    pred_k = kPredLine;
  }
  int rc = check_query_args(pred_k, q, n, points, handler, arg);
  if (rc) return rc;
  if (n == 0) return RTS_OK;
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  if (ix->n == 0) return RTS_OK;
  // Building or refitting can allocate (the tree grows with the boxes), so
  // the index must already be up to date.
  RTS_CHECK_ARG(!ix->stale && !ix->refit_needed && ix->n_tree == ix->n,
                "rts_index_query_static: the index has mutations pending; call rts_index_build "
                "first");
  char* ext = nullptr;
  rc = static_scratch(ix->device, &ext);
  if (rc) return rc;
  return run_plan(ix, plan_query(ix, pred_k, points, n), pred_k, q, n, points, ix->f64(), handler,
                  arg, stream, ext, kStaticScratchBytes);
}

uint64_t rts_index_size(const rts_index* ix) { return ix ? ix->n : 0; }

int rts_index_query_plan(rts_index* ix, rts_predicate pred, int points, uint64_t n,
                         rts_query_plan* out) {
  RTS_CHECK_ARG(ix, "rts_index_query_plan: null index");
  RTS_CHECK_ARG(out, "rts_index_query_plan: null plan");
  RTS_CHECK_ARG(pred >= RTS_INTERSECTS && pred <= RTS_WITHIN, "rts_index_query_plan: unknown predicate");
  RTS_CHECK_ARG(points >= RTS_QUERY_ENVELOPES && points <= RTS_QUERY_LINES,
                "rts_index_query_plan: unknown query kind");
  std::lock_guard<std::mutex> lk(ix->mu);
  const Plan p = points == RTS_QUERY_LINES ? plan_query(ix, kPredLine, false, n)
                                           : plan_query(ix, pred, points == RTS_QUERY_POINTS, n);
  out->traversal = (uint32_t)p.traversal;
  out->sort_queries = p.sort ? 1u : 2u;
  out->leaf_size = p.leaf_size;
  out->reserved = 0;
  return RTS_OK;
}

int rts_index_order_queries(rts_index* ix, const void* queries, uint64_t n, int points,
                            uint32_t* order, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_order_queries: null index");
  RTS_CHECK_ARG(n < (1ull << 32), "rts_index_order_queries: n must be below 2^32");
  RTS_CHECK_ARG(points >= RTS_QUERY_ENVELOPES && points <= RTS_QUERY_LINES,
                "rts_index_order_queries: unknown query kind");
  if (n == 0) return RTS_OK;
  RTS_CHECK_ARG(queries && order, "rts_index_order_queries: null array");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  char* block = nullptr;
  const size_t keys_bytes = align_up(4 * n * sizeof(uint32_t));
  RTS_HIP(scratch_alloc(reinterpret_cast<void**>(&block), keys_bytes + sort_scratch_bytes(n),
                        stream));
  const uint32_t* perm = nullptr;
  // Envelopes and segments share a layout; a segment is ordered by the
  // centre of its bounding box.
  int rc = order_queries(ix, queries, n, points == RTS_QUERY_POINTS, ix->f64(), block,
                         block + keys_bytes, stream, &perm);
  if (!rc) RTS_HIP(hipMemcpyAsync(order, perm, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
  RTS_HIP(hipFreeAsync(block, stream));
  return rc;
}

int rts_index_tree_view(rts_index* ix, rts_stream_t stream, rts_tree_view* out) {
  RTS_CHECK_ARG(ix, "rts_index_tree_view: null index");
  RTS_CHECK_ARG(out, "rts_index_tree_view: null view");
  std::memset(out, 0, sizeof(*out));
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  out->coord = ix->cfg.coord;
  out->leaf_size = ix->cfg.leaf_size;
  if (ix->n == 0) return RTS_OK;
  int rc = ensure_built(ix, stream, true);  // one tree: a delta tree is merged first
  if (rc) return rc;
  const TreeView t = tree_view(ix);
  out->leaf_boxes = t.leaf_boxes;
  out->leaf_ids = t.leaf_ids;
  out->nodes = t.nodes;
  out->root_box = t.root_box;
  out->root_live = t.root_live;
  out->error = t.error;
  out->boxes64 = t.boxes64;
  out->n = t.n;
  return RTS_OK;
}

int rts_index_build(rts_index* ix, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_build: null index");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  return ensure_built(ix, stream);
}

int rts_index_optimize(rts_index* ix, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_optimize: null index");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  return ensure_built(ix, stream, true);
}

int rts_index_sync(rts_index* ix, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_sync: null index");
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  RTS_HIP(hipStreamSynchronize(stream));
  poll_cost(ix);
  return check_device_error(ix);
}

int rts_index_get_stats(rts_index* ix, rts_index_stats* out) {
  RTS_CHECK_ARG(ix && out, "rts_index_get_stats: null argument");
  std::lock_guard<std::mutex> lk(ix->mu);
  poll_cost(ix);
  out->n_boxes = ix->n;
  out->n_tree = ix->n_tree;
  out->delta_boxes = (uint32_t)ix->n_delta;
  out->builds = ix->builds;
  out->refits = ix->refits;
  out->sah_at_build = ix->sah_at_build;
  out->sah_current = ix->sah_current;
  out->stale = (ix->stale || ix->n > ix->n_tree) ? 1 : 0;
  return RTS_OK;
}

int rts_index_memory(rts_index* ix, rts_memory_usage* out) {
  RTS_CHECK_ARG(ix && out, "rts_index_memory: null argument");
  std::lock_guard<std::mutex> lk(ix->mu);
  const uint64_t cap = ix->cap, t = ix->tcap;
  out->geometries = cap * (sizeof(float4) + sizeof(uint32_t)) + (ix->boxes64 ? cap * sizeof(Box64) : 0);
  out->aabbs = t * sizeof(float4);
  out->bvh = t ? (t * sizeof(uint32_t) * 4                      // leaf ids, keys, keys/vals alt
                  + t * sizeof(NodeRec) + 3 * t * sizeof(uint32_t) // nodes, parents, far ends
                  + kTopPairs * sizeof(NodeRec)                     // the tree's top
                  + (t + refit_boundary_words(t)) * sizeof(uint32_t) // flags + boundary list
                  + sort_scratch_bytes(t) + ix->n_partials_cap * sizeof(float4))
               : 0;
  out->total = out->geometries + out->aabbs + out->bvh + 256;
  return RTS_OK;
}

int rts_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                   uint32_t* vals_out, uint64_t n, int key_bits, rts_stream_t stream) {
  RTS_CHECK_ARG(keys_out && vals_out && (keys_in || n == 0), "rts_sort_pairs: null pointer");
  RTS_CHECK_ARG(key_bits >= 1 && key_bits <= 32, "rts_sort_pairs: key_bits must be in [1, 32]");
  RTS_CHECK_ARG(n < (1ull << 30), "rts_sort_pairs: n must be below 2^30");
  if (n == 0) return RTS_OK;
  const size_t scratch = sort_scratch_bytes(n);
  char* base = nullptr;
  RTS_HIP(scratch_alloc(reinterpret_cast<void**>(&base), 2 * n * sizeof(uint32_t) + scratch + 256,
                        stream));
  uint32_t* kb = reinterpret_cast<uint32_t*>(base);
  uint32_t* vb = kb + n;
  char* sbase = reinterpret_cast<char*>(
      (reinterpret_cast<uintptr_t>(vb + n) + 255) & ~uintptr_t(255));
  SortScratch ss;
  sort_scratch_carve(sbase, n, &ss);
  RTS_HIP(hipMemcpyAsync(keys_out, keys_in, n * sizeof(uint32_t), hipMemcpyDefault, stream));
  if (vals_in)
    RTS_HIP(hipMemcpyAsync(vals_out, vals_in, n * sizeof(uint32_t), hipMemcpyDefault, stream));
  bool in_b = false;
  RTS_HIP(radix_sort_pairs(keys_out, vals_out, kb, vb, (uint32_t)n, key_bits, vals_in == nullptr, ss,
                           stream, &in_b));
  if (in_b) {
    RTS_HIP(hipMemcpyAsync(keys_out, kb, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    RTS_HIP(hipMemcpyAsync(vals_out, vb, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
  }
  uint32_t err = 0;
  RTS_HIP(hipMemcpyAsync(&err, ss.error, sizeof(err), hipMemcpyDeviceToHost, stream));
  RTS_HIP(hipFreeAsync(base, stream));
  RTS_HIP(hipStreamSynchronize(stream));
  if (err) {
    set_error("rts_sort_pairs: look-back exceeded its wait bound");
    return RTS_ERR_DEVICE;
  }
  return RTS_OK;
}

void rts_tune_config_default(rts_tune_config* c) {
  if (!c) return;
  c->sample = 32768;
  c->leaf_sizes = (1u << 2) | (1u << 3) | (1u << 4) | (1u << 5);
  c->reps = 2;
  c->apply = 1;
}

int rts_index_autotune(rts_index* ix, rts_predicate pred, const void* queries, uint64_t n,
                       int points, rts_handler handler, const rts_tune_config* cfg_in,
                       rts_tune_result* out, rts_stream_t stream) {
  RTS_CHECK_ARG(ix, "rts_index_autotune: null index");
  RTS_CHECK_ARG(out, "rts_index_autotune: null result");
  std::memset(out, 0, sizeof(*out));
  rts_tune_config cfg;
  rts_tune_config_default(&cfg);
  if (cfg_in) cfg = *cfg_in;
  RTS_CHECK_ARG(cfg.sample > 0 && cfg.reps > 0, "rts_index_autotune: sample and reps must be > 0");
  RTS_CHECK_ARG((cfg.leaf_sizes & 0x7Eu) != 0,
                "rts_index_autotune: leaf_sizes needs a bit in 1..6 (leaf sizes 2..64)");
  const bool pts = points != 0;
  const bool f64 = ix->cfg.coord == RTS_COORD_F64;  // kernels take doubles
  int rc = check_query_args(pred, queries, n, pts, handler, ix);
  if (rc || n == 0) return rc;
  std::lock_guard<std::mutex> lk(ix->mu);
  RTS_HIP(hipSetDevice(ix->device));
  if (ix->n == 0) return RTS_OK;
  rc = ensure_built(ix, stream);
  if (rc) return rc;

  // An evenly strided sample, shrunk until its results fit kMaxPairs (512 MiB
  // of pairs: one timed run then writes at most that much, about 2 ms on the
  // 10M-box index, which keeps a predicate's tuning under 50 ms). A double index's
  // sample is taken in doubles; one that rounds doubles to floats times the
  // rounded sample.
  const size_t qbytes = query_bytes(pts, ix->f64());
  const uint64_t kMaxPairs = 1ull << 26;
  TuneBuffers tb;
  RTS_HIP(hipMalloc(&tb.small, 64));
  RTS_HIP(hipEventCreate(&tb.ev[0]));
  RTS_HIP(hipEventCreate(&tb.ev[1]));
  unsigned long long* counter = reinterpret_cast<unsigned long long*>(tb.small + 32);
  const Plan base = plan_query(ix, pred, pts, n);
  uint64_t m = std::min<uint64_t>(cfg.sample, n);
  uint64_t total = 0;
  for (int attempt = 0;; ++attempt) {
    if (tb.sample) RTS_HIP(hipFree(tb.sample));
    tb.sample = nullptr;
    RTS_HIP(hipMalloc(&tb.sample, m * qbytes));
    RTS_HIP(launch_gather_strided(queries, (uint32_t)qbytes, n / m, tb.sample, (uint32_t)m, stream));
    if (ix->cfg.coord == RTS_COORD_F64_AS_F32) {
      void* rounded = nullptr;
      RTS_HIP(hipMalloc(&rounded, m * query_bytes(pts, false)));
      const hipError_t e = launch_round_queries64(tb.sample, pts, m, rounded, stream);
      (void)hipStreamSynchronize(stream);
      (void)hipFree(tb.sample);
      tb.sample = rounded;
      RTS_HIP(e);
    }
    RTS_HIP(hipMemsetAsync(counter, 0, sizeof(*counter), stream));
    rc = run_plan(ix, base, pred, tb.sample, m, pts, f64, RTS_HANDLER_COUNT, counter, stream);
    if (rc) return rc;
    RTS_HIP(hipMemcpyAsync(&total, counter, sizeof(total), hipMemcpyDeviceToHost, stream));
    RTS_HIP(hipStreamSynchronize(stream));
    if (handler == RTS_HANDLER_COUNT || total <= kMaxPairs || m <= 4096 || attempt == 3) break;
    m = std::max<uint64_t>(4096, (uint64_t)((double)m * (double)kMaxPairs / (double)total * 0.9));
  }
  void* arg = counter;
  if (handler == RTS_HANDLER_COLLECT) {
    const uint64_t cap = std::min(total, kMaxPairs) + 1;
    RTS_HIP(hipMalloc(reinterpret_cast<void**>(&tb.pairs), cap * sizeof(rts_pair)));
    rts_dev_queue q{tb.pairs, cap, counter};
    RTS_HIP(hipMemcpyAsync(tb.small, &q, sizeof(q), hipMemcpyHostToDevice, stream));
    arg = tb.small;
  }

  // Time one plan: the fastest of `reps` runs, after an untimed one when
  // the plan's traversal mode has not run yet (its kernels then load and
  // warm up). A plan whose first run is already 3x slower than the best so
  // far is not repeated.
  double best = 1e300;
  bool warm[3] = {false, false, false};
  auto time_on = [&](const Plan& p, const void* qs, uint64_t qn, void* qarg,
                     double* ms_out) -> int {
    float ms = 0.0f;
    double fastest = 1e300;
    const uint32_t r0 = warm[p.traversal] ? 1u : 0u;
    warm[p.traversal] = true;
    for (uint32_t r = r0; r <= cfg.reps; ++r) {
      RTS_HIP(hipMemsetAsync(counter, 0, sizeof(*counter), stream));
      RTS_HIP(hipEventRecord(tb.ev[0], stream));
      int e = run_plan(ix, p, pred, qs, qn, pts, f64, handler, qarg, stream);
      if (e) return e;
      RTS_HIP(hipEventRecord(tb.ev[1], stream));
      RTS_HIP(hipEventSynchronize(tb.ev[1]));
      RTS_HIP(hipEventElapsedTime(&ms, tb.ev[0], tb.ev[1]));
      if (r == r0 && ms > 3.0 * best) {
        fastest = ms;
        break;
      }
      if (r > 0) fastest = std::min(fastest, (double)ms);
    }
    *ms_out = fastest;
    return RTS_OK;
  };
  auto time_plan = [&](const Plan& p, double* ms_out) -> int {
    return time_on(p, tb.sample, m, arg, ms_out);
  };

  double default_ms = 0;
  rc = time_plan(base, &default_ms);
  if (rc) return rc;
  best = default_ms;
  Plan winner = base;
  uint32_t plans = 1;
  // Traversal modes in the outer loop: a mode whose first plan is already 4x
  // slower than the best is not tried with the other leaf sizes or orders
  // (per-lane traversal of queries with thousands of results each, say).
  // Query ordering is only tuned when the sample is the whole query set: a
  // strided sample has none of the set's locality, which is what Morton
  // ordering exploits, so otherwise the order stays with the call's size rule.
  const bool whole = m == n;
  const int other =
      base.traversal == RTS_TRAVERSAL_WAVE ? (int)RTS_TRAVERSAL_THREAD : (int)RTS_TRAVERSAL_WAVE;
  // A lane emits its query's results one per step, while a wavefront scans
  // 512 leaves per step: queries averaging more than kLaneMaxResults results
  // are not timed one per lane (on the 10M-box index the per-lane probe of
  // intersects queries with ~1e4 results each took 446 ms against 6 ms).
  constexpr double kLaneMaxResults = 256.0;
  const bool lane_hopeless = (double)total > kLaneMaxResults * (double)m;
  for (int trav : {base.traversal, other}) {
    if (trav == RTS_TRAVERSAL_THREAD && lane_hopeless) continue;
    bool first = true;
    for (uint32_t k = 1; k <= 6; ++k) {
      if (!(cfg.leaf_sizes & (1u << k))) continue;
      for (bool sort : {false, true}) {
        if (!whole && sort != base.sort) continue;
        Plan p{trav, sort, 1u << k};
        if (p.traversal == base.traversal && p.sort == base.sort && p.leaf_size == base.leaf_size)
          continue;
        double ms = 0;
        rc = time_plan(p, &ms);
        if (rc) return rc;
        ++plans;
        if (ms < best) {
          best = ms;
          winner = p;
        }
        if (first && ms > 4.0 * best) goto next_mode;
        first = false;
      }
    }
  next_mode:;
  }
  // Query ordering pays through the walks' cache locality, which a strided
  // sample does not have. When the whole set is cheap to run -- at most
  // kOrderTuneQueries queries and, going by the sample, at most kMaxPairs
  // results -- the winning plan is timed on the whole set in both orders.
  // (On the 100k x 100k point query, ordering cost 0.088 ms and saved
  // 0.022 ms of traversal: 0.164 against 0.100 ms per call.)
  bool order_tuned = whole, order = winner.sort;
  const double est_all = (double)total * (double)n / (double)std::max<uint64_t>(m, 1);
  if (!whole && n <= kOrderTuneQueries && ix->cfg.coord != RTS_COORD_F64_AS_F32 &&
      est_all <= (double)kMaxPairs) {
    void* all_arg = counter;
    uint64_t cap = 0;
    if (handler == RTS_HANDLER_COLLECT) {
      cap = (uint64_t)(est_all * 1.5) + 4096;
      RTS_HIP(hipMalloc(reinterpret_cast<void**>(&tb.all_pairs), cap * sizeof(rts_pair)));
      rts_dev_queue q{tb.all_pairs, cap, counter};
      RTS_HIP(hipMemcpyAsync(tb.small + 40, &q, sizeof(q), hipMemcpyHostToDevice, stream));
      all_arg = tb.small + 40;
    }
    double ms_in[2] = {0, 0};
    bool fits = true;
    // Both orders get their full repetitions (the early exit compares
    // against the sample's best time).
    const double sample_best = best;
    best = 1e300;
    for (int o = 0; o < 2 && fits; ++o) {
      Plan p = winner;
      p.sort = o == 1;
      rc = time_on(p, queries, n, all_arg, &ms_in[o]);
      if (rc) return rc;
      if (handler == RTS_HANDLER_COLLECT) {
        unsigned long long got = 0;
        RTS_HIP(hipMemcpyAsync(&got, counter, sizeof(got), hipMemcpyDeviceToHost, stream));
        RTS_HIP(hipStreamSynchronize(stream));
        fits = got <= cap;
      }
    }
    best = sample_best;
    if (fits) {
      order_tuned = true;
      order = ms_in[1] < ms_in[0];
      plans += 2;
    }
  }
  rc = check_device_error(ix);
  if (rc) return rc;
  if (cfg.apply) {
    rts_index::TunedPlan& t = ix->tuned[pred][pts ? 1 : 0];
    t.valid = true;
    t.traversal = (uint32_t)winner.traversal;
    t.sort = order;
    t.sort_tuned = order_tuned;
    // The ordering measured at n queries moves the size rule's threshold:
    // calls of at least min(n, the rule's size) are ordered when ordering
    // won, calls of at most max(n, ...) are not when it lost.
    t.sort_min_n = order ? std::min<uint64_t>(n, kSortMinQueries)
                         : std::max<uint64_t>(n + 1, kSortMinQueries);
    t.leaf_size = winner.leaf_size;
  }
  out->leaf_size = winner.leaf_size;
  out->traversal = (uint32_t)winner.traversal;
  out->sort_queries = order_tuned ? (order ? 1 : 2) : 0;
  out->plans = plans;
  out->sample = m;
  out->sample_results = total;
  out->best_ms = best;
  out->default_ms = default_ms;
  return RTS_OK;
}

}  // extern "C"
