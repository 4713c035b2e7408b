/*
 * rts_c.h -- C ABI of the MI355X-native LibRTS spatial index.
 *
 * This is the drop-in boundary for rtspatial::SpatialIndex<float, 2> and
 * SpatialIndex<double, 2> (reference: include/rtspatial/spatial_index.cuh:62-931). Every entry point
 * is extern "C", takes plain pointers and sizes, returns an int status
 * (RTS_OK = 0) and never throws; rts_last_error() gives the message of the
 * calling thread's last failure. The C++ facade in include/rtspatial/ wraps
 * these calls with the reference's class and method names.
 *
 * Memory: box, id, update and query arrays may live in device memory or in
 * pinned (hipHostMalloc) host memory, as the reference's ArrayView allows
 * (test/test_point_queries.h:62-90 passes pinned_vector views). Insert copies
 * the boxes into index-owned device memory; every other array is borrowed and
 * must stay valid until the work queued on `stream` has completed.
 *
 * Streams: all work is ordered on the caller's stream. Kernel parameters are
 * passed by value per launch, so queries on different streams may run
 * concurrently against an index that is not being mutated.
 */
#ifndef RTS_C_H
#define RTS_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same underlying type as hipStream_t; a hipStream_t converts implicitly. */
typedef struct ihipStream_t* rts_stream_t;

/* Status codes. */
enum {
  RTS_OK = 0,
  RTS_ERR_INVALID = 1,  /* bad argument (null pointer, bad enum, size limit) */
  RTS_ERR_HIP = 2,      /* a HIP runtime call failed */
  RTS_ERR_DEVICE = 3,   /* a kernel reported an internal error (bounded wait expired) */
  RTS_ERR_MISSING = 4   /* no compatible GPU */
};

/* Point<float,2> (geom/point.cuh:11-42): 8 bytes. */
typedef struct rts_point {
  float x, y;
} rts_point;

/* Envelope<Point<float,2>> (geom/envelope.cuh:13-109): {min, max}, 16 bytes. */
typedef struct rts_envelope {
  float x0, y0, x1, y1;
} rts_envelope;

/* thrust::pair<size_t, envelope_t> as taken by Update (spatial_index.cuh:280-281): 24 bytes. */
typedef struct rts_update {
  uint64_t id;
  rts_envelope box;
} rts_update;

/* Point<double,2> (geom/point.cuh:44-75): 16 bytes. */
typedef struct rts_point_d {
  double x, y;
} rts_point_d;

/* Envelope<Point<double,2>>: {min, max}, 32 bytes. */
typedef struct rts_envelope_d {
  double x0, y0, x1, y1;
} rts_envelope_d;

/* thrust::pair<size_t, Envelope<Point<double,2>>>: 40 bytes. */
typedef struct rts_update_d {
  uint64_t id;
  rts_envelope_d box;
} rts_update_d;

/* Line<Point<float,2>> (geom/line.cuh:10-27): the segment p1-p2, 16 bytes. */
typedef struct rts_line {
  rts_point p1, p2;
} rts_line;

/* Line<Point<double,2>>: 32 bytes. */
typedef struct rts_line_d {
  rts_point_d p1, p2;
} rts_line_d;

/* One result of the collecting handler, in its order
 * (default_handlers/collecting_handlers.h:9-15): geom_id first. */
typedef struct rts_pair {
  uint32_t geom_id;
  uint32_t query_id;
} rts_pair;

/* Device-side result queue, the layout of dev::Queue<pair<uint32,uint32>>
 * (utils/queue.h:12-58) with a 64-bit counter: the reference's uint32 counter
 * (queue.h:112) cannot hold the ~1e10 pairs of a 10M x 1M intersects join.
 * Results whose slot is >= capacity are dropped, but the counter still counts
 * them, so after a launch `*counter` is the exact number of results and
 * `*counter > capacity` means the launch overflowed (the reference leaves
 * overflow undefined in release builds). */
typedef struct rts_dev_queue {
  rts_pair* data;
  uint64_t capacity;
  unsigned long long* counter;
} rts_dev_queue;

/* Predicate (geom/predicates.h:7-10) plus one addition.
 *   RTS_INTERSECTS: closed overlap            (envelope.cuh:55-69)
 *   RTS_CONTAINS:   data box contains query   (envelope.cuh:29-53; the
 *                   reference's kContains for points and envelopes,
 *                   shaders_contains_envelope_query_2d.cu:22)
 *   RTS_WITHIN:     data box lies inside the query envelope (the converse of
 *                   RTS_CONTAINS; no counterpart in the reference). */
typedef enum rts_predicate {
  RTS_INTERSECTS = 0,
  RTS_CONTAINS = 1,
  RTS_WITHIN = 2
} rts_predicate;

/* Result handler: how `arg` of a query is interpreted. The reference selects
 * its handler at shader compile time (CMakeLists.txt:73-104); here it is a
 * run-time argument.
 *   RTS_HANDLER_COUNT:   arg is a device `unsigned long long*`; the number of
 *                        results is added to it (counting_handlers.h:7-25).
 *   RTS_HANDLER_COLLECT: arg is a device pointer to an rts_dev_queue; results
 *                        are appended (collecting_handlers.h:9-31). */
typedef enum rts_handler {
  RTS_HANDLER_COUNT = 0,
  RTS_HANDLER_COLLECT = 1
} rts_handler;

/* Rebuild policy for mutations. */
typedef enum rts_rebuild_policy {
  RTS_REBUILD_SAH = 0,     /* refit; rebuild when SAH cost > sah_rebuild_ratio x cost at build,
                              or when updates since the last build cover >= 1/200 of the boxes */
  RTS_REBUILD_NEVER = 1,   /* refit only (the reference's behaviour, spatial_index.cuh:208-336) */
  RTS_REBUILD_ALWAYS = 2   /* rebuild after every mutation */
} rts_rebuild_policy;

/* Traversal mode. */
typedef enum rts_traversal {
  RTS_TRAVERSAL_AUTO = 0,
  RTS_TRAVERSAL_THREAD = 1, /* one query per lane */
  RTS_TRAVERSAL_WAVE = 2    /* one query per wavefront, lanes share its nodes and leaves */
} rts_traversal;

/* Coordinate type of an index: SpatialIndex<float,2> or SpatialIndex<double,2>. */
typedef enum rts_coord {
  RTS_COORD_F32 = 0,
  /* Double coordinates, exact double predicates. The tree holds FP32 bounds
   * rounded outward (the reference's 2-ulp widening, geom/envelope.cuh:131-143,
   * utils/helpers.h:26-41); a leaf test that FP32 cannot decide reads the
   * double box. */
  RTS_COORD_F64 = 1,
  /* Double coordinates rounded to the nearest float, then FP32 predicates: the
   * results of the reference's SpatialIndex<double,2>, whose
   * Point<double,2>::get_coordinate returns float (geom/point.cuh:68). */
  RTS_COORD_F64_AS_F32 = 2
} rts_coord;

/* Config (spatial_index.cuh:22-37). OptiX-only fields are accepted and
 * ignored; the sampling fields drive the selectivity probe of the tuner. */
typedef struct rts_config {
  uint64_t max_geometries;       /* initial box capacity; grows on demand (10M) */
  uint64_t max_queries;          /* queries per launch hint (1M) */
  int32_t preallocate;           /* allocate max_geometries at init */
  int32_t prefer_fast_build_geom;  /* ignored (OptiX build flag) */
  int32_t prefer_fast_build_query; /* ignored (OptiX build flag) */
  int32_t compact;               /* ignored (OptiX compaction) */
  float geom_sample_rate;        /* 0.2 */
  float query_sample_rate;       /* 0.2 */
  float intersect_cost_weight;   /* 0.95 */
  uint32_t max_geom_samples;     /* 10000 */
  uint32_t max_query_samples;    /* 1000 */
  uint32_t max_parallelism;      /* 512 */
  /* MI355X additions */
  uint32_t leaf_size;            /* ranges of <= leaf_size boxes are scanned, not descended (8) */
  uint32_t rebuild_policy;       /* rts_rebuild_policy (RTS_REBUILD_SAH) */
  float sah_rebuild_ratio;       /* 1.5 */
  uint32_t traversal;            /* rts_traversal (AUTO) */
  uint32_t sort_queries;         /* 0 = auto, 1 = Morton-order queries, 2 = keep input order */
  uint32_t coord;                /* rts_coord (RTS_COORD_F32). A float index takes the rts_index_*
                                    calls below, a double one the rts_index_*_d calls. */
  uint32_t insert_delta_pct;     /* boxes inserted since the last full build go to a second,
                                    small tree (built alone, queried with the main one) while they
                                    are at most this percentage of the main tree (2); 0: every
                                    Insert rebuilds the whole tree. RTS_REBUILD_ALWAYS always
                                    rebuilds. */
} rts_config;

typedef struct rts_index rts_index;

/* Fills `cfg` with the reference's defaults plus the additions' defaults. */
void rts_config_default(rts_config* cfg);

/* SpatialIndex::Init (spatial_index.cuh:83-110). cfg may be NULL (defaults).
 * The index lives on the current HIP device. */
int rts_index_create(const rts_config* cfg, rts_index** out);
int rts_index_destroy(rts_index* index);

/* SpatialIndex::Clear (spatial_index.cuh:118-129). Synchronous. */
int rts_index_clear(rts_index* index);

/* SpatialIndex::Insert (spatial_index.cuh:142-196). Boxes get the implicit ids
 * size() .. size()+n-1 (spatial_index.cuh:134-136); ids are never reused. */
int rts_index_insert(rts_index* index, const rts_envelope* boxes, uint64_t n, rts_stream_t stream);

/* SpatialIndex::Delete (spatial_index.cuh:208-265). `ids` are size_t; ids >= size()
 * are ignored (spatial_index.cuh:219). A deleted box never matches again
 * unless Update gives it a new envelope. */
int rts_index_delete(rts_index* index, const uint64_t* ids, uint64_t n, rts_stream_t stream);

/* SpatialIndex::Update (spatial_index.cuh:280-336): box[id] = envelope. Ids >=
 * size() are ignored. */
int rts_index_update(rts_index* index, const rts_update* updates, uint64_t n, rts_stream_t stream);

/* SpatialIndex::Query(Predicate, ArrayView<point_t>, void*, stream)
 * (spatial_index.cuh:352-391). Only RTS_CONTAINS is defined for points; any
 * other predicate is RTS_ERR_INVALID (the reference aborts, :387-389). */
int rts_index_query_points(rts_index* index, rts_predicate pred, const rts_point* queries,
                           uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);

/* SpatialIndex::Query(Predicate, ArrayView<envelope_t>, void*, stream)
 * (spatial_index.cuh:407-515). RTS_INTERSECTS, RTS_CONTAINS, RTS_WITHIN. */
int rts_index_query_envelopes(rts_index* index, rts_predicate pred, const rts_envelope* queries,
                              uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);

/* The same four calls for an index created with coord RTS_COORD_F64 or
 * RTS_COORD_F64_AS_F32 (SpatialIndex<double,2>: spatial_index.cuh:142-515
 * instantiated with COORD_T = double). The float calls on a double index, and
 * these on a float index, return RTS_ERR_INVALID. rts_index_delete takes ids
 * and serves both. */
int rts_index_insert_d(rts_index* index, const rts_envelope_d* boxes, uint64_t n,
                       rts_stream_t stream);
int rts_index_update_d(rts_index* index, const rts_update_d* updates, uint64_t n,
                       rts_stream_t stream);
int rts_index_query_points_d(rts_index* index, rts_predicate pred, const rts_point_d* queries,
                             uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);
int rts_index_query_envelopes_d(rts_index* index, rts_predicate pred,
                                const rts_envelope_d* queries, uint64_t n, rts_handler handler,
                                void* arg, rts_stream_t stream);

/* SpatialIndex::Query(Predicate, ArrayView<line_t>, void*, stream)
 * (spatial_index.cuh:517-547), line segments against data boxes.
 *   RTS_INTERSECTS: the closed segment p1-p2 meets the closed data box,
 *                   decided exactly (any float coordinates; doubles of
 *                   magnitude 0 or 2^-450 .. 2^500). The reference declares
 *                   this query but its intersection program reports nothing
 *                   (shaders_intersects_line_query_2d.cu:14-41; README.md:59
 *                   lists line queries as not yet implemented).
 *   RTS_CONTAINS:   no results, as the reference (spatial_index.cuh:524-526).
 *   RTS_WITHIN:     RTS_ERR_INVALID.
 * A segment with a NaN coordinate meets nothing. _d: a double index. */
int rts_index_query_lines(rts_index* index, rts_predicate pred, const rts_line* queries,
                          uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);
int rts_index_query_lines_d(rts_index* index, rts_predicate pred, const rts_line_d* queries,
                            uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);

/* The kind of the queries of rts_index_query_plan / rts_index_order_queries
 * (their `points` argument). */
typedef enum rts_query_kind {
  RTS_QUERY_ENVELOPES = 0,
  RTS_QUERY_POINTS = 1,
  RTS_QUERY_LINES = 2
} rts_query_kind;

/* The index's tree, for kernels a caller compiles around its own result
 * handler (include/rtspatial/device/query.h). The reference compiles its
 * query shaders around a handler header of rtspatial_handle_* functions
 * (default_handlers/counting_handlers.h, test/CMakeLists.txt:8); here the
 * caller's HIP code instantiates the traversal kernels with its handler and
 * runs them on this view. rts_index_tree_view brings the tree up to date on
 * `stream` (a pending build or refit is queued there, and a small tree of
 * recent inserts is merged into the main one) and fills `out`; the view stays
 * valid until the next call that changes the index (insert, delete, update,
 * clear, build). All pointers are device pointers. */
typedef struct rts_tree_view {
  const void* leaf_boxes;    /* float (x0, y0, x1, y1) per leaf, in tree order */
  const uint32_t* leaf_ids;  /* box id per leaf */
  const void* nodes;         /* n - 1 node records of 64 bytes */
  const void* root_box;      /* float (x0, y0, x1, y1) of the whole tree */
  const uint32_t* root_live; /* live leaves in the tree */
  uint32_t* error;           /* the index's device error word */
  const void* boxes64;       /* RTS_COORD_F64: rts_envelope_d per id; otherwise NULL */
  uint32_t n;                /* leaves (0: nothing to query) */
  uint32_t leaf_size;        /* leaf-range size of the index's config */
  uint32_t coord;            /* rts_coord of the index */
  uint32_t reserved;
} rts_tree_view;

int rts_index_tree_view(rts_index* index, rts_stream_t stream, rts_tree_view* out);

/* The plan the library runs a query call of this kind (`points`: an
 * rts_query_kind) and size under: the config's settings, else the
 * autotuner's stored plan (not for segments), else the defaults.
 * Kernels compiled by the caller follow it to run as the library's do. */
typedef struct rts_query_plan {
  uint32_t traversal;    /* RTS_TRAVERSAL_THREAD or RTS_TRAVERSAL_WAVE */
  uint32_t sort_queries; /* 1 = Morton-ordered, 2 = input order */
  uint32_t leaf_size;    /* leaf-range size */
  uint32_t reserved;
} rts_query_plan;

int rts_index_query_plan(rts_index* index, rts_predicate pred, int points, uint64_t n,
                         rts_query_plan* out);

/* order[i] = index of the query that comes i-th in Morton order of the query
 * points / envelope and segment bounding-box centres (`points`: an
 * rts_query_kind; over the index's scene bounds; the tree must be
 * current, e.g. after rts_index_tree_view on the same stream). `queries` are
 * floats for a float index and doubles for a double one; n < 2^32. Queued on
 * `stream`. */
int rts_index_order_queries(rts_index* index, const void* queries, uint64_t n, int points,
                            uint32_t* order, rts_stream_t stream);

/* Allocation-free query. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The same query as
 * rts_index_query_points / _envelopes / _lines (`kind`: an rts_query_kind;
 * queries are floats on a float index, doubles on an RTS_COORD_F64 index),
 * with the same results, but its scratch (query order, per-query counts and
 * offsets, scan and ticket words) lives in a fixed 32 MiB array that the
 * library holds on each GPU it runs on until the process ends, and it calls
 * no hipMalloc, hipFree or other allocating routine. Limits, each refused
 * with RTS_ERR_INVALID before anything is queued:
 *   - n <= 1048576 queries per call;
 *   - the index is up to date (after rts_index_build, or a query, following
 *     the last Insert / Delete / Update), since a build can allocate;
 *   - not on an RTS_COORD_F64_AS_F32 index (its queries are rounded into an
 *     allocated buffer).
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (on different streams, or from different threads with
 * different indexes) race on it and must be serialised by the caller; calls
 * queued on one stream are ordered and safe. */
int rts_index_query_static(rts_index* index, rts_predicate pred, int kind, const void* queries,
                           uint64_t n, rts_handler handler, void* arg, rts_stream_t stream);

/* Number of ids handed out so far (live and deleted). */
uint64_t rts_index_size(const rts_index* index);

/* Brings the tree up to date now (the index otherwise rebuilds lazily at the
 * next query after an Insert). No reference counterpart: the reference builds
 * inside Insert (spatial_index.cuh:161-190). */
int rts_index_build(rts_index* index, rts_stream_t stream);

/* SpatialIndex::Optimize (spatial_index.cuh:766-826 merges the per-Insert
 * BVHs): brings the tree up to date and merges a small tree of recent inserts
 * into the main one, so queries walk one tree. */
int rts_index_optimize(rts_index* index, rts_stream_t stream);

/* Waits for `stream` and reports any error a kernel of this index flagged. */
int rts_index_sync(rts_index* index, rts_stream_t stream);

typedef struct rts_index_stats {
  uint64_t n_boxes;        /* ids handed out */
  uint64_t n_tree;         /* leaves in the current tree */
  uint64_t builds;         /* full LBVH builds */
  uint64_t refits;         /* refits after Update/Delete */
  double sah_at_build;     /* SAH-style cost right after the last build */
  double sah_current;      /* the last measured cost (after the last refit) */
  int32_t stale;           /* 1 if the next query will build (the whole tree or the small one) */
  uint32_t delta_boxes;    /* boxes in the small tree of recent inserts (0: one tree) */
} rts_index_stats;

int rts_index_get_stats(rts_index* index, rts_index_stats* out);

/* Device memory the index holds, in bytes (SpatialIndex::PrintMemoryUsage,
 * spatial_index.cuh:893-903, reports BVH / geometries / AABBs in MB). */
typedef struct rts_memory_usage {
  uint64_t geometries;  /* boxes by id (float; a double index adds its double boxes) */
  uint64_t aabbs;       /* float boxes in tree order (the leaves) */
  uint64_t bvh;         /* nodes, leaf ids, parents, build keys and scratch */
  uint64_t total;
} rts_memory_usage;

int rts_index_memory(rts_index* index, rts_memory_usage* out);

/* Autotuner. No reference counterpart: it stands where the reference's
 * CalculateBestParallelism sampling (spatial_index.cuh:41-57,690-760) picks a
 * ray parallelism. It times the query over an evenly strided sample of
 * `queries` (shrunk until its results are at most 2^26 pairs) under every
 * candidate plan -- leaf size, one query per lane or per wavefront (per lane
 * only when the sample averages at most 256 results per query) -- with the
 * given handler, and then Morton-ordered against input-ordered queries on the
 * whole set when that is cheap (at most 4M queries and, going by the sample,
 * 2^26 results; a strided sample has none of the set's locality). The
 * ordering measured at n queries moves the size rule's threshold (65536
 * queries): calls of at least min(n, 65536) queries are ordered when it won,
 * calls of at most max(n, 65535) are not when it lost. With `apply` set, it
 * stores the fastest plan in the index for this
 * (predicate, query kind); later queries of that kind, and the batches of
 * rts_serve_run, use it for every setting their config leaves on auto. The
 * plan survives rts_index_clear. Results are identical under every plan;
 * only the time differs. `queries` is rts_point* when `points` is 1, else
 * rts_envelope* (rts_point_d* / rts_envelope_d* for a double index), in
 * device or pinned host memory. Synchronous. */
typedef struct rts_tune_config {
  uint32_t sample;      /* queries probed (32768; all of them if fewer) */
  uint32_t leaf_sizes;  /* bit k set: try leaf size 2^k (bits 2..5: 4, 8, 16, 32) */
  uint32_t reps;        /* timed runs per plan; the fastest counts (2) */
  int32_t apply;        /* store the winner in the index (1) */
} rts_tune_config;

typedef struct rts_tune_result {
  uint32_t leaf_size;      /* the winner */
  uint32_t traversal;      /* RTS_TRAVERSAL_THREAD or RTS_TRAVERSAL_WAVE */
  uint32_t sort_queries;   /* 1 = Morton-ordered, 2 = input order; 0 = not tuned (the whole
                              set was too large to time), the call's size decides */
  uint32_t plans;          /* plans timed */
  uint64_t sample;         /* queries in the sample */
  uint64_t sample_results; /* results of the sample */
  double best_ms;          /* the winner's time over the sample */
  double default_ms;       /* the untuned plan's time over the sample */
} rts_tune_result;

void rts_tune_config_default(rts_tune_config* cfg);
int rts_index_autotune(rts_index* index, rts_predicate pred, const void* queries, uint64_t n,
                       int points, rts_handler handler, const rts_tune_config* cfg,
                       rts_tune_result* out, rts_stream_t stream);

/* Standalone LDS-staged onesweep radix sort of (uint32 key, uint32 value)
 * pairs on bits [0, key_bits) of the key, stable. `vals_in` NULL means values
 * 0..n-1. Outputs must not alias inputs. Scratch is allocated on `stream`. */
int rts_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                   uint32_t* vals_out, uint64_t n, int key_bits, rts_stream_t stream);

/* Checksum of a result-pair array, for checking result sets too large to
 * compare pair by pair (no reference counterpart). For each query id q <
 * n_queries: counts[q] = pairs with that query id, gsums[q] = sum over them
 * of ghash(geom_id) mod 2^64, where ghash(b) = splitmix64(b ^
 * 0xD1B54A32D192ED03) (splitmix64: Steele, Lea and Flood, OOPSLA 2014).
 * Pairs whose query id is >= n_queries are skipped. counts and gsums are
 * device arrays of n_queries words, overwritten. A request's checksum is the
 * sum over its queries of qmult(q) * gsums[q], qmult(q) = splitmix64(q) | 1
 * (rts_request_result.hash). Queued on `stream`. */
int rts_pairs_checksum(const rts_pair* pairs, uint64_t n, uint64_t n_queries, uint64_t* counts,
                       uint64_t* gsums, rts_stream_t stream);

/* ------------------------------------------------------------------------
 * Served / batched execution (no reference counterpart: the reference answers
 * one Query call at a time, spatial_index.cuh:352-515).
 *
 * rts_serve_run drives one request trace against a resident index and reports
 * each request's admission and completion time. Requests become eligible at
 * their arrival time (seconds from the start of the run); update batches are
 * applied at their arrival time, and every request is answered on the index
 * state of its admission (its epoch = update batches applied before it).
 *
 *   RTS_SERVE_BATCHED: continuous batching. Up to `slots` batches are in flight
 *     on separate streams; whenever a slot is free, every arrived request
 *     (split into chunks when large) is admitted while its predicted result
 *     slots fit the slot's arena, and the batch runs as one persistent
 *     work-stealing launch sequence (count, scan, write). Chunks whose slots
 *     overflow the arena are re-launched.
 *     Segment requests, and every request on a double index, are walked by
 *     one count and one write launch each inside their batch (the plain
 *     path's wave kernels), into the batch's counts, scan and arena.
 *   RTS_SERVE_PLAIN: the per-request path: each request (split the same way)
 *     is one rts_index_query_* call on one stream, in arrival order.
 *
 * The index holds float coordinates (RTS_COORD_F32) or exact doubles
 * (RTS_COORD_F64; requests then hold rts_point_d / rts_envelope_d /
 * rts_line_d arrays); RTS_COORD_F64_AS_F32 is RTS_ERR_INVALID here.
 * ------------------------------------------------------------------------ */
typedef enum rts_serve_mode {
  RTS_SERVE_BATCHED = 0,
  RTS_SERVE_PLAIN = 1
} rts_serve_mode;

typedef struct rts_request {
  double arrival_s;      /* eligible from this time (seconds after the run starts) */
  const void* queries;   /* device (or pinned) rts_point or rts_envelope array */
  uint64_t n;            /* queries */
  int32_t points;        /* query kind: 0 = envelopes, 1 = points (predicate RTS_CONTAINS),
                            2 = line segments (RTS_INTERSECTS, or RTS_CONTAINS: none) */
  int32_t predicate;     /* rts_predicate */
  rts_pair* out;         /* optional device buffer for the request's pairs, (geom_id,
                            query index within the request); NULL = results are
                            dropped once complete */
  uint64_t out_capacity; /* pairs `out` holds */
} rts_request;

typedef struct rts_update_batch {
  double arrival_s;
  const rts_update* updates; /* device (or pinned); rts_update_d records on a double index */
  uint64_t n;
} rts_update_batch;

typedef struct rts_request_result {
  double admit_s;        /* first launch of the request */
  double done_s;         /* all of its results complete */
  uint64_t count;        /* results */
  uint64_t hash;         /* with hash_results: checksum of the request's pairs, the sum
                            over them of qmult(q) * ghash(geom_id) mod 2^64, q = query
                            index within the request (see rts_pairs_checksum) */
  uint32_t epoch;        /* update batches applied before the request ran */
  uint32_t relaunches;   /* chunk re-launches after an arena overflow */
  int32_t truncated;     /* `out` was too small for the results */
  int32_t pad;
} rts_request_result;

typedef struct rts_serve_config {
  uint32_t mode;               /* rts_serve_mode */
  uint32_t slots;              /* batches in flight (batched mode) (3) */
  uint32_t max_batch_requests; /* request chunks per batch (4096) */
  uint32_t use_graphs;         /* must be 0: hipGraph replay of a slot's launches is refused
                                  (RTS_ERR_INVALID), because a replay run did not complete
                                  on the MI355X and the cause is not established */
  uint64_t max_batch_queries;  /* queries per batch (4M). Each slot also holds a range log of
                                  192 8-byte entries per query (6.4 GB per slot at 4M) */
  uint64_t chunk_queries;      /* largest chunk of one request (256K) */
  uint64_t arena_pairs;        /* result arena per slot (or of the plain path), pairs. A
                                  batched envelope query takes slots for a bound on its
                                  results (its leaf ranges that need tests counted whole,
                                  about 1.2x its results on the served mix) */
  uint32_t hash_results;       /* compute each request's pair hash (verification runs) */
  uint32_t traversal;          /* rts_traversal for batched tiles: AUTO picks by predicate */
  uint32_t head_share_pct;     /* while later requests wait, the largest share of a batch's
                                  remaining result budget one request may take, in percent
                                  (50); requests that do not fit are skipped for later ones */
  uint32_t sort_min_tiles;     /* batched: Morton-order the wave tiles of a batch that has at
                                  least this many (4096; larger batches gain L2 locality from
                                  the order, smaller ones would mostly pay the sort's launches) */
} rts_serve_config;

typedef struct rts_serve_stats {
  double wall_s;          /* first arrival to last completion */
  uint64_t requests;
  uint64_t queries;
  uint64_t results;
  uint64_t batches;       /* launches (batches, or plain chunk queries) */
  uint64_t relaunches;
  uint64_t updates;
} rts_serve_stats;

void rts_serve_config_default(rts_serve_config* cfg);

/* Runs the trace to completion (synchronous). `results` has n_requests
 * entries; `stats` may be NULL. Requests and updates must be sorted by
 * arrival time. */
int rts_serve_run(rts_index* index, const rts_serve_config* cfg, const rts_request* requests,
                  uint64_t n_requests, const rts_update_batch* updates, uint64_t n_updates,
                  rts_request_result* results, rts_serve_stats* stats);

/* Per-phase GPU timing: HIP events recorded on the caller's stream around
 * each phase's kernels. No reference counterpart (the reference's
 * IntersectsWhatQueryProfiling, spatial_index.cuh:566-683, times with host
 * stopwatches and stream syncs). */
enum {
  RTS_PHASE_BUILD_MORTON = 0,  /* scene bounds + Morton codes */
  RTS_PHASE_BUILD_SORT = 1,    /* onesweep radix sort of the codes */
  RTS_PHASE_BUILD_TREE = 2,    /* leaf gather + Karras hierarchy + refit */
  RTS_PHASE_REFIT = 3,         /* refit after Update/Delete */
  RTS_PHASE_MUTATE = 4,        /* Update/Delete scatter */
  RTS_PHASE_QUERY_ORDER = 5,   /* Morton ordering of the queries */
  RTS_PHASE_TRAVERSE = 6,      /* the traversal kernel that emits the results */
  RTS_PHASE_QUERY_COUNT = 7,   /* count pass + scan + queue reservation (two-pass queries) */
  RTS_PHASE_BATCH_COUNT = 8,   /* batched: count pass + scan + request ranges */
  RTS_PHASE_BATCH_WRITE = 9,   /* batched: write pass */
  RTS_PHASE_COUNT = 12
};

typedef struct rts_profile {
  uint64_t count[RTS_PHASE_COUNT];
  double ms[RTS_PHASE_COUNT];
} rts_profile;

/* Turns phase timing on or off (process-wide; off by default). */
void rts_profile_enable(int on);

/* Waits for the recorded events, adds them to the totals, copies the totals
 * to `out` (may be NULL) and clears them if `reset`. */
int rts_profile_read(rts_profile* out, int reset);

/* Thread-local message of the last failing call ("" if none). */
const char* rts_last_error(void);

/* Library version string. */
const char* rts_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RTS_C_H */
