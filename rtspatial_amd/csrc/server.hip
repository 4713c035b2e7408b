// Served and plain execution of a request trace against a resident index
// (rts_serve_run). No reference counterpart: the reference answers one
// Query call at a time (spatial_index.cuh:352-515).
//
// Batched mode is continuous batching over `slots` streams. Each slot owns
// device buffers for one batch (descriptors, per-query result bounds, exact
// counts and offsets, result arena) and a pinned staging area; its launch
// sequence (descriptor upload, count pass, scan, request ranges, write pass,
// per-request results, range download) is
// launched directly for every batch. (Replaying that sequence from a hipGraph
// did not complete on the MI355X, with the cause not established, so
// rts_serve_run refuses use_graphs.)
//
// Admission: arrived requests are taken in arrival order and split into
// chunks of at most chunk_queries queries; chunks join the batch while their
// predicted arena slots (an estimate of slots per query for each predicate,
// refreshed from every completed chunk; a query's slots are a bound on its
// results) fit the slot's arena. While later
// requests wait, one request takes at most head_share_pct percent of the
// remaining budget, and
// requests that do not fit are skipped for later ones that do (backfill). A
// chunk whose slots overflow the arena is re-launched in a later batch.
//
// Updates: once an update batch has arrived, no new request is admitted; when
// the requests already started are fully admitted and their batches are done
// (none can need a re-launch any more), the update is applied on an update
// stream and later batches wait for it. So each request runs entirely on the
// index state of its admission (its epoch).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "index_state.h"
#include "rts_internal.h"
#include "rtspatial/device/traverse.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

#define SRV_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(std::string("serve: ") + #call + ": " + hipGetErrorString(e_));    \
      return RTS_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

using Clock = std::chrono::steady_clock;

// RTS_SERVE_TRACE=1 prints each step of the host loop to stderr.
bool trace_on() {
  static const bool on = [] {
    const char* v = std::getenv("RTS_SERVE_TRACE");
    return v && v[0] == '1';
  }();
  return on;
}
#define SRV_TRACE(...)                          \
  do {                                          \
    if (trace_on()) {                           \
      std::fprintf(stderr, "[rts_serve] " __VA_ARGS__); \
      std::fputc('\n', stderr);                \
      std::fflush(stderr);                      \
    }                                           \
  } while (0)

constexpr uint64_t kUnknown = ~0ull;

struct Chunk {
  uint32_t req;
  uint64_t q0, n;
  uint64_t known;        // arena slots it takes, when known (a re-launch), else kUnknown
  uint64_t known_exact;  // its results, when known
};

struct ReqState {
  uint64_t next_q = 0;      // first query not yet admitted
  uint64_t chunks_out = 0;  // admitted chunks not yet complete
  uint64_t filled = 0;      // results copied to `out`
  bool admitted = false;
  bool done = false;
};

// Results-per-query estimate for one kind of request.
struct Estimate {
  double per_query = 1.0;
  bool seeded = false;
  void observe(uint64_t results, uint64_t queries) {
    if (!queries) return;
    const double v = (double)results / (double)queries;
    per_query = seeded ? 0.75 * per_query + 0.25 * v : v;
    seeded = true;
  }
};

// rts_request.points: the query kind.
constexpr int kEnvelopes = 0, kPoints = 1, kSegments = 2;

// Estimate classes: envelope predicates 0..2, points 3, segments 4.
constexpr int kKinds = 5;
int kind_of(const rts_request& r) {
  return r.points == kPoints ? 3 : r.points == kSegments ? 4 : (r.predicate & 3);
}

// Bytes of one query of r: float or double (a double index takes doubles).
size_t query_bytes_of(const rts_index* ix, const rts_request& r) {
  return (r.points == kPoints ? 8 : 16) * (ix->f64() ? 2 : 1);
}

// Requests the batch kernels do not walk: segments, and every query on a
// double index (BatchExt).
bool is_ext(const rts_index* ix, const rts_request& r) {
  return r.points == kSegments || ix->f64();
}

// One query call of the plain path for n queries of r starting at qp.
int query_request(rts_index* ix, const rts_request& r, const void* qp, uint64_t n,
                  rts_handler handler, void* arg, hipStream_t stream) {
  const rts_predicate pred = (rts_predicate)r.predicate;
  if (r.points == kPoints)
    return ix->f64() ? rts_index_query_points_d(ix, pred, static_cast<const rts_point_d*>(qp), n,
                                                handler, arg, stream)
                     : rts_index_query_points(ix, pred, static_cast<const rts_point*>(qp), n,
                                              handler, arg, stream);
  if (r.points == kSegments)
    return ix->f64() ? rts_index_query_lines_d(ix, pred, static_cast<const rts_line_d*>(qp), n,
                                               handler, arg, stream)
                     : rts_index_query_lines(ix, pred, static_cast<const rts_line*>(qp), n, handler,
                                             arg, stream);
  return ix->f64() ? rts_index_query_envelopes_d(ix, pred, static_cast<const rts_envelope_d*>(qp),
                                                 n, handler, arg, stream)
                   : rts_index_query_envelopes(ix, pred, static_cast<const rts_envelope*>(qp), n,
                                               handler, arg, stream);
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  char* h_desc = nullptr;                // pinned: [TreeView | BatchHeader | BatchRequest x max]
  uint64_t* h_ranges = nullptr;          // pinned: 3 x max
  unsigned long long* h_hash = nullptr;  // pinned: max
  char* d_desc = nullptr;
  uint32_t* d_counters = nullptr;
  uint32_t* d_giants = nullptr;  // kGiantBuckets x max_batch_queries tile ids
  uint2* d_log = nullptr;        // range-log entries (kServeLogPerQuery per query)
  uint32_t* d_log_off = nullptr; // per query: first range-log entry
  uint32_t* d_log_n = nullptr;   // per query: range-log entries
  uint32_t* d_counts = nullptr;  // per query: result bound (BatchArgs)
  uint32_t* d_exact = nullptr;   // per query: results
  uint64_t* d_offsets = nullptr;
  uint64_t* d_exact_off = nullptr;
  uint64_t* d_seg = nullptr;
  uint64_t* d_total = nullptr;
  uint64_t* d_ranges = nullptr;
  unsigned long long* d_hash = nullptr;
  rts_pair* d_arena = nullptr;
  uint32_t* d_sort = nullptr;  // Morton ordering of the wave tiles: 4 key/value arrays + scratch
  // Extension requests: [0, 24) rts_dev_queue over the arena, [32, 40) a 0,
  // [64, 64 + 4 kWaveTicketWords) wave tickets.
  char* d_ext = nullptr;
  std::vector<BatchExt> ext;
  std::vector<Chunk> chunks;
  bool busy = false;
};

constexpr size_t kViewOff = 0, kHeaderOff = 128, kReqsOff = 256;
static_assert(sizeof(TreeView) <= kHeaderOff, "tree view fits its slot");

struct Engine {
  rts_index* ix = nullptr;
  rts_serve_config cfg{};
  const rts_request* reqs = nullptr;
  uint64_t n_reqs = 0;
  const rts_update_batch* ups = nullptr;
  uint64_t n_ups = 0;
  rts_request_result* res = nullptr;
  std::vector<ReqState> st;
  Estimate est[kKinds];
  Clock::time_point t0;
  size_t desc_bytes = 0;
  uint32_t epoch = 0;
  uint64_t batches = 0, relaunches = 0, applied_updates = 0, done = 0;
  std::deque<uint32_t> queue;  // requests with queries not yet admitted, in arrival order
  std::deque<Chunk> retry;     // overflowed chunks, with their exact counts
  uint64_t partial = 0;        // requests admitted in part
  bool hold_new = false;       // an update is waiting: admit new requests only whole
  double last_progress = 0;    // time of the last completion (watchdog)
  hipStream_t upd_stream = nullptr;
  hipEvent_t upd_event = nullptr;
  bool have_upd_event = false;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

double predicted(const Engine& E, const Chunk& c) {
  if (c.known != kUnknown) return (double)c.known;
  return E.est[kind_of(E.reqs[c.req])].per_query * 1.5 * (double)c.n + 16.0;
}

// A chunk's results are final.
// `slots`: the arena slots its results took (the batched path sizes each
// query's slots by a bound on its results); the estimates predict slots.
void finish_chunk(Engine& E, const Chunk& c, uint64_t count, uint64_t slots, uint64_t hash,
                  double t) {
  ReqState& s = E.st[c.req];
  rts_request_result& R = E.res[c.req];
  R.count += count;
  R.hash += hash;
  E.last_progress = t;
  E.est[kind_of(E.reqs[c.req])].observe(slots, c.n);
  if (--s.chunks_out == 0 && s.next_q >= E.reqs[c.req].n) {
    s.done = true;
    R.done_s = t;
    E.done++;
  }
}

// Fills `out` with re-launches first, then chunks of arrived requests, while
// they fit: at most max_req chunks, max_q queries and `budget` predicted
// results. With `force`, one chunk is taken even if it does not fit. With
// `fifo` (the plain path), requests are taken strictly in arrival order.
void admit(Engine& E, double t, uint32_t max_req, uint64_t max_q, double budget, bool force,
           bool fifo, std::vector<Chunk>& out) {
  out.clear();
  uint64_t q_used = 0;
  double used = 0;
  while (!E.retry.empty() && out.size() < max_req) {
    Chunk c = E.retry.front();
    if (c.known != kUnknown && c.known > E.cfg.arena_pairs && c.n == 1) {
      // One query that needs more slots than an arena holds: report its
      // exact count, mark the request truncated and drop its pairs.
      E.retry.pop_front();
      E.res[c.req].truncated = 1;
      finish_chunk(E, c, c.known_exact, c.known, 0, t);
      continue;
    }
    if (c.known != kUnknown && (double)c.known > (double)E.cfg.arena_pairs && c.n > 1) {
      // It can never fit one arena: split it; the halves' counts are unknown.
      E.retry.pop_front();
      const uint64_t h = c.n / 2;
      E.retry.push_front(Chunk{c.req, c.q0 + h, c.n - h, kUnknown, kUnknown});
      E.retry.push_front(Chunk{c.req, c.q0, h, kUnknown, kUnknown});
      E.st[c.req].chunks_out++;
      continue;
    }
    const double p = predicted(E, c);
    if ((used + p > budget || q_used + c.n > max_q) && !(force && out.empty())) return;
    out.push_back(c);
    E.retry.pop_front();
    used += p;
    q_used += c.n;
  }
  // Arrived requests in arrival order. Unless `fifo`, a request may take at
  // most head_share_pct percent of the batch's remaining result budget while
  // later requests wait behind it, and a request that does not fit is skipped rather than
  // blocking the ones behind it (backfill): small requests that arrive during
  // a large one run beside its chunks instead of after all of them.
  for (size_t j = 0; j < E.queue.size() && out.size() < max_req;) {
    const uint32_t ri = E.queue[j];
    const rts_request& r = E.reqs[ri];
    if (r.arrival_s > t) break;
    ReqState& s = E.st[ri];
    // While an update waits, a new request is admitted only whole (it then
    // completes on the current epoch) and only while requests admitted
    // before it are still being admitted: no request is left partly admitted
    // across the update, and the hold ends when those requests are in.
    // (Holding every new request: 30.1-30.6 Mq/s saturated under the 10 Hz
    // update stream; admitting whole ones: 31.8-32.3, against 32.7 with no
    // updates.)
    const bool whole_only = E.hold_new && !s.admitted;
    if (whole_only && (fifo || E.partial == 0)) {
      if (fifo) break;
      ++j;
      continue;
    }
    const uint64_t left = r.n - s.next_q;
    if (left == 0) {  // an empty request completes on admission
      E.queue.erase(E.queue.begin() + j);
      if (!s.done) {
        s.admitted = s.done = true;
        E.res[ri].admit_s = E.res[ri].done_s = t;
        E.res[ri].epoch = E.epoch;
        E.done++;
      }
      continue;
    }
    const bool others =
        !fifo && j + 1 < E.queue.size() && E.reqs[E.queue[j + 1]].arrival_s <= t;
    const double per = E.est[kind_of(r)].per_query * 1.5;
    uint64_t k = std::min<uint64_t>(left, E.cfg.chunk_queries);
    k = std::min<uint64_t>(k, max_q > q_used ? max_q - q_used : 0);
    const double room = (budget - used - 16.0) * (others ? E.cfg.head_share_pct / 100.0 : 1.0);
    if (per > 0) k = std::min<uint64_t>(k, (uint64_t)std::max(0.0, room / per));
    if (whole_only && k < left) {
      ++j;
      continue;
    }
    if (k == 0) {
      if (force && out.empty()) {
        k = 1;
      } else if (fifo) {
        break;
      } else {
        ++j;
        continue;
      }
    }
    const Chunk c{ri, s.next_q, k, kUnknown, kUnknown};
    out.push_back(c);
    if (!s.admitted) {
      s.admitted = true;
      E.res[ri].admit_s = t;
      E.res[ri].epoch = E.epoch;
      E.partial++;
    }
    s.next_q += k;
    s.chunks_out++;
    q_used += k;
    used += predicted(E, c);
    if (s.next_q >= r.n) {
      E.queue.erase(E.queue.begin() + j);
      E.partial--;
    } else {
      ++j;
    }
  }
}

void overflowed(Engine& E, const Chunk& c, uint64_t slots, uint64_t count) {
  E.retry.push_back(Chunk{c.req, c.q0, c.n, slots, count});
  E.res[c.req].relaunches++;
  E.relaunches++;
  E.est[kind_of(E.reqs[c.req])].observe(slots, c.n);
}

// Copies a finished chunk's cnt pairs to the request's `out`: with `batch`,
// those of the batch's queries [qbase, qbase + c.n) (each query's pairs lie
// at the start of its slots), else the contiguous src[0, cnt).
int copy_out(Engine& E, const Chunk& c, const rts_pair* src, const BatchArgs* batch,
             uint64_t qbase, uint64_t cnt, hipStream_t stream) {
  const rts_request& r = E.reqs[c.req];
  if (!r.out || !cnt) return RTS_OK;
  ReqState& st = E.st[c.req];
  uint64_t take = cnt;
  if (st.filled + take > r.out_capacity) {
    take = r.out_capacity > st.filled ? r.out_capacity - st.filled : 0;
    E.res[c.req].truncated = 1;
  }
  if (take) {
    if (batch)
      SRV_HIP(launch_copy_queries(*batch, qbase, c.n, r.out + st.filled, take, stream));
    else
      SRV_HIP(hipMemcpyAsync(r.out + st.filled, src, take * sizeof(rts_pair),
                             hipMemcpyDeviceToDevice, stream));
    SRV_HIP(hipStreamSynchronize(stream));
  }
  st.filled += take;
  return RTS_OK;
}

// Tile mode and leaf-range size of a request: the serve config's traversal
// if it names one, else the plan stored by rts_index_autotune for this
// (predicate, query kind), else one query per wavefront for envelope
// intersects / within and 64 per tile otherwise.
uint32_t tile_flags(const Engine& E, const rts_request& r) {
  if (is_ext(E.ix, r)) return kReqExt;
  const bool pts = r.points == kPoints;
  uint32_t f = (pts ? kReqPoints : 0u) | ((uint32_t)(r.predicate & 3) << kReqPredShift);
  const rts_index::TunedPlan& tp = E.ix->tuned[r.predicate & 3][pts ? 1 : 0];
  const uint32_t trav = E.cfg.traversal;
  bool wave;
  if (trav == RTS_TRAVERSAL_WAVE || trav == RTS_TRAVERSAL_THREAD)
    wave = trav == RTS_TRAVERSAL_WAVE;
  else if (tp.valid)
    wave = tp.traversal == RTS_TRAVERSAL_WAVE;
  else
    wave = !pts && r.predicate != RTS_CONTAINS;
  if (wave) f |= kReqWaveTiles;
  if (tp.valid && tp.leaf_size >= 1 && tp.leaf_size <= 64) {
    uint32_t lg = 0;
    while ((2u << lg) <= tp.leaf_size) lg++;
    f |= (lg + 1) << kReqLeafShift;
  }
  return f;
}

// Range-log entries of a slot (8 bytes each): 192 per query of a full batch
// (envelope queries of the served mix log a few hundred leaf ranges each; a
// query past its wave's share is walked again), at most 2^30.
uint64_t serve_log_entries(uint64_t max_batch_queries) {
  return std::min<uint64_t>(max_batch_queries * 192, 1ull << 30);
}

BatchArgs slot_args(const Engine& E, const Slot& s) {
  BatchArgs a;
  a.header = reinterpret_cast<const BatchHeader*>(s.d_desc + kHeaderOff);
  a.requests = reinterpret_cast<const BatchRequest*>(s.d_desc + kReqsOff);
  a.tile_counters = s.d_counters;
  a.counts = s.d_counts;
  a.exact = s.d_exact;
  a.offsets = s.d_offsets;
  a.exact_offsets = s.d_exact_off;
  a.seg_sums = s.d_seg;
  a.total = s.d_total;
  a.ranges = s.d_ranges;
  a.arena = s.d_arena;
  a.arena_cap = E.cfg.arena_pairs;
  a.order = nullptr;
  a.giants = s.d_giants;
  a.giant_cap = (uint32_t)E.cfg.max_batch_queries;
  a.log.entries = s.d_log;
  a.log.n_entries = serve_log_entries(E.cfg.max_batch_queries);
  a.log.off = s.d_log_off;
  a.log.n = s.d_log_n;
  return a;
}

// One batch: descriptors up, count / scan / write, ranges (and hashes) down.
int record_slot_work(Engine& E, Slot& s, size_t up_bytes, size_t n_req) {
  SRV_HIP(hipMemcpyAsync(s.d_desc, s.h_desc, up_bytes, hipMemcpyHostToDevice, s.stream));
  const BatchArgs a = slot_args(E, s);
  // Morton-order the wave tiles of large batches.
  const BatchHeader* h = reinterpret_cast<const BatchHeader*>(s.h_desc + kHeaderOff);
  const uint32_t wave_tiles = h->class_end[kClassWavePoints];
  BatchSort sort{};
  const uint64_t maxq = E.cfg.max_batch_queries;
  sort.keys_a = s.d_sort;
  sort.vals_a = s.d_sort + maxq;
  sort.keys_b = s.d_sort + 2 * maxq;
  sort.vals_b = s.d_sort + 3 * maxq;
  sort_scratch_carve(s.d_sort + 4 * maxq, maxq, &sort.ss);
  sort.ss.error = E.ix->error;  // a look-back that gives up is reported after the run
  sort.scene = E.ix->scene;
  sort.n = wave_tiles >= E.cfg.sort_min_tiles ? wave_tiles : 0;
  const BatchExtArgs ea{reinterpret_cast<uint32_t*>(s.d_ext + 64),
                        reinterpret_cast<const uint64_t*>(s.d_ext + 32), s.d_ext};
  // Requests that copy their pairs out need the compacted offsets.
  bool collect = false;
  for (const Chunk& c : s.chunks) collect = collect || E.reqs[c.req].out != nullptr;
  SRV_HIP(launch_batch(reinterpret_cast<const TreeView*>(s.d_desc + kViewOff), a, sort,
                       s.ext.data(), s.ext.size(), ea, collect, s.stream));
  SRV_HIP(hipMemcpyAsync(s.h_ranges, s.d_ranges, 3 * n_req * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, s.stream));
  if (E.cfg.hash_results) {
    SRV_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(s.d_hash), 0u, 2 * n_req, s.stream));
    SRV_HIP(launch_hash_queries(a, s.d_hash, s.stream));
    SRV_HIP(hipMemcpyAsync(s.h_hash, s.d_hash, n_req * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s.stream));
  }
  return RTS_OK;
}

int slot_init(Engine& E, Slot& s) {
  const uint32_t maxr = E.cfg.max_batch_requests;
  const uint64_t maxq = E.cfg.max_batch_queries;
  SRV_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  SRV_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_desc), E.desc_bytes));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_ranges), 3 * maxr * sizeof(uint64_t)));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_hash), maxr * sizeof(unsigned long long)));
  SRV_HIP(hipMalloc(&s.d_desc, E.desc_bytes));
  SRV_HIP(hipMalloc(&s.d_counters, kTileCounterWords * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_giants, kGiantBuckets * maxq * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_log, serve_log_entries(maxq) * sizeof(uint2)));
  SRV_HIP(hipMalloc(&s.d_log_off, maxq * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_log_n, maxq * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_counts, maxq * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_exact, maxq * sizeof(uint32_t)));
  SRV_HIP(hipMalloc(&s.d_offsets, maxq * sizeof(uint64_t)));
  SRV_HIP(hipMalloc(&s.d_exact_off, maxq * sizeof(uint64_t)));
  SRV_HIP(hipMalloc(&s.d_seg, batch_seg_sums_words() * sizeof(uint64_t)));
  // [0] bound total, [1] the extension launches' queue counter, [2] exact total
  SRV_HIP(hipMalloc(&s.d_total, 4 * sizeof(uint64_t)));
  SRV_HIP(hipMalloc(&s.d_ranges, 3 * maxr * sizeof(uint64_t)));
  SRV_HIP(hipMalloc(&s.d_hash, maxr * sizeof(unsigned long long)));
  SRV_HIP(hipMalloc(&s.d_arena, E.cfg.arena_pairs * sizeof(rts_pair)));
  SRV_HIP(hipMalloc(&s.d_sort, 4 * maxq * sizeof(uint32_t) + sort_scratch_bytes(maxq)));
  SRV_HIP(hipMalloc(&s.d_ext, 64 + kWaveTicketWords * sizeof(uint32_t)));
  {
    // The arena as a result queue for the extension launches' write pass
    // (which stores at exact offsets and never touches the counter).
    const rts_dev_queue q{s.d_arena, E.cfg.arena_pairs,
                          reinterpret_cast<unsigned long long*>(s.d_total + 1)};
    SRV_HIP(hipMemset(s.d_ext, 0, 64));
    SRV_HIP(hipMemcpy(s.d_ext, &q, sizeof(q), hipMemcpyHostToDevice));
  }
  std::memset(s.h_desc, 0, E.desc_bytes);
  SRV_TRACE("slot buffers ready (arena %llu pairs)", (unsigned long long)E.cfg.arena_pairs);
  return RTS_OK;
}

void slot_free(Slot& s) {
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  if (s.h_desc) (void)hipHostFree(s.h_desc);
  if (s.h_ranges) (void)hipHostFree(s.h_ranges);
  if (s.h_hash) (void)hipHostFree(s.h_hash);
  for (void* p : {(void*)s.d_desc, (void*)s.d_counters, (void*)s.d_giants, (void*)s.d_log,
                  (void*)s.d_log_off, (void*)s.d_log_n, (void*)s.d_counts,
                  (void*)s.d_exact, (void*)s.d_offsets, (void*)s.d_exact_off,
                  (void*)s.d_seg, (void*)s.d_total, (void*)s.d_ranges, (void*)s.d_hash,
                  (void*)s.d_arena, (void*)s.d_sort, (void*)s.d_ext})
    if (p) (void)hipFree(p);
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = Slot();
}

// Seeds the per-kind estimates with a counting query over a sample of the
// first request of each kind.
int probe_estimates(Engine& E, hipStream_t stream) {
  unsigned long long* d_cnt = nullptr;
  unsigned long long* h_cnt = nullptr;
  SRV_HIP(hipMalloc(&d_cnt, sizeof(unsigned long long)));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_cnt), sizeof(unsigned long long)));
  bool seen[kKinds] = {false, false, false, false, false};
  int rc = RTS_OK;
  for (uint64_t i = 0; i < E.n_reqs && !rc; i++) {
    const rts_request& r = E.reqs[i];
    const int k = kind_of(r);
    if (seen[k] || r.n == 0) continue;
    seen[k] = true;
    const uint64_t m = std::min<uint64_t>(r.n, 4096);
    if (launch_fill_u32(reinterpret_cast<uint32_t*>(d_cnt), 0u, 2, stream) != hipSuccess) {
      rc = RTS_ERR_HIP;
      break;
    }
    rc = query_request(E.ix, r, r.queries, m, RTS_HANDLER_COUNT, d_cnt, stream);
    if (rc) break;
    if (hipMemcpyAsync(h_cnt, d_cnt, sizeof(*h_cnt), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      rc = RTS_ERR_HIP;
      break;
    }
    E.est[k].observe(*h_cnt, m);
  }
  (void)hipFree(d_cnt);
  (void)hipHostFree(h_cnt);
  return rc;
}

// An arrived update waits until no request is admitted in part, nothing is in
// flight (an in-flight chunk may still overflow and need a re-launch) and no
// chunk waits for a re-launch, so that every request runs on one index state;
// meanwhile new requests are admitted only whole, and none once every earlier
// request is fully admitted (see admit), so the wait ends.
bool update_ready(Engine& E, uint64_t next_up, double t, bool idle) {
  if (next_up >= E.n_ups || E.ups[next_up].arrival_s > t) {
    E.hold_new = false;
    return false;
  }
  E.hold_new = true;
  return idle && E.partial == 0 && E.retry.empty();
}

// An update batch on the index: rts_update records, or rts_update_d ones on
// a double index.
int update_index(Engine& E, const rts_update_batch& u, hipStream_t stream) {
  return E.ix->f64()
             ? rts_index_update_d(E.ix, reinterpret_cast<const rts_update_d*>(u.updates), u.n,
                                  stream)
             : rts_index_update(E.ix, u.updates, u.n, stream);
}

int apply_update(Engine& E, const rts_update_batch& u, const std::vector<Slot>& slots) {
  for (const Slot& s : slots)
    if (s.busy) SRV_HIP(hipStreamWaitEvent(E.upd_stream, s.done, 0));
  int rc = update_index(E, u, E.upd_stream);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(E.ix->mu);
    rc = ensure_built(E.ix, E.upd_stream, true);
  }
  if (rc) return rc;
  SRV_HIP(hipEventRecord(E.upd_event, E.upd_stream));
  E.have_upd_event = true;
  E.epoch++;
  E.applied_updates++;
  return RTS_OK;
}

// The host loop gives up if work is in flight but nothing has completed for
// this long, so that rts_serve_run always returns.
constexpr double kStallSeconds = 120.0;

bool stalled(Engine& E, bool in_flight) {
  const double t = E.now();
  if (!in_flight) {
    E.last_progress = t;
    return false;
  }
  if (t - E.last_progress > kStallSeconds) {
    set_error("rts_serve_run: no batch completed for 120 s");
    return true;
  }
  return false;
}

// ------------------------------------------------------------ batched ----
int complete_slot(Engine& E, Slot& s) {
  const double tc = E.now();
  SRV_TRACE("batch of %zu chunks complete at %.6f s", s.chunks.size(), tc);
  for (size_t i = 0; i < s.chunks.size(); i++) {
    const Chunk& c = s.chunks[i];
    // {first slot, slots (the queries' result bounds), results}
    const uint64_t start = s.h_ranges[3 * i], slots = s.h_ranges[3 * i + 1];
    const uint64_t cnt = s.h_ranges[3 * i + 2];
    if (start + slots > E.cfg.arena_pairs) {
      overflowed(E, c, slots, cnt);
      continue;
    }
    const BatchArgs a = slot_args(E, s);
    const uint64_t qbase = reinterpret_cast<const BatchRequest*>(s.h_desc + kReqsOff)[i].qbase;
    const int rc = copy_out(E, c, nullptr, &a, qbase, cnt, s.stream);
    if (rc) return rc;
    finish_chunk(E, c, cnt, slots, E.cfg.hash_results ? s.h_hash[i] : 0, tc);
  }
  s.busy = false;
  return RTS_OK;
}

int launch_slot(Engine& E, Slot& s, std::vector<Chunk> chunks) {
  // Requests class by class: each tile class has its own kernel over its
  // own tile range.
  std::stable_sort(chunks.begin(), chunks.end(), [&](const Chunk& x, const Chunk& y) {
    return tile_class(tile_flags(E, E.reqs[x.req])) < tile_class(tile_flags(E, E.reqs[y.req]));
  });
  const TreeView tv = tree_view(E.ix);
  *reinterpret_cast<TreeView*>(s.h_desc + kViewOff) = tv;
  BatchHeader* h = reinterpret_cast<BatchHeader*>(s.h_desc + kHeaderOff);
  BatchRequest* br = reinterpret_cast<BatchRequest*>(s.h_desc + kReqsOff);
  uint64_t qb = 0;
  uint32_t tb = 0;
  uint32_t class_end[kClasses] = {0, 0, 0, 0, 0};
  s.ext.clear();
  for (size_t i = 0; i < chunks.size(); i++) {
    const Chunk& c = chunks[i];
    const rts_request& r = E.reqs[c.req];
    br[i].queries = static_cast<const char*>(r.queries) + c.q0 * query_bytes_of(E.ix, r);
    br[i].n = (uint32_t)c.n;
    br[i].flags = tile_flags(E, r);
    br[i].qbase = qb;
    br[i].tbase = tb;
    br[i].qid_base = (uint32_t)c.q0;  // pairs report query indices within the request
    if (br[i].flags & kReqExt) {
      // Walked by its own launches (after every tile class: no tiles).
      BatchExt x{};
      x.ql.tree = tv;
      x.ql.queries = br[i].queries;
      x.ql.n = (uint32_t)c.n;
      x.ql.qid_base = (uint32_t)c.q0;
      x.ql.order = nullptr;
      const bool seg = r.points == kSegments;
      x.ql.pred = seg ? kPredLine : r.predicate;
      x.ql.point_queries = r.points == kPoints ? 1 : 0;
      x.ql.f64 = E.ix->f64() ? 1 : 0;
      x.ql.sink = SinkArgs{RTS_HANDLER_COUNT, nullptr};
      x.ql.traversal = RTS_TRAVERSAL_WAVE;
      const rts_index::TunedPlan& tp = E.ix->tuned[r.predicate & 3][r.points == kPoints ? 1 : 0];
      x.ql.tree.leaf_size =
          (!seg && tp.valid && tp.leaf_size >= 1) ? tp.leaf_size : E.ix->cfg.leaf_size;
      x.qbase = qb;
      x.none = seg && r.predicate != RTS_INTERSECTS;  // no box lies inside a segment
      s.ext.push_back(x);
      qb += c.n;
      continue;
    }
    qb += c.n;
    tb += (br[i].flags & kReqWaveTiles) ? (uint32_t)c.n : (uint32_t)((c.n + kWave - 1) / kWave);
    class_end[tile_class(br[i].flags)] = tb;
  }
  for (int c = 1; c < kClasses; c++)  // classes without requests end where the previous ends
    if (class_end[c] < class_end[c - 1]) class_end[c] = class_end[c - 1];
  h->n_requests = (uint32_t)chunks.size();
  h->n_tiles = tb;
  for (int c = 0; c < kClasses; c++) h->class_end[c] = class_end[c];
  h->pad = 0;
  h->n_queries = qb;
  if (E.have_upd_event) SRV_HIP(hipStreamWaitEvent(s.stream, E.upd_event, 0));
  s.chunks = chunks;
  {
    const int rc = record_slot_work(E, s, kReqsOff + chunks.size() * sizeof(BatchRequest),
                                    chunks.size());
    if (rc) return rc;
  }
  SRV_HIP(hipEventRecord(s.done, s.stream));
  s.busy = true;
  E.batches++;
  SRV_TRACE("batch %llu launched: %zu chunks, %u tiles, %llu queries", (unsigned long long)E.batches,
            chunks.size(), h->n_tiles, (unsigned long long)h->n_queries);
  return RTS_OK;
}

int run_batched(Engine& E) {
  std::vector<Slot> slots(std::max<uint32_t>(1, E.cfg.slots));
  int rc = RTS_OK;
  for (auto& s : slots)
    if ((rc = slot_init(E, s))) break;
  uint64_t next_up = 0;
  std::vector<Chunk> chunks;
  E.t0 = Clock::now();
  while (!rc && E.done < E.n_reqs) {
    for (auto& s : slots)
      if (!rc && s.busy && hipEventQuery(s.done) == hipSuccess) rc = complete_slot(E, s);
    const double t = E.now();
    bool idle = true;
    for (const auto& s : slots) idle = idle && !s.busy;
    if (stalled(E, !idle)) {
      rc = RTS_ERR_DEVICE;
      break;
    }
    while (!rc && update_ready(E, next_up, t, idle)) rc = apply_update(E, E.ups[next_up++], slots);
    for (auto& s : slots) {
      if (rc || s.busy) continue;
      admit(E, t, E.cfg.max_batch_requests, E.cfg.max_batch_queries, (double)E.cfg.arena_pairs,
            true, false, chunks);
      if (chunks.empty()) break;
      rc = launch_slot(E, s, chunks);
    }
  }
  for (auto& s : slots) slot_free(s);
  return rc;
}

// -------------------------------------------------------------- plain ----
struct PlainChunk {
  Chunk c;
  uint32_t log;  // ring slot of its counter span, hash and event
};

int run_plain(Engine& E) {
  constexpr uint32_t kRing = 1024;
  hipStream_t stream = nullptr;
  rts_pair* arena = nullptr;
  unsigned long long* d_counter = nullptr;
  rts_dev_queue* d_queue = nullptr;
  uint64_t* d_spans = nullptr;  // per launch: queue counter before and after
  uint64_t* h_spans = nullptr;  // pinned copy
  unsigned long long* d_hash = nullptr;
  unsigned long long* h_hash = nullptr;
  std::vector<hipEvent_t> evs(kRing, nullptr);
  SRV_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  SRV_HIP(hipMalloc(&arena, E.cfg.arena_pairs * sizeof(rts_pair)));
  SRV_HIP(hipMalloc(&d_counter, sizeof(unsigned long long)));
  SRV_HIP(hipMalloc(&d_queue, sizeof(rts_dev_queue)));
  SRV_HIP(hipMalloc(&d_spans, 2 * kRing * sizeof(uint64_t)));
  SRV_HIP(hipMalloc(&d_hash, kRing * sizeof(unsigned long long)));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_spans), 2 * kRing * sizeof(uint64_t)));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_hash), kRing * sizeof(unsigned long long)));
  for (auto& e : evs) SRV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // Set up on the loop's own stream: it does not wait for the null stream,
  // so a null-stream memset could still be pending when the first query
  // reads the counter.
  const rts_dev_queue q{arena, E.cfg.arena_pairs, d_counter};
  SRV_HIP(hipMemcpyAsync(d_queue, &q, sizeof(q), hipMemcpyHostToDevice, stream));
  SRV_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(d_counter), 0u, 2, stream));
  SRV_HIP(hipStreamSynchronize(stream));

  std::deque<PlainChunk> inflight;
  uint64_t next_up = 0, ring = 0;
  double used = 0;  // predicted results in the arena since its last reset
  int rc = RTS_OK;
  std::vector<Chunk> one;
  auto reset_arena = [&]() -> int {
    SRV_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(d_counter), 0u, 2, stream));
    used = 0;
    return RTS_OK;
  };
  E.t0 = Clock::now();
  while (!rc && E.done < E.n_reqs) {
    // completions, in launch order
    while (!rc && !inflight.empty() && hipEventQuery(evs[inflight.front().log]) == hipSuccess) {
      const PlainChunk pc = inflight.front();
      inflight.pop_front();
      const uint64_t before = h_spans[2 * pc.log], after = h_spans[2 * pc.log + 1];
      if (after > E.cfg.arena_pairs) {
        overflowed(E, pc.c, after - before, after - before);
        continue;
      }
      rc = copy_out(E, pc.c, arena + before, nullptr, 0, after - before, stream);
      finish_chunk(E, pc.c, after - before, after - before,
                   E.cfg.hash_results ? h_hash[pc.log] : 0, E.now());
    }
    if (rc) break;
    if (stalled(E, !inflight.empty())) {
      rc = RTS_ERR_DEVICE;
      break;
    }
    const double t = E.now();
    while (!rc && update_ready(E, next_up, t, inflight.empty())) {
      rc = update_index(E, E.ups[next_up], stream);
      next_up++;
      E.epoch++;
      E.applied_updates++;
    }
    if (rc || inflight.size() >= kRing - 1) continue;
    // A re-launch waits until everything launched before it is done, then
    // runs on an emptied arena.
    if (!E.retry.empty() && used > 0) {
      if (inflight.empty()) rc = reset_arena();
      continue;
    }
    admit(E, t, 1, E.cfg.chunk_queries, (double)E.cfg.arena_pairs - used, used == 0, true, one);
    if (one.empty()) {
      // Arrived work does not fit the rest of the arena: empty it once idle.
      const bool waiting = !E.queue.empty() && E.reqs[E.queue.front()].arrival_s <= t;
      if (waiting && inflight.empty() && used > 0) rc = reset_arena();
      continue;
    }
    const Chunk c = one[0];
    const rts_request& r = E.reqs[c.req];
    used += predicted(E, c);
    const uint32_t li = (uint32_t)(ring++ % kRing);
    SRV_HIP(hipMemcpyAsync(d_spans + 2 * li, d_counter, sizeof(uint64_t), hipMemcpyDeviceToDevice,
                           stream));
    const void* qp = static_cast<const char*>(r.queries) + c.q0 * query_bytes_of(E.ix, r);
    rc = query_request(E.ix, r, qp, c.n, RTS_HANDLER_COLLECT, d_queue, stream);
    if (rc) break;
    SRV_HIP(hipMemcpyAsync(d_spans + 2 * li + 1, d_counter, sizeof(uint64_t),
                           hipMemcpyDeviceToDevice, stream));
    // Query ids come back relative to the launch: make them relative to the request.
    if (c.q0 || E.cfg.hash_results) {
      if (E.cfg.hash_results)
        SRV_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(d_hash + li), 0u, 2, stream));
      SRV_HIP(launch_rebase_hash_span(arena, E.cfg.arena_pairs, d_spans + 2 * li, (uint32_t)c.q0,
                                      E.cfg.hash_results ? 1 : 0, d_hash + li, stream));
    }
    SRV_HIP(hipMemcpyAsync(h_spans + 2 * li, d_spans + 2 * li, 2 * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, stream));
    if (E.cfg.hash_results)
      SRV_HIP(hipMemcpyAsync(h_hash + li, d_hash + li, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, stream));
    SRV_HIP(hipEventRecord(evs[li], stream));
    inflight.push_back(PlainChunk{c, li});
    E.batches++;
  }
  (void)hipStreamSynchronize(stream);
  for (auto& e : evs)
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(arena);
  (void)hipFree(d_counter);
  (void)hipFree(d_queue);
  (void)hipFree(d_spans);
  (void)hipFree(d_hash);
  (void)hipHostFree(h_spans);
  (void)hipHostFree(h_hash);
  (void)hipStreamDestroy(stream);
  return rc;
}

}  // namespace
}  // namespace rts

using namespace rts;

extern "C" {

void rts_serve_config_default(rts_serve_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->mode = RTS_SERVE_BATCHED;
  c->slots = 3;
  c->max_batch_requests = 4096;
  // hipGraph replay of the batch sequence is refused (see rts_serve_run).
  c->use_graphs = 0;
  c->max_batch_queries = 4ull << 20;
  c->chunk_queries = 256ull << 10;
  c->arena_pairs = 1ull << 30;
  c->hash_results = 0;
  c->traversal = RTS_TRAVERSAL_AUTO;
  c->head_share_pct = 50;
  c->sort_min_tiles = 4096;
}

int rts_serve_run(rts_index* ix, const rts_serve_config* cfg, const rts_request* reqs,
                  uint64_t n_reqs, const rts_update_batch* ups, uint64_t n_ups,
                  rts_request_result* results, rts_serve_stats* stats) {
  if (!ix || !results || (n_reqs && !reqs) || (n_ups && !ups)) {
    set_error("rts_serve_run: null argument");
    return RTS_ERR_INVALID;
  }
  if (ix->cfg.coord == RTS_COORD_F64_AS_F32) {
    set_error("rts_serve_run: the served path takes float or exact double indexes "
              "(RTS_COORD_F32, RTS_COORD_F64), not RTS_COORD_F64_AS_F32");
    return RTS_ERR_INVALID;
  }
  Engine E;
  E.ix = ix;
  if (cfg) E.cfg = *cfg; else rts_serve_config_default(&E.cfg);
  if (E.cfg.max_batch_requests == 0 || E.cfg.max_batch_requests > 65536 ||
      E.cfg.max_batch_queries == 0 || E.cfg.max_batch_queries >= (1ull << 31) ||
      E.cfg.chunk_queries == 0 || E.cfg.arena_pairs == 0 || E.cfg.head_share_pct == 0 ||
      E.cfg.head_share_pct > 100) {
    set_error("rts_serve_run: invalid configuration");
    return RTS_ERR_INVALID;
  }
  // hipGraph replay of a slot's launch sequence is refused: a run with it did
  // not complete on the MI355X (the host stayed blocked inside a HIP call)
  // and the cause has not been established.
  if (E.cfg.use_graphs != 0) {
    set_error("rts_serve_run: use_graphs must be 0 (hipGraph replay of the batch sequence "
              "did not complete on the MI355X; the cause is not established)");
    return RTS_ERR_INVALID;
  }
  for (uint64_t i = 0; i < n_reqs; i++) {
    const rts_request& r = reqs[i];
    if ((r.n && !r.queries) || r.predicate < 0 || r.predicate > 2 || r.points < kEnvelopes ||
        r.points > kSegments || (r.points == kPoints && r.predicate != RTS_CONTAINS) ||
        (r.points == kSegments && r.predicate == RTS_WITHIN) || r.n >= (1ull << 31) ||
        (i && r.arrival_s < reqs[i - 1].arrival_s)) {
      set_error("rts_serve_run: bad request " + std::to_string(i));
      return RTS_ERR_INVALID;
    }
  }
  E.reqs = reqs;
  E.n_reqs = n_reqs;
  E.ups = ups;
  E.n_ups = n_ups;
  E.res = results;
  std::memset(results, 0, n_reqs * sizeof(rts_request_result));
  E.st.assign(n_reqs, ReqState());
  E.desc_bytes = kReqsOff + (size_t)E.cfg.max_batch_requests * sizeof(BatchRequest);
  for (uint32_t i = 0; i < n_reqs; i++) E.queue.push_back(i);
  int rc = hipSetDevice(ix->device) == hipSuccess ? RTS_OK : RTS_ERR_HIP;
  if (!rc && hipStreamCreateWithFlags(&E.upd_stream, hipStreamNonBlocking) != hipSuccess)
    rc = RTS_ERR_HIP;
  if (!rc && hipEventCreateWithFlags(&E.upd_event, hipEventDisableTiming) != hipSuccess)
    rc = RTS_ERR_HIP;
  if (!rc) {
    std::lock_guard<std::mutex> lk(ix->mu);
    rc = ensure_built(ix, E.upd_stream, true);
  }
  if (!rc && hipStreamSynchronize(E.upd_stream) != hipSuccess) rc = RTS_ERR_HIP;
  if (!rc) rc = probe_estimates(E, E.upd_stream);
  if (!rc) rc = (E.cfg.mode == RTS_SERVE_PLAIN) ? run_plain(E) : run_batched(E);
  if (E.upd_stream) {
    (void)hipStreamSynchronize(E.upd_stream);
    (void)hipStreamDestroy(E.upd_stream);
  }
  if (!rc) {
    std::lock_guard<std::mutex> lk(ix->mu);
    rc = device_error_status(ix);
  }
  if (E.upd_event) (void)hipEventDestroy(E.upd_event);
  if (stats) {
    stats->requests = n_reqs;
    stats->queries = 0;
    stats->results = 0;
    double first = n_reqs ? reqs[0].arrival_s : 0, last = 0;
    for (uint64_t i = 0; i < n_reqs; i++) {
      stats->queries += reqs[i].n;
      stats->results += results[i].count;
      last = std::max(last, results[i].done_s);
    }
    stats->wall_s = last - first;
    stats->batches = E.batches;
    stats->relaunches = E.relaunches;
    stats->updates = E.applied_updates;
  }
  return rc;
}

}  // extern "C"
