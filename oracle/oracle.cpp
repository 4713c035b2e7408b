// CPU oracle for the LibRTS spatial-index hot path.
//
// TEST INFRASTRUCTURE ONLY. This library checks results: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
// product (rtspatial_amd) never links or calls it.
//
// It restates, in plain C++ on the host, what the reference computes:
//   * exact closed-interval predicates      include/rtspatial/geom/envelope.cuh:29-69
//   * pair order (geom_id, query_id)         include/rtspatial/default_handlers/collecting_handlers.h:9-31
//   * the reference tests' input generators  test/test_commons.h:22-67,
//                                            test/test_envelope_queries.h:121-134
//   * a brute-force join and an STR R-tree join over those predicates
//   * the same predicates and brute-force join in double precision
//     (SpatialIndex<double, 2>), exact
//   * line-segment intersects (the query of spatial_index.cuh:517-547,
//     whose reference program reports nothing), exact
//   * the reference's ray-slab formulation of intersects
//                                            include/rtspatial/details/ray_params.h:53-167,
//                                            src/shaders/shaders_intersects_envelope_query_2d.cu:14-128
//     (only used to reproduce the reference's own golden count, which carries one
//      padded-slab false positive; the exact oracle does not use it).
//
// Pinning (what checks each function against something other than itself):
//   predicates + brute force : known answers 4224111 (test_point_queries.h:27,59),
//                              5 (:89), 2 and 2->1 (test_envelope_queries.h:35,190,203);
//                              closed-form tiny cases; R-tree == brute force.
//   ray formulation          : reproduces 6549771 (test_envelope_queries.h:57,90) and
//                              differs from the exact join by exactly one pair.
//   R-tree                   : equality with brute force on seeded inputs.
//   double predicates/join   : known answer 5 (test_point_queries.h:92-120);
//                              equality with the float join on float-exact inputs;
//                              tiny cases whose answer flips under float rounding.
//   generators               : the known answers above are only reached with
//                              the exact libstdc++ mt19937 streams.
//   hash/count summaries     : equality with summaries recomputed from full pair sets.
//   segment predicate        : equality with the exact rational restatement in
//                              oracle.py, itself checked against an exact
//                              Liang-Barsky clip and closed-form cases.
//
// Deleted boxes are modelled by an optional `alive` mask (nullptr = all alive).
// Predicate codes: 0 = intersects, 1 = contains (data box contains the query,
// the reference's kContains), 2 = within (data box inside the query envelope).
// Query kinds: 0 = points (x, y), 1 = envelopes (x0, y0, x1, y1), 2 = line
// segments (x1, y1, x2, y2): intersects only (contains and within match
// nothing, as the reference's kContains line query reports nothing,
// spatial_index.cuh:524-526).

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

struct Box {
  float x0, y0, x1, y1;
};

enum { P_INTERSECTS = 0, P_CONTAINS = 1, P_WITHIN = 2 };
enum { Q_POINT = 0, Q_ENVELOPE = 1, Q_LINE = 2 };

// ------------------------------------------------------ line segments ----
// Line<Point<T,2>> queries (geom/line.cuh:10-27) against closed data boxes:
// the closed segment (x1, y1, x2, y2) meets the closed box iff their bounding
// boxes overlap and the four box corners are not all strictly on one side of
// the segment's line. The side of corner c is the sign of
//   (x2 - x1) (cy - y1) - (y2 - y1) (cx - x1),
// taken from a double evaluation when it clears a generous error margin and
// otherwise computed exactly in integers: every double is m 2^k with an
// integer m, so after scaling by 2^-kmin the differences and products are
// integer arithmetic (sign-magnitude, 32-bit limbs).
struct BigInt {
  bool neg = false;
  std::vector<uint32_t> mag;  // little-endian, no leading zero limbs
};
void big_trim(BigInt& a) {
  while (!a.mag.empty() && a.mag.back() == 0) a.mag.pop_back();
  if (a.mag.empty()) a.neg = false;
}
int mag_cmp(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
  if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
  for (size_t i = a.size(); i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}
std::vector<uint32_t> mag_add(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
  std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1, 0);
  uint64_t carry = 0;
  for (size_t i = 0; i < r.size(); i++) {
    uint64_t v = carry + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
    r[i] = (uint32_t)v;
    carry = v >> 32;
  }
  return r;
}
std::vector<uint32_t> mag_sub(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
  // a >= b
  std::vector<uint32_t> r(a.size(), 0);
  int64_t borrow = 0;
  for (size_t i = 0; i < a.size(); i++) {
    int64_t v = (int64_t)a[i] - borrow - (int64_t)(i < b.size() ? b[i] : 0);
    borrow = v < 0;
    r[i] = (uint32_t)(v + (borrow << 32));
  }
  return r;
}
BigInt big_add(const BigInt& a, const BigInt& b) {
  BigInt r;
  if (a.neg == b.neg) {
    r.neg = a.neg;
    r.mag = mag_add(a.mag, b.mag);
  } else if (mag_cmp(a.mag, b.mag) >= 0) {
    r.neg = a.neg;
    r.mag = mag_sub(a.mag, b.mag);
  } else {
    r.neg = b.neg;
    r.mag = mag_sub(b.mag, a.mag);
  }
  big_trim(r);
  return r;
}
BigInt big_neg(BigInt a) {
  if (!a.mag.empty()) a.neg = !a.neg;
  return a;
}
BigInt big_mul(const BigInt& a, const BigInt& b) {
  BigInt r;
  r.mag.assign(a.mag.size() + b.mag.size() + 1, 0);
  for (size_t i = 0; i < a.mag.size(); i++) {
    uint64_t carry = 0;
    for (size_t j = 0; j < b.mag.size(); j++) {
      uint64_t v = (uint64_t)a.mag[i] * b.mag[j] + r.mag[i + j] + carry;
      r.mag[i + j] = (uint32_t)v;
      carry = v >> 32;
    }
    for (size_t k = i + b.mag.size(); carry; k++) {
      uint64_t v = (uint64_t)r.mag[k] + carry;
      r.mag[k] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  r.neg = a.neg != b.neg;
  big_trim(r);
  return r;
}
// The integer d * 2^-kmin, for a finite d whose exponent k (d = m 2^k, m an
// integer of at most 53 bits) is >= kmin.
BigInt big_from(double d, int kmin) {
  BigInt r;
  if (d == 0.0) return r;
  int e;
  const double f = std::frexp(std::fabs(d), &e);     // |d| = f 2^e, f in [0.5, 1)
  uint64_t m = (uint64_t)std::ldexp(f, 53);           // exact
  const int shift = e - 53 - kmin;
  r.mag.assign((size_t)(shift / 32) + 3, 0);
  const int w = shift / 32, b = shift % 32;
  const unsigned __int128 v = (unsigned __int128)m << b;
  for (int i = 0; i < 3; i++) r.mag[w + i] = (uint32_t)(v >> (32 * i));
  r.neg = d < 0.0;
  big_trim(r);
  return r;
}
int exponent_of(double d) {  // k with d = m 2^k, m a 53-bit integer
  int e;
  std::frexp(d, &e);
  return e - 53;
}
// Sign of (x2 - x1)(cy - y1) - (y2 - y1)(cx - x1), exactly.
int orient_sign(const double* s, double cx, double cy) {
  const double a = (s[2] - s[0]) * (cy - s[1]), b = (s[3] - s[1]) * (cx - s[0]);
  const double o = a - b, margin = 1e-9 * (std::fabs(a) + std::fabs(b)) + 1e-290;
  if (o > margin) return 1;
  if (o < -margin) return -1;
  const double v[6] = {s[0], s[1], s[2], s[3], cx, cy};
  int kmin = INT32_MAX;
  for (double x : v)
    if (x != 0.0) kmin = std::min(kmin, exponent_of(x));
  if (kmin == INT32_MAX) return 0;
  BigInt x1 = big_from(s[0], kmin), y1 = big_from(s[1], kmin), x2 = big_from(s[2], kmin),
         y2 = big_from(s[3], kmin), X = big_from(cx, kmin), Y = big_from(cy, kmin);
  const BigInt r = big_add(big_mul(big_add(x2, big_neg(x1)), big_add(Y, big_neg(y1))),
                           big_neg(big_mul(big_add(y2, big_neg(y1)), big_add(X, big_neg(x1)))));
  return r.mag.empty() ? 0 : (r.neg ? -1 : 1);
}
// The closed segment s meets the closed box (x0, y0, x1, y1); a NaN
// coordinate of either meets nothing.
bool segment_meets(const double* s, double x0, double y0, double x1, double y1) {
  for (int k = 0; k < 4; k++)
    if (std::isnan(s[k])) return false;
  if (!(std::min(s[0], s[2]) <= x1 && std::max(s[0], s[2]) >= x0 &&
        std::min(s[1], s[3]) <= y1 && std::max(s[1], s[3]) >= y0))
    return false;
  const double cx[4] = {x0, x1, x1, x0}, cy[4] = {y0, y0, y1, y1};
  int pos = 0, neg = 0;
  for (int k = 0; k < 4; k++) {
    const int o = orient_sign(s, cx[k], cy[k]);
    pos += o > 0;
    neg += o < 0;
  }
  return pos < 4 && neg < 4;
}

// envelope.cuh:29-38 -- Envelope::Contains(point), closed on both ends.
inline bool box_contains_point(const Box& b, float px, float py) {
  return px >= b.x0 && px <= b.x1 && py >= b.y0 && py <= b.y1;
}
// envelope.cuh:40-53 -- Envelope::Contains(envelope): `a` contains `o`.
inline bool box_contains_box(const Box& a, const Box& o) {
  return o.x0 >= a.x0 && o.x1 <= a.x1 && o.y0 >= a.y0 && o.y1 <= a.y1;
}
// envelope.cuh:55-69 -- Envelope::Intersects, closed overlap on both axes.
inline bool box_intersects(const Box& a, const Box& o) {
  return o.x0 <= a.x1 && o.x1 >= a.x0 && o.y0 <= a.y1 && o.y1 >= a.y0;
}

// The predicate evaluated for (data box g, query q). Point queries only have
// `contains` (spatial_index.cuh:352-391 aborts on anything else); the oracle
// evaluates the same closed form for every predicate code on points.
inline bool eval(int pred, int kind, const Box& g, const float* q) {
  if (kind == Q_POINT) return box_contains_point(g, q[0], q[1]);
  if (kind == Q_LINE) {
    const double sd[4] = {q[0], q[1], q[2], q[3]};
    return pred == P_INTERSECTS && segment_meets(sd, g.x0, g.y0, g.x1, g.y1);
  }
  Box qb{q[0], q[1], q[2], q[3]};
  switch (pred) {
    case P_INTERSECTS: return box_intersects(g, qb);
    case P_CONTAINS: return box_contains_box(g, qb);  // reference kContains: g contains q
    default: return box_contains_box(qb, g);          // within: q contains g
  }
}

// The same closed forms in double precision, for SpatialIndex<double, 2>
// (envelope.cuh:29-69 with COORD_T = double). The reference's double Point
// returns float coordinates to these predicates (point.cuh:68); this oracle
// evaluates them exactly, which is what a double index promises. The
// reference's rounded behaviour is the float oracle on rounded inputs.
struct Box64 {
  double x0, y0, x1, y1;
};
inline bool eval64(int pred, int kind, const Box64& g, const double* q) {
  if (kind == Q_POINT) return q[0] >= g.x0 && q[0] <= g.x1 && q[1] >= g.y0 && q[1] <= g.y1;
  if (kind == Q_LINE) return pred == P_INTERSECTS && segment_meets(q, g.x0, g.y0, g.x1, g.y1);
  switch (pred) {
    case P_INTERSECTS: return q[0] <= g.x1 && q[2] >= g.x0 && q[1] <= g.y1 && q[3] >= g.y0;
    case P_CONTAINS: return q[0] >= g.x0 && q[2] <= g.x1 && q[1] >= g.y0 && q[3] <= g.y1;
    default: return g.x0 >= q[0] && g.x1 <= q[2] && g.y0 >= q[1] && g.y1 <= q[3];
  }
}
inline bool eval_t(int pred, int kind, const Box& g, const float* q) { return eval(pred, kind, g, q); }
inline bool eval_t(int pred, int kind, const Box64& g, const double* q) {
  return eval64(pred, kind, g, q);
}

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

int resolve_threads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// Runs fn(begin, end, thread_index) over [0, n) split into `nt` contiguous chunks.
void parallel_ranges(uint64_t n, int nt, const std::function<void(uint64_t, uint64_t, int)>& fn) {
  if (nt <= 1 || n < 2) {
    fn(0, n, 0);
    return;
  }
  std::vector<std::thread> ths;
  uint64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    uint64_t b = std::min<uint64_t>(n, chunk * t), e = std::min<uint64_t>(n, b + chunk);
    ths.emplace_back(fn, b, e, t);
  }
  for (auto& th : ths) th.join();
}

// Dynamic work distribution: blocks of `grain` queries handed out by an atomic.
void parallel_dynamic(uint64_t n, int nt, uint64_t grain,
                      const std::function<void(uint64_t, uint64_t, int)>& fn) {
  std::atomic<uint64_t> next{0};
  auto worker = [&](int t) {
    for (;;) {
      uint64_t b = next.fetch_add(grain);
      if (b >= n) break;
      fn(b, std::min(n, b + grain), t);
    }
  };
  if (nt <= 1) {
    worker(0);
    return;
  }
  std::vector<std::thread> ths;
  for (int t = 0; t < nt; t++) ths.emplace_back(worker, t);
  for (auto& th : ths) th.join();
}

// ---------------------------------------------------------------- R-tree ----
// Sort-Tile-Recursive bulk-loaded R-tree with exact FP32 node bounds. Node
// tests are necessary conditions for the leaf predicate, so pruning never
// drops a true pair. Every node also holds the number of its matchable
// entries (boxes without a NaN coordinate) and the sum of their id hashes
// ghash(id) (below), so a node whose bounds lie inside an envelope query
// answers intersects and within for its whole subtree at once. Entries can
// be given new boxes in place (rtree_update), after which every node's
// bounds and sums are recomputed (a refit: same tree, exact bounds).
struct RTree {
  int fanout = 16;
  std::vector<Box> nodes;              // bounds of every node (leaves first)
  std::vector<uint32_t> child_begin;   // per node: first child index (or first entry)
  std::vector<uint32_t> child_count;
  std::vector<uint8_t> is_leaf;
  std::vector<uint64_t> node_cnt;      // matchable entries below the node
  std::vector<uint64_t> node_gsum;     // sum of ghash(id) over them
  std::vector<uint32_t> entries;       // leaf entry -> box id
  std::vector<Box> entry_boxes;        // leaf entry -> box (copied for locality)
  std::vector<uint32_t> pos_of_id;     // box id -> entry (UINT32_MAX: not in the tree)
  uint32_t root = 0;
  bool empty = true;
  int threads = 0;
};

// Factored pair checksum. A pair (query q, box b) hashes to
// qmult(q) * ghash(b) mod 2^64, with qmult odd; the checksum of a query's
// result set is the sum of ghash(b) over it, and of a request the sum over
// its queries of qmult(q) times that. Because the hash factors, a subtree
// that matches whole contributes its stored sum. A missing, extra or
// substituted pair changes a query's sum by ghash differences (splitmix64 is
// a bijection, so distinct ids hash apart), which qmult (odd, so invertible
// mod 2^64) carries into the request's sum.
inline uint64_t ghash(uint64_t b) { return splitmix64(b ^ 0xD1B54A32D192ED03ull); }
inline uint64_t qmult(uint64_t q) { return splitmix64(q) | 1ull; }

Box unite(const Box& a, const Box& b) {
  return Box{std::min(a.x0, b.x0), std::min(a.y0, b.y0), std::max(a.x1, b.x1), std::max(a.y1, b.y1)};
}

const Box kEmptyBox{INFINITY, INFINITY, -INFINITY, -INFINITY};

inline bool box_has_nan(const Box& x) {
  return std::isnan(x.x0) || std::isnan(x.y0) || std::isnan(x.x1) || std::isnan(x.y1);
}

// The extent a box covers, corners in either order.
inline Box normalized(const Box& x) {
  return Box{std::min(x.x0, x.x1), std::min(x.y0, x.y1), std::max(x.x0, x.x1), std::max(x.y0, x.y1)};
}

// Ascending-order integer image of a float; NaN sorts last.
inline uint32_t float_key(float v) {
  if (std::isnan(v)) return 0xFFFFFFFFu;
  uint32_t u;
  std::memcpy(&u, &v, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Sorts `v` ascending: chunks sorted in parallel, then merged pairwise.
void parallel_sort_u64(std::vector<uint64_t>& v, int nt) {
  const size_t n = v.size();
  int parts = 1;
  while (parts * 2 <= nt && n / (size_t)(parts * 2) >= (1u << 15)) parts *= 2;
  if (parts == 1) {
    std::sort(v.begin(), v.end());
    return;
  }
  std::vector<size_t> cut(parts + 1);
  for (int i = 0; i <= parts; i++) cut[i] = n * (size_t)i / (size_t)parts;
  parallel_ranges((uint64_t)parts, parts, [&](uint64_t b, uint64_t e, int) {
    for (uint64_t i = b; i < e; i++) std::sort(v.begin() + cut[i], v.begin() + cut[i + 1]);
  });
  std::vector<uint64_t> tmp(n);
  for (int w = 1; w < parts; w *= 2) {
    const int pairs = parts / (2 * w);
    parallel_ranges((uint64_t)pairs, pairs, [&](uint64_t b, uint64_t e, int) {
      for (uint64_t k = b; k < e; k++) {
        const size_t lo = cut[2 * w * k], mid = cut[2 * w * k + w], hi = cut[2 * w * (k + 1)];
        std::merge(v.begin() + lo, v.begin() + mid, v.begin() + mid, v.begin() + hi,
                   tmp.begin() + lo);
      }
    });
    v.swap(tmp);
  }
}

// STR packing of `items` (indices into `cb`, the item bounds) into groups of
// `fanout`: items sorted by centre x (ties by index), then each slab of
// ~sqrt(groups) groups sorted by centre y. NaN centres sort last.
void str_order(std::vector<uint32_t>& items, const std::vector<Box>& cb, int fanout, int nt) {
  const size_t n = items.size();
  if (n <= (size_t)fanout) return;
  const size_t leaves = (n + fanout - 1) / fanout;
  const size_t slabs = (size_t)std::ceil(std::sqrt((double)leaves));
  const size_t per_slab = slabs ? ((leaves + slabs - 1) / slabs) * fanout : n;
  std::vector<uint64_t> key(n);
  parallel_ranges(n, nt, [&](uint64_t b, uint64_t e, int) {
    for (uint64_t i = b; i < e; i++) {
      const Box& x = cb[items[i]];
      key[i] = ((uint64_t)float_key(0.5f * (x.x0 + x.x1)) << 32) | items[i];
    }
  });
  parallel_sort_u64(key, nt);
  const uint64_t n_slabs = (n + per_slab - 1) / per_slab;
  parallel_dynamic(n_slabs, nt, 1, [&](uint64_t sb, uint64_t se, int) {
    for (uint64_t sl = sb; sl < se; sl++) {
      const size_t s = sl * per_slab, e = std::min(n, s + per_slab);
      for (size_t i = s; i < e; i++) {
        const uint32_t it = (uint32_t)key[i];
        const Box& x = cb[it];
        key[i] = ((uint64_t)float_key(0.5f * (x.y0 + x.y1)) << 32) | it;
      }
      std::sort(key.begin() + s, key.begin() + e);
    }
  });
  for (size_t i = 0; i < n; i++) items[i] = (uint32_t)key[i];
}

// Bounds and sums of every node from the entries: leaf nodes in parallel,
// then inner nodes in index order (children always precede their parent).
void rtree_refit(RTree* t) {
  const size_t m = t->nodes.size();
  parallel_ranges(m, resolve_threads(t->threads), [&](uint64_t b, uint64_t e, int) {
    for (uint64_t v = b; v < e; v++) {
      if (!t->is_leaf[v]) continue;
      Box bx = kEmptyBox;
      uint64_t c = 0, g = 0;
      const uint32_t cb = t->child_begin[v], cc = t->child_count[v];
      for (uint32_t i = cb; i < cb + cc; i++) {
        const Box& x = t->entry_boxes[i];
        if (box_has_nan(x)) continue;
        bx = unite(bx, normalized(x));
        c++;
        g += ghash(t->entries[i]);
      }
      t->nodes[v] = bx;
      t->node_cnt[v] = c;
      t->node_gsum[v] = g;
    }
  });
  for (size_t v = 0; v < m; v++) {
    if (t->is_leaf[v]) continue;
    Box bx = kEmptyBox;
    uint64_t c = 0, g = 0;
    const uint32_t cb = t->child_begin[v], cc = t->child_count[v];
    for (uint32_t i = cb; i < cb + cc; i++) {
      bx = unite(bx, t->nodes[i]);
      c += t->node_cnt[i];
      g += t->node_gsum[i];
    }
    t->nodes[v] = bx;
    t->node_cnt[v] = c;
    t->node_gsum[v] = g;
  }
}

RTree* rtree_build(const Box* boxes, uint64_t n, const uint8_t* alive, int fanout, int threads) {
  auto* t = new RTree();
  t->fanout = fanout < 2 ? 16 : fanout;
  t->threads = threads;
  const int nt = resolve_threads(threads);
  t->pos_of_id.assign(n, UINT32_MAX);
  std::vector<uint32_t> ids;
  ids.reserve(n);
  for (uint64_t i = 0; i < n; i++)
    if (!alive || alive[i]) ids.push_back((uint32_t)i);
  if (ids.empty()) return t;
  t->empty = false;
  {
    std::vector<Box> cb(boxes, boxes + n);
    // Dead boxes stay out of the tree; NaN centres sort last but are kept
    // (their predicates are false anyway).
    str_order(ids, cb, t->fanout, nt);
  }
  t->entries = ids;
  t->entry_boxes.resize(ids.size());
  for (size_t i = 0; i < ids.size(); i++) {
    t->entry_boxes[i] = boxes[ids[i]];
    t->pos_of_id[ids[i]] = (uint32_t)i;
  }
  // Level 0: leaves over entries (bounds and sums come from rtree_refit).
  std::vector<uint32_t> level;
  for (size_t s = 0; s < ids.size(); s += t->fanout) {
    const size_t e = std::min(ids.size(), s + (size_t)t->fanout);
    Box b = kEmptyBox;
    for (size_t i = s; i < e; i++)
      if (!box_has_nan(t->entry_boxes[i])) b = unite(b, normalized(t->entry_boxes[i]));
    level.push_back((uint32_t)t->nodes.size());
    t->nodes.push_back(b);
    t->child_begin.push_back((uint32_t)s);
    t->child_count.push_back((uint32_t)(e - s));
    t->is_leaf.push_back(1);
  }
  while (level.size() > 1) {
    std::vector<uint32_t> order = level;
    str_order(order, t->nodes, t->fanout, nt);
    // Children of one parent must be contiguous in `nodes`: re-emit them.
    std::vector<uint32_t> next;
    std::vector<uint32_t> copies;
    for (size_t s = 0; s < order.size(); s += t->fanout) {
      const size_t e = std::min(order.size(), s + (size_t)t->fanout);
      const uint32_t first = (uint32_t)t->nodes.size();
      for (size_t i = s; i < e; i++) {
        const uint32_t c = order[i];
        const Box cbox = t->nodes[c];
        const uint32_t cbeg = t->child_begin[c], ccnt = t->child_count[c];
        const uint8_t cleaf = t->is_leaf[c];
        t->nodes.push_back(cbox);
        t->child_begin.push_back(cbeg);
        t->child_count.push_back(ccnt);
        t->is_leaf.push_back(cleaf);
      }
      copies.push_back(first);
      copies.push_back((uint32_t)(e - s));
    }
    for (size_t k = 0; k < copies.size(); k += 2) {
      const uint32_t first = copies[k], cnt = copies[k + 1];
      Box b = kEmptyBox;
      for (uint32_t i = 0; i < cnt; i++) b = unite(b, t->nodes[first + i]);
      next.push_back((uint32_t)t->nodes.size());
      t->nodes.push_back(b);
      t->child_begin.push_back(first);
      t->child_count.push_back(cnt);
      t->is_leaf.push_back(0);
    }
    level.swap(next);
  }
  t->root = level[0];
  t->node_cnt.assign(t->nodes.size(), 0);
  t->node_gsum.assign(t->nodes.size(), 0);
  rtree_refit(t);
  return t;
}

// Node bounds inside the query envelope (closed): every matchable box below
// then intersects the query and lies within it.
inline bool node_inside(const Box& n, const float* q) {
  return q[0] <= n.x0 && q[1] <= n.y0 && n.x1 <= q[2] && n.y1 <= q[3];
}

// Necessary condition on a node's bounds for any descendant box to satisfy
// the predicate against q.
inline bool node_may_match(int pred, int kind, const Box& n, const float* q) {
  if (kind == Q_POINT) return box_contains_point(n, q[0], q[1]);
  Box qb{q[0], q[1], q[2], q[3]};
  switch (pred) {
    case P_CONTAINS: return box_contains_box(n, qb);  // g contains q and n contains g
    default: return box_intersects(n, qb);            // intersects, within
  }
}

template <typename Emit>
void rtree_query_one(const RTree* t, int pred, int kind, const float* q, Emit&& emit,
                     std::vector<uint32_t>& stack) {
  if (t->empty) return;
  stack.clear();
  if (!node_may_match(pred, kind, t->nodes[t->root], q)) return;
  stack.push_back(t->root);
  while (!stack.empty()) {
    uint32_t v = stack.back();
    stack.pop_back();
    uint32_t b = t->child_begin[v], c = t->child_count[v];
    if (t->is_leaf[v]) {
      for (uint32_t i = 0; i < c; i++) {
        if (eval(pred, kind, t->entry_boxes[b + i], q)) emit(t->entries[b + i]);
      }
    } else {
      for (uint32_t i = 0; i < c; i++) {
        if (node_may_match(pred, kind, t->nodes[b + i], q)) stack.push_back(b + i);
      }
    }
  }
}

// Count and ghash sum of the boxes that satisfy pred against q. Envelope
// intersects and within take a node whose bounds lie inside q whole, from its
// stored count and sum; everything else is tested box by box.
template <int PRED, int KIND>
void rtree_checksum_t(const RTree* t, const float* qp, uint64_t* cnt, uint64_t* gsum,
                      std::vector<uint32_t>& stack) {
  uint64_t c = 0, g = 0;
  constexpr bool kWhole = KIND == Q_ENVELOPE && (PRED == P_INTERSECTS || PRED == P_WITHIN);
  const Box* nodes = t->nodes.data();
  const uint32_t* cbeg = t->child_begin.data();
  const uint32_t* ccnt = t->child_count.data();
  const uint8_t* leaf = t->is_leaf.data();
  const uint64_t* ncnt = t->node_cnt.data();
  const uint64_t* ngs = t->node_gsum.data();
  const Box* eb = t->entry_boxes.data();
  const uint32_t* ent = t->entries.data();
  const float q[4] = {qp[0], qp[1], KIND == Q_POINT ? qp[0] : qp[2], KIND == Q_POINT ? qp[1] : qp[3]};
  stack.resize(64 * (size_t)t->fanout + 16);
  uint32_t* st = stack.data();
  size_t sp = 0;
  if (t->empty || !node_may_match(PRED, KIND, nodes[t->root], q)) {
    *cnt = 0;
    *gsum = 0;
    return;
  }
  st[sp++] = t->root;
  while (sp) {
    const uint32_t v = st[--sp];
    if (kWhole && node_inside(nodes[v], q)) {
      c += ncnt[v];
      g += ngs[v];
      continue;
    }
    const uint32_t b = cbeg[v], n = ccnt[v];
    if (leaf[v]) {
      for (uint32_t i = b; i < b + n; i++) {
        if (eval(PRED, KIND, eb[i], q)) {
          c++;
          g += ghash(ent[i]);
        }
      }
    } else {
      for (uint32_t i = b; i < b + n; i++)
        if (node_may_match(PRED, KIND, nodes[i], q)) st[sp++] = i;
    }
  }
  *cnt = c;
  *gsum = g;
}

void rtree_checksum_one(const RTree* t, int pred, int kind, const float* q, uint64_t* cnt,
                        uint64_t* gsum, std::vector<uint32_t>& stack) {
  if (kind == Q_POINT) return rtree_checksum_t<P_CONTAINS, Q_POINT>(t, q, cnt, gsum, stack);
  if (kind == Q_ENVELOPE) {
    if (pred == P_INTERSECTS) return rtree_checksum_t<P_INTERSECTS, Q_ENVELOPE>(t, q, cnt, gsum, stack);
    if (pred == P_WITHIN) return rtree_checksum_t<P_WITHIN, Q_ENVELOPE>(t, q, cnt, gsum, stack);
    return rtree_checksum_t<P_CONTAINS, Q_ENVELOPE>(t, q, cnt, gsum, stack);
  }
  return rtree_checksum_t<P_INTERSECTS, Q_LINE>(t, q, cnt, gsum, stack);
}

// ------------------------------------------- reference ray formulation ----
// ray_params.h:67-84 (Compute(envelope, diagonal)) and :137-166 (IsHit), written
// out in IEEE single precision, operation for operation.
struct Ray {
  float ox, oy, dx, dy;
};

inline Ray ray_of(const Box& e, bool anti_diagonal) {
  float p1x = e.x0, p1y = e.y0, p2x = e.x1, p2y = e.y1;
  if (anti_diagonal) {
    p1x = e.x1;
    p1y = e.y0;
    p2x = e.x0;
    p2y = e.y1;
  }
  return Ray{p1x, p1y, p2x - p1x, p2y - p1y};
}

inline bool ray_hits(const Ray& r, const Box& e) {
  const float eps = FLT_EPSILON;
  const float gamma3 = (3 * eps) / (1 - 3 * eps);
  const float pad = 1 + 2 * gamma3;
  float t0 = 0, t1 = std::nextafter(1.0f, FLT_MAX);
  const float o[2] = {r.ox, r.oy}, d[2] = {r.dx, r.dy};
  const float lo[2] = {e.x0, e.y0}, hi[2] = {e.x1, e.y1};
  for (int i = 0; i < 2; i++) {
    volatile float inv = 1 / d[i];  // keep the division a separate rounding
    float tn = (lo[i] - o[i]) * inv;
    float tf = (hi[i] - o[i]) * inv;
    if (tn > tf) std::swap(tn, tf);
    tf *= pad;
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return false;
  }
  return true;
}

// shaders_intersects_envelope_query_2d.cu:14-41 (forward IS) and :75-128 (backward).
inline bool ref_ray_pair(const Box& g, const Box& q) {
  bool fwd = ray_hits(ray_of(q, true), g) && !ray_hits(ray_of(g, false), q);
  bool valid = g.x0 < g.x1 && g.y0 < g.y1;  // envelope.cuh:103-105
  bool bwd = valid && ray_hits(ray_of(g, false), q);
  return fwd || bwd;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ generators ----
// test/test_commons.h:22-49 -- GenerateUniformBoxes<float>(n, max_w, max_h):
// std::mt19937(n), min ~ U[0,1), max = min(min + U[0, max_side], 1).
void orc_ref_uniform_boxes_f32(uint64_t n, float max_w, float max_h, float* out) {
  std::mt19937 mt((std::mt19937::result_type)n);
  std::uniform_real_distribution<float> min_x(0.0, 1.0), min_y(0.0, 1.0);
  std::uniform_real_distribution<float> w_dist(0.0, max_w), h_dist(0.0, max_h);
  for (uint64_t i = 0; i < n; i++) {
    float x = min_x(mt), y = min_y(mt), w = w_dist(mt), h = h_dist(mt);
    out[4 * i + 0] = x;
    out[4 * i + 1] = y;
    out[4 * i + 2] = std::min(x + w, 1.0f);
    out[4 * i + 3] = std::min(y + h, 1.0f);
  }
}

// test/test_commons.h:51-67 -- GenerateUniformPoints<float>(n).
void orc_ref_uniform_points_f32(uint64_t n, float* out) {
  std::mt19937 mt((std::mt19937::result_type)n);
  std::uniform_real_distribution<float> gx(0.0, 1.0), gy(0.0, 1.0);
  for (uint64_t i = 0; i < n; i++) {
    float x = gx(mt), y = gy(mt);
    out[2 * i + 0] = x;
    out[2 * i + 1] = y;
  }
}

// test/test_envelope_queries.h:121-134 -- the distinct update ids drawn with
// std::mt19937(n_updates) and uniform_int_distribution<size_t>(0, n1 - 1).
void orc_ref_update_ids(uint64_t n1, uint64_t n_updates, uint64_t* out) {
  std::unordered_set<size_t> ids;
  std::uniform_int_distribution<size_t> id_dist(0, n1 - 1);
  std::mt19937 mt((std::mt19937::result_type)n_updates);
  uint64_t k = 0;
  while (ids.size() < n_updates) {
    size_t id = id_dist(mt);
    if (ids.find(id) == ids.end()) {
      out[k++] = id;
      ids.insert(id);
    }
  }
}

// ---------------------------------------------------------- predicates ----
int orc_eval(int pred, int kind, const float* box, const float* query) {
  Box g{box[0], box[1], box[2], box[3]};
  return eval(pred, kind, g, query) ? 1 : 0;
}

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }

}  // extern "C"

namespace {

// ---------------------------------------------------------- brute force ----
// Counts pairs; per_query (nullable) receives each query's count, hsum/hxor
// (nullable) each query's sum / xor of splitmix64((q << 32) | g).
template <typename BoxT, typename T>
uint64_t bf_summary(int pred, int kind, const T* boxes, uint64_t nb, const uint8_t* alive,
                    const T* queries, uint64_t nq, int threads, uint64_t* per_query,
                    uint64_t* hsum, uint64_t* hxor) {
  const BoxT* b = reinterpret_cast<const BoxT*>(boxes);
  int qs = kind == Q_POINT ? 2 : 4;
  int nt = resolve_threads(threads);
  std::vector<uint64_t> part(nt, 0);
  parallel_dynamic(nq, nt, 64, [&](uint64_t qb, uint64_t qe, int t) {
    uint64_t tot = 0;
    for (uint64_t q = qb; q < qe; q++) {
      const T* qq = queries + q * qs;
      uint64_t c = 0, s = 0, x = 0;
      for (uint64_t g = 0; g < nb; g++) {
        if (alive && !alive[g]) continue;
        if (eval_t(pred, kind, b[g], qq)) {
          c++;
          if (hsum) {
            uint64_t h = splitmix64((q << 32) | g);
            s += h;
            x ^= h;
          }
        }
      }
      if (per_query) per_query[q] = c;
      if (hsum) hsum[q] = s;
      if (hxor) hxor[q] = x;
      tot += c;
    }
    part[t] += tot;
  });
  uint64_t total = 0;
  for (auto v : part) total += v;
  return total;
}

// Fills out[2*i] = query_id, out[2*i+1] = box_id sorted by (query, box).
// Returns the total number of pairs; writes at most `cap` pairs.
template <typename BoxT, typename T>
uint64_t bf_pairs(int pred, int kind, const T* boxes, uint64_t nb, const uint8_t* alive,
                  const T* queries, uint64_t nq, int threads, uint32_t* out, uint64_t cap) {
  const BoxT* b = reinterpret_cast<const BoxT*>(boxes);
  int qs = kind == Q_POINT ? 2 : 4;
  int nt = resolve_threads(threads);
  std::vector<std::vector<uint32_t>> local(nt);
  parallel_ranges(nq, nt, [&](uint64_t qb, uint64_t qe, int t) {
    auto& v = local[t];
    for (uint64_t q = qb; q < qe; q++) {
      const T* qq = queries + q * qs;
      for (uint64_t g = 0; g < nb; g++) {
        if (alive && !alive[g]) continue;
        if (eval_t(pred, kind, b[g], qq)) {
          v.push_back((uint32_t)q);
          v.push_back((uint32_t)g);
        }
      }
    }
  });
  uint64_t k = 0;
  for (auto& v : local) {
    for (size_t i = 0; i < v.size(); i += 2, k++) {
      if (k < cap) {
        out[2 * k] = v[i];
        out[2 * k + 1] = v[i + 1];
      }
    }
  }
  return k;
}

}  // namespace

extern "C" {

uint64_t orc_bf_summary(int pred, int kind, const float* boxes, uint64_t nb, const uint8_t* alive,
                        const float* queries, uint64_t nq, int threads, uint64_t* per_query,
                        uint64_t* hsum, uint64_t* hxor) {
  return bf_summary<Box>(pred, kind, boxes, nb, alive, queries, nq, threads, per_query, hsum,
                         hxor);
}
uint64_t orc_bf_pairs(int pred, int kind, const float* boxes, uint64_t nb, const uint8_t* alive,
                      const float* queries, uint64_t nq, int threads, uint32_t* out, uint64_t cap) {
  return bf_pairs<Box>(pred, kind, boxes, nb, alive, queries, nq, threads, out, cap);
}
// Double coordinates, exact double predicates.
uint64_t orc_bf_summary_f64(int pred, int kind, const double* boxes, uint64_t nb,
                            const uint8_t* alive, const double* queries, uint64_t nq, int threads,
                            uint64_t* per_query, uint64_t* hsum, uint64_t* hxor) {
  return bf_summary<Box64>(pred, kind, boxes, nb, alive, queries, nq, threads, per_query, hsum,
                           hxor);
}
uint64_t orc_bf_pairs_f64(int pred, int kind, const double* boxes, uint64_t nb,
                          const uint8_t* alive, const double* queries, uint64_t nq, int threads,
                          uint32_t* out, uint64_t cap) {
  return bf_pairs<Box64>(pred, kind, boxes, nb, alive, queries, nq, threads, out, cap);
}
int orc_eval_f64(int pred, int kind, const double* box, const double* query) {
  Box64 g{box[0], box[1], box[2], box[3]};
  return eval64(pred, kind, g, query) ? 1 : 0;
}

// -------------------------------------------------------------- R-tree ----
void* orc_rtree_build(const float* boxes, uint64_t nb, const uint8_t* alive, int fanout,
                      int threads) {
  return rtree_build(reinterpret_cast<const Box*>(boxes), nb, alive, fanout, threads);
}

// Gives the boxes `ids` new envelopes (a NaN box matches nothing: a delete)
// and refits the tree. Ids that are not in the tree (dead at build time, or
// beyond it) are ignored. Returns the number applied.
uint64_t orc_rtree_update(void* tree, const uint64_t* ids, const float* boxes, uint64_t n) {
  RTree* t = static_cast<RTree*>(tree);
  uint64_t applied = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (ids[i] >= t->pos_of_id.size()) continue;
    const uint32_t pos = t->pos_of_id[ids[i]];
    if (pos == UINT32_MAX) continue;
    t->entry_boxes[pos] = Box{boxes[4 * i], boxes[4 * i + 1], boxes[4 * i + 2], boxes[4 * i + 3]};
    applied++;
  }
  if (!t->empty) rtree_refit(t);
  return applied;
}

// Per query: count and ghash sum of its results (factored checksum); returns
// the total count. qmult / ghash are exported for the checkers.
uint64_t orc_rtree_checksums(void* tree, int pred, int kind, const float* queries, uint64_t nq,
                             int threads, uint64_t* per_query, uint64_t* gsum) {
  const RTree* t = static_cast<const RTree*>(tree);
  const int qs = kind == Q_POINT ? 2 : 4;
  const int nt = resolve_threads(threads);
  // Queries in Z-order of their centres (over the tree's root bounds): a
  // thread's consecutive queries then walk nearby nodes, which stay cached.
  std::vector<uint64_t> order(nq);
  {
    const Box r = t->empty ? Box{0, 0, 1, 1} : t->nodes[t->root];
    const float sx = r.x1 > r.x0 ? 65535.0f / (r.x1 - r.x0) : 0.0f;
    const float sy = r.y1 > r.y0 ? 65535.0f / (r.y1 - r.y0) : 0.0f;
    parallel_ranges(nq, nt, [&](uint64_t b, uint64_t e, int) {
      for (uint64_t i = b; i < e; i++) {
        const float* q = queries + i * qs;
        const float cx = kind == Q_POINT ? q[0] : 0.5f * (q[0] + q[2]);
        const float cy = kind == Q_POINT ? q[1] : 0.5f * (q[1] + q[3]);
        const float fx = std::min(65535.0f, std::max(0.0f, (cx - r.x0) * sx));
        const float fy = std::min(65535.0f, std::max(0.0f, (cy - r.y0) * sy));
        // NaN centres land at 0 (the comparisons above fail): any order is fine.
        const uint32_t ix = std::isnan(fx) ? 0u : (uint32_t)fx, iy = std::isnan(fy) ? 0u : (uint32_t)fy;
        uint64_t z = 0;
        for (int k = 0; k < 16; k++)
          z |= (uint64_t)((ix >> k) & 1u) << (2 * k) | (uint64_t)((iy >> k) & 1u) << (2 * k + 1);
        order[i] = (z << 32) | i;
      }
    });
    parallel_sort_u64(order, nt);
  }
  std::vector<uint64_t> part(nt, 0);
  parallel_dynamic(nq, nt, 64, [&](uint64_t qb, uint64_t qe, int th) {
    std::vector<uint32_t> stack;
    uint64_t tot = 0;
    for (uint64_t k = qb; k < qe; k++) {
      const uint64_t q = (uint32_t)order[k];
      uint64_t c = 0, g = 0;
      rtree_checksum_one(t, pred, kind, queries + q * qs, &c, &g, stack);
      if (per_query) per_query[q] = c;
      if (gsum) gsum[q] = g;
      tot += c;
    }
    part[th] += tot;
  });
  uint64_t total = 0;
  for (auto v : part) total += v;
  return total;
}

uint64_t orc_ghash(uint64_t b) { return ghash(b); }
uint64_t orc_qmult(uint64_t q) { return qmult(q); }

void orc_rtree_free(void* t) { delete static_cast<RTree*>(t); }

uint64_t orc_rtree_summary(void* tree, int pred, int kind, const float* queries, uint64_t nq,
                           int threads, uint64_t* per_query, uint64_t* hsum, uint64_t* hxor) {
  const RTree* t = static_cast<const RTree*>(tree);
  int qs = kind == Q_POINT ? 2 : 4;
  int nt = resolve_threads(threads);
  std::vector<uint64_t> part(nt, 0);
  parallel_dynamic(nq, nt, 16, [&](uint64_t qb, uint64_t qe, int th) {
    std::vector<uint32_t> stack;
    uint64_t tot = 0;
    for (uint64_t q = qb; q < qe; q++) {
      uint64_t c = 0, s = 0, x = 0;
      rtree_query_one(t, pred, kind, queries + q * qs, [&](uint32_t g) {
        c++;
        if (hsum) {
          uint64_t h = splitmix64((q << 32) | g);
          s += h;
          x ^= h;
        }
      }, stack);
      if (per_query) per_query[q] = c;
      if (hsum) hsum[q] = s;
      if (hxor) hxor[q] = x;
      tot += c;
    }
    part[th] += tot;
  });
  uint64_t total = 0;
  for (auto v : part) total += v;
  return total;
}

// Pairs (query_id, box_id) sorted by (query, box); returns the total count.
uint64_t orc_rtree_pairs(void* tree, int pred, int kind, const float* queries, uint64_t nq,
                         int threads, uint32_t* out, uint64_t cap) {
  const RTree* t = static_cast<const RTree*>(tree);
  int qs = kind == Q_POINT ? 2 : 4;
  int nt = resolve_threads(threads);
  std::vector<std::vector<uint32_t>> local(nt);
  parallel_ranges(nq, nt, [&](uint64_t qb, uint64_t qe, int th) {
    std::vector<uint32_t> stack, hits;
    auto& v = local[th];
    for (uint64_t q = qb; q < qe; q++) {
      hits.clear();
      rtree_query_one(t, pred, kind, queries + q * qs, [&](uint32_t g) { hits.push_back(g); }, stack);
      std::sort(hits.begin(), hits.end());
      for (uint32_t g : hits) {
        v.push_back((uint32_t)q);
        v.push_back(g);
      }
    }
  });
  uint64_t k = 0;
  for (auto& v : local) {
    for (size_t i = 0; i < v.size(); i += 2, k++) {
      if (k < cap) {
        out[2 * k] = v[i];
        out[2 * k + 1] = v[i + 1];
      }
    }
  }
  return k;
}

// ------------------------------------------- reference ray formulation ----
// Counts the pairs the reference's two-pass ray formulation reports for
// envelope intersects; writes up to `cap` (query, box) pairs that it reports
// but the exact predicate rejects (extra) -- and, separately, pairs the exact
// predicate accepts that it misses (missing).
uint64_t orc_refray_intersects(const float* boxes, uint64_t nb, const float* queries, uint64_t nq,
                               int threads, uint32_t* extra, uint64_t extra_cap, uint64_t* n_extra,
                               uint32_t* missing, uint64_t missing_cap, uint64_t* n_missing) {
  const Box* b = reinterpret_cast<const Box*>(boxes);
  const Box* qb = reinterpret_cast<const Box*>(queries);
  int nt = resolve_threads(threads);
  std::vector<uint64_t> cnt(nt, 0);
  std::vector<std::vector<uint32_t>> ex(nt), mi(nt);
  parallel_ranges(nq, nt, [&](uint64_t s, uint64_t e, int t) {
    for (uint64_t q = s; q < e; q++) {
      for (uint64_t g = 0; g < nb; g++) {
        bool r = ref_ray_pair(b[g], qb[q]);
        bool x = box_intersects(b[g], qb[q]);
        if (r) cnt[t]++;
        if (r && !x) {
          ex[t].push_back((uint32_t)q);
          ex[t].push_back((uint32_t)g);
        }
        if (!r && x) {
          mi[t].push_back((uint32_t)q);
          mi[t].push_back((uint32_t)g);
        }
      }
    }
  });
  uint64_t total = 0, ke = 0, km = 0;
  for (int t = 0; t < nt; t++) {
    total += cnt[t];
    for (size_t i = 0; i < ex[t].size(); i += 2, ke++)
      if (ke < extra_cap) {
        extra[2 * ke] = ex[t][i];
        extra[2 * ke + 1] = ex[t][i + 1];
      }
    for (size_t i = 0; i < mi[t].size(); i += 2, km++)
      if (km < missing_cap) {
        missing[2 * km] = mi[t][i];
        missing[2 * km + 1] = mi[t][i + 1];
      }
  }
  if (n_extra) *n_extra = ke;
  if (n_missing) *n_missing = km;
  return total;
}

}  // extern "C"
