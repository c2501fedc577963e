// knn_index.hpp — header-only C++ adapter that re-exposes the reference's
// template entry points over the C-ABI of ddg.h, templated over the caller's
// own types, so a caller of include/index.h / include/utils/beamSearch.h
// keeps its call sites and its Graph / PointRange objects:
//
//   ddg::knn_index<Point, PointRange, indexType>{BP}                 index.h:43-55
//       .build_index(G, Points, BuildStats)                          index.h:125-137
//       .batch_insert(inserts, G, Points, BuildStats, alpha, ...)    index.h:213-341
//       .robustPrune(p, cand, G, Points, alpha, add)                 index.h:63-123
//       .lazy_delete(p / deletes, G)                                 index.h:139-164
//   ddg::beam_search<Point, PointRange, indexType>(p, G, Points, start(s), QP)
//                                                                    beamSearch.h:41-184
//   ddg::searchAll<Point, PointRange, indexType>(Query, G, Base, stats, start, QP)
//                                                                    beamSearch.h:314-346
//
// Graph, PointRange, Point, BuildParams, QueryParams and stats are template
// parameters used only through the reference's public accessors:
//   Graph (graph.h:119-209)        size(), max_degree(), G[i] -> an edge range
//                                  with size(), [j], begin(), update_neighbors()
//   PointRange (point_range.h:53-113) size(), dimension(), aligned_dimension(),
//                                  Points[i] -> a Point
//   Point (euclidian_point.h:68-107, mips_point.h:68-109)  [j], id(),
//                                  static is_metric() (true: squared L2,
//                                  false: negated inner product)
//   BuildParams / QueryParams (types.h:103-156)  R, L, alpha, two_pass /
//                                  k, beamSize, cut, limit, degree_limit
//   stats (stats.h:53-88)          increment_visited(i, v), increment_dist(i, v)
// so the reference's own Graph<unsigned>, PointRange<T, Euclidian_Point<T>>
// etc. work unchanged (tests/cpp/dropin_demo.cpp compiles against them), and
// so do the standalone types of this header (ddg::Graph, ddg::PointRange,
// ddg::Euclidian_Point, ddg::Mips_Point) for callers without the reference.
// Quantized_Mips_Point (mips_point.h:111-187) has no traversal here; its
// distance is ddg_quantized_mips_distances.
//
// One device handle is kept per PointRange object (and degree bound): the
// vectors are uploaded once; before a call that reads the graph, the rows
// that differ from the graph the device last saw are uploaded (a host copy of
// that graph is kept for the comparison); after build_index / batch_insert
// the device graph is written back into the caller's Graph.  Calls on one
// handle are serialised by a mutex, so the per-point calls of a parallel_for
// are correct, but the batched forms (beam_search_batch, robustPrune_batch)
// are the ones to call from a hot loop.  A non-OK status from the library is
// reported like the reference reports its errors: a message on stdout and
// abort().
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <type_traits>
#include <utility>
#include <vector>

#include "../ddg.h"

namespace ddg {

inline void check(ddg_status st, const ddg_index* h) {
  if (st != DDG_OK) {
    std::cout << "ERROR: " << ddg_last_error(h) << std::endl;
    abort();
  }
}

inline long dim_round_up(long dim, long tp_size) {  // point_range.h:45-50
  long qt = (dim * tp_size) / 64, rem = (dim * tp_size) % 64;
  return rem == 0 ? dim : ((qt + 1) * 64) / tp_size;
}

template <class T> struct dtype_of;
template <> struct dtype_of<uint8_t> { static constexpr ddg_dtype v = DDG_U8; };
template <> struct dtype_of<int8_t> { static constexpr ddg_dtype v = DDG_I8; };
template <> struct dtype_of<float> { static constexpr ddg_dtype v = DDG_F32; };

// ============================================================ standalone types
// The reference's in-memory layouts (point_range.h:45-50, 80: rows padded to
// 64 bytes; graph.h:126-128: n x (R+1) u32 rows [deg, nbrs...]).
template <class T>
struct Euclidian_Point {  // euclidian_point.h:68-107
  using distanceType = float;
  static bool is_metric() { return true; }
  Euclidian_Point() = default;
  Euclidian_Point(const T* v, unsigned d_, unsigned ad_, long id) : values(v), d(d_), aligned_d(ad_), id_(id) {}
  T operator[](long i) const { return values[i]; }
  long id() const { return id_; }
  const T* values = nullptr;
  unsigned d = 0, aligned_d = 0;
  long id_ = -1;
};

template <class T>
struct Mips_Point {  // mips_point.h:68-109
  using distanceType = float;
  static bool is_metric() { return false; }
  Mips_Point() = default;
  Mips_Point(const T* v, unsigned d_, unsigned ad_, long id) : values(v), d(d_), aligned_d(ad_), id_(id) {}
  T operator[](long i) const { return values[i]; }
  long id() const { return id_; }
  const T* values = nullptr;
  unsigned d = 0, aligned_d = 0;
  long id_ = -1;
};

// PointRange<T, Point> (point_range.h:53-113).
template <class T, class Point = Euclidian_Point<T>>
struct PointRange {
  std::vector<T> values;
  unsigned dims = 0, aligned_dims = 0;
  size_t n = 0;
  PointRange() = default;
  PointRange(const T* rows, size_t n_, unsigned d_) : dims(d_), aligned_dims((unsigned)dim_round_up(d_, sizeof(T))), n(n_) {
    values.assign(n * aligned_dims, T(0));
    for (size_t i = 0; i < n; i++) std::memcpy(values.data() + i * aligned_dims, rows + i * dims, dims * sizeof(T));
  }
  explicit PointRange(const char* filename) {
    FILE* f = fopen(filename, "rb");
    if (!f) { std::cout << "ERROR: cannot open " << filename << std::endl; abort(); }
    unsigned hdr[2];
    if (fread(hdr, 4, 2, f) != 2) { std::cout << "ERROR: short vector file" << std::endl; abort(); }
    n = hdr[0];
    dims = hdr[1];
    aligned_dims = (unsigned)dim_round_up(dims, sizeof(T));
    values.assign(n * aligned_dims, T(0));
    std::vector<T> row(dims);
    for (size_t i = 0; i < n; i++) {
      if (fread(row.data(), sizeof(T), dims, f) != dims) { std::cout << "ERROR: short vector file" << std::endl; abort(); }
      std::memcpy(values.data() + i * aligned_dims, row.data(), dims * sizeof(T));
    }
    fclose(f);
  }
  size_t size() const { return n; }
  long dimension() const { return dims; }
  long aligned_dimension() const { return aligned_dims; }
  unsigned get_dims() const { return dims; }
  Point operator[](long i) const { return Point(values.data() + i * aligned_dims, dims, aligned_dims, i); }
  const T* row(size_t i) const { return values.data() + i * aligned_dims; }
};

// edgeRange (graph.h:43-117): a view of one row.
struct edgeRange {
  uint32_t* edges = nullptr;  // [deg, nbr...]
  long maxDeg = 0;
  size_t size() const { return edges[0]; }
  uint32_t operator[](uint32_t j) const {
    if (j > edges[0]) { std::cout << "ERROR: tried to exceed range" << std::endl; abort(); }
    return edges[j + 1];
  }
  uint32_t* begin() const { return edges + 1; }
  uint32_t* end() const { return edges + 1 + edges[0]; }
  template <class Rng>
  void update_neighbors(const Rng& r) {
    if ((long)r.size() > maxDeg) { std::cout << "ERROR in update_neighbors: cannot exceed max degree " << maxDeg << std::endl; abort(); }
    edges[0] = (uint32_t)r.size();
    for (size_t i = 0; i < r.size(); i++) edges[i + 1] = r[i];
  }
  template <class Rng>
  void append_neighbors(const Rng& r) {
    if ((long)(r.size() + edges[0]) > maxDeg) { std::cout << "ERROR in append_neighbors: cannot exceed max degree " << maxDeg << std::endl; abort(); }
    for (size_t i = 0; i < r.size(); i++) edges[edges[0] + i + 1] = r[i];
    edges[0] += (uint32_t)r.size();
  }
  void clear_neighbors() { edges[0] = 0; }
};

// Graph<indexType> (graph.h:119-209).
struct Graph {
  long maxDeg = 0;
  size_t n = 0;
  std::vector<uint32_t> graph;
  Graph() = default;
  Graph(long maxDeg_, size_t n_) : maxDeg(maxDeg_), n(n_), graph(n_ * (maxDeg_ + 1), 0) {}
  explicit Graph(const char* gFile) {
    FILE* f = fopen(gFile, "rb");
    if (!f) { std::cout << "ERROR: cannot open " << gFile << std::endl; abort(); }
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) { std::cout << "ERROR: short graph file" << std::endl; abort(); }
    n = hdr[0];
    maxDeg = hdr[1];
    std::vector<uint32_t> deg(n);
    if (fread(deg.data(), 4, n, f) != n) { std::cout << "ERROR: short graph file" << std::endl; abort(); }
    graph.assign(n * (maxDeg + 1), 0);
    for (size_t i = 0; i < n; i++) {
      if (deg[i] > (uint32_t)maxDeg) { std::cout << "ERROR: degree exceeds max degree" << std::endl; abort(); }
      uint32_t* r = graph.data() + i * (maxDeg + 1);
      r[0] = deg[i];
      if (fread(r + 1, 4, deg[i], f) != deg[i]) { std::cout << "ERROR: short graph file" << std::endl; abort(); }
    }
    fclose(f);
  }
  void save(const char* oFile) const {
    FILE* f = fopen(oFile, "wb");
    if (!f) { std::cout << "ERROR: cannot write " << oFile << std::endl; abort(); }
    uint32_t hdr[2] = {(uint32_t)n, (uint32_t)maxDeg};
    fwrite(hdr, 4, 2, f);
    for (size_t i = 0; i < n; i++) fwrite(&graph[i * (maxDeg + 1)], 4, 1, f);
    for (size_t i = 0; i < n; i++) fwrite(&graph[i * (maxDeg + 1) + 1], 4, graph[i * (maxDeg + 1)], f);
    fclose(f);
  }
  long max_degree() const { return maxDeg; }
  size_t size() const { return n; }
  edgeRange operator[](size_t i) { return edgeRange{graph.data() + i * (maxDeg + 1), maxDeg}; }
};

struct BuildParams {  // types.h:103-142 (Vamana fields)
  long L = 0, R = 0;
  double alpha = 1.2;
  bool two_pass = false;
  BuildParams() = default;
  BuildParams(long R_, long L_, double a, bool tp) : L(L_), R(R_), alpha(a), two_pass(tp) {}
};

struct QueryParams {  // types.h:145-156
  long k = 0, beamSize = 0;
  double cut = 1.35;
  long limit = 0, degree_limit = 0;
  QueryParams() = default;
  QueryParams(long k_, long Q, double c, long lim, long dg) : k(k_), beamSize(Q), cut(c), limit(lim), degree_limit(dg) {}
};

template <class indexType = uint32_t>
struct stats {  // stats.h:53-88
  std::vector<indexType> visited, distances;
  stats() = default;
  explicit stats(size_t n) : visited(n, 0), distances(n, 0) {}
  void increment_dist(indexType i, indexType j) { distances[i] += j; }
  void increment_visited(indexType i, indexType j) { visited[i] += j; }
};

// ==================================================== access to caller types
namespace detail {

template <class PR> using point_t = std::decay_t<decltype(std::declval<PR&>()[0])>;
template <class PR> using elem_t = std::decay_t<decltype(std::declval<point_t<PR>&>()[0])>;

template <class PR>
ddg_metric metric_of() { return point_t<PR>::is_metric() ? DDG_L2 : DDG_MIPS; }

// One aligned row (zero padding) of a Point into out[0, ad).
template <class P, class T>
void point_row(const P& p, long d, long ad, T* out) {
  for (long j = 0; j < d; j++) out[j] = p[j];
  for (long j = d; j < ad; j++) out[j] = T(0);
}

// The rows of a point range as one n x aligned_d array: the standalone
// PointRange hands out its storage, any other type is read point by point.
template <class T, class P>
const T* range_rows(PointRange<T, P>& pr, std::vector<T>&) { return pr.values.data(); }
template <class PR>
const elem_t<PR>* range_rows(PR& pr, std::vector<elem_t<PR>>& tmp) {
  const long n = (long)pr.size(), d = pr.dimension(), ad = pr.aligned_dimension();
  tmp.assign((size_t)n * ad, elem_t<PR>(0));
  for (long i = 0; i < n; i++) point_row(pr[i], d, ad, tmp.data() + (size_t)i * ad);
  return tmp.data();
}

// The graph's row array when its rows are one contiguous n x (R+1) block (the
// reference's Graph keeps them in one sequence), else nullptr.
template <class G>
uint32_t* contiguous_rows(G& g) {
  const size_t n = g.size();
  const long R1 = g.max_degree() + 1;
  if (n == 0) return nullptr;
  uint32_t* base = (uint32_t*)&*g[0].begin() - 1;
  for (size_t i : {(size_t)1, n / 2, n - 1}) {
    if (i >= n) continue;
    if ((uint32_t*)&*g[(uint32_t)i].begin() - 1 != base + i * R1) return nullptr;
  }
  return base[0] == g[0].size() ? base : nullptr;
}

template <class G>
void read_row(G& g, size_t i, uint32_t* out, long R1) {
  auto e = g[(uint32_t)i];
  const uint32_t deg = (uint32_t)e.size();
  out[0] = deg;
  for (uint32_t j = 0; j < deg; j++) out[1 + j] = e[j];
  for (long j = 1 + deg; j < R1; j++) out[j] = 0;
}

template <class G>
void write_rows(G& g, const uint32_t* rows) {
  const long R1 = g.max_degree() + 1;
  if (uint32_t* base = contiguous_rows(g)) {
    std::memcpy(base, rows, g.size() * R1 * 4);
    return;
  }
  std::vector<uint32_t> nb;
  for (size_t i = 0; i < g.size(); i++) {
    const uint32_t* r = rows + i * R1;
    nb.assign(r + 1, r + 1 + r[0]);
    g[(uint32_t)i].update_neighbors(nb);
  }
}

// A stats object sized for n points (stats.h:56-59), or one without a
// visited sequence (then it is trusted).
template <class S>
auto stats_fits(S& s, size_t n, int) -> decltype(s.visited.size(), bool()) { return s.visited.size() == n; }
template <class S>
bool stats_fits(S&, size_t, long) { return true; }

template <class E> struct is_pair : std::false_type {};
template <class A, class B> struct is_pair<std::pair<A, B>> : std::true_type {};

inline uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  return h;
}

}  // namespace detail

// ============================================================ device handles
// The device copy of one PointRange and of the graph the device last saw.
class Device {
 public:
  ddg_index* h = nullptr;
  size_t n = 0;
  long R = 0, d = 0, ad = 0;
  uint64_t fp = 0;
  std::vector<uint32_t> seen;  // the graph rows the device holds (n x (R+1))
  std::mutex mu;

  template <class PR>
  static uint64_t fingerprint(PR& P) {
    const size_t n = P.size();
    const long d = P.dimension();
    uint64_t h = detail::mix(detail::mix(n, (uint64_t)d), (uint64_t)P.aligned_dimension());
    for (size_t s = 0; s < 64 && n; s++) {
      const size_t i = (size_t)((s * 0x9e3779b97f4a7c15ull) % n);
      auto p = P[(long)i];
      for (long j = 0; j < d; j += std::max<long>(1, d / 8)) {
        auto v = p[j];
        uint64_t b = 0;
        std::memcpy(&b, &v, sizeof(v));
        h = detail::mix(h, b);
      }
    }
    return h;
  }

  template <class PR>
  Device(PR& P, long R_) : n(P.size()), R(R_), d(P.dimension()), ad(P.aligned_dimension()), fp(fingerprint(P)) {
    using T = detail::elem_t<PR>;
    std::vector<T> tmp;
    const T* rows = detail::range_rows(P, tmp);
    check(ddg_create(dtype_of<T>::v, detail::metric_of<PR>(), n, (uint32_t)d, (uint32_t)ad, rows, (uint32_t)R, &h),
          nullptr);
    seen.assign(n * (R + 1), 0);  // ddg_create starts from an empty graph
  }
  ~Device() { ddg_destroy(h); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;

  // Uploads the rows of g that differ from the device's graph.
  template <class G>
  void sync(G& g) {
    const long R1 = R + 1;
    std::vector<uint32_t> ids, rows, row(R1);
    uint32_t* base = detail::contiguous_rows(g);
    for (size_t i = 0; i < n; i++) {
      const uint32_t* r = base ? base + i * R1 : row.data();
      if (!base) detail::read_row(g, i, row.data(), R1);
      const uint32_t* s = seen.data() + i * R1;
      if (r[0] == s[0] && std::memcmp(r + 1, s + 1, r[0] * 4) == 0) continue;
      ids.push_back((uint32_t)i);
      rows.insert(rows.end(), r, r + R1);
      std::memcpy(seen.data() + i * R1, r, R1 * 4);
    }
    if (!ids.empty()) check(ddg_graph_set_rows(h, ids.data(), ids.size(), rows.data()), h);
  }

  // Uploads rows ids[0, m) of g (the rows a prune with add = true reads).
  template <class G>
  void sync_rows(G& g, const uint32_t* ids, size_t m) {
    const long R1 = R + 1;
    std::vector<uint32_t> rows(m * R1);
    for (size_t t = 0; t < m; t++) {
      detail::read_row(g, ids[t], rows.data() + t * R1, R1);
      std::memcpy(seen.data() + (size_t)ids[t] * R1, rows.data() + t * R1, R1 * 4);
    }
    if (m) check(ddg_graph_set_rows(h, ids, m, rows.data()), h);
  }

  // Writes the device graph back into g (and remembers it as seen).
  template <class G>
  void pull(G& g) {
    check(ddg_graph_get(h, seen.data()), h);
    detail::write_rows(g, seen.data());
  }
};

namespace detail {
struct Registry {
  std::mutex mu;
  std::map<std::pair<const void*, long>, std::unique_ptr<Device>> slots;
};
inline Registry& registry() {
  static Registry r;
  return r;
}
}  // namespace detail

// The cached handle of (Points, R); recreated when the object at that address
// no longer holds the same points.
template <class PR>
Device& device(PR& P, long R) {
  detail::Registry& r = detail::registry();
  std::lock_guard<std::mutex> lk(r.mu);
  auto& slot = r.slots[{(const void*)&P, R}];
  if (!slot || slot->n != P.size() || slot->d != P.dimension() || slot->fp != Device::fingerprint(P)) {
    slot.reset();
    slot.reset(new Device(P, R));
  }
  return *slot;
}

// Releases the cached handles of Points (for instance before it is destroyed).
template <class PR>
void release(PR& P) {
  detail::Registry& r = detail::registry();
  std::lock_guard<std::mutex> lk(r.mu);
  for (auto it = r.slots.begin(); it != r.slots.end();)
    it = it->first.first == (const void*)&P ? r.slots.erase(it) : std::next(it);
}

// ============================================================== beam search
template <class indexType = uint32_t>
using pid = std::pair<indexType, float>;

template <class indexType>
using beam_result = std::pair<std::pair<std::vector<pid<indexType>>, std::vector<pid<indexType>>>, size_t>;

// beam_search for many points at once (no reference counterpart: the batched
// form of the per-point calls of index.h:274-279 / bsp_index.hpp:259-266).
template <class Point, class PR, class indexType, class GraphT, class PointSeq, class QPT>
std::vector<beam_result<indexType>> beam_search_batch(const PointSeq& ps, GraphT& G, PR& Points,
                                                      const std::vector<indexType>& starts, QPT& QP) {
  using T = detail::elem_t<PR>;
  Device& dv = device(Points, G.max_degree());
  std::lock_guard<std::mutex> lk(dv.mu);
  dv.sync(G);
  const size_t nq = ps.size();
  std::vector<T> q(nq * dv.ad);
  std::vector<uint32_t> qid(nq);
  for (size_t i = 0; i < nq; i++) {
    detail::point_row(ps[i], dv.d, dv.ad, q.data() + i * dv.ad);
    const long id = ps[i].id();
    qid[i] = id < 0 ? 0xFFFFFFFFu : (uint32_t)id;
  }
  std::vector<uint32_t> st(starts.begin(), starts.end());
  ddg_query_params qp{(uint32_t)QP.k, (uint32_t)QP.beamSize, QP.cut, (uint64_t)std::max<long>(QP.limit, 0),
                      (uint32_t)std::max<long>(QP.degree_limit, 0), DDG_TRAV_NATIVE, 0};
  const uint32_t fcap = std::max<uint32_t>(qp.beam, (uint32_t)st.size());
  uint32_t vcap = 2 * qp.beam + 64;
  std::vector<uint32_t> fi(nq * fcap), fl(nq), vi, vl(nq);
  std::vector<float> fd(nq * fcap), vd;
  std::vector<uint64_t> cm(nq);
  for (;;) {
    vi.assign(nq * (size_t)vcap, 0);
    vd.assign(nq * (size_t)vcap, 0.f);
    ddg_status s = ddg_beam_search(dv.h, q.data(), nq, qid.data(), st.data(), (uint32_t)st.size(), &qp, fcap,
                                   fi.data(), fd.data(), fl.data(), vcap, vi.data(), vd.data(), vl.data(), cm.data());
    if (s == DDG_EINVAL && nq && *std::max_element(vl.begin(), vl.end()) > vcap) {
      vcap = *std::max_element(vl.begin(), vl.end());  // retry with room for the longest list
      continue;
    }
    check(s, dv.h);
    break;
  }
  std::vector<beam_result<indexType>> out(nq);
  for (size_t i = 0; i < nq; i++) {
    auto& fr = out[i].first.first;
    auto& vis = out[i].first.second;
    for (uint32_t j = 0; j < fl[i]; j++) fr.push_back({(indexType)fi[i * fcap + j], fd[i * fcap + j]});
    for (uint32_t j = 0; j < vl[i]; j++) vis.push_back({(indexType)vi[i * vcap + j], vd[i * vcap + j]});
    out[i].second = (size_t)cm[i];
  }
  return out;
}

// beam_search (beamSearch.h:50-184): ((frontier, visited), dist_cmps).
template <class Point, class PR, class indexType, class GraphT, class Seq, class QPT,
          class = decltype(std::declval<Seq&>().begin())>
beam_result<indexType> beam_search(Point p, GraphT& G, PR& Points, Seq starting_points, QPT& QP) {
  std::vector<indexType> st(starting_points.begin(), starting_points.end());
  std::vector<Point> one{p};
  return beam_search_batch<Point, PR, indexType>(one, G, Points, st, QP)[0];
}

// beam_search from one start point (beamSearch.h:41-48).
template <class Point, class PR, class indexType, class GraphT, class QPT>
beam_result<indexType> beam_search(Point p, GraphT& G, PR& Points, indexType starting_point, QPT& QP) {
  std::vector<indexType> st{starting_point};
  return beam_search<Point, PR, indexType>(p, G, Points, st, QP);
}

// searchAll (beamSearch.h:314-346): query i uses p.id() = i.
template <class Point, class PR, class indexType, class GraphT, class StatsT, class QPT>
std::vector<std::vector<indexType>> searchAll(PR& Query_Points, GraphT& G, PR& Base_Points, StatsT& QueryStats,
                                              indexType starting_point, QPT& QP) {
  using T = detail::elem_t<PR>;
  if (QP.k > QP.beamSize) {
    std::cout << "Error: beam search parameter Q = " << QP.beamSize << " same size or smaller than k = " << QP.k
              << std::endl;
    abort();
  }
  Device& dv = device(Base_Points, G.max_degree());
  std::lock_guard<std::mutex> lk(dv.mu);
  dv.sync(G);
  const size_t nq = Query_Points.size();
  std::vector<T> tmp;
  const T* q = detail::range_rows(Query_Points, tmp);
  ddg_query_params qp{(uint32_t)QP.k, (uint32_t)QP.beamSize, QP.cut, (uint64_t)std::max<long>(QP.limit, 0),
                      (uint32_t)std::max<long>(QP.degree_limit, 0), DDG_TRAV_NATIVE, 0};
  std::vector<uint32_t> ids(nq * QP.k), vis(nq);
  std::vector<uint64_t> cmps(nq);
  uint32_t start = (uint32_t)starting_point;
  check(ddg_search(dv.h, q, nq, nullptr, &start, 1, &qp, nullptr, ids.data(), nullptr, cmps.data(), vis.data()),
        dv.h);
  std::vector<std::vector<indexType>> all(nq);
  const bool count = detail::stats_fits(QueryStats, nq, 0);
  for (size_t i = 0; i < nq; i++) {
    all[i].assign(ids.begin() + i * QP.k, ids.begin() + (i + 1) * QP.k);
    if (!count) continue;
    QueryStats.increment_visited((indexType)i, (indexType)vis[i]);
    QueryStats.increment_dist((indexType)i, (indexType)cmps[i]);
  }
  return all;
}

// ============================================================== knn_index
template <class Point, class PR, class indexType = uint32_t>
struct knn_index {
  long R = 0, L = 0;
  double alpha = 1.2;
  bool two_pass = false;
  indexType start_point = 0;
  std::set<indexType> delete_set;  // index.h:52; nothing on the build or search path reads it

  template <class BPT>
  explicit knn_index(BPT& BP) : R(BP.R), L(BP.L), alpha(BP.alpha), two_pass(BP.two_pass) {}
  indexType get_start() { return start_point; }

  // index.h:139-164: marks points deleted.  An id above G.size() aborts (the
  // reference's own bound, which lets G.size() through); the start point is
  // refused with a message.
  template <class GraphT>
  void lazy_delete(indexType p, GraphT& G) {
    if ((size_t)p > G.size()) {
      std::cout << "ERROR: invalid point " << p << " given to lazy_delete" << std::endl;
      abort();
    }
    if (p == start_point) {
      std::cout << "Deleting start_point not permitted; continuing" << std::endl;
      return;
    }
    delete_set.insert(p);
  }
  template <class Seq, class GraphT>
  void lazy_delete(const Seq& deletes, GraphT& G) {
    for (indexType p : deletes) lazy_delete(p, G);
  }

  template <class GraphT, class StatsT>
  void build_index(GraphT& G, PR& Points, StatsT& BuildStats) {
    start_point = 0;
    Device& dv = device(Points, G.max_degree());
    std::lock_guard<std::mutex> lk(dv.mu);
    dv.sync(G);
    ddg_build_params bp{(uint32_t)R, (uint32_t)L, alpha, two_pass ? 1 : 0, 0, 2.0, 1.0, 0, 1, 0};
    std::vector<uint32_t> vc(G.size(), 0);
    check(ddg_build(dv.h, &bp, vc.data(), nullptr), dv.h);
    dv.pull(G);
    if (detail::stats_fits(BuildStats, G.size(), 0))
      for (size_t i = 0; i < G.size(); i++) BuildStats.increment_visited((indexType)i, (indexType)vc[i]);
  }

  template <class Seq, class GraphT, class StatsT>
  void batch_insert(Seq& inserts, GraphT& G, PR& Points, StatsT& /*BuildStats*/, double alpha_,
                    bool random_order = false, double base = 2, double max_fraction = .02, bool /*print*/ = true) {
    Device& dv = device(Points, G.max_degree());
    std::lock_guard<std::mutex> lk(dv.mu);
    dv.sync(G);
    std::vector<uint32_t> ins(inserts.begin(), inserts.end());
    check(ddg_batch_insert(dv.h, ins.data(), ins.size(), alpha_, (uint32_t)L, start_point, random_order ? 1 : 0,
                           base, max_fraction), dv.h);
    dv.pull(G);
  }

  // robustPrune for many points at once (no reference counterpart: the
  // batched form of the per-point calls of index.h:281 / 314 and
  // bsp_index.hpp:272, 392): cands[t] are (id, distance to pts[t]) pairs.
  template <class GraphT>
  std::vector<std::vector<indexType>> robustPrune_batch(const std::vector<indexType>& pts,
                                                        const std::vector<std::vector<pid<indexType>>>& cands,
                                                        GraphT& G, PR& Points, double alpha_, bool add = true) {
    Device& dv = device(Points, G.max_degree());
    std::lock_guard<std::mutex> lk(dv.mu);
    std::vector<uint32_t> p32(pts.begin(), pts.end());
    if (add) dv.sync_rows(G, p32.data(), p32.size());
    std::vector<uint64_t> off(pts.size() + 1, 0);
    std::vector<uint32_t> ids;
    std::vector<float> ds;
    for (size_t t = 0; t < pts.size(); t++) {
      for (auto& c : cands[t]) { ids.push_back((uint32_t)c.first); ds.push_back(c.second); }
      off[t + 1] = ids.size();
    }
    std::vector<uint32_t> out(pts.size() * R), len(pts.size());
    check(ddg_robust_prune(dv.h, pts.size(), p32.data(), off.data(), ids.data(), ds.data(), alpha_, add ? 1 : 0,
                           (uint32_t)R, out.data(), len.data()), dv.h);
    std::vector<std::vector<indexType>> res(pts.size());
    for (size_t t = 0; t < pts.size(); t++) res[t].assign(out.begin() + t * R, out.begin() + t * R + len[t]);
    return res;
  }

  // index.h:63-123: `cand` is any sequence of (id, distance to p) pairs
  // (index.h:63-110) or of ids whose distances are computed on the device
  // (index.h:114-123).
  template <class CandSeq, class GraphT>
  std::vector<indexType> robustPrune(indexType p, CandSeq&& cand, GraphT& G, PR& Points, double alpha_,
                                     bool add = true) {
    using E = std::decay_t<decltype(*std::begin(cand))>;
    if constexpr (detail::is_pair<E>::value) {
      std::vector<std::vector<pid<indexType>>> c(1);
      for (auto& x : cand) c[0].push_back({(indexType)x.first, (float)x.second});
      return robustPrune_batch(std::vector<indexType>{p}, c, G, Points, alpha_, add)[0];
    } else {
      Device& dv = device(Points, G.max_degree());
      std::lock_guard<std::mutex> lk(dv.mu);
      uint32_t p32 = (uint32_t)p;
      if (add) dv.sync_rows(G, &p32, 1);
      std::vector<uint32_t> ids;
      for (auto& x : cand) ids.push_back((uint32_t)x);
      uint64_t off[2] = {0, ids.size()};
      std::vector<uint32_t> out(R);
      uint32_t len = 0;
      check(ddg_robust_prune(dv.h, 1, &p32, off, ids.data(), nullptr, alpha_, add ? 1 : 0, (uint32_t)R, out.data(),
                             &len), dv.h);
      return std::vector<indexType>(out.begin(), out.begin() + len);
    }
  }
};

}  // namespace ddg
