// knn_index.hpp — header-only C++ adapter that re-exposes the reference's
// template entry points over the C-ABI of ddg.h, so a caller of
// include/index.h / include/utils/beamSearch.h keeps its call sites:
//
//   ddg::knn_index<T, Metric>{BP}.build_index(G, Points, stats)   index.h:125-137
//   ddg::knn_index<...>::batch_insert(inserts, G, Points, ...)     index.h:213-341
//   ddg::knn_index<...>::robustPrune(p, cand, G, Points, alpha)     index.h:63-123
//   ddg::searchAll(Query, G, Base, stats, start, QP)                beamSearch.h:314-346
//
// Graph and PointRange keep the reference's in-memory layouts
// (graph.h:126-128: n x (R+1) u32 rows; point_range.h:45-50: rows padded
// to 64 bytes), so conversion is a plain copy.  A non-OK status from the
// library is reported like the reference reports its errors: a message on
// stdout and abort().
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <set>
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

struct Euclidian { static constexpr ddg_metric v = DDG_L2; };  // Euclidian_Point
struct Mips { static constexpr ddg_metric v = DDG_MIPS; };      // Mips_Point

// PointRange (point_range.h:53-113): n rows of aligned_dims elements.
template <class T, class Metric = Euclidian>
struct PointRange {
  std::vector<T> values;
  unsigned dims = 0, aligned_dims = 0;
  size_t n = 0;
  PointRange() = default;
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
  const T* row(size_t i) const { return values.data() + i * aligned_dims; }
};

// Graph (graph.h:119-209).
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
  uint32_t degree(size_t i) const { return graph[i * (maxDeg + 1)]; }
  const uint32_t* neighbors(size_t i) const { return &graph[i * (maxDeg + 1) + 1]; }
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
};

// One device handle per (PointRange, Graph) pair.
template <class T, class Metric>
struct Device {
  ddg_index* h = nullptr;
  Device(const PointRange<T, Metric>& P, const Graph& G) {
    check(ddg_create(dtype_of<T>::v, Metric::v, P.n, P.dims, P.aligned_dims, P.values.data(),
                     (uint32_t)G.maxDeg, &h), nullptr);
    check(ddg_graph_set(h, G.graph.data()), h);
  }
  ~Device() { ddg_destroy(h); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
};

template <class T, class Metric = Euclidian, class indexType = uint32_t>
struct knn_index {
  BuildParams BP;
  indexType start_point = 0;
  std::set<indexType> delete_set;  // index.h:52; nothing on the build or search path reads it
  explicit knn_index(BuildParams& bp) : BP(bp) {}
  indexType get_start() { return start_point; }

  // index.h:139-164: marks points deleted.  An id above G.size() aborts (the
  // reference's own bound, which lets G.size() through); the start point is
  // refused with a message.
  void lazy_delete(indexType p, Graph& G) {
    if (p > (long)G.size()) {
      std::cout << "ERROR: invalid point " << p << " given to lazy_delete" << std::endl;
      abort();
    }
    if (p == start_point) {
      std::cout << "Deleting start_point not permitted; continuing" << std::endl;
      return;
    }
    delete_set.insert(p);
  }
  template <typename Seq>
  void lazy_delete(const Seq& deletes, Graph& G) {
    for (indexType p : deletes) lazy_delete(p, G);
  }

  void build_index(Graph& G, PointRange<T, Metric>& Points, stats<indexType>& BuildStats) {
    start_point = 0;
    Device<T, Metric> d(Points, G);
    ddg_build_params bp{(uint32_t)BP.R, (uint32_t)BP.L, BP.alpha, BP.two_pass ? 1 : 0, 0, 2.0, 1.0, 0, 1, 0};
    std::vector<uint32_t> vc(G.n, 0);
    check(ddg_build(d.h, &bp, vc.data(), nullptr), d.h);
    check(ddg_graph_get(d.h, G.graph.data()), d.h);
    if (BuildStats.visited.size() == G.n)
      for (size_t i = 0; i < G.n; i++) BuildStats.visited[i] += vc[i];
  }

  void batch_insert(std::vector<indexType>& inserts, Graph& G, PointRange<T, Metric>& Points,
                    stats<indexType>& /*BuildStats*/, double alpha, bool random_order = false, double base = 2,
                    double max_fraction = .02, bool /*print*/ = true) {
    Device<T, Metric> d(Points, G);
    check(ddg_batch_insert(d.h, inserts.data(), inserts.size(), alpha, (uint32_t)BP.L, start_point,
                           random_order ? 1 : 0, base, max_fraction), d.h);
    check(ddg_graph_get(d.h, G.graph.data()), d.h);
  }

  std::vector<indexType> robustPrune(indexType p, std::vector<std::pair<indexType, float>>& cand, Graph& G,
                                     PointRange<T, Metric>& Points, double alpha, bool add = true) {
    Device<T, Metric> d(Points, G);
    std::vector<uint32_t> ids(cand.size());
    std::vector<float> ds(cand.size());
    for (size_t i = 0; i < cand.size(); i++) { ids[i] = cand[i].first; ds[i] = cand[i].second; }
    uint64_t off[2] = {0, cand.size()};
    std::vector<uint32_t> out(BP.R);
    uint32_t len = 0;
    check(ddg_robust_prune(d.h, 1, &p, off, ids.data(), ds.data(), alpha, add ? 1 : 0, (uint32_t)BP.R,
                           out.data(), &len), d.h);
    out.resize(len);
    return out;
  }
};

// searchAll (beamSearch.h:314-346): query i uses p.id() = i.
template <class T, class Metric, class indexType = uint32_t>
std::vector<std::vector<indexType>> searchAll(PointRange<T, Metric>& Query_Points, Graph& G,
                                              PointRange<T, Metric>& Base_Points, stats<indexType>& QueryStats,
                                              indexType starting_point, QueryParams& QP) {
  if (QP.k > QP.beamSize) {
    std::cout << "Error: beam search parameter Q = " << QP.beamSize << " same size or smaller than k = " << QP.k
              << std::endl;
    abort();
  }
  Device<T, Metric> d(Base_Points, G);
  const size_t nq = Query_Points.size();
  ddg_query_params qp{(uint32_t)QP.k, (uint32_t)QP.beamSize, QP.cut, (uint64_t)QP.limit,
                      (uint32_t)QP.degree_limit, DDG_TRAV_NATIVE, 0};
  std::vector<uint32_t> ids(nq * QP.k);
  std::vector<uint64_t> cmps(nq);
  std::vector<uint32_t> vis(nq);
  check(ddg_search(d.h, Query_Points.values.data(), nq, nullptr, &starting_point, 1, &qp, nullptr, ids.data(),
                   nullptr, cmps.data(), vis.data()), d.h);
  std::vector<std::vector<indexType>> all(nq);
  for (size_t i = 0; i < nq; i++) {
    all[i].assign(ids.begin() + i * QP.k, ids.begin() + (i + 1) * QP.k);
    if (QueryStats.visited.size() == nq) {
      QueryStats.visited[i] += vis[i];
      QueryStats.distances[i] += (indexType)cmps[i];
    }
  }
  return all;
}

}  // namespace ddg
