// Test infrastructure: include/ddg/knn_index.hpp used as a drop-in on the
// REFERENCE's own types.  Compiled by oracle/Makefile (target `ref`) against
// the reference's headers where they lie (graph.h, point_range.h,
// euclidian_point.h, beamSearch.h, types.h, stats.h, parlay) into
// oracle/_ref/dropin_demo, linked with libddg.so.  In one process it runs the
// reference's own beam_search / searchAll (CPU) and the adapter's (GPU) on the
// same Graph<unsigned> and PointRange<float, Euclidian_Point<float>> objects
// and compares them bit for bit; tests/test_adapter.py runs it on the GPU.
//
//   dropin_demo <base> <queries> <graph> <out_prefix> <Ls> <R> <L> <alpha>
//
// base / queries: reference vector files ([u32 n][u32 d][n*d f32]); graph:
// Graph::save format.  Writes <out_prefix>.prune (robustPrune results of the
// build-mode searches of points 0..199: per point [u32 p][u32 len][len ids])
// and <out_prefix>.graph (ddg::knn_index::build_index into a reference
// Graph).  Exit status 0 when every comparison matched.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "utils/beamSearch.h"
#include "utils/euclidian_point.h"
#include "utils/graph.h"
#include "utils/point_range.h"
#include "utils/stats.h"
#include "utils/types.h"

#include "ddg/knn_index.hpp"

using Point = Euclidian_Point<float>;
using PR = PointRange<float, Point>;
using pidf = std::pair<unsigned, float>;

template <class A, class B>
static bool same_pairs(const A& a, const B& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++) {
    if (a[i].first != b[i].first) return false;
    if (std::memcmp(&a[i].second, &b[i].second, 4) != 0) return false;
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc != 9) {
    std::fprintf(stderr, "usage: dropin_demo base queries graph out_prefix Ls R L alpha\n");
    return 2;
  }
  PR Base(argv[1]);
  PR Query(argv[2]);
  Graph<unsigned> G(argv[3]);
  const std::string out = argv[4];
  const long Ls = atol(argv[5]), R = atol(argv[6]), L = atol(argv[7]);
  const double alpha = atof(argv[8]);
  int bad = 0;

  // 1. beam_search, one point per call and batched, against the reference's own
  QueryParams QP(10, Ls, 1.35, (long)G.size(), G.max_degree());
  std::vector<Point> qs;
  for (size_t i = 0; i < Query.size(); i++) qs.push_back(Query[i]);
  auto batch = ddg::beam_search_batch<Point, PR, unsigned>(qs, G, Base, std::vector<unsigned>{0u}, QP);
  for (size_t i = 0; i < Query.size(); i++) {
    auto ref = beam_search<Point, PR, unsigned>(Query[i], G, Base, 0u, QP);
    const auto& b = batch[i];
    bool ok = same_pairs(ref.first.first, b.first.first) && same_pairs(ref.first.second, b.first.second) &&
              ref.second == b.second;
    if (i < 40) {
      auto one = ddg::beam_search<Point, PR, unsigned>(Query[i], G, Base, 0u, QP);
      ok = ok && same_pairs(one.first.first, b.first.first) && same_pairs(one.first.second, b.first.second) &&
           one.second == b.second;
    }
    if (!ok && bad++ < 5) std::printf("beam_search differs on query %zu\n", i);
  }
  std::printf("beam_search: %zu queries compared\n", Query.size());

  // 2. searchAll against the reference's own
  stats<unsigned> s1(Query.size()), s2(Query.size());
  auto r1 = searchAll<Point, PR, unsigned>(Query, G, Base, s1, 0u, QP);
  auto r2 = ddg::searchAll<Point, PR, unsigned>(Query, G, Base, s2, 0u, QP);
  for (size_t i = 0; i < Query.size(); i++) {
    bool ok = r1[i].size() == r2[i].size() && s1.visited[i] == s2.visited[i] && s1.distances[i] == s2.distances[i];
    for (size_t j = 0; ok && j < r1[i].size(); j++) ok = r1[i][j] == r2[i][j];
    if (!ok && bad++ < 10) std::printf("searchAll differs on query %zu\n", i);
  }

  // 3. robustPrune of build-mode visited lists (index.h:277-281), one call and batched
  BuildParams BP(R, L, alpha, false);
  ddg::knn_index<Point, PR, unsigned> I(BP);
  QueryParams QB((long)0, L, 0.0, (long)Base.size(), G.max_degree());
  FILE* fp = std::fopen((out + ".prune").c_str(), "wb");
  std::vector<unsigned> pts;
  std::vector<std::vector<ddg::pid<unsigned>>> cands;
  std::vector<std::vector<unsigned>> singles;
  for (unsigned p = 0; p < 200; p++) {
    auto visited = beam_search<Point, PR, unsigned>(Base[p], G, Base, 0u, QB).first.second;
    auto o = I.robustPrune(p, visited, G, Base, alpha);
    uint32_t h[2] = {p, (uint32_t)o.size()};
    std::fwrite(h, 4, 2, fp);
    std::fwrite(o.data(), 4, o.size(), fp);
    pts.push_back(p);
    cands.emplace_back(visited.begin(), visited.end());
    singles.push_back(o);
  }
  std::fclose(fp);
  auto batched = I.robustPrune_batch(pts, cands, G, Base, alpha);
  for (size_t t = 0; t < pts.size(); t++)
    if (batched[t] != singles[t] && bad++ < 15) std::printf("robustPrune_batch differs on point %u\n", pts[t]);

  // 4. a graph row changed on the host reaches the device before the next search
  {
    auto e = G[7];
    std::vector<unsigned> nb;
    for (size_t j = 0; j < e.size(); j++) nb.push_back(e[j]);
    std::reverse(nb.begin(), nb.end());
    if (!nb.empty()) nb.pop_back();
    G[7].update_neighbors(nb);
    G[0].update_neighbors(std::vector<unsigned>{7});  // every search now goes through row 7
    for (size_t i = 0; i < 20; i++) {
      auto ref = beam_search<Point, PR, unsigned>(Query[i], G, Base, 0u, QP);
      auto one = ddg::beam_search<Point, PR, unsigned>(Query[i], G, Base, 0u, QP);
      if (!(same_pairs(ref.first.first, one.first.first) && same_pairs(ref.first.second, one.first.second) &&
            ref.second == one.second) && bad++ < 20)
        std::printf("beam_search after a host graph change differs on query %zu\n", i);
    }
  }

  // 5. build_index into a reference Graph
  Graph<unsigned> G2(R, Base.size());
  stats<unsigned> bs(Base.size());
  I.build_index(G2, Base, bs);
  G2.save(const_cast<char*>((out + ".graph").c_str()));
  size_t vsum = 0;
  for (size_t i = 0; i < Base.size(); i++) vsum += bs.visited[i];
  if (vsum == 0 && bad++ < 25) std::printf("build_index counted no visits\n");

  std::printf("dropin_demo: %d mismatches\n", bad);
  return bad ? 5 : 0;
}
