// Test program for include/ddg/knn_index.hpp with its standalone types: the
// build and query calls of src/diskann.cpp:41-81 and check_nn_recall.h:52-67,
// written against the header-only adapter exactly as a caller of the
// reference would write them.
//
//   adapter_demo <base.u8bin> <queries.u8bin> <out_graph> <out_ids> R L alpha k beam
//
// Builds with knn_index::build_index, saves the graph (Graph::save format),
// runs searchAll from start point 0 and writes the ids (u32, nq x k), then
// checks lazy_delete's bookkeeping.  Exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ddg/knn_index.hpp"

int main(int argc, char** argv) {
  if (argc != 10) {
    std::fprintf(stderr, "usage: adapter_demo base queries out_graph out_ids R L alpha k beam\n");
    return 2;
  }
  const long R = atol(argv[5]), L = atol(argv[6]), k = atol(argv[8]), beam = atol(argv[9]);
  const double alpha = atof(argv[7]);
  using Point = ddg::Euclidian_Point<uint8_t>;
  using PR = ddg::PointRange<uint8_t, Point>;

  ddg::BuildParams BP(R, L, alpha, false);
  PR Points(argv[1]);
  ddg::Graph G(BP.R, Points.size());
  ddg::stats<unsigned int> BuildStats(G.size());
  ddg::knn_index<Point, PR, unsigned int> I(BP);
  I.build_index(G, Points, BuildStats);
  G.save(argv[3]);

  PR Query(argv[2]);
  ddg::stats<unsigned int> QueryStats(Query.size());
  ddg::QueryParams QP(k, beam, 1.35, (long)G.size(), G.max_degree());
  auto res = ddg::searchAll<Point, PR, unsigned int>(Query, G, Points, QueryStats, 0u, QP);
  FILE* f = std::fopen(argv[4], "wb");
  if (!f) return 3;
  for (auto& r : res) std::fwrite(r.data(), 4, r.size(), f);
  std::fclose(f);

  std::vector<unsigned int> del = {1, 2, 0};  // the start point is refused
  I.lazy_delete(del, G);
  if (I.delete_set.size() != 2) return 4;
  return 0;
}
