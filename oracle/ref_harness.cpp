// Test infrastructure: a command-line driver around the REFERENCE's own
// search-path headers (include/utils/{beamSearch,graph,point_range,
// euclidian_point,mips_point,NSGDist,types,stats}.h and the vendored parlay),
// compiled from where they lie by oracle/Makefile into oracle/_ref/.  Only
// tests/golden/make_golden.py runs it, to write golden vectors, and the CPU
// tests compare the oracle against it when it is present.  Nothing here is
// shipped or timed.
//
// include/index.h (build path) is NOT compiled: it includes <mpi.h> and
// <fmt/core.h>, which this image does not have.
//
// Modes (all binary files little-endian):
//   hash <count> <out>                         u64 hash64_2(i), i < count
//   dist <dtype> <metric> <A> <B> <out>        f32 Point(A[i]).distance(Point(B[i]))
//   search <dtype> <metric> <base> <queries> <graph> <k> <beam> <cut> <limit>
//          <degree_limit> <start> <out>        beam_search(Query[i], ...) per query
//   insert <dtype> <metric> <base> <graph> <ids> <L> <start> <out>
//                                              build-mode beam_search(Base[id]) traces
//   groupby <pairs> <out>                      parlay::group_by_key
//   perm <n> <out>                             parlay::random_permutation<int>(n)
//                                              (batch_insert's random_order)
//   qmips <dtype> <codes> <queries> <qq> <max> <min> <out>
//                                              f32 Quantized_Mips_Point<T>(codes[j])
//                                              .distance(query i) for every (i, j),
//                                              the query a Mips_Point<float> (qq 0)
//                                              or a Quantized_Mips_Point<T> (qq 1)
// Vector files use the reference format [u32 n][u32 d][n*d T]; <ids> and
// <pairs> are raw u32 arrays.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "utils/beamSearch.h"
#include "utils/euclidian_point.h"
#include "utils/graph.h"
#include "utils/mips_point.h"
#include "utils/point_range.h"
#include "utils/types.h"
#include "parlay/primitives.h"
#include "parlay/random.h"
#include "parlay/utilities.h"

static std::vector<uint32_t> read_u32(const char* fn) {
  std::ifstream f(fn, std::ios::binary | std::ios::ate);
  size_t sz = f.tellg();
  f.seekg(0);
  std::vector<uint32_t> v(sz / 4);
  f.read((char*)v.data(), sz);
  return v;
}

template <class PointT, class T>
static int run_search(int argc, char** argv) {
  using PR = PointRange<T, PointT>;
  PR Base(argv[4]);
  PR Query(argv[5]);
  Graph<unsigned int> G(argv[6]);
  QueryParams QP(atol(argv[7]), atol(argv[8]), atof(argv[9]), atol(argv[10]), atol(argv[11]));
  unsigned start = (unsigned)atol(argv[12]);
  FILE* out = fopen(argv[13], "wb");
  size_t nq = Query.size();
  std::vector<std::vector<char>> rec(nq);
  parlay::parallel_for(0, nq, [&](size_t i) {
    auto [pr, cmps] = beam_search<PointT, PR, unsigned>(Query[i], G, Base, start, QP);
    auto& fr = pr.first;
    auto& vis = pr.second;
    uint32_t h[2] = {(uint32_t)fr.size(), (uint32_t)vis.size()};
    uint64_t c = cmps;
    auto& r = rec[i];
    auto put = [&](const void* p, size_t n) { r.insert(r.end(), (const char*)p, (const char*)p + n); };
    put(h, 8);
    put(&c, 8);
    for (auto& e : fr) put(&e.first, 4);
    for (auto& e : fr) put(&e.second, 4);
  });
  for (auto& r : rec) fwrite(r.data(), 1, r.size(), out);
  fclose(out);
  return 0;
}

template <class PointT, class T>
static int run_insert(int argc, char** argv) {
  using PR = PointRange<T, PointT>;
  PR Base(argv[4]);
  Graph<unsigned int> G(argv[5]);
  std::vector<uint32_t> ids = read_u32(argv[6]);
  long L = atol(argv[7]);
  unsigned start = (unsigned)atol(argv[8]);
  FILE* out = fopen(argv[9], "wb");
  QueryParams QP((long)0, L, (double)0.0, (long)Base.size(), (long)G.max_degree());
  for (uint32_t id : ids) {
    auto [pr, cmps] = beam_search<PointT, PR, unsigned>(Base[id], G, Base, start, QP);
    auto& fr = pr.first;
    auto& vis = pr.second;
    uint32_t h[2] = {(uint32_t)fr.size(), (uint32_t)vis.size()};
    uint64_t c = cmps;
    fwrite(h, 4, 2, out);
    fwrite(&c, 8, 1, out);
    for (auto& e : fr) fwrite(&e.first, 4, 1, out);
    for (auto& e : fr) fwrite(&e.second, 4, 1, out);
    for (auto& e : vis) fwrite(&e.first, 4, 1, out);
    for (auto& e : vis) fwrite(&e.second, 4, 1, out);
  }
  fclose(out);
  return 0;
}

template <class PointT, class T>
static int run_dist(int argc, char** argv) {
  using PR = PointRange<T, PointT>;
  PR A(argv[4]);
  PR B(argv[5]);
  FILE* out = fopen(argv[6], "wb");
  for (size_t i = 0; i < A.size(); i++) {
    float d = A[i].distance(B[i]);
    fwrite(&d, 4, 1, out);
  }
  fclose(out);
  return 0;
}

// [u32 n][u32 d][n*d T] -> rows (unpadded), n, d
template <class T>
static std::vector<T> read_vecs(const char* fn, uint32_t& n, uint32_t& d) {
  std::ifstream f(fn, std::ios::binary);
  f.read((char*)&n, 4);
  f.read((char*)&d, 4);
  std::vector<T> v((size_t)n * d);
  f.read((char*)v.data(), v.size() * sizeof(T));
  return v;
}

template <class T>
static int run_qmips(char** argv) {
  uint32_t n, d, nq, dq;
  std::vector<T> codes = read_vecs<T>(argv[3], n, d);
  const bool qq = atoi(argv[5]) != 0;
  const float mx = (float)atof(argv[6]), mn = (float)atof(argv[7]);
  FILE* out = fopen(argv[8], "wb");
  if (qq) {
    std::vector<T> q = read_vecs<T>(argv[4], nq, dq);
    for (uint32_t i = 0; i < nq; i++) {
      Quantized_Mips_Point<T> qp(q.data() + (size_t)i * d, d, d, i, mx, mn);
      for (uint32_t j = 0; j < n; j++) {
        Quantized_Mips_Point<T> cp(codes.data() + (size_t)j * d, d, d, j, mx, mn);
        float r = cp.distance(qp);
        fwrite(&r, 4, 1, out);
      }
    }
  } else {
    std::vector<float> q = read_vecs<float>(argv[4], nq, dq);
    for (uint32_t i = 0; i < nq; i++) {
      Mips_Point<float> qp(q.data() + (size_t)i * d, d, d, i);
      for (uint32_t j = 0; j < n; j++) {
        Quantized_Mips_Point<T> cp(codes.data() + (size_t)j * d, d, d, j, mx, mn);
        float r = cp.distance(qp);
        fwrite(&r, 4, 1, out);
      }
    }
  }
  fclose(out);
  return 0;
}

template <template <class> class PT>
static int dispatch_dtype(const std::string& mode, const std::string& dtype, int argc, char** argv) {
  if (dtype == "uint8") {
    if (mode == "search") return run_search<PT<uint8_t>, uint8_t>(argc, argv);
    if (mode == "insert") return run_insert<PT<uint8_t>, uint8_t>(argc, argv);
    return run_dist<PT<uint8_t>, uint8_t>(argc, argv);
  }
  if (dtype == "int8") {
    if (mode == "search") return run_search<PT<int8_t>, int8_t>(argc, argv);
    if (mode == "insert") return run_insert<PT<int8_t>, int8_t>(argc, argv);
    return run_dist<PT<int8_t>, int8_t>(argc, argv);
  }
  if (mode == "search") return run_search<PT<float>, float>(argc, argv);
  if (mode == "insert") return run_insert<PT<float>, float>(argc, argv);
  return run_dist<PT<float>, float>(argc, argv);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string mode = argv[1];
  if (mode == "hash") {
    uint64_t cnt = strtoull(argv[2], nullptr, 10);
    FILE* out = fopen(argv[3], "wb");
    for (uint64_t i = 0; i < cnt; i++) {
      uint64_t h = parlay::hash64_2(i);
      fwrite(&h, 8, 1, out);
    }
    fclose(out);
    return 0;
  }
  if (mode == "perm") {
    const int n = atoi(argv[2]);
    auto p = parlay::random_permutation<int>(n);
    FILE* out = fopen(argv[3], "wb");
    fwrite(p.data(), 4, p.size(), out);
    fclose(out);
    return 0;
  }
  if (mode == "groupby") {
    std::vector<uint32_t> raw = read_u32(argv[2]);
    parlay::sequence<std::pair<unsigned, unsigned>> pairs(raw.size() / 2);
    for (size_t i = 0; i < pairs.size(); i++) pairs[i] = {raw[2 * i], raw[2 * i + 1]};
    auto g = parlay::group_by_key(pairs);
    FILE* out = fopen(argv[3], "wb");
    uint64_t ng = g.size();
    fwrite(&ng, 8, 1, out);
    for (auto& [k, vs] : g) {
      uint32_t h[2] = {k, (uint32_t)vs.size()};
      fwrite(h, 4, 2, out);
      fwrite(vs.data(), 4, vs.size(), out);
    }
    fclose(out);
    return 0;
  }
  if (mode == "qmips") {
    std::string dtype = argv[2];
    if (dtype == "uint8") return run_qmips<uint8_t>(argv);
    if (dtype == "int8") return run_qmips<int8_t>(argv);
    return 2;
  }
  if (mode == "search" || mode == "insert" || mode == "dist") {
    std::string dtype = argv[2], metric = argv[3];
    if (metric == "mips") return dispatch_dtype<Mips_Point>(mode, dtype, argc, argv);
    return dispatch_dtype<Euclidian_Point>(mode, dtype, argc, argv);
  }
  return 2;
}
