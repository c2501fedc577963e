// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference's Vamana/DiskANN build and search path,
// used ONLY as the checker in tests/, by __graft_entry__.smoke() and as the
// timed CPU baseline leg of bench.py.  The product (distributed_diskann_amd)
// never links, loads or calls this file.
//
// Compiled like the reference (Makefile:1): g++ -O2 -std=c++17 -pthread with
// no -march, plus -ffp-contract=off, so the FP32 summation order of
// NSGDist.h:113-131 and libstdc++'s std::sort tie order are the reference's.
//
// Function → reference anchor → what pins it (tests/test_oracle*.py):
//   dist_*            euclidian_point.h:44-65, NSGDist.h:113-131,
//                     mips_point.h:44-66          golden vectors from oracle/_ref
//                                                 (real reference headers), int64
//                                                 closed form for integer types
//   hash64_2          parlay/utilities.h:145-150  SplitMix64 textbook constants +
//                                                 golden vectors from oracle/_ref
//   beam_search       beamSearch.h:50-184         bit-exact vs oracle/_ref (real
//                                                 beam_search) on the random-xs
//                                                 and small-config fixtures
//   search_all        beamSearch.h:322-346        same fixtures (per query)
//   robust_prune      index.h:63-123              special cases: huge alpha keeps
//                                                 the R nearest distinct ids; alpha=0
//                                                 keeps only the nearest; brute-force
//                                                 re-statement on tiny inputs; the
//                                                 pruning invariant on every output
//   batch schedule    index.h:233-261,            closed form (2^i-1 doubling then
//                     bsp_index.hpp:162-166,238-246  max_batch strides) at n=1e4..1e7
//   group_by          parlay group_by_key         vs oracle/_ref (real parlay
//                     (group_by.h:155-163)        group_by_key) on random pairs
//   random_permutation parlay/random.h:89-159,    vs oracle/_ref (real parlay
//                     utilities.h:131-140         random_permutation) at n from 0
//                     (hash64), counting_sort.h   to 1M, both sides of the
//                     (stable bucket sort)        sequential threshold
//   batch_insert /    index.h:125-137, 213-341;   closed-form builds (hand-worked
//   build             bsp_index.hpp:141-434 (P=1)  1-D cases at |S|+|G[t]| = R and
//                                                 R+1, a reverse prune that keeps
//                                                 old edges, a two_pass build; huge
//                                                 alpha and alpha = 0 closed forms),
//                                                 each of three planted mistakes
//                                                 caught (tests/test_oracle_mutants.py);
//                                                 recall on random-xs against the
//                                                 reference's own GT file; graph
//                                                 invariants; BSP(P=1) == build_index
//                                                 without the final sort
//   ground_truth      (sklearn brute force in the  numpy float64 brute force
//                     reference's dataset tools)
//   quantized_mips    mips_point.h:111-187         golden vectors from oracle/_ref
//                                                 (the reference's own
//                                                 Quantized_Mips_Point::distance)
//
// Tie mode: ORC_TIE_REFERENCE sorts prune candidates and the final rows with
// std::sort by distance only (index.h:78-79, 133-136), exactly as the
// reference does.  ORC_TIE_ID sorts by (distance, id), which is the
// deterministic order the GPU build uses; everything else is identical.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <utility>
#include <vector>

namespace orc {

enum { U8 = 0, I8 = 1, F32 = 2 };
enum { L2 = 0, MIPS = 1 };
enum { TIE_REFERENCE = 0, TIE_ID = 1 };

// ---------------------------------------------------------------- distances
// euclidian_point.h:44-60 — integer squared L2, int32 accumulation, cast.
template <class T>
static inline float l2_int(const T* p, const T* q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) {
    int32_t t = (int32_t)((int16_t)q[i] - (int16_t)p[i]);
    r += t * t;
  }
  return (float)r;
}

// NSGDist.h:113-131 — the scalar path the reference's flags select:
// per group of four, ((d0^2 + d1^2) + d2^2) + d3^2 is added to the running
// sum; the last 0-3 elements are added one by one.
static inline float l2_f32(const float* a, const float* b, unsigned d) {
  float r = 0.f;
  unsigned i = 0;
  for (; i + 4 <= d; i += 4) {
    float e0 = a[i] - b[i], e1 = a[i + 1] - b[i + 1];
    float e2 = a[i + 2] - b[i + 2], e3 = a[i + 3] - b[i + 3];
    float g = e0 * e0;
    g = g + e1 * e1;
    g = g + e2 * e2;
    g = g + e3 * e3;
    r = r + g;
  }
  for (; i < d; i++) {
    float e = a[i] - b[i];
    r = r + e * e;
  }
  return r;
}

// mips_point.h:44-58 — integer negated inner product.
template <class T>
static inline float ip_int(const T* p, const T* q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) r += (int32_t)q[i] * (int32_t)p[i];
  return -((float)r);
}

// mips_point.h:60-66 — sequential float sum, negated.
static inline float ip_f32(const float* p, const float* q, unsigned d) {
  float r = 0.f;
  for (unsigned i = 0; i < d; i++) r = r + q[i] * p[i];
  return -r;
}

static inline float distance(int dtype, int metric, const void* a, const void* b, unsigned d) {
  if (metric == L2) {
    if (dtype == U8) return l2_int((const uint8_t*)a, (const uint8_t*)b, d);
    if (dtype == I8) return l2_int((const int8_t*)a, (const int8_t*)b, d);
    return l2_f32((const float*)a, (const float*)b, d);
  }
  if (dtype == U8) return ip_int((const uint8_t*)a, (const uint8_t*)b, d);
  if (dtype == I8) return ip_int((const int8_t*)a, (const int8_t*)b, d);
  return ip_f32((const float*)a, (const float*)b, d);
}

static inline size_t elem_size(int dtype) { return dtype == F32 ? 4 : 1; }

// parlay/utilities.h:145-150 (SplitMix64 finaliser).
static inline uint64_t hash64_2(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

// ------------------------------------------------------------------- views
struct Points {
  int dtype, metric;
  uint64_t n;
  uint32_t d, ad;
  const void* rows;
  const void* row(uint64_t i) const {
    return (const char*)rows + i * (uint64_t)ad * elem_size(dtype);
  }
  // Points[a].distance(Points[b]) == euclidian_distance(a.values, b.values)
  float dist(const void* a, const void* b) const { return distance(dtype, metric, a, b, d); }
  float dist_id(uint64_t a, uint64_t b) const { return dist(row(a), row(b)); }
};

struct GraphView {
  uint64_t n;
  uint32_t R;
  uint32_t* rows;  // n * (R+1): [deg, nbr_0 .. nbr_{R-1}]   (graph.h:126-128)
  uint32_t* row(uint64_t i) const { return rows + i * (uint64_t)(R + 1); }
  uint32_t deg(uint64_t i) const { return row(i)[0]; }
};

using pid = std::pair<uint32_t, float>;

// --------------------------------------------------------------- threading
static int num_threads(int requested) {
  if (requested > 0) return requested;
  const char* e = getenv("ORACLE_NUM_THREADS");
  if (!e) e = getenv("PARLAY_NUM_THREADS");
  if (e && atoi(e) > 0) return atoi(e);
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class F>
static void parallel_for(uint64_t lo, uint64_t hi, int nthreads, F f) {
  if (hi <= lo) return;
  uint64_t total = hi - lo;
  int t = num_threads(nthreads);
  if (t <= 1 || total < 2) {
    for (uint64_t i = lo; i < hi; i++) f(i);
    return;
  }
  if ((uint64_t)t > total) t = (int)total;
  std::atomic<uint64_t> next(lo);
  const uint64_t chunk = std::max<uint64_t>(1, total / ((uint64_t)t * 16));
  auto worker = [&]() {
    for (;;) {
      uint64_t s = next.fetch_add(chunk);
      if (s >= hi) break;
      uint64_t e = std::min(hi, s + chunk);
      for (uint64_t i = s; i < e; i++) f(i);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < t; i++) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

// ------------------------------------------------------------- beam search
struct QP {
  int64_t k, beam;
  double cut;
  int64_t limit, degree_limit;
};

struct SearchResult {
  std::vector<pid> frontier, visited;
  uint64_t dist_cmps = 0;
};

static inline bool less_pid(const pid& a, const pid& b) {
  return a.second < b.second || (a.second == b.second && a.first < b.first);
}

// beamSearch.h:50-184.  `query` is the query vector, `self_id` is p.id()
// (the row index for queries, the point id for inserts, -1 for none).
static void beam_search(const Points& P, const GraphView& G, const void* query, int64_t self_id,
                        const uint32_t* starts, uint32_t nstarts, const QP& qp, SearchResult& out) {
  const int64_t beam = qp.beam;
  // beamSearch.h:66 — lossy visited filter of 2^bits slots.
  int bits = std::max<int>(10, (int)std::ceil(std::log2((double)(beam * beam))) - 2);
  const uint32_t mask = (1u << bits) - 1u;
  std::vector<uint32_t> table((size_t)1 << bits, 0xFFFFFFFFu);
  auto seen = [&](uint32_t a) {
    uint32_t s = (uint32_t)(hash64_2(a) & mask);
    if (table[s] == a) return true;
    table[s] = a;
    return false;
  };
  auto dq = [&](uint32_t a) { return P.dist(P.row(a), query); };

  std::vector<pid> frontier;
  frontier.reserve(beam);
  for (uint32_t i = 0; i < nstarts; i++) frontier.push_back(pid(starts[i], dq(starts[i])));
  std::sort(frontier.begin(), frontier.end(), less_pid);

  std::vector<pid> unvisited((size_t)std::max<int64_t>(beam, (int64_t)frontier.size()));
  unvisited[0] = frontier[0];
  std::vector<pid>& visited = out.visited;
  visited.clear();
  uint64_t dist_cmps = nstarts;
  int64_t remain = 1, num_visited = 0;

  std::vector<pid> merged(beam + G.R + frontier.size());
  std::vector<pid> cand;
  std::vector<uint32_t> keep;
  while (remain > 0 && num_visited < qp.limit) {
    pid cur = unvisited[0];
    visited.insert(std::upper_bound(visited.begin(), visited.end(), cur, less_pid), cur);
    num_visited++;
    cand.clear();
    keep.clear();
    const uint32_t* row = G.row(cur.first);
    int64_t m = std::min<int64_t>(row[0], qp.degree_limit);
    for (int64_t i = 0; i < m; i++) {
      uint32_t a = row[1 + i];
      if ((int64_t)a == self_id || seen(a)) continue;  // short-circuit: self never touches the table
      keep.push_back(a);
    }
    float cutoff = ((int64_t)frontier.size() < beam) ? (float)std::numeric_limits<int>::max()
                                                      : frontier.back().second;
    for (uint32_t a : keep) {
      float dd = dq(a);
      dist_cmps++;
      if (dd >= cutoff) continue;
      cand.push_back(pid(a, dd));
    }
    std::sort(cand.begin(), cand.end(), less_pid);
    if (merged.size() < frontier.size() + cand.size()) merged.resize(frontier.size() + cand.size());
    int64_t nsz = std::set_union(frontier.begin(), frontier.end(), cand.begin(), cand.end(),
                                 merged.begin(), less_pid) - merged.begin();
    nsz = std::min<int64_t>(beam, nsz);
    if (qp.k > 0 && nsz > qp.k && P.metric == L2) {
      // beamSearch.h:162-167: the probe (0, cut * d_k) is a pair<int,double>
      // converted to pair<u32,float> by the comparator.
      pid probe(0u, (float)(qp.cut * (double)merged[qp.k].second));
      nsz = std::upper_bound(merged.begin(), merged.begin() + nsz, probe, less_pid) - merged.begin();
    }
    frontier.assign(merged.begin(), merged.begin() + nsz);
    if (unvisited.size() < frontier.size()) unvisited.resize(frontier.size());
    remain = std::set_difference(frontier.begin(), frontier.end(), visited.begin(), visited.end(),
                                 unvisited.begin(), less_pid) - unvisited.begin();
  }
  out.frontier = frontier;
  out.dist_cmps = dist_cmps;
}

// ------------------------------------------------------------ robustPrune
// index.h:63-110 (candidates with distances) and 114-123 (ids only).
static void robust_prune(const Points& P, const GraphView& G, uint32_t p, std::vector<pid>& cands,
                         double alpha, bool add, uint32_t R, int tie_mode, std::vector<uint32_t>& out) {
  if (add) {
    const uint32_t* row = G.row(p);
    for (uint32_t i = 0; i < row[0]; i++) cands.push_back(pid(row[1 + i], P.dist_id(row[1 + i], p)));
  }
  if (tie_mode == TIE_REFERENCE)
    std::sort(cands.begin(), cands.end(), [](const pid& a, const pid& b) { return a.second < b.second; });
  else
    std::sort(cands.begin(), cands.end(), less_pid);
  out.clear();
  size_t idx = 0;
  while (out.size() < R && idx < cands.size()) {
    uint32_t ps = cands[idx].first;
    idx++;
    if (ps == p || ps == 0xFFFFFFFFu) continue;
    out.push_back(ps);
    for (size_t i = idx; i < cands.size(); i++) {
      uint32_t pp = cands[i].first;
      if (pp == 0xFFFFFFFFu) continue;
      float dsp = P.dist_id(ps, pp);
      if (alpha * (double)dsp <= (double)cands[i].second) cands[i].first = 0xFFFFFFFFu;
    }
  }
}

// ------------------------------------------------------------ batch schedule
// index.h:233-261 / bsp_index.hpp:162-166, 238-246.
static size_t max_batch_size(double max_fraction, uint64_t count) {
  size_t mb = std::min<size_t>((size_t)(max_fraction * (double)(float)count), 1000000ul);
  return mb == 0 ? (size_t)count : mb;
}

static void schedule(uint64_t m, double base, size_t max_batch, std::vector<std::pair<uint64_t, uint64_t>>& out) {
  out.clear();
  size_t inc = 0, count = 0;
  while (count < m) {
    size_t floor_, ceiling;
    if (std::pow(base, (double)inc) <= (double)max_batch) {
      floor_ = (size_t)std::pow(base, (double)inc) - 1;
      ceiling = std::min<size_t>((size_t)std::pow(base, (double)(inc + 1)), m) - 1;
      count = std::min<size_t>((size_t)std::pow(base, (double)(inc + 1)), m) - 1;
    } else {
      floor_ = count;
      ceiling = std::min<size_t>(count + max_batch, m);
      count += max_batch;
    }
    out.push_back({floor_, ceiling});
    inc++;
  }
}

// ------------------------------------------------------------- group by key
// parlay::group_by_key semantics as used by batch_insert: values of one key
// keep their input order; groups are emitted here in increasing key order
// (group order does not matter to the caller: each group owns its own row).
static void group_by(const std::vector<std::pair<uint32_t, uint32_t>>& pairs,
                     std::vector<std::pair<uint32_t, std::vector<uint32_t>>>& groups) {
  std::vector<uint32_t> order(pairs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return pairs[a].first < pairs[b].first; });
  groups.clear();
  for (size_t i = 0; i < order.size(); i++) {
    const auto& pr = pairs[order[i]];
    if (groups.empty() || groups.back().first != pr.first) groups.push_back({pr.first, {}});
    groups.back().second.push_back(pr.second);
  }
}

// ------------------------------------------------------------- batch insert
struct BuildStats {
  uint64_t dist_cmps = 0, visited = 0, prunes_fwd = 0, prunes_rev = 0, appends = 0, pairs = 0;
};

// index.h:213-341 (random_order = false).  `max_batch_base` is n for
// knn_index::batch_insert (index.h:233-234) and m for the BSP variant
// (bsp_index.hpp:164-165).
static int batch_insert(const Points& P, GraphView& G, const uint32_t* inserts, uint64_t m, double alpha,
                        double base, double max_fraction, uint64_t max_batch_base, uint32_t start,
                        uint32_t L, uint32_t R, int tie_mode, int nthreads, uint32_t* visited_counts,
                        BuildStats* st) {
  for (uint64_t i = 0; i < m; i++)  // index.h:217-223 (which lets p == n through; that row would be
    if (inserts[i] >= G.n) return -2;  // read out of bounds, so it is rejected here)
  std::vector<std::pair<uint64_t, uint64_t>> batches;
  schedule(m, base, max_batch_size(max_fraction, max_batch_base), batches);
  std::vector<std::vector<uint32_t>> new_out;
  std::vector<uint64_t> vis_sz;
  std::vector<uint64_t> cmps;
  int err = 0;
  for (auto [lo, hi] : batches) {
    uint64_t B = hi - lo;
    new_out.assign(B, {});
    vis_sz.assign(B, 0);
    cmps.assign(B, 0);
    QP qp{0, (int64_t)L, 0.0, (int64_t)P.n, (int64_t)G.R};
    parallel_for(lo, hi, nthreads, [&](uint64_t i) {
      uint32_t index = inserts[i];
      SearchResult sr;
      beam_search(P, G, P.row(index), (int64_t)index, &start, 1, qp, sr);
      vis_sz[i - lo] = sr.visited.size();
      cmps[i - lo] = sr.dist_cmps;
      robust_prune(P, G, index, sr.visited, alpha, true, R, tie_mode, new_out[i - lo]);
    });
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint64_t i = 0; i < B; i++) {
      uint32_t index = inserts[i + lo];
      if (visited_counts) visited_counts[index] += (uint32_t)vis_sz[i];
      if (st) { st->visited += vis_sz[i]; st->dist_cmps += cmps[i]; st->prunes_fwd++; }
      for (uint32_t t : new_out[i]) pairs.push_back({t, index});
    }
    for (uint64_t i = 0; i < B; i++) {  // update_neighbors (graph.h:72-81)
      uint32_t* row = G.row(inserts[i + lo]);
      if (new_out[i].size() > G.R) { err = -3; continue; }
      row[0] = (uint32_t)new_out[i].size();
      for (size_t j = 0; j < new_out[i].size(); j++) row[1 + j] = new_out[i][j];
    }
    if (err) return err;
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> groups;
    group_by(pairs, groups);
    if (st) st->pairs += pairs.size();
    std::atomic<int> gerr(0);
    std::atomic<uint64_t> nprune(0), napp(0);
    parallel_for(0, groups.size(), nthreads, [&](uint64_t j) {
      uint32_t t = groups[j].first;
      std::vector<uint32_t>& S = groups[j].second;
      uint32_t* row = G.row(t);
      if (S.size() + row[0] <= R) {  // append_neighbors (graph.h:84-93)
        if (S.size() + row[0] > G.R) { gerr = -3; return; }
        for (size_t i = 0; i < S.size(); i++) row[1 + row[0] + i] = S[i];
        row[0] += (uint32_t)S.size();
        napp++;
      } else {
        std::vector<pid> cc;
        cc.reserve(S.size() + row[0]);
        for (uint32_t c : S) cc.push_back(pid(c, P.dist_id(c, t)));
        std::vector<uint32_t> o;
        robust_prune(P, G, t, cc, alpha, true, R, tie_mode, o);
        if (o.size() > G.R) { gerr = -3; return; }
        row[0] = (uint32_t)o.size();
        for (size_t i = 0; i < o.size(); i++) row[1 + i] = o[i];
        nprune++;
      }
    });
    if (st) { st->prunes_rev += nprune; st->appends += napp; }
    if (gerr) return gerr;
  }
  return 0;
}

// index.h:133-136 — sort every row by distance to its owner.
// parlay::random_permutation<int>(n) with the default generator (state 0),
// which batch_insert draws for random_order (index.h:237-241):
// hash64 (utilities.h:131-140); random::ith_rand / fork (random.h:33-45);
// seq_random_shuffle_ (random.h:52-60): Fisher-Yates from the back with
// ith_rand(i) % (i + 1); random_shuffle_ (random.h:62-98): below 8192
// elements one sequential shuffle, otherwise a stable bucket sort of the
// positions by ith_rand(i) & (2^bits - 1), bits = (log2_up(n) - 7) / 2 below
// 2^27 elements and log2_up(n) - 17 above, then bucket b shuffled with fork(b).
static inline uint64_t phash64(uint64_t u) {
  uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
  v ^= v >> 21;
  v ^= v << 37;
  v ^= v >> 4;
  v *= 4768777513237032717ull;
  v ^= v << 20;
  v ^= v >> 41;
  v ^= v << 5;
  return v;
}

static void shuffle_seq(int32_t* a, uint64_t n, uint64_t state) {
  for (uint64_t i = n; i-- > 1;) std::swap(a[i], a[phash64(i + state) % (i + 1)]);
}

static void random_permutation(uint64_t n, int32_t* out) {
  if (n < 8192) {
    for (uint64_t i = 0; i < n; i++) out[i] = (int32_t)i;
    shuffle_seq(out, n, 0);
    return;
  }
  uint64_t lg = 0;
  for (uint64_t b = n - 1; b > 0; b >>= 1) lg++;
  const uint64_t bits = n < (1ull << 27) ? (lg - 7) / 2 : lg - 17;
  const uint64_t nb = 1ull << bits, mask = nb - 1;
  std::vector<uint64_t> start(nb + 1, 0);
  for (uint64_t i = 0; i < n; i++) start[(phash64(i) & mask) + 1]++;
  for (uint64_t b = 0; b < nb; b++) start[b + 1] += start[b];
  std::vector<uint64_t> pos(start.begin(), start.end() - 1);
  for (uint64_t i = 0; i < n; i++) out[pos[phash64(i) & mask]++] = (int32_t)i;
  for (uint64_t b = 0; b < nb; b++) shuffle_seq(out + start[b], start[b + 1] - start[b], phash64(phash64(b)));
}

static void sort_rows(const Points& P, GraphView& G, int tie_mode, int nthreads) {
  parallel_for(0, G.n, nthreads, [&](uint64_t i) {
    uint32_t* row = G.row(i);
    if (tie_mode == TIE_REFERENCE) {
      std::sort(row + 1, row + 1 + row[0], [&](uint32_t j, uint32_t k) {
        return P.dist_id(i, j) < P.dist_id(i, k);
      });
    } else {
      std::vector<pid> v;
      for (uint32_t j = 0; j < row[0]; j++) v.push_back(pid(row[1 + j], P.dist_id(i, row[1 + j])));
      std::sort(v.begin(), v.end(), less_pid);
      for (uint32_t j = 0; j < row[0]; j++) row[1 + j] = v[j].first;
    }
  });
}

// Quantized_Mips_Point<T>::distance (mips_point.h:111-141, 160-162): code c
// decodes to (float)c * mult + min_coord with mult = (max - min) / maxval and
// maxval = (float)(T)(2^(8 sizeof T) - 1) (255 for uint8; the cast makes it
// -1 for int8); the products with the query (float, or decoded the same way
// when qq) are added in dimension order, and the sum is returned unnegated.
// out is nq x n.
template <class T>
static void qmips(uint32_t d, float mx, float mn, const void* queries, int qq, uint64_t nq, const T* codes,
                  uint64_t n, float* out) {
  const float maxval = static_cast<float>(static_cast<T>((((size_t)1) << (sizeof(T) * 8)) - 1));
  const float delta = mx - mn;
  const float mult = delta / maxval;
  for (uint64_t i = 0; i < nq; i++)
    for (uint64_t j = 0; j < n; j++) {
      const T* p = codes + j * d;
      float r = 0;
      for (uint32_t k = 0; k < d; k++) {
        const float dc = static_cast<float>(p[k]) * mult + mn;
        const float qk = qq ? static_cast<float>(((const T*)queries)[i * d + k]) * mult + mn
                            : ((const float*)queries)[i * d + k];
        r += dc * qk;
      }
      out[i * n + j] = r;
    }
}

}  // namespace orc

// ================================================================ C entry points
extern "C" {

typedef struct {
  int dtype, metric;
  uint64_t n;
  uint32_t d, ad;
  const void* rows;
} orc_points;

typedef struct {
  uint64_t n;
  uint32_t R;
  uint32_t* rows;
} orc_graph;

typedef struct {
  uint32_t R, L;
  double alpha;
  int two_pass;
  uint32_t start_point;
  double base, max_fraction;
  uint64_t max_build;
  int final_sort;
  int bsp;  // 1: BSP_index::insert (P=1) batch sizing over m; 0: knn_index over n
} orc_build_params;

static orc::Points mk(const orc_points* p) {
  return orc::Points{p->dtype, p->metric, p->n, p->d, p->ad, p->rows};
}
static orc::GraphView mkg(const orc_graph* g) { return orc::GraphView{g->n, g->R, g->rows}; }

uint64_t orc_hash64_2(uint64_t x) { return orc::hash64_2(x); }

void orc_distances(int dtype, int metric, const void* A, const void* B, uint64_t npairs, uint32_t d,
                   uint32_t ad, float* out) {
  size_t es = orc::elem_size(dtype);
  for (uint64_t i = 0; i < npairs; i++)
    out[i] = orc::distance(dtype, metric, (const char*)A + i * ad * es, (const char*)B + i * ad * es, d);
}

// One beam_search call; returns 0, or -1 if a caller buffer is too small.
int orc_beam_search(const orc_points* P, const orc_graph* G, const void* query, int64_t self_id,
                    const uint32_t* starts, uint32_t nstarts, int64_t k, int64_t beam, double cut,
                    int64_t limit, int64_t degree_limit, uint32_t* fr_ids, float* fr_dists,
                    uint32_t* fr_len, uint32_t fr_cap, uint32_t* vis_ids, float* vis_dists,
                    uint32_t* vis_len, uint32_t vis_cap, uint64_t* dist_cmps) {
  orc::Points p = mk(P);
  orc::GraphView g = mkg(G);
  orc::SearchResult sr;
  orc::beam_search(p, g, query, self_id, starts, nstarts, orc::QP{k, beam, cut, limit, degree_limit}, sr);
  *dist_cmps = sr.dist_cmps;
  *fr_len = (uint32_t)sr.frontier.size();
  *vis_len = (uint32_t)sr.visited.size();
  if (sr.frontier.size() > fr_cap || sr.visited.size() > vis_cap) return -1;
  for (size_t i = 0; i < sr.frontier.size(); i++) { fr_ids[i] = sr.frontier[i].first; fr_dists[i] = sr.frontier[i].second; }
  for (size_t i = 0; i < sr.visited.size(); i++) { vis_ids[i] = sr.visited[i].first; vis_dists[i] = sr.visited[i].second; }
  return 0;
}

// searchAll (beamSearch.h:322-346): query i uses p.id() = i.  Writes the
// first k frontier entries (fewer if the frontier is shorter: then the tail
// is id 0xFFFFFFFF / +inf), dist_cmps and |visited| per query.
int orc_search_all(const orc_points* Base, const orc_graph* G, const void* queries, uint64_t nq,
                   const uint32_t* starts, uint32_t nstarts, int64_t k, int64_t beam, double cut,
                   int64_t limit, int64_t degree_limit, const uint32_t* per_query_beam,
                   uint32_t* out_ids, float* out_dists, uint64_t* dist_cmps, uint32_t* visited,
                   int nthreads) {
  if (k > beam && !per_query_beam) return -1;  // beamSearch.h:327-331
  orc::Points p = mk(Base);
  orc::GraphView g = mkg(G);
  size_t qstride = (size_t)Base->ad * orc::elem_size(Base->dtype);
  std::atomic<int> err(0);
  orc::parallel_for(0, nq, nthreads, [&](uint64_t i) {
    int64_t b = per_query_beam ? (int64_t)per_query_beam[i] : beam;
    if (k > b) { err = -1; return; }
    orc::SearchResult sr;
    orc::beam_search(p, g, (const char*)queries + i * qstride, (int64_t)i, starts, nstarts,
                     orc::QP{k, b, cut, limit, degree_limit}, sr);
    for (int64_t j = 0; j < k; j++) {
      bool ok = j < (int64_t)sr.frontier.size();
      out_ids[i * k + j] = ok ? sr.frontier[j].first : 0xFFFFFFFFu;
      if (out_dists) out_dists[i * k + j] = ok ? sr.frontier[j].second : std::numeric_limits<float>::infinity();
    }
    if (dist_cmps) dist_cmps[i] = sr.dist_cmps;
    if (visited) visited[i] = (uint32_t)sr.visited.size();
  });
  return err.load();
}

int orc_robust_prune(const orc_points* P, const orc_graph* G, uint32_t p, const uint32_t* cand_ids,
                     const float* cand_dists, uint32_t ncand, double alpha, int add, uint32_t R,
                     int tie_mode, uint32_t* out, uint32_t* out_len) {
  orc::Points pp = mk(P);
  orc::GraphView g = mkg(G);
  std::vector<orc::pid> c;
  for (uint32_t i = 0; i < ncand; i++)
    c.push_back(orc::pid(cand_ids[i], cand_dists ? cand_dists[i] : pp.dist_id(cand_ids[i], p)));
  std::vector<uint32_t> o;
  orc::robust_prune(pp, g, p, c, alpha, add != 0, R, tie_mode, o);
  *out_len = (uint32_t)o.size();
  for (size_t i = 0; i < o.size(); i++) out[i] = o[i];
  return 0;
}

int orc_schedule(uint64_t m, double base, double max_fraction, uint64_t max_batch_base,
                 uint64_t* floors, uint64_t* ceilings, uint64_t cap, uint64_t* nbatches) {
  std::vector<std::pair<uint64_t, uint64_t>> b;
  orc::schedule(m, base, orc::max_batch_size(max_fraction, max_batch_base), b);
  *nbatches = b.size();
  if (b.size() > cap) return -1;
  for (size_t i = 0; i < b.size(); i++) { floors[i] = b[i].first; ceilings[i] = b[i].second; }
  return 0;
}

// Groups (key, value) pairs; out_keys[g], out_offsets[g..g+1] delimit
// out_vals.  Groups in increasing key order, values in input order.
int orc_group_by_key(const uint32_t* keys, const uint32_t* vals, uint64_t npairs, uint32_t* out_keys,
                     uint64_t* out_offsets, uint32_t* out_vals, uint64_t* ngroups) {
  std::vector<std::pair<uint32_t, uint32_t>> pairs(npairs);
  for (uint64_t i = 0; i < npairs; i++) pairs[i] = {keys[i], vals[i]};
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> groups;
  orc::group_by(pairs, groups);
  uint64_t off = 0;
  for (size_t g = 0; g < groups.size(); g++) {
    out_keys[g] = groups[g].first;
    out_offsets[g] = off;
    for (uint32_t v : groups[g].second) out_vals[off++] = v;
  }
  out_offsets[groups.size()] = off;
  *ngroups = groups.size();
  return 0;
}

// knn_index::batch_insert with explicit start point.
int orc_batch_insert(const orc_points* P, orc_graph* G, const uint32_t* inserts, uint64_t m, double alpha,
                     double base, double max_fraction, uint32_t start, uint32_t L, uint32_t R,
                     int tie_mode, int nthreads, uint32_t* visited_counts, uint64_t* stats6) {
  orc::Points pp = mk(P);
  orc::GraphView g = mkg(G);
  orc::BuildStats st;
  int r = orc::batch_insert(pp, g, inserts, m, alpha, base, max_fraction, g.n, start, L, R, tie_mode,
                            nthreads, visited_counts, &st);
  if (stats6) {
    stats6[0] = st.dist_cmps; stats6[1] = st.visited; stats6[2] = st.prunes_fwd;
    stats6[3] = st.prunes_rev; stats6[4] = st.appends; stats6[5] = st.pairs;
  }
  return r;
}

// knn_index::build_index (index.h:125-137) when bsp == 0, and
// BSP_index::insert at P=1 (bsp_index.hpp:141-434) when bsp == 1.
int orc_build(const orc_points* P, orc_graph* G, const orc_build_params* bp, int tie_mode, int nthreads,
              uint32_t* visited_counts, uint64_t* stats6) {
  orc::Points pp = mk(P);
  orc::GraphView g = mkg(G);
  uint64_t n = P->n;
  if (g.n != n || bp->R > g.R) return -1;
  if (bp->start_point >= n) return -2;
  uint64_t m = n;
  if (bp->bsp && bp->max_build) m = std::min<uint64_t>(bp->max_build, n);
  std::vector<uint32_t> inserts(n);
  for (uint64_t i = 0; i < n; i++) inserts[i] = (uint32_t)i;
  orc::BuildStats st;
  int r = 0;
  uint64_t mbb = bp->bsp ? m : n;
  if (bp->two_pass && !bp->bsp)
    r = orc::batch_insert(pp, g, inserts.data(), m, 1.0, bp->base, bp->max_fraction, mbb, bp->start_point,
                          bp->L, bp->R, tie_mode, nthreads, visited_counts, &st);
  if (!r)
    r = orc::batch_insert(pp, g, inserts.data(), m, bp->alpha, bp->base, bp->max_fraction, mbb,
                          bp->start_point, bp->L, bp->R, tie_mode, nthreads, visited_counts, &st);
  if (!r && bp->final_sort) orc::sort_rows(pp, g, tie_mode, nthreads);
  if (stats6) {
    stats6[0] = st.dist_cmps; stats6[1] = st.visited; stats6[2] = st.prunes_fwd;
    stats6[3] = st.prunes_rev; stats6[4] = st.appends; stats6[5] = st.pairs;
  }
  return r;
}

void orc_random_permutation(uint64_t n, int32_t* out) { orc::random_permutation(n, out); }

int orc_quantized_mips(int dtype, uint32_t d, float mx, float mn, const void* queries, int qq, uint64_t nq,
                       const void* codes, uint64_t n, float* out) {
  if (dtype == orc::U8) orc::qmips<uint8_t>(d, mx, mn, queries, qq, nq, (const uint8_t*)codes, n, out);
  else if (dtype == orc::I8) orc::qmips<int8_t>(d, mx, mn, queries, qq, nq, (const int8_t*)codes, n, out);
  else return -1;
  return 0;
}

int orc_sort_rows(const orc_points* P, orc_graph* G, int tie_mode, int nthreads) {
  orc::Points pp = mk(P);
  orc::GraphView g = mkg(G);
  orc::sort_rows(pp, g, tie_mode, nthreads);
  return 0;
}

// FP64 brute-force top-k (integer types: exact int64), ties by id.
int orc_ground_truth(const orc_points* Base, const void* queries, uint64_t nq, uint32_t k, uint32_t* ids,
                     double* dists, int nthreads) {
  orc::Points B = mk(Base);
  if (k > B.n) return -1;
  size_t es = orc::elem_size(B.dtype);
  size_t qstride = (size_t)B.ad * es;
  orc::parallel_for(0, nq, nthreads, [&](uint64_t qi) {
    const char* q = (const char*)queries + qi * qstride;
    std::vector<std::pair<double, uint32_t>> v(B.n);
    for (uint64_t i = 0; i < B.n; i++) {
      const char* p = (const char*)B.row(i);
      double s = 0;
      if (B.dtype == orc::F32) {
        const float* a = (const float*)q; const float* b = (const float*)p;
        for (uint32_t j = 0; j < B.d; j++) {
          if (B.metric == orc::L2) { double t = (double)a[j] - (double)b[j]; s += t * t; }
          else s -= (double)a[j] * (double)b[j];
        }
      } else {
        int64_t acc = 0;
        for (uint32_t j = 0; j < B.d; j++) {
          int64_t x = B.dtype == orc::U8 ? (int64_t)((const uint8_t*)q)[j] : (int64_t)((const int8_t*)q)[j];
          int64_t y = B.dtype == orc::U8 ? (int64_t)((const uint8_t*)p)[j] : (int64_t)((const int8_t*)p)[j];
          acc += B.metric == orc::L2 ? (x - y) * (x - y) : -(x * y);
        }
        s = (double)acc;
      }
      v[i] = {s, (uint32_t)i};
    }
    std::partial_sort(v.begin(), v.begin() + k, v.end());
    for (uint32_t j = 0; j < k; j++) { ids[qi * k + j] = v[j].second; dists[qi * k + j] = v[j].first; }
  });
  return 0;
}

}  // extern "C"
