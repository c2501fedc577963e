// Brute-force top-k ground truth over the index's vectors: FP64 accumulation
// for float32 rows, exact int64 for uint8/int8; order (distance, id).
// The reference has no brute-force code: its ground-truth files come from
// sklearn NearestNeighbors(brute) in its dataset tools.
//
// One wave per query: each lane scans a strided share of the base and keeps
// a sorted top-k list in LDS; the 64 lists are then merged by rank.
#include <algorithm>

#include "device.hpp"
#include "internal.hpp"

namespace ddg {

constexpr uint32_t kGtMaxK = 100;  // the reference's ground-truth files hold 100 neighbours

template <typename T, int METRIC>
__device__ __forceinline__ double exact_dist(const T* row, const T* q, uint32_t d) {
  if constexpr (sizeof(T) == 4) {
    double s = 0.0;
    for (uint32_t j = 0; j < d; j++) {
      const double a = (double)q[j], b = (double)row[j];
      if (METRIC == L2) { const double t = a - b; s += t * t; }
      else s -= a * b;
    }
    return s;
  } else {
    int64_t s = 0;
    for (uint32_t j = 0; j < d; j++) {
      const int64_t a = (int64_t)q[j], b = (int64_t)row[j];
      s += METRIC == L2 ? (a - b) * (a - b) : -(a * b);
    }
    return (double)s;
  }
}

__device__ __forceinline__ bool less_gt(double da, uint32_t ia, double db, uint32_t ib) {
  return da < db || (da == db && ia < ib);
}

template <typename T, int METRIC>
__global__ __launch_bounds__(256) void k_ground_truth(const T* base, uint64_t n, const T* queries, uint64_t nq,
                                                      uint32_t d, uint32_t ad, uint32_t k, uint32_t* ids,
                                                      double* dists) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t rb16 = (ad * (uint32_t)sizeof(T) + 15u) & ~15u;
  const uint32_t per_wave = rb16 + kWave * k * 12u;
  unsigned char* my = smem + w * per_wave;
  T* qv = (T*)my;
  double* ld = (double*)(my + rb16) + (uint64_t)lane * k;
  uint32_t* li = (uint32_t*)((double*)(my + rb16) + (uint64_t)kWave * k) + (uint64_t)lane * k;
  double* all_d = (double*)(my + rb16);
  uint32_t* all_i = (uint32_t*)(all_d + (uint64_t)kWave * k);

  for (uint64_t qi = (uint64_t)blockIdx.x * (blockDim.x >> 6) + w; qi < nq;
       qi += (uint64_t)gridDim.x * (blockDim.x >> 6)) {
    stage_row(qv, queries + qi * ad, ad * (uint32_t)sizeof(T));
    for (uint32_t j = 0; j < k; j++) { ld[j] = __builtin_inf(); li[j] = 0xFFFFFFFFu; }
    wave_sync();
    uint32_t cnt = 0;
    double worst = __builtin_inf();
    uint32_t worst_id = 0xFFFFFFFFu;
    for (uint64_t i = lane; i < n; i += kWave) {
      const double s = exact_dist<T, METRIC>(base + i * ad, qv, d);
      if (cnt < k || less_gt(s, (uint32_t)i, worst, worst_id)) {
        // insertion into the lane's sorted list
        uint32_t pos = cnt < k ? cnt : k - 1;
        while (pos > 0 && less_gt(s, (uint32_t)i, ld[pos - 1], li[pos - 1])) {
          ld[pos] = ld[pos - 1];
          li[pos] = li[pos - 1];
          pos--;
        }
        ld[pos] = s;
        li[pos] = (uint32_t)i;
        if (cnt < k) cnt++;
        worst = ld[k - 1];
        worst_id = li[k - 1];
      }
    }
    wave_sync();
    // merge: rank every (lane, slot) entry among all 64*k entries
    const uint32_t total = kWave * k;
    for (uint32_t e = lane; e < total; e += kWave) {
      const double de = all_d[e];
      const uint32_t ie = all_i[e];
      if (ie == 0xFFFFFFFFu) continue;
      uint32_t rank = 0;
      for (uint32_t t = 0; t < total && rank < k; t++) {
        if (less_gt(all_d[t], all_i[t], de, ie)) rank++;
      }
      if (rank < k) {
        ids[qi * k + rank] = ie;
        if (dists) dists[qi * k + rank] = de;
      }
    }
    wave_sync();
  }
}

// ------------------------------------------------------------ MFMA path
// The distance matrix of a 64-query block against 64-point tiles of one
// slice of the base, on matrix cores:
//   float32 rows: v_mfma_f64_16x16x4_f64 on FP64 copies (products of float32
//                 values are exact in FP64), distance = |q|^2 + |x|^2 - 2 q.x
//                 (L2) or -q.x (MIPS) in FP64 with FP64 norms;
//   uint8/int8:   v_mfma_i32_16x16x64_i8, exact; uint8 values u enter as
//                 u - 128 and q.x = S + 128 (sum q + sum x) - 16384 P over
//                 the P = 4Q dimensions the tiles cover (padding u = 0).
// Each wave owns 16 queries (the MFMA rows) and 4 blocks of 16 points.  Lane
// group g = lane / 16 supplies dimensions [g Q, g Q + Q) of every K step (the
// k index of a product is a relabelling of dimensions, shared by both
// operands).  Per query, a threshold (the current k-th (distance, id)) filters
// the tile into an LDS candidate list; a list nearing capacity is cut back to
// its k smallest by rank.  Slices write sorted partial top-k lists, which a
// second kernel merges.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kGtRows = 16;    // queries per wave
constexpr uint32_t kGtTile = 32;    // points per tile
constexpr uint32_t kGtWavesF = 4;   // waves per block, FP64 path
constexpr uint32_t kGtWavesI8 = 8;  // waves per block, int8 path: 128 queries share each tile (k <= 32)
constexpr uint32_t kGtWavesI8Big = 4;  // k > 32: longer candidate lists, 64 queries per tile
constexpr uint32_t kGtSlabs = 3;   // candidate-list entries per lane in a compaction (cap <= 192)
static_assert(kGtTile + (kGtMaxK + 15) / 16 * 16 <= kWave * kGtSlabs, "a candidate list must fit one compaction");
constexpr uint32_t kGtMaxQf = 64;   // FP64 path: dims per lane group (d <= 256)
constexpr uint32_t kGtMaxSteps = 8; // int8 path: K steps of 64 (d <= 512)

struct GtArgs {
  const void* base;
  const void* queries;
  uint64_t n, nq, per_slice;
  uint32_t d, ad, k, cap, Q, bstride, nslices;
  const double* pn;    // per point: |x|^2 (FP64 / exact), then per point sum of values (uint8)
  const double* psum;
  const double* qn;
  const double* qsum;
  double* part_d;      // nq x nslices x k
  uint32_t* part_i;
};

// per-wave candidate lists: cd[16][cap] f64 | thr_d[16] f64 | ci[16][cap] u32 | thr_i[16] u32 | cnt[16] u32
__host__ __device__ inline uint32_t gt_wave_bytes(uint32_t cap) { return 192u * cap + 256u; }
__host__ __device__ inline uint32_t gt_tile_off(uint32_t cap, uint32_t waves) { return 1024u + waves * gt_wave_bytes(cap); }

struct GtTopk {
  double* cd;
  double* thr_d;
  uint32_t* ci;
  uint32_t* thr_i;
  uint32_t* cnt;
  uint32_t cap, k, tile;  // tile: points offered between two maintain() calls
  __device__ GtTopk(unsigned char* base, uint32_t cap_, uint32_t k_, uint32_t tile_)
      : cap(cap_), k(k_), tile(tile_) {
    cd = (double*)base;
    thr_d = cd + kGtRows * cap;
    ci = (uint32_t*)(thr_d + kGtRows);
    thr_i = ci + kGtRows * cap;
    cnt = thr_i + kGtRows;
  }
  __device__ void init() {
    const uint32_t l = lane_id();
    if (l < kGtRows) {
      thr_d[l] = __builtin_inf();
      thr_i[l] = 0xFFFFFFFFu;
      cnt[l] = 0;
    }
    wave_sync();
  }
  __device__ void offer(uint32_t row, double dist, uint32_t id) {
    if (less_gt(dist, id, thr_d[row], thr_i[row])) {
      const uint32_t pos = atomicAdd(&cnt[row], 1u);
      cd[row * cap + pos] = dist;
      ci[row * cap + pos] = id;
    }
  }
  // keeps the k smallest of row's list, sorted; the k-th becomes the threshold.
  // A list holds at most cap <= kWave * kGtSlabs entries, kGtSlabs per lane;
  // its ids are distinct (a tile offers each point once per query), so the
  // ranks are a permutation.
  __device__ void compact(uint32_t row) {
    const uint32_t l = lane_id();
    const uint32_t c = cnt[row];
    double* D = cd + row * cap;
    uint32_t* I = ci + row * cap;
    double dv[kGtSlabs];
    uint32_t iv[kGtSlabs], rv[kGtSlabs];
#pragma unroll
    for (uint32_t s = 0; s < kGtSlabs; s++) {
      const uint32_t e = l + s * kWave;
      dv[s] = e < c ? D[e] : 0.0;
      iv[s] = e < c ? I[e] : 0u;
      rv[s] = 0;
    }
    for (uint32_t t = 0; t < c; t++) {
      const double dt = D[t];
      const uint32_t it = I[t];
#pragma unroll
      for (uint32_t s = 0; s < kGtSlabs; s++) rv[s] += less_gt(dt, it, dv[s], iv[s]) ? 1u : 0u;
    }
    wave_sync();
#pragma unroll
    for (uint32_t s = 0; s < kGtSlabs; s++)
      if (l + s * kWave < c && rv[s] < k) { D[rv[s]] = dv[s]; I[rv[s]] = iv[s]; }
    wave_sync();
    if (l == 0) {
      const uint32_t nc = c < k ? c : k;
      cnt[row] = nc;
      if (nc == k) { thr_d[row] = D[k - 1]; thr_i[row] = I[k - 1]; }
    }
    wave_sync();
  }
  // after a tile: cut back every list that could overflow on the next one
  __device__ void maintain() {
    wave_sync();
    const uint32_t l = lane_id();
    uint64_t need = ballot(l < kGtRows && cnt[l] + tile > cap);
    while (need) {
      const int r = first_lane(need);
      need &= need - 1;
      compact((uint32_t)r);
    }
  }
  __device__ void finish(const GtArgs& g, uint64_t q0, uint32_t slice) {
    wave_sync();
    for (uint32_t r = 0; r < kGtRows; r++)
      if (cnt[r] > 0) compact(r);
    for (uint32_t e = lane_id(); e < kGtRows * k; e += kWave) {
      const uint32_t row = e / k, j = e - row * k;
      const uint64_t q = q0 + row;
      if (q >= g.nq) continue;
      const uint64_t o = (q * g.nslices + slice) * k + j;
      const bool ok = j < cnt[row];
      g.part_d[o] = ok ? cd[row * cap + j] : __builtin_inf();
      g.part_i[o] = ok ? ci[row * cap + j] : 0xFFFFFFFFu;
    }
  }
};

template <int METRIC>
__global__ __launch_bounds__(256, 2) void k_gt_f64(GtArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tile_pn = (double*)smem;
  float* Bs = (float*)(smem + gt_tile_off(g.cap, kGtWavesF));
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t g4 = lane >> 4, c16 = lane & 15u;
  GtTopk tk(smem + 1024u + wave * gt_wave_bytes(g.cap), g.cap, g.k, kGtTile);
  tk.init();
  const uint64_t q0 = (uint64_t)blockIdx.x * (kGtWavesF * kGtRows) + wave * kGtRows;
  const uint32_t Q = g.Q;

  // A operand: query q0 + c16, dimensions g4 Q + s, held in registers as
  // float (widened to double exactly at each use: half the registers, so two
  // blocks fit a CU)
  float af[kGtMaxQf];
  {
    const uint64_t qa = q0 + c16 < g.nq ? q0 + c16 : g.nq - 1;
    const float* qr = (const float*)g.queries + qa * g.ad;
#pragma unroll
    for (uint32_t s = 0; s < kGtMaxQf; s++) {
      const uint32_t j = g4 * Q + s;
      af[s] = (s < Q && j < g.d) ? qr[j] : 0.f;
    }
  }
  // v_mfma_f64_16x16x4_f64 leaves row 4 r + lane / 16, column lane % 16 in
  // accumulator register r of each lane
  double qn_r[4];
#pragma unroll
  for (uint32_t r = 0; r < 4; r++) {
    const uint64_t q = q0 + 4u * r + g4;
    qn_r[r] = (METRIC == L2 && q < g.nq) ? g.qn[q] : 0.0;
  }

  const uint64_t p_begin = (uint64_t)blockIdx.y * g.per_slice;
  const uint64_t p_end = p_begin + g.per_slice < g.n ? p_begin + g.per_slice : g.n;
  const float* X = (const float*)g.base;
  for (uint64_t t0 = p_begin; t0 < p_end; t0 += kGtTile) {
    __syncthreads();
    // 16-byte chunks of the tile rows (the device rows are zero beyond d),
    // eight loads in flight per thread before the LDS stores
    const uint32_t total = kGtTile * Q;
    for (uint32_t e0 = threadIdx.x; e0 < total; e0 += 8u * blockDim.x) {
      float4 v[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        const uint32_t e = e0 + u * blockDim.x;
        v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < total) {
          const uint32_t row = e / Q, c = e - row * Q;
          const uint64_t p = t0 + row;
          if (p < p_end) v[u] = *(const float4*)(X + p * g.ad + 4u * c);
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        const uint32_t e = e0 + u * blockDim.x;
        if (e < total) {
          const uint32_t row = e / Q, c = e - row * Q;
          float* o = Bs + row * g.bstride + 4u * c;
          o[0] = v[u].x;
          o[1] = v[u].y;
          o[2] = v[u].z;
          o[3] = v[u].w;
        }
      }
    }
    if (threadIdx.x < kGtTile) {
      const uint64_t p = t0 + threadIdx.x;
      tile_pn[threadIdx.x] = (METRIC == L2 && p < p_end) ? g.pn[p] : 0.0;
    }
    __syncthreads();
    v4d acc[kGtTile / 16];
#pragma unroll
    for (uint32_t b = 0; b < kGtTile / 16; b++) acc[b] = v4d{0.0, 0.0, 0.0, 0.0};
    const float* bp = Bs + c16 * g.bstride + g4 * Q;
#pragma unroll
    for (uint32_t s = 0; s < kGtMaxQf; s++) {
      if (s < Q) {
#pragma unroll
        for (uint32_t b = 0; b < kGtTile / 16; b++)
          acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)af[s], (double)bp[16u * b * g.bstride + s], acc[b], 0, 0, 0);
      }
    }
#pragma unroll
    for (uint32_t b = 0; b < kGtTile / 16; b++) {
      const uint32_t pl = 16u * b + c16;
      const uint64_t p = t0 + pl;
#pragma unroll
      for (uint32_t r = 0; r < 4; r++) {
        const uint32_t row = 4u * r + g4;
        const double dot = acc[b][r];
        const double dist = METRIC == L2 ? qn_r[r] + tile_pn[pl] - 2.0 * dot : -dot;
        if (q0 + row < g.nq && p < p_end) tk.offer(row, dist, (uint32_t)p);
      }
    }
    tk.maintain();
  }
  tk.finish(g, q0, blockIdx.y);
}

template <typename T, int METRIC, uint32_t WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_gt_i8(GtArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool U8 = sizeof(T) == 1 && T(-1) > T(0);
  constexpr uint32_t kFlip = U8 ? 0x80808080u : 0u;
  double* tile_pn = (double*)smem;
  double* tile_ps = tile_pn + kGtTile;
  unsigned char* Bs = smem + gt_tile_off(g.cap, WAVES);
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t g4 = lane >> 4, c16 = lane & 15u;
  GtTopk tk(smem + 1024u + wave * gt_wave_bytes(g.cap), g.cap, g.k, kGtTile);
  tk.init();
  const uint64_t q0 = (uint64_t)blockIdx.x * (WAVES * kGtRows) + wave * kGtRows;
  const uint32_t Q = g.Q, Q4 = 4u * Q, nsteps = Q / 16u, cpr = Q4 / 16u;
  const double P = (double)Q4;

  v4i af[kGtMaxSteps];
  {
    const uint64_t qa = q0 + c16 < g.nq ? q0 + c16 : g.nq - 1;
    const unsigned char* qr = (const unsigned char*)g.queries + qa * g.ad;
#pragma unroll
    for (uint32_t s = 0; s < kGtMaxSteps; s++) {
      const uint32_t o = g4 * Q + 16u * s;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (s < nsteps && o + 16u <= g.ad) v = *(const uint4*)(qr + o);
      af[s] = v4i{(int)(v.x ^ kFlip), (int)(v.y ^ kFlip), (int)(v.z ^ kFlip), (int)(v.w ^ kFlip)};
    }
  }
  // v_mfma_i32_16x16x64_i8 leaves row 4 (lane / 16) + r, column lane % 16 in
  // accumulator register r of each lane
  double qn_r[4], qs_r[4];
#pragma unroll
  for (uint32_t r = 0; r < 4; r++) {
    const uint64_t q = q0 + 4u * g4 + r;
    qn_r[r] = q < g.nq ? g.qn[q] : 0.0;
    qs_r[r] = q < g.nq ? g.qsum[q] : 0.0;
  }

  const uint64_t p_begin = (uint64_t)blockIdx.y * g.per_slice;
  const uint64_t p_end = p_begin + g.per_slice < g.n ? p_begin + g.per_slice : g.n;
  const unsigned char* X = (const unsigned char*)g.base;
  for (uint64_t t0 = p_begin; t0 < p_end; t0 += kGtTile) {
    __syncthreads();
    const uint32_t total = kGtTile * cpr;
    for (uint32_t e0 = threadIdx.x; e0 < total; e0 += 4u * blockDim.x) {
      uint4 v[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t e = e0 + u * blockDim.x;
        v[u] = make_uint4(0u, 0u, 0u, 0u);
        if (e < total) {
          const uint32_t row = e / cpr, c = e - row * cpr;
          const uint64_t p = t0 + row;
          if (p < p_end && 16u * c + 16u <= g.ad) v[u] = *(const uint4*)(X + p * g.ad + 16u * c);
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; u++) {
        const uint32_t e = e0 + u * blockDim.x;
        if (e < total) {
          const uint32_t row = e / cpr, c = e - row * cpr;
          *(uint4*)(Bs + row * g.bstride + 16u * c) =
              make_uint4(v[u].x ^ kFlip, v[u].y ^ kFlip, v[u].z ^ kFlip, v[u].w ^ kFlip);
        }
      }
    }
    if (threadIdx.x < kGtTile) {
      const uint64_t p = t0 + threadIdx.x;
      tile_pn[threadIdx.x] = p < p_end ? g.pn[p] : 0.0;
      tile_ps[threadIdx.x] = p < p_end ? g.psum[p] : 0.0;
    }
    __syncthreads();
    v4i acc[kGtTile / 16];
#pragma unroll
    for (uint32_t b = 0; b < kGtTile / 16; b++) acc[b] = v4i{0, 0, 0, 0};
    const unsigned char* bp = Bs + c16 * g.bstride + g4 * Q;
#pragma unroll
    for (uint32_t s = 0; s < kGtMaxSteps; s++) {
      if (s < nsteps) {
#pragma unroll
        for (uint32_t b = 0; b < kGtTile / 16; b++) {
          const v4i bv = *(const v4i*)(bp + 16u * b * g.bstride + 16u * s);
          acc[b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[s], bv, acc[b], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (uint32_t b = 0; b < kGtTile / 16; b++) {
      const uint32_t pl = 16u * b + c16;
      const uint64_t p = t0 + pl;
#pragma unroll
      for (uint32_t r = 0; r < 4; r++) {
        const uint32_t row = 4u * g4 + r;
        double dot = (double)acc[b][r];
        if (U8) dot += 128.0 * (qs_r[r] + tile_ps[pl]) - 16384.0 * P;
        const double dist = METRIC == L2 ? qn_r[r] + tile_pn[pl] - 2.0 * dot : -dot;
        if (q0 + row < g.nq && p < p_end) tk.offer(row, dist, (uint32_t)p);
      }
    }
    tk.maintain();
  }
  tk.finish(g, q0, blockIdx.y);
}

// |x|^2 (FP64 for float32, exact for integers) and the sum of the values
template <typename T>
__global__ __launch_bounds__(256) void k_gt_norms(const T* rows, uint64_t m, uint32_t d, uint32_t ad, double* nrm,
                                                  double* sum) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T* r = rows + i * ad;
  if constexpr (sizeof(T) == 4) {
    double s = 0.0;
    for (uint32_t j = 0; j < d; j++) { const double v = (double)r[j]; s += v * v; }
    nrm[i] = s;
    sum[i] = 0.0;
  } else {
    int64_t s = 0, t = 0;
    for (uint32_t j = 0; j < d; j++) { const int64_t v = (int64_t)r[j]; s += v * v; t += v; }
    nrm[i] = (double)s;
    sum[i] = (double)t;
  }
}

// final top-k of each query over its nslices sorted partial lists (one wave per query)
__global__ __launch_bounds__(256) void k_gt_merge(const double* pd, const uint32_t* pi, uint64_t nq, uint32_t E,
                                                  uint32_t k, uint32_t* ids, double* dists) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  double* D = (double*)(smem + (size_t)w * ((E * 12u + 15u) & ~15u));
  uint32_t* I = (uint32_t*)(D + E);
  const uint64_t q = (uint64_t)blockIdx.x * (blockDim.x >> 6) + w;
  if (q >= nq) return;  // whole waves leave together
  for (uint32_t e = lane; e < E; e += kWave) {
    D[e] = pd[q * E + e];
    I[e] = pi[q * E + e];
  }
  wave_sync();
  for (uint32_t e = lane; e < E; e += kWave) {
    const double de = D[e];
    const uint32_t ie = I[e];
    uint32_t rank = 0;
    for (uint32_t t = 0; t < E; t++) {
      const double dt = D[t];
      const uint32_t it = I[t];
      rank += (less_gt(dt, it, de, ie) || (dt == de && it == ie && t < e)) ? 1u : 0u;
    }
    if (rank < k) {
      ids[q * k + rank] = ie;
      if (dists) dists[q * k + rank] = de;
    }
  }
}

struct GtPlan {
  bool mfma;
  uint32_t Q, bstride, cap, S, lds, waves;
  uint64_t per_slice;
};

static GtPlan gt_plan(int dtype, uint64_t n, uint64_t nq, uint32_t d, uint32_t k) {
  GtPlan p{};
  p.cap = kGtTile + (k + 15u) / 16u * 16u;  // the k kept entries plus one tile's offers
  const uint32_t qd = (d + 3) / 4;
  if (dtype == F32) {
    p.mfma = qd <= kGtMaxQf;
    p.Q = qd;
    p.bstride = 4 * qd + 1;
    p.waves = kGtWavesF;
    p.lds = gt_tile_off(p.cap, kGtWavesF) + kGtTile * p.bstride * 4;
  } else {
    p.Q = (qd + 15) / 16 * 16;
    p.mfma = p.Q / 16 <= kGtMaxSteps;
    p.bstride = 4 * p.Q + 16;
    p.waves = k > 32 ? kGtWavesI8Big : kGtWavesI8;
    p.lds = gt_tile_off(p.cap, p.waves) + kGtTile * p.bstride;
  }
  if (p.lds > 160u * 1024u) p.mfma = false;  // beyond a CU's LDS: the per-query kernel
  const uint32_t qb = p.waves * kGtRows;  // queries per block
  const uint64_t qtiles = (nq + qb - 1) / qb;
  const uint64_t tiles = (n + kGtTile - 1) / kGtTile;
  uint64_t S = (2048 + qtiles - 1) / qtiles;
  S = std::min<uint64_t>(S, std::max<uint64_t>(1, 1024 / k));
  S = std::max<uint64_t>(1, std::min<uint64_t>(S, tiles));
  const uint64_t per_tiles = (tiles + S - 1) / S;
  p.S = (uint32_t)((tiles + per_tiles - 1) / per_tiles);
  p.per_slice = per_tiles * kGtTile;
  return p;
}

size_t ground_truth_tmp_bytes(int dtype, uint64_t n, uint64_t nq, uint32_t d, uint32_t k) {
  const GtPlan p = gt_plan(dtype, n, nq, d, k);
  if (!p.mfma) return 256;
  return (2 * n + 2 * nq) * 8 + nq * (uint64_t)p.S * k * 12 + 1024;
}

template <typename T, int M>
static hipError_t gt_t(const void* base, uint64_t n, const void* queries, uint64_t nq, uint32_t d, uint32_t ad,
                       uint32_t k, uint32_t* ids, double* dists, void* tmp, size_t tmp_bytes, hipStream_t s) {
  const GtPlan p = gt_plan(sizeof(T) == 4 ? F32 : (T(-1) > T(0) ? U8 : I8), n, nq, d, k);
  if (!p.mfma) {  // rows beyond the MFMA register budget: one wave per query
    const uint32_t rb16 = (ad * (uint32_t)sizeof(T) + 15u) & ~15u;
    const uint32_t per_wave = rb16 + kWave * k * 12u;
    uint32_t wpb = std::max<uint32_t>(1u, std::min<uint32_t>(4u, (160u * 1024u) / per_wave));
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((nq + wpb - 1) / wpb, 65535);
    (void)hipFuncSetAttribute((const void*)&k_ground_truth<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(wpb * per_wave));
    hipLaunchKernelGGL((k_ground_truth<T, M>), dim3(blocks), dim3(wpb * kWave), wpb * per_wave, s, (const T*)base,
                       n, (const T*)queries, nq, d, ad, k, ids, dists);
    return hipGetLastError();
  }
  if (tmp_bytes < ground_truth_tmp_bytes(sizeof(T) == 4 ? F32 : (T(-1) > T(0) ? U8 : I8), n, nq, d, k))
    return hipErrorInvalidValue;
  GtArgs g{};
  double* f = (double*)tmp;
  g.base = base;
  g.queries = queries;
  g.n = n;
  g.nq = nq;
  g.per_slice = p.per_slice;
  g.d = d;
  g.ad = ad;
  g.k = k;
  g.cap = p.cap;
  g.Q = p.Q;
  g.bstride = p.bstride;
  g.nslices = p.S;
  g.pn = f;
  g.psum = f + n;
  g.qn = f + 2 * n;
  g.qsum = f + 2 * n + nq;
  g.part_d = f + 2 * n + 2 * nq;
  g.part_i = (uint32_t*)(g.part_d + nq * (uint64_t)p.S * k);
  hipLaunchKernelGGL((k_gt_norms<T>), dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, (const T*)base, n, d, ad,
                     (double*)g.pn, (double*)g.psum);
  hipLaunchKernelGGL((k_gt_norms<T>), dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, s, (const T*)queries, nq, d,
                     ad, (double*)g.qn, (double*)g.qsum);
  const uint32_t waves = p.waves;
  const dim3 grid((uint32_t)((nq + waves * kGtRows - 1) / (waves * kGtRows)), p.S);
  if constexpr (sizeof(T) == 4) {
    (void)hipFuncSetAttribute((const void*)&k_gt_f64<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
    hipLaunchKernelGGL((k_gt_f64<M>), grid, dim3(64 * kGtWavesF), p.lds, s, g);
  } else {
    if (waves == kGtWavesI8) {
      (void)hipFuncSetAttribute((const void*)&k_gt_i8<T, M, kGtWavesI8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)p.lds);
      hipLaunchKernelGGL((k_gt_i8<T, M, kGtWavesI8>), grid, dim3(64 * kGtWavesI8), p.lds, s, g);
    } else {
      (void)hipFuncSetAttribute((const void*)&k_gt_i8<T, M, kGtWavesI8Big>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
      hipLaunchKernelGGL((k_gt_i8<T, M, kGtWavesI8Big>), grid, dim3(64 * kGtWavesI8Big), p.lds, s, g);
    }
  }
  const uint32_t E = p.S * k;
  const uint32_t wpb = 4, mlds = wpb * ((E * 12u + 15u) & ~15u);
  (void)hipFuncSetAttribute((const void*)&k_gt_merge, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlds);
  hipLaunchKernelGGL(k_gt_merge, dim3((uint32_t)((nq + wpb - 1) / wpb)), dim3(wpb * kWave), mlds, s,
                     (const double*)g.part_d, (const uint32_t*)g.part_i, nq, E, k, ids, dists);
  return hipGetLastError();
}

hipError_t launch_ground_truth(int dtype, int metric, const void* base, uint64_t n, const void* queries,
                               uint64_t nq, uint32_t d, uint32_t ad, uint32_t k, uint32_t* ids, double* dists,
                               void* tmp, size_t tmp_bytes, hipStream_t s) {
  if (k == 0 || k > kGtMaxK || k > n) return hipErrorInvalidValue;
  if (dtype == U8)
    return metric == L2 ? gt_t<uint8_t, L2>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s)
                        : gt_t<uint8_t, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s);
  if (dtype == I8)
    return metric == L2 ? gt_t<int8_t, L2>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s)
                        : gt_t<int8_t, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s);
  return metric == L2 ? gt_t<float, L2>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s)
                      : gt_t<float, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, tmp, tmp_bytes, s);
}

}  // namespace ddg
