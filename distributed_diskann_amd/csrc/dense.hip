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

constexpr uint32_t kGtMaxK = 32;

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

size_t ground_truth_tmp_bytes(uint64_t, uint64_t, uint32_t) { return 0; }

template <typename T, int M>
static hipError_t gt_t(const void* base, uint64_t n, const void* queries, uint64_t nq, uint32_t d, uint32_t ad,
                       uint32_t k, uint32_t* ids, double* dists, hipStream_t s) {
  const uint32_t rb16 = (ad * (uint32_t)sizeof(T) + 15u) & ~15u;
  const uint32_t per_wave = rb16 + kWave * k * 12u;
  uint32_t wpb = std::max<uint32_t>(1u, std::min<uint32_t>(4u, (160u * 1024u) / per_wave));
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((nq + wpb - 1) / wpb, 65535);
  (void)hipFuncSetAttribute((const void*)&k_ground_truth<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(wpb * per_wave));
  hipLaunchKernelGGL((k_ground_truth<T, M>), dim3(blocks), dim3(wpb * kWave), wpb * per_wave, s, (const T*)base, n,
                     (const T*)queries, nq, d, ad, k, ids, dists);
  return hipGetLastError();
}

hipError_t launch_ground_truth(int dtype, int metric, const void* base, uint64_t n, const void* queries,
                               uint64_t nq, uint32_t d, uint32_t ad, uint32_t k, uint32_t* ids, double* dists,
                               void*, size_t, hipStream_t s) {
  if (k == 0 || k > kGtMaxK || k > n) return hipErrorInvalidValue;
  if (dtype == U8)
    return metric == L2 ? gt_t<uint8_t, L2>(base, n, queries, nq, d, ad, k, ids, dists, s)
                        : gt_t<uint8_t, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, s);
  if (dtype == I8)
    return metric == L2 ? gt_t<int8_t, L2>(base, n, queries, nq, d, ad, k, ids, dists, s)
                        : gt_t<int8_t, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, s);
  return metric == L2 ? gt_t<float, L2>(base, n, queries, nq, d, ad, k, ids, dists, s)
                      : gt_t<float, MIPS>(base, n, queries, nq, d, ad, k, ids, dists, s);
}

}  // namespace ddg
