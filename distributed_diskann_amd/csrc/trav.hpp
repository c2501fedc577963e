// Traversal distance policies of the search kernels.
//
// TravDist<T, METRIC, TRAV> computes the distance between the query staged in
// LDS and row `id` of the vector copy the search traverses:
//   TRAV_NATIVE  the index's own rows with the exact reference arithmetic
//                (Dist<T, METRIC>: int32-exact for u8/i8, the reference's
//                summation order for f32);
//   TRAV_BF16    a BF16 copy of float32 rows (round to nearest even), the
//                query kept in FP32, FP32 fused multiply-add accumulation;
//   TRAV_FP8     an FP8-e4m3 copy of float32 rows with one scale per
//                dimension, x_ij ~ s_j * e_ij with s_j = max_i |x_ij| / 240
//                (every code lies in [-240, 240], so the copy decodes the
//                same way in the OCP and the FNUZ variants of e4m3), FP32
//                accumulation.  MIPS stages q_j * s_j once per query; L2
//                accumulates (q_j - s_j * e_ij)^2.
// Reduced copies store rows of `trav_stride` bytes (`trav_chunks` 16-byte
// chunks), zero beyond d; the staged query is zero beyond d too, so a
// distance walks whole rows and the padding adds nothing.
#pragma once

#include "device.hpp"
#include "internal.hpp"

namespace ddg {

enum TravMode : int { TRAV_NATIVE = 0, TRAV_BF16 = 1, TRAV_FP8 = 2 };

template <typename T, int METRIC, int TRAV>
struct TravDist;

// Every policy offers: the row of point `id` as 16-byte chunks (row_ptr, for
// a cooperative gather into LDS), and the distance to a row given by pointer
// (in HBM or in LDS) or by id.
template <typename T, int METRIC>
struct TravDist<T, METRIC, TRAV_NATIVE> {
  using D = Dist<T, METRIC>;
  using Q = typename D::Q;
  __device__ static Q prep(const SearchArgs& a, const unsigned char* ql) {
    return D::prep((const T*)ql, (a.ad * (uint32_t)sizeof(T)) >> 4);
  }
  __device__ static const uint4* row_ptr(const SearchArgs& a, uint32_t id) {
    return (const uint4*)((const T*)a.base + (uint64_t)id * a.ad);
  }
  __device__ static float dist_row(const SearchArgs& a, const uint4* row, const unsigned char* ql, const Q& qs) {
    return D::dist((const T*)row, (const T*)ql, qs, a.d, (a.ad * (uint32_t)sizeof(T)) >> 4);
  }
  __device__ static float dist(const SearchArgs& a, uint32_t id, const unsigned char* ql, const Q& qs) {
    return dist_row(a, row_ptr(a, id), ql, qs);
  }
};

// The sum of a reduced row as split_dist forms it: the partial sums of its
// a.trav_split blocks of a.trav_cps chunks, added pairwise (block 0 + 1,
// 2 + 3, then those pairs, ...), so one lane computing a whole row returns
// the same bits as a group of lanes splitting it.
template <typename P>
__device__ __forceinline__ float tree_sum(const SearchArgs& a, const uint4* r, const unsigned char* ql) {
  float p[8];
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t cb = k * a.trav_cps;
    const uint32_t ce = cb + a.trav_cps < a.trav_chunks ? cb + a.trav_cps : a.trav_chunks;
    p[k] = (k < a.trav_split && cb < ce) ? P::part(a, r, ql, cb, ce) : 0.f;
  }
#pragma unroll
  for (uint32_t w = 1; w < 8; w <<= 1)
#pragma unroll
    for (uint32_t k = 0; k + w < 8; k += 2 * w)
      if (w < a.trav_split) p[k] = p[k] + p[k + w];
  return p[0];
}

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

template <int METRIC>
__device__ __forceinline__ void acc_term(float& s, float q, float e) {
  if (METRIC == L2) {
    const float t = q - e;
    s = __builtin_fmaf(t, t, s);
  } else {
    s = __builtin_fmaf(q, e, s);
  }
}

template <int METRIC>
struct TravDist<float, METRIC, TRAV_BF16> {
  struct Q { int unused; };
  __device__ static Q prep(const SearchArgs&, const unsigned char*) { return Q{0}; }
  __device__ static const uint4* row_ptr(const SearchArgs& a, uint32_t id) {
    return (const uint4*)((const unsigned char*)a.trav_rows + (uint64_t)id * a.trav_stride);
  }
  __device__ static float dist(const SearchArgs& a, uint32_t id, const unsigned char* ql, const Q& qs) {
    return dist_row(a, row_ptr(a, id), ql, qs);
  }
  __device__ static float dist_row(const SearchArgs& a, const uint4* r, const unsigned char* ql, const Q&) {
    return finish(tree_sum<TravDist>(a, r, ql));
  }
  __device__ static float finish(float s) { return METRIC == L2 ? s : -s; }
  // sum of the terms of chunks [cb, ce) of row r (cb < ce)
  __device__ static float part(const SearchArgs& a, const uint4* r, const unsigned char* ql, uint32_t cb,
                               uint32_t ce) {
    const float4* q = (const float4*)ql;
    float s0 = 0.f, s1 = 0.f;
    for_chunks(r, ce, [&](uint4 x, uint32_t c, bool ok) {
      const float4 qa = q[2 * c];
      const float4 qb = q[2 * c + 1];
      float t0 = s0, t1 = s1;
      acc_term<METRIC>(t0, qa.x, bf16_lo(x.x));
      acc_term<METRIC>(t1, qa.y, bf16_hi(x.x));
      acc_term<METRIC>(t0, qa.z, bf16_lo(x.y));
      acc_term<METRIC>(t1, qa.w, bf16_hi(x.y));
      acc_term<METRIC>(t0, qb.x, bf16_lo(x.z));
      acc_term<METRIC>(t1, qb.y, bf16_hi(x.z));
      acc_term<METRIC>(t0, qb.z, bf16_lo(x.w));
      acc_term<METRIC>(t1, qb.w, bf16_hi(x.w));
      s0 = ok ? t0 : s0;
      s1 = ok ? t1 : s1;
      // one chunk's staged query words at a time: hoisting them all costs
      // more occupancy than their LDS latency
      __builtin_amdgcn_sched_barrier(0);
    }, cb);
    return s0 + s1;
  }
};

// four codes of one 32-bit word, element order = byte order
template <int METRIC>
__device__ __forceinline__ void fp8_word(float& s0, float& s1, uint32_t w, const float4& q, const float4& sc) {
  const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, false);
  const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, true);
  if (METRIC == L2) {
    const float t0 = __builtin_fmaf(-sc.x, lo[0], q.x);
    const float t1 = __builtin_fmaf(-sc.y, lo[1], q.y);
    const float t2 = __builtin_fmaf(-sc.z, hi[0], q.z);
    const float t3 = __builtin_fmaf(-sc.w, hi[1], q.w);
    s0 = __builtin_fmaf(t0, t0, s0);
    s1 = __builtin_fmaf(t1, t1, s1);
    s0 = __builtin_fmaf(t2, t2, s0);
    s1 = __builtin_fmaf(t3, t3, s1);
  } else {  // sc holds q_j * s_j
    s0 = __builtin_fmaf(sc.x, lo[0], s0);
    s1 = __builtin_fmaf(sc.y, lo[1], s1);
    s0 = __builtin_fmaf(sc.z, hi[0], s0);
    s1 = __builtin_fmaf(sc.w, hi[1], s1);
  }
}

template <int METRIC>
struct TravDist<float, METRIC, TRAV_FP8> {
  struct Q { int unused; };
  __device__ static Q prep(const SearchArgs&, const unsigned char*) { return Q{0}; }
  __device__ static const uint4* row_ptr(const SearchArgs& a, uint32_t id) {
    return (const uint4*)((const unsigned char*)a.trav_rows + (uint64_t)id * a.trav_stride);
  }
  __device__ static float dist(const SearchArgs& a, uint32_t id, const unsigned char* ql, const Q& qs) {
    return dist_row(a, row_ptr(a, id), ql, qs);
  }
  __device__ static float dist_row(const SearchArgs& a, const uint4* r, const unsigned char* ql, const Q&) {
    return finish(tree_sum<TravDist>(a, r, ql));
  }
  __device__ static float finish(float s) { return METRIC == L2 ? s : -s; }
  // sum of the terms of chunks [cb, ce) of row r (cb < ce)
  __device__ static float part(const SearchArgs& a, const uint4* r, const unsigned char* ql, uint32_t cb,
                               uint32_t ce) {
    const float4* q = (const float4*)ql;
    const float4* sc = (const float4*)(ql + a.off_qaux);
    float s0 = 0.f, s1 = 0.f;
    for_chunks(r, ce, [&](uint4 x, uint32_t c, bool ok) {
      float t0 = s0, t1 = s1;
      fp8_word<METRIC>(t0, t1, x.x, q[4 * c + 0], sc[4 * c + 0]);
      fp8_word<METRIC>(t0, t1, x.y, q[4 * c + 1], sc[4 * c + 1]);
      fp8_word<METRIC>(t0, t1, x.z, q[4 * c + 2], sc[4 * c + 2]);
      fp8_word<METRIC>(t0, t1, x.w, q[4 * c + 3], sc[4 * c + 3]);
      s0 = ok ? t0 : s0;
      s1 = ok ? t1 : s1;
      // one chunk's staged query words at a time: hoisting them all costs
      // more occupancy than their LDS latency
      __builtin_amdgcn_sched_barrier(0);
    }, cb);
    return s0 + s1;
  }
};

// Distance of the row of `id` split over groups of a.trav_split lanes: lane
// `sub` of a group sums the terms of chunks [sub * cps, (sub + 1) * cps) of
// the reduced row (cps = a.trav_cps chunks, one load batch), and the group's
// partial sums are added in a fixed butterfly order, so every lane of the
// group returns the same value.  All lanes of the wave must call it; `valid`
// lanes contribute their part, the others zero.
template <typename T, int METRIC, int TRAV>
__device__ __forceinline__ float split_dist(const SearchArgs& a, uint32_t id, bool valid, const unsigned char* ql,
                                            uint32_t sub) {
  using P = TravDist<T, METRIC, TRAV>;
  const uint32_t cb = sub * a.trav_cps;
  const uint32_t ce = cb + a.trav_cps < a.trav_chunks ? cb + a.trav_cps : a.trav_chunks;
  float s = 0.f;
  if (valid && cb < ce) s = P::part(a, P::row_ptr(a, id), ql, cb, ce);
  for (uint32_t o = 1; o < a.trav_split; o <<= 1) s += __shfl_xor(s, (int)o);
  return P::finish(s);
}

// Stages one query row into LDS: the first d_bytes bytes of src (whole
// 16-byte chunks are read; a row is a multiple of 16 bytes long), zeros up to
// dst_bytes (a multiple of 16).
__device__ __forceinline__ void stage_query(void* dst, const void* src, uint32_t d_bytes, uint32_t dst_bytes) {
  uint4* o = (uint4*)dst;
  const uint4* s = (const uint4*)src;
  for (uint32_t c = lane_id(); c < (dst_bytes >> 4); c += kWave) {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t b0 = c * 16u;
    if (b0 < d_bytes) {
      v = s[c];
      if (b0 + 16u > d_bytes) {
        const uint32_t keep = d_bytes - b0;  // 1..15 bytes of this chunk belong to the row
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
        for (uint32_t i = 0; i < 4; i++) {
          const uint32_t lo = 4u * i;
          if (lo >= keep) w[i] = 0u;
          else if (lo + 4u > keep) w[i] &= (1u << ((keep - lo) * 8u)) - 1u;
        }
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    o[c] = v;
  }
}

// Stages the query of a TRAV search: the row itself (zero beyond d), and for
// FP8 the per-dimension factors (s_j for L2, q_j * s_j for MIPS).
template <typename T, int METRIC, int TRAV>
__device__ __forceinline__ void stage_traversal_query(const SearchArgs& a, unsigned char* ql, const T* src) {
  stage_query(ql, src, a.d * (uint32_t)sizeof(T), a.qrow_bytes);
  if constexpr (TRAV == TRAV_FP8) {
    wave_sync();
    const float* q = (const float*)ql;
    float* aux = (float*)(ql + a.off_qaux);
    for (uint32_t j = lane_id(); j < (a.qrow_bytes >> 2); j += kWave) {
      const float s = j < a.d ? a.trav_scale[j] : 0.f;
      aux[j] = METRIC == L2 ? s : q[j] * s;
    }
  }
}

}  // namespace ddg
