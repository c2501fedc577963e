// Device-side building blocks shared by the kernels: wave64 primitives, the
// visited-filter hash, (distance, id) ordering and the per-dtype distance
// functions.  Everything here is written for gfx950 (CDNA4, wave64).
//
// The whole library is compiled with -ffp-contract=off: the FP32 distance
// functions below reproduce the reference's summation order exactly and must
// not be contracted into FMAs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ddg {

constexpr uint32_t kWave = 64;
constexpr uint32_t kFlag = 0x80000000u;   // frontier entry already visited
constexpr uint32_t kIdMask = 0x7FFFFFFFu;
constexpr uint32_t kPruned = 0xFFFFFFFFu; // index.h:102 sentinel (-1)
constexpr uint32_t kNoTask = 0xFFFFFFFFu; // queue slot without a task

enum Dtype : int { U8 = 0, I8 = 1, F32 = 2 };
enum Metric : int { L2 = 0, MIPS = 1 };

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }

// number of set bits of m below this lane
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }

// Orders this wave's LDS accesses: everything before is visible to every lane
// of the wave after.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Like wave_sync, for global memory written by one lane and read by another
// lane of the same wave (workgroup-scope release/acquire).
__device__ __forceinline__ void wave_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ uint32_t bcast0(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Next task index of a persistent wave.  The wave barrier first reconverges
// every lane: without it the compiler may thread a preceding `if (lane == 0)`
// block straight into this one, so that lanes 1..63 re-read a stale index
// while lane 0 takes the atomic (observed as an endless re-run of one task).
__device__ __forceinline__ uint32_t fetch_task(uint32_t* counter) {
  wave_sync();
  uint32_t v = 0;
  if (lane_id() == 0) v = atomicAdd(counter, 1u);
  return bcast0(v);
}

// SplitMix64 finaliser (parlay/utilities.h:145-150).
__device__ __forceinline__ uint64_t hash64_2(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

// beamSearch.h:66: bits = max(10, ceil(log2(beam^2)) - 2), in integers.
__host__ __device__ __forceinline__ uint32_t filter_bits(uint64_t beam) {
  uint64_t x = beam * beam;
  uint32_t c = 0;
  while (c < 63 && (1ull << c) < x) c++;  // ceil(log2(x)) for x >= 1
  int b = (int)c - 2;
  return (uint32_t)(b < 10 ? 10 : b);
}

// (distance, id) order of beamSearch.h:59-61.
__device__ __forceinline__ bool less_di(float da, uint32_t ia, float db, uint32_t ib) {
  // both sides evaluated (no short circuit): compares and mask operations,
  // no branch around the second test
  return (da < db) | ((da == db) & (ia < ib));
}

// #entries of the sorted list (d[], id[] & kIdMask) strictly below (x, xi)
__device__ __forceinline__ uint32_t lower_bound_di(const float* d, const uint32_t* id, uint32_t n, float x,
                                                   uint32_t xi) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (less_di(d[mid], id[mid] & kIdMask, x, xi)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// #entries of the sorted list not above (x, xi)
__device__ __forceinline__ uint32_t upper_bound_di(const float* d, const uint32_t* id, uint32_t n, float x,
                                                   uint32_t xi) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (less_di(x, xi, d[mid], id[mid] & kIdMask)) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// ------------------------------------------------------------------ distances
// A distance walks its row in batches of kChunkBatch 16-byte chunks whose
// loads are all in flight before the arithmetic (one memory round trip per
// batch).  Whole batches carry no per-chunk conditions; the last, partial
// batch re-reads in-bounds chunks and its extra terms are masked out, so no
// chunk load or term sits behind a branch.
constexpr uint32_t kChunkBatch = 8;

// body(x, c, ok) for the chunks c = cb .. n-1 of row r (cb < n); ok is false
// for the masked tail slots, whose c repeats the last chunk.  Whole rows of
// the library's copies are multiples of four chunks (64-byte alignment), so
// they never reach the masked tail.
template <typename F>
__device__ __forceinline__ void for_chunks(const uint4* r, uint32_t n, F&& body, uint32_t cb = 0) {
  uint32_t c0 = cb;
  for (; c0 + kChunkBatch <= n; c0 += kChunkBatch) {
    uint4 xs[kChunkBatch];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch; u++) xs[u] = r[c0 + u];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch; u++) body(xs[u], c0 + u, true);
  }
  if (c0 + kChunkBatch / 2 <= n) {
    uint4 xs[kChunkBatch / 2];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch / 2; u++) xs[u] = r[c0 + u];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch / 2; u++) body(xs[u], c0 + u, true);
    c0 += kChunkBatch / 2;
  }
  if (c0 < n) {
    uint4 xs[kChunkBatch];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch; u++) xs[u] = r[c0 + u < n ? c0 + u : n - 1u];
#pragma unroll
    for (uint32_t u = 0; u < kChunkBatch; u++) body(xs[u], c0 + u < n ? c0 + u : n - 1u, c0 + u < n);
  }
}

__device__ __forceinline__ uint4 zero_unless(uint4 x, bool ok) {
  return ok ? x : make_uint4(0u, 0u, 0u, 0u);
}

// One lane computes one distance between a row in HBM and a reference vector
// staged in LDS.  Rows are aligned_d elements long, a multiple of 64 bytes,
// with zero padding (the library zeroes the padding of every device copy).
//
// Integer types: squared L2 = |q|^2 + |p|^2 - 2 q.p and the inner product are
// exact in int32 with the dot4 instructions (|values| < 2^24 at d <= 256), so
// the float cast gives the reference's bits (euclidian_point.h:44-60,
// mips_point.h:44-58).
template <typename T, int METRIC>
struct Dist;

template <int METRIC>
struct Dist<uint8_t, METRIC> {
  struct Q { uint32_t qq; };
  __device__ static Q prep(const uint8_t* q, uint32_t nchunks) {
    uint32_t s = 0;
    const uint4* v = (const uint4*)q;
    for (uint32_t c = 0; c < nchunks; c++) {
      uint4 x = v[c];
      s = __builtin_amdgcn_udot4(x.x, x.x, s, false);
      s = __builtin_amdgcn_udot4(x.y, x.y, s, false);
      s = __builtin_amdgcn_udot4(x.z, x.z, s, false);
      s = __builtin_amdgcn_udot4(x.w, x.w, s, false);
    }
    return Q{s};
  }
  __device__ static float dist(const uint8_t* row, const uint8_t* q, const Q& st, uint32_t /*d*/,
                               uint32_t nchunks) {
    const uint4* v = (const uint4*)q;
    uint32_t pp = 0, qp = 0;
    for_chunks((const uint4*)row, nchunks, [&](uint4 x, uint32_t c, bool ok) {
      x = zero_unless(x, ok);
      const uint4 y = v[c];
      qp = __builtin_amdgcn_udot4(x.x, y.x, qp, false);
      qp = __builtin_amdgcn_udot4(x.y, y.y, qp, false);
      qp = __builtin_amdgcn_udot4(x.z, y.z, qp, false);
      qp = __builtin_amdgcn_udot4(x.w, y.w, qp, false);
      if (METRIC == L2) {
        pp = __builtin_amdgcn_udot4(x.x, x.x, pp, false);
        pp = __builtin_amdgcn_udot4(x.y, x.y, pp, false);
        pp = __builtin_amdgcn_udot4(x.z, x.z, pp, false);
        pp = __builtin_amdgcn_udot4(x.w, x.w, pp, false);
      }
    });
    if (METRIC == L2) return (float)(int32_t)(st.qq + pp - 2u * qp);
    return -((float)(int32_t)qp);
  }
};

template <int METRIC>
struct Dist<int8_t, METRIC> {
  struct Q { int32_t qq; };
  __device__ static Q prep(const int8_t* q, uint32_t nchunks) {
    int32_t s = 0;
    const uint4* v = (const uint4*)q;
    for (uint32_t c = 0; c < nchunks; c++) {
      uint4 x = v[c];
      s = __builtin_amdgcn_sdot4((int)x.x, (int)x.x, s, false);
      s = __builtin_amdgcn_sdot4((int)x.y, (int)x.y, s, false);
      s = __builtin_amdgcn_sdot4((int)x.z, (int)x.z, s, false);
      s = __builtin_amdgcn_sdot4((int)x.w, (int)x.w, s, false);
    }
    return Q{s};
  }
  __device__ static float dist(const int8_t* row, const int8_t* q, const Q& st, uint32_t /*d*/,
                               uint32_t nchunks) {
    const uint4* v = (const uint4*)q;
    int32_t pp = 0, qp = 0;
    for_chunks((const uint4*)row, nchunks, [&](uint4 x, uint32_t c, bool ok) {
      x = zero_unless(x, ok);
      const uint4 y = v[c];
      qp = __builtin_amdgcn_sdot4((int)x.x, (int)y.x, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.y, (int)y.y, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.z, (int)y.z, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.w, (int)y.w, qp, false);
      if (METRIC == L2) {
        pp = __builtin_amdgcn_sdot4((int)x.x, (int)x.x, pp, false);
        pp = __builtin_amdgcn_sdot4((int)x.y, (int)x.y, pp, false);
        pp = __builtin_amdgcn_sdot4((int)x.z, (int)x.z, pp, false);
        pp = __builtin_amdgcn_sdot4((int)x.w, (int)x.w, pp, false);
      }
    });
    if (METRIC == L2) return (float)(st.qq + pp - 2 * qp);
    return -((float)qp);
  }
};

// FP32: the reference's own order.  L2 follows NSGDist.h:113-131 (per group
// of four ((e0^2 + e1^2) + e2^2) + e3^2 added to the running sum, then the
// 0-3 tail elements one by one); MIPS follows mips_point.h:60-66 (sequential
// sum, negated).  A masked group leaves the sum as it was.  The zero padding
// of rows and staged query adds exact zeros, so MIPS, and L2 when d is a
// multiple of four, walk the whole aligned row (nchunks groups, no tail).
template <int METRIC>
struct Dist<float, METRIC> {
  struct Q { int unused; };
  __device__ static Q prep(const float*, uint32_t) { return Q{0}; }
  __device__ static float dist(const float* row, const float* q, const Q&, uint32_t d, uint32_t nchunks) {
    const float4* q4 = (const float4*)q;
    float acc = 0.f;
    const bool whole = METRIC == MIPS || (d & 3u) == 0u;
    const uint32_t ng = whole ? nchunks : d >> 2;  // groups of four, in order, then the tail
    if (ng > 0)
      for_chunks((const uint4*)row, ng, [&](uint4 xr, uint32_t g, bool ok) {
        const float4 a = make_float4(__uint_as_float(xr.x), __uint_as_float(xr.y), __uint_as_float(xr.z),
                                     __uint_as_float(xr.w));
        const float4 b = q4[g];
        float t;
        if (METRIC == L2) {
          float e0 = a.x - b.x, e1 = a.y - b.y, e2 = a.z - b.z, e3 = a.w - b.w;
          float gs = e0 * e0;
          gs = gs + e1 * e1;
          gs = gs + e2 * e2;
          gs = gs + e3 * e3;
          t = acc + gs;
        } else {
          t = acc + b.x * a.x;
          t = t + b.y * a.y;
          t = t + b.z * a.z;
          t = t + b.w * a.w;
        }
        acc = ok ? t : acc;
      });
    for (uint32_t i = whole ? d : ng << 2; i < d; i++) {
      if (METRIC == L2) {
        float e = row[i] - q[i];
        acc = acc + e * e;
      } else {
        acc = acc + q[i] * row[i];
      }
    }
    return METRIC == L2 ? acc : -acc;
  }
};

// Dist<float, METRIC>::dist of one row split over a group of S lanes (S a
// power of two, groups aligned in the wave), for rows that walk whole
// groups of four (MIPS, or L2 with d % 4 == 0) in ng chunks.  Lane `sub` of
// the group loads chunks [sub * cps, sub * cps + cps) together (one round
// trip for the whole row instead of one per batch of eight), forms its
// products (MIPS) or per-group sums of squares (L2) with the same roundings
// as the one-lane walk, and then the running sum is handed along the group:
// lane sub adds its terms, in order, to the sum lane sub - 1 passes on.  The
// adds are the one-lane walk's, in its order, so the bits are the same.  All
// lanes of the wave call it with the same S / cps; every lane's row pointer
// must be valid (lanes without a row pass any valid row, their result is
// discarded).  The last lane of each group returns the distance.  MAXC is
// the number of chunk slots a lane walks (cps <= MAXC): instantiated for the
// share itself, the chain adds nothing for slots past it.
constexpr uint32_t kSplitChunks = 8;  // at most this many chunks per lane

template <int METRIC, uint32_t MAXC = kSplitChunks>
__device__ __forceinline__ float split_row_dist(const float* row, const float* q, uint32_t ng, uint32_t S,
                                                uint32_t cps, uint32_t sub) {
  const float4* q4 = (const float4*)q;
  const uint4* r = (const uint4*)row;
  const uint32_t cb = sub * cps;
  uint4 xs[MAXC];
#pragma unroll
  for (uint32_t u = 0; u < MAXC; u++) {
    const uint32_t c = cb + u;
    xs[u] = r[(u < cps && c < ng) ? c : 0u];
  }
  // the terms this lane adds: four products per chunk (MIPS), or one sum of
  // squares per chunk (L2: ((e0^2 + e1^2) + e2^2) + e3^2, NSGDist.h:113-131).
  // A slot past the lane's chunks holds +0.0 terms: the running sum starts
  // at +0.0 and never becomes -0.0 (an exact zero sum is +0.0), so adding
  // +0.0 leaves its bits as they are, and the chain below needs no selects.
  float t[MAXC][4];
  bool ok[MAXC];
#pragma unroll
  for (uint32_t u = 0; u < MAXC; u++) {
    const uint32_t c = cb + u;
    ok[u] = u < cps && c < ng;
    const float4 b = q4[ok[u] ? c : 0u];
    const float4 a = make_float4(__uint_as_float(xs[u].x), __uint_as_float(xs[u].y), __uint_as_float(xs[u].z),
                                 __uint_as_float(xs[u].w));
    if (METRIC == L2) {
      const float e0 = a.x - b.x, e1 = a.y - b.y, e2 = a.z - b.z, e3 = a.w - b.w;
      float gs = e0 * e0;
      gs = gs + e1 * e1;
      gs = gs + e2 * e2;
      gs = gs + e3 * e3;
      t[u][0] = ok[u] ? gs : 0.f;
    } else {
      t[u][0] = ok[u] ? b.x * a.x : 0.f;
      t[u][1] = ok[u] ? b.y * a.y : 0.f;
      t[u][2] = ok[u] ? b.z * a.z : 0.f;
      t[u][3] = ok[u] ? b.w * a.w : 0.f;
    }
  }
  float acc = 0.f;
  for (uint32_t s = 0; s < S; s++) {
    const float in = __shfl_up(acc, 1);
    if (sub == s) {
      float v = s == 0 ? 0.f : in;
#pragma unroll
      for (uint32_t u = 0; u < MAXC; u++) {
        if (METRIC == L2) {
          v = v + t[u][0];
        } else {
          v = v + t[u][0];
          v = v + t[u][1];
          v = v + t[u][2];
          v = v + t[u][3];
        }
      }
      acc = v;
    }
  }
  return METRIC == L2 ? acc : -acc;
}

// Copies one aligned row (bytes a multiple of 16) into LDS with the wave.
__device__ __forceinline__ void stage_row(void* dst, const void* src, uint32_t bytes) {
  uint4* o = (uint4*)dst;
  const uint4* s = (const uint4*)src;
  for (uint32_t c = lane_id(); c < (bytes >> 4); c += kWave) o[c] = s[c];
}

}  // namespace ddg
