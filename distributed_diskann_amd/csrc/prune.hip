// robustPrune (index.h:63-123) — one wavefront per point, persistent launch.
//
// Task t prunes point p = pts[t] over candidates A_t (ids, with distances to
// p or without), plus G[p] when `add` (index.h:70-75).  Candidates are
// ordered by (distance, id) with a bitonic sort in LDS; the reference's
// std::sort orders by distance only, so equal-distance candidates may come
// out in another order there (its tie order is libstdc++'s, not a contract).
// Then the greedy sweep of index.h:84-106: take the next candidate that is
// neither p nor pruned, and prune every later candidate p' with
// alpha * d(p*, p') <= d(p, p') evaluated in double.
//
// Modes: forward (write up to Rsel ids per task to out_ids/out_len), and
// reverse (append_if_fits: if |A| + deg(p) <= Rsel append A to G[p]
// (graph.h:84-93), else replace G[p] by the pruned list; index.h:309-322).
//
// A task whose candidate list exceeds the LDS capacity is queued in
// big_tasks; a second launch (big_pass = 1) prunes the queued tasks in global
// scratch: a slot of gcap entries per wave when the host knows the longest
// list, otherwise a slice from a device bump allocator (big_off) over a pool
// the host sized so it cannot run out.  Both launches read their task counts
// on the device, so a round can be recorded before its sizes exist.
//
// When a task's candidate rows fit the per-wave LDS row cache (rrows rows),
// they are staged there once (STAGED): d(p, c) and every d(p*, p') of the
// greedy sweep then read LDS instead of HBM, with the same arithmetic.  The
// cache stride is an odd number of 16-byte chunks, so the lanes of a wave
// reading the same chunk of different rows spread over all LDS banks.
#include "device.hpp"
#include "internal.hpp"

namespace ddg {

// One robustPrune task (index.h:63-110) with the wave: candidates, sort,
// greedy selection, write-back.
template <typename T, int METRIC, bool BIG, bool STAGED>
__device__ __forceinline__ void prune_task(const PruneArgs& a, uint32_t t, uint32_t p, uint32_t* grow, uint64_t a0,
                                           uint64_t alen, uint32_t gdeg, T* pv, T* sv, uint32_t* sel, float* Cd,
                                           uint32_t* Ci, uint16_t* Cs, T* rows) {
  {
    using D = Dist<T, METRIC>;
    const uint32_t lane = lane_id();
    const uint32_t row_bytes = a.ad * (uint32_t)sizeof(T);
    const uint32_t nchunks = row_bytes >> 4;
    const T* base = (const T*)a.base;
    const uint32_t M = (uint32_t)(alen + gdeg);
    stage_row(pv, base + (uint64_t)p * a.ad, row_bytes);
    // ---- candidates: A (given or computed distances) then G[p]
    for (uint32_t j = lane; j < M; j += kWave)
      Ci[j] = j < alen ? a.a_ids[a0 + j] : grow[1 + (j - (uint32_t)alen)];
    if (BIG) __threadfence_block();
    wave_sync();
    if constexpr (STAGED) {  // candidate j's row to rows[j], eight 16-byte loads in flight per lane
      const uint32_t total = M * nchunks;
      for (uint32_t e0 = lane; e0 < total; e0 += 8u * kWave) {
        uint4 v[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) {
          const uint32_t e = e0 + u * kWave;
          if (e < total) {
            const uint32_t j = e / nchunks, c = e - j * nchunks;
            v[u] = ((const uint4*)(base + (uint64_t)Ci[j] * a.ad))[c];
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) {
          const uint32_t e = e0 + u * kWave;
          if (e < total) {
            const uint32_t j = e / nchunks, c = e - j * nchunks;
            ((uint4*)(rows + (uint64_t)j * a.rstride))[c] = v[u];
          }
        }
      }
      wave_sync();
    }
    const typename D::Q ps = D::prep(pv, nchunks);
    for (uint32_t j = lane; j < M; j += kWave) {
      const uint32_t id = Ci[j];
      const T* r = STAGED ? rows + (uint64_t)j * a.rstride : base + (uint64_t)id * a.ad;
      Cd[j] = (j < alen && a.a_dists) ? a.a_dists[a0 + j] : D::dist(r, pv, ps, a.d, nchunks);
      if (STAGED) Cs[j] = (uint16_t)j;
    }
    uint32_t P2 = 1;
    while (P2 < M) P2 <<= 1;
    for (uint32_t j = M + lane; j < P2; j += kWave) {
      Cd[j] = __builtin_inff();
      Ci[j] = kPruned;
      if (STAGED) Cs[j] = 0;
    }
    if (BIG) __threadfence_block();
    wave_sync();

    // ---- bitonic sort of P2 entries by (distance, id)
    for (uint32_t kk = 2; kk <= P2; kk <<= 1) {
      for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
        for (uint32_t i = lane; i < (P2 >> 1); i += kWave) {
          const uint32_t lo = ((i / jj) * (jj << 1)) + (i % jj);
          const uint32_t hi = lo + jj;
          const bool up = (lo & kk) == 0;
          const float dl = Cd[lo], dh = Cd[hi];
          const uint32_t il = Ci[lo], ih = Ci[hi];
          const bool sw = up ? less_di(dh, ih, dl, il) : less_di(dl, il, dh, ih);
          if (sw) {
            Cd[lo] = dh; Cd[hi] = dl; Ci[lo] = ih; Ci[hi] = il;
            if (STAGED) { const uint16_t sl = Cs[lo]; Cs[lo] = Cs[hi]; Cs[hi] = sl; }
          }
        }
        if (BIG) __threadfence_block();
        wave_sync();
      }
    }

    // ---- greedy selection (index.h:84-106)
    uint32_t nsel = 0, idx = 0;
    while (nsel < a.Rsel && idx < M) {
      int found = -1;
      for (uint32_t b0 = idx; b0 < M; b0 += kWave) {
        const uint32_t i = b0 + lane;
        bool ok = false;
        if (i < M) {
          const uint32_t c = Ci[i];
          ok = c != kPruned && c != p;
        }
        const uint64_t m = ballot(ok);
        if (m) { found = (int)(b0 + (uint32_t)first_lane(m)); break; }
      }
      if (found < 0) break;
      const uint32_t pstar = Ci[found];
      if (lane == 0) sel[nsel] = pstar;
      nsel++;
      idx = (uint32_t)found + 1u;
      if (nsel >= a.Rsel) break;  // the reference's remaining prunes cannot change the result
      const T* ps_row = sv;
      if constexpr (STAGED) {
        ps_row = rows + (uint64_t)Cs[found] * a.rstride;
      } else {
        stage_row(sv, base + (uint64_t)pstar * a.ad, row_bytes);
        wave_sync();
      }
      const typename D::Q ss = D::prep(ps_row, nchunks);
      for (uint32_t j = idx + lane; j < M; j += kWave) {
        const uint32_t c = Ci[j];
        if (c != kPruned) {
          const T* r = STAGED ? rows + (uint64_t)Cs[j] * a.rstride : base + (uint64_t)c * a.ad;
          const float dsp = D::dist(r, ps_row, ss, a.d, nchunks);
          if (a.alpha * (double)dsp <= (double)Cd[j]) Ci[j] = kPruned;
        }
      }
      if (BIG) __threadfence_block();
      wave_sync();
    }
    wave_sync();
    if (a.write_graph) {
      for (uint32_t j = lane; j < nsel; j += kWave) grow[1 + j] = sel[j];
      wave_sync();
      if (lane == 0) grow[0] = nsel;
    } else {
      for (uint32_t j = lane; j < nsel; j += kWave) a.out_ids[(uint64_t)t * a.Rsel + j] = sel[j];
      if (lane == 0) a.out_len[t] = nsel;
    }
  }
}

template <typename T, int METRIC, bool BIG>
__global__ __launch_bounds__(256) void k_prune(PruneArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = lane_id();
  const uint32_t wave = bcast0(threadIdx.x >> 6);  // scalar: per-wave pointers stay in SGPRs
  unsigned char* my = smem + wave * a.lds_per_wave;
  T* pv = (T*)my;                                   // p's row
  T* sv = (T*)(my + ((a.ad * sizeof(T) + 15u) & ~15u));  // p*'s row
  uint32_t* sel = (uint32_t*)(my + 2u * ((a.ad * sizeof(T) + 15u) & ~15u));
  float* Cd = (float*)(my + a.off_cand);                       // LDS layout (not BIG)
  uint32_t* Ci = (uint32_t*)(Cd + a.ccap);
  uint16_t* Cs = (uint16_t*)(my + a.off_cand + 8u * a.ccap);
  T* rows = (T*)(my + a.off_rows);
  const uint32_t cap = a.ccap;
  uint64_t ntasks = a.np;
  if (BIG) ntasks = a.nbig[0];
  else if (a.np_dev && a.np_dev[0] < ntasks) ntasks = a.np_dev[0];

  for (;;) {
    const uint32_t tq = fetch_task(a.work);
    if (tq >= ntasks) break;
    const uint32_t t = BIG ? a.big_tasks[tq] : tq;
    if (t == kNoTask) continue;  // fetch_task reconverges the wave first
    if (BIG) {
      const uint64_t off = a.big_off ? a.big_off[tq]
                                     : ((uint64_t)blockIdx.x * (blockDim.x >> 6) + wave) * a.gcap;
      Cd = a.gscratch_d + off;
      Ci = a.gscratch_i + off;
    }
    const uint32_t p = a.pts[t];
    uint32_t* grow = a.graph + (uint64_t)p * a.R1;
    uint64_t a0, alen;
    if (a.a_off) { a0 = a.a_off[t]; alen = a.a_off[t + 1] - a0; }
    else { a0 = (uint64_t)t * a.a_stride; alen = a.a_len[t]; }
    const uint32_t deg = grow[0];
    const uint32_t gdeg = a.add ? deg : 0u;

    const uint64_t M64 = alen + gdeg;
    if (a.append_if_fits && alen + deg <= a.Rsel) {
      // append_neighbors (graph.h:84-93)
      for (uint32_t j = lane; j < alen; j += kWave) grow[1 + deg + j] = a.a_ids[a0 + j];
      wave_sync();
      if (lane == 0) grow[0] = deg + (uint32_t)alen;
    } else if (!BIG && M64 > cap) {
      if (lane == 0) {
        const uint32_t s = atomicAdd(a.nbig, 1u);
        if (a.big_off) {
          uint64_t need = 1;
          while (need < M64) need <<= 1;  // the bitonic sort pads to a power of two
          const uint64_t off = atomicAdd((unsigned long long*)a.gused, (unsigned long long)need);
          if (off + need <= a.gpool) {
            a.big_tasks[s] = t;
            a.big_off[s] = off;
          } else {
            a.big_tasks[s] = kNoTask;
            atomicOr(a.status, 2u);
          }
        } else if (M64 <= a.gcap) {
          a.big_tasks[s] = t;
        } else {
          a.big_tasks[s] = kNoTask;
          atomicOr(a.status, 2u);
        }
      }
    } else if (!BIG && M64 <= a.rrows) {
      prune_task<T, METRIC, false, true>(a, t, p, grow, a0, alen, gdeg, pv, sv, sel, Cd, Ci, Cs, rows);
    } else {
      prune_task<T, METRIC, BIG, false>(a, t, p, grow, a0, alen, gdeg, pv, sv, sel, Cd, Ci, Cs, rows);
    }
    wave_sync();
  }
}

template <typename T, int M>
static void set_attrs_t(int lds) {
  (void)hipFuncSetAttribute((const void*)&k_prune<T, M, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void*)&k_prune<T, M, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

void prune_set_attributes(int max_lds) {
  set_attrs_t<uint8_t, L2>(max_lds);
  set_attrs_t<uint8_t, MIPS>(max_lds);
  set_attrs_t<int8_t, L2>(max_lds);
  set_attrs_t<int8_t, MIPS>(max_lds);
  set_attrs_t<float, L2>(max_lds);
  set_attrs_t<float, MIPS>(max_lds);
}

// (the dynamic-LDS limit is raised once by prune_set_attributes, outside any
// stream capture)
template <typename T, int M>
static hipError_t launch_t(const PruneArgs& a, const Launch& l, hipStream_t s) {
  if (a.big_pass)
    hipLaunchKernelGGL((k_prune<T, M, true>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s, a);
  else
    hipLaunchKernelGGL((k_prune<T, M, false>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s,
                       a);
  return hipGetLastError();
}

hipError_t launch_prune(int dtype, int metric, const PruneArgs& a, const Launch& l, hipStream_t s) {
  if (dtype == U8) return metric == L2 ? launch_t<uint8_t, L2>(a, l, s) : launch_t<uint8_t, MIPS>(a, l, s);
  if (dtype == I8) return metric == L2 ? launch_t<int8_t, L2>(a, l, s) : launch_t<int8_t, MIPS>(a, l, s);
  return metric == L2 ? launch_t<float, L2>(a, l, s) : launch_t<float, MIPS>(a, l, s);
}

}  // namespace ddg
