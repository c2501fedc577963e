// robustPrune (index.h:63-123) — one wavefront per point, persistent launch.
//
// Task t prunes point p = pts[t] over candidates A_t (ids, with distances to
// p or without), plus G[p] when `add` (index.h:70-75).  The greedy sweep of
// index.h:76-106 over candidates sorted by (distance, id) — take the next
// candidate that is neither p nor pruned, and prune every later candidate p'
// with alpha * d(p*, p') <= d(p, p') evaluated in double — runs without the
// sort: each step takes the smallest (distance, id) of the live candidates.
// The reference's std::sort orders by distance only, so equal-distance
// candidates may come out in another order there (its tie order is
// libstdc++'s, not a contract).
//
// Modes: forward (write up to Rsel ids per task to out_ids/out_len), and
// reverse (append_if_fits: if |A| + deg(p) <= Rsel append A to G[p]
// (graph.h:84-93), else replace G[p] by the pruned list; index.h:309-322).
//
// A task whose candidate list exceeds the LDS capacity is queued in
// big_tasks; a second launch (big_pass = 1, k_prune_big) prunes the queued
// tasks in global scratch, one workgroup per task: a slot of gcap entries per
// block when the host knows the longest list, otherwise a slice from a device
// bump allocator (big_off) over a pool the host sized so it cannot run out.
// Both launches read their task counts on the device, so a round can be
// recorded before its sizes exist.
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
// greedy selection, write-back.  SPLIT (float32 rows of more than one load
// batch) spreads each distance over a group of a.nat_split lanes
// (split_row_dist: all of the row's loads in flight at once, the same
// additions in the same order); a pass then covers 64 / nat_split
// candidates and the group's last lane stands for its candidate.
template <typename T, int METRIC, bool STAGED, bool SPLIT = false>
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
    const uint32_t lg = SPLIT ? (uint32_t)__builtin_ctz(a.nat_split) : 0u;
    const uint32_t G = kWave >> lg;       // candidates per pass
    const uint32_t li = lane >> lg;       // this lane's candidate in the pass
    const uint32_t sub = lane & ((1u << lg) - 1u);
    const bool rep = sub == (1u << lg) - 1u;  // the lane that holds the distance
    auto row_dist = [&](const T* r, const T* q, const typename D::Q& qs) -> float {
      if constexpr (SPLIT) return split_row_dist<METRIC>((const float*)r, (const float*)q, nchunks, a.nat_split, a.nat_cps, sub);
      else return D::dist(r, q, qs, a.d, nchunks);
    };
    stage_row(pv, base + (uint64_t)p * a.ad, row_bytes);
    // ---- candidates: A (given or computed distances) then G[p]
    for (uint32_t j = lane; j < M; j += kWave)
      Ci[j] = j < alen ? a.a_ids[a0 + j] : grow[1 + (j - (uint32_t)alen)];
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
    // given distances first, then the computed ones (every lane of a pass
    // takes part in a split distance; lanes past the list repeat its last row)
    const uint32_t given = a.a_dists ? (uint32_t)alen : 0u;
    for (uint32_t j = lane; j < given; j += kWave) Cd[j] = a.a_dists[a0 + j];
    for (uint32_t b0 = given; b0 < M; b0 += G) {
      const uint32_t j = b0 + li < M ? b0 + li : M - 1u;
      const T* r = STAGED ? rows + (uint64_t)j * a.rstride : base + (uint64_t)Ci[j] * a.ad;
      const float dj = row_dist(r, pv, ps);
      if (rep && b0 + li < M) Cd[j] = dj;
    }
    if (STAGED)
      for (uint32_t j = lane; j < M; j += kWave) Cs[j] = (uint16_t)j;
    wave_sync();

    // ---- greedy selection (index.h:76-106) without a sort.  Sorting by
    // (distance, id) and taking the first candidate that is neither pruned
    // nor p is the same as taking, each step, the smallest (distance, id) of
    // the live candidates: every candidate the sorted sweep passed is either
    // selected, pruned or p itself.  So candidates equal to p are dropped
    // first, then each step selects the live minimum, prunes the other live
    // candidates against it (they all come after it in sorted order; an equal
    // copy of it too), compacts the survivors to the front in their order and
    // finds the next minimum in the same pass.
    float bd = __builtin_inff();
    uint32_t bi = kPruned, bp = kPruned;  // this lane's smallest (distance, id, position)
    auto offer = [&](float d, uint32_t i, uint32_t pos) {
      if (d < bd || (d == bd && (i < bi || (i == bi && pos < bp)))) { bd = d; bi = i; bp = pos; }
    };
    auto wave_min = [&]() {
      for (int o = 1; o < (int)kWave; o <<= 1) {
        const float od = __shfl_xor(bd, o);
        const uint32_t oi = __shfl_xor(bi, o), op = __shfl_xor(bp, o);
        offer(od, oi, op);
      }
      bp = bcast0(bp);
      bi = bcast0(bi);
    };
    uint32_t live = 0;
    for (uint32_t b0 = 0; b0 < M; b0 += kWave) {
      const uint32_t j = b0 + lane;
      bool keep = false;
      uint32_t c = 0, cs = 0;
      float cd = 0.f;
      if (j < M) {
        c = Ci[j];
        cd = Cd[j];
        if (STAGED) cs = Cs[j];
        keep = c != p;
      }
      const uint64_t km = ballot(keep);
      wave_sync();  // every lane has read its entry before any is overwritten
      if (keep) {
        const uint32_t o = live + mbcnt(km);  // o <= j: entries move left
        Ci[o] = c;
        Cd[o] = cd;
        if (STAGED) Cs[o] = (uint16_t)cs;
        offer(cd, c, o);
      }
      live += popc(km);
      wave_sync();
    }
    wave_min();

    uint32_t nsel = 0;
    while (bp != kPruned) {
      const uint32_t pstar = bi, at = bp;
      if (lane == 0) sel[nsel] = pstar;
      nsel++;
      if (nsel >= a.Rsel) break;  // the reference's remaining prunes cannot change the result
      const T* ps_row = sv;
      if constexpr (STAGED) {
        ps_row = rows + (uint64_t)Cs[at] * a.rstride;
      } else {
        stage_row(sv, base + (uint64_t)pstar * a.ad, row_bytes);
        wave_sync();
      }
      const typename D::Q ss = D::prep(ps_row, nchunks);
      bd = __builtin_inff();
      bi = bp = kPruned;
      uint32_t kept = 0;
      for (uint32_t b0 = 0; b0 < live; b0 += G) {
        const uint32_t j = b0 + li;
        const uint32_t jr = j < live ? j : live - 1u;  // a valid row for every lane of a split
        const uint32_t c = Ci[jr];
        const float cd = Cd[jr];
        uint32_t cs = 0;
        if (STAGED) cs = Cs[jr];
        const T* r = STAGED ? rows + (uint64_t)cs * a.rstride : base + (uint64_t)c * a.ad;
        const float dsp = row_dist(r, ps_row, ss);
        const bool keep = rep && j < live && j != at && !(a.alpha * (double)dsp <= (double)cd);
        const uint64_t km = ballot(keep);
        wave_sync();  // every lane has read its entry (and p*'s row slot) before any is overwritten
        if (keep) {
          const uint32_t o = kept + mbcnt(km);
          Ci[o] = c;
          Cd[o] = cd;
          if (STAGED) Cs[o] = (uint16_t)cs;
          offer(cd, c, o);
        }
        kept += popc(km);
        wave_sync();
      }
      live = kept;
      wave_min();
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

typedef int v4i_t __attribute__((ext_vector_type(4)));
constexpr uint32_t kGramMax = 256;  // candidates of a Gram-matrix task
constexpr uint32_t kGramWords = kGramMax / 32;

// LDS bytes of the Gram-matrix path: norms and sums (int32 each) and the
// prune bit matrix (kGramMax rows of kGramMax bits)
__host__ __device__ constexpr uint32_t gram_lds_bytes() { return kGramMax * 8u + kGramMax * kGramWords * 4u; }

// robustPrune of one task over integer rows (uint8 / int8, d <= 256) from a
// Gram matrix on the matrix cores.  Candidates are sorted by (distance, id)
// (candidates equal to p last, never selected); for every pair a < b in that
// order the exact integer distance d(a, b) comes from one 16 x 16 x 64
// v_mfma_i32_16x16x64_i8 tile (|a|^2 + |b|^2 - 2 a.b for L2, -a.b for MIPS;
// uint8 enters as x - 128, corrected by 128 (sum a + sum b) - 16384 ad), and
// bit b of row a records whether a would prune b: alpha * d(a, b) <= d(p, b)
// in double (index.h:97-103).  The greedy sweep (index.h:84-106) is then a
// walk over bit rows: take the lowest live position, clear its row's bits.
// Same integers, same float casts and the same double test as the per-pair
// distances of prune_task, so the selections are identical.
// One pass over an integer candidate row r: its distance to the staged row q
// (the bits of Dist<T, METRIC>::dist), its squared norm nn and, for uint8,
// its byte sum ss; all sums exact in int32.
template <typename T, int METRIC>
__device__ __forceinline__ float int_row_terms(const uint4* r, const T* q, const typename Dist<T, METRIC>::Q& qs,
                                               uint32_t nchunks, int32_t& nn, int32_t& ss) {
  const uint4* v = (const uint4*)q;
  if constexpr (sizeof(T) != 1) {  // never reached: the Gram path takes integer rows only
    nn = ss = 0;
    return 0.f;
  } else if constexpr (T(-1) > T(0)) {
    uint32_t qp = 0, pp = 0, su = 0;
    for_chunks(r, nchunks, [&](uint4 x, uint32_t c, bool ok) {
      x = zero_unless(x, ok);
      const uint4 y = v[c];
      qp = __builtin_amdgcn_udot4(x.x, y.x, qp, false);
      qp = __builtin_amdgcn_udot4(x.y, y.y, qp, false);
      qp = __builtin_amdgcn_udot4(x.z, y.z, qp, false);
      qp = __builtin_amdgcn_udot4(x.w, y.w, qp, false);
      pp = __builtin_amdgcn_udot4(x.x, x.x, pp, false);
      pp = __builtin_amdgcn_udot4(x.y, x.y, pp, false);
      pp = __builtin_amdgcn_udot4(x.z, x.z, pp, false);
      pp = __builtin_amdgcn_udot4(x.w, x.w, pp, false);
      su = __builtin_amdgcn_udot4(x.x, 0x01010101u, su, false);
      su = __builtin_amdgcn_udot4(x.y, 0x01010101u, su, false);
      su = __builtin_amdgcn_udot4(x.z, 0x01010101u, su, false);
      su = __builtin_amdgcn_udot4(x.w, 0x01010101u, su, false);
    });
    nn = (int32_t)pp;
    ss = (int32_t)su;
    if (METRIC == L2) return (float)(int32_t)(qs.qq + pp - 2u * qp);
    return -((float)(int32_t)qp);
  } else {
    int32_t qp = 0, pp = 0;
    for_chunks(r, nchunks, [&](uint4 x, uint32_t c, bool ok) {
      x = zero_unless(x, ok);
      const uint4 y = v[c];
      qp = __builtin_amdgcn_sdot4((int)x.x, (int)y.x, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.y, (int)y.y, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.z, (int)y.z, qp, false);
      qp = __builtin_amdgcn_sdot4((int)x.w, (int)y.w, qp, false);
      pp = __builtin_amdgcn_sdot4((int)x.x, (int)x.x, pp, false);
      pp = __builtin_amdgcn_sdot4((int)x.y, (int)x.y, pp, false);
      pp = __builtin_amdgcn_sdot4((int)x.z, (int)x.z, pp, false);
      pp = __builtin_amdgcn_sdot4((int)x.w, (int)x.w, pp, false);
    });
    nn = pp;
    ss = 0;
    if (METRIC == L2) return (float)(qs.qq + pp - 2 * qp);
    return -((float)qp);
  }
}

// Sortable key of a candidate: the distance's bits sign-folded over the id
// (zero taken as +0.0), so the unsigned order of keys is the (distance, id)
// order of less_di.
__device__ __forceinline__ uint64_t prune_key(float d, uint32_t id) {
  uint32_t b = __float_as_uint(d);
  if ((b << 1) == 0u) b = 0u;
  b ^= (uint32_t)((int32_t)b >> 31) | 0x80000000u;
  return ((uint64_t)b << 32) | id;
}

// rk[u] += the number of keys of K[0, M4) below ku[u]; K is padded to M4 (a
// multiple of 4) with keys above every candidate's.
template <uint32_t NS>
__device__ __forceinline__ void count_below_keys(const uint64_t* K, uint32_t M4, const uint64_t* ku, uint32_t* rk) {
  for (uint32_t jj = 0; jj < M4; jj += 2) {
    const ulonglong2 e = *(const ulonglong2*)(K + jj);
#pragma unroll
    for (uint32_t u = 0; u < NS; u++) rk[u] += (uint32_t)(e.x < ku[u]) + (uint32_t)(e.y < ku[u]);
  }
}

// The largest integer x with alpha * (double)x <= c (alpha > 0), clamped to
// +-2^30: the prune test of index.h:97-103 for an integer distance x, which
// (float)x holds exactly below 2^24.
__device__ __forceinline__ int32_t prune_threshold(double alpha, double c) {
  constexpr double kLim = 1073741824.0;
  double f = floor(c / alpha);
  f = f < -kLim ? -kLim : (f > kLim ? kLim : f);
  while (f < kLim && alpha * (f + 1.0) <= c) f += 1.0;
  while (f > -kLim && !(alpha * f <= c)) f -= 1.0;
  return (int32_t)f;
}

// The upper triangle of the Gram matrix of the sorted candidates in 16 x 16
// tiles: v_mfma_i32_16x16x64_i8 leaves row 4 (lane / 16) + r, column
// lane % 16 in accumulator register r; lane (g4, c16) supplies the 16 bytes
// at chunk g4 NST + s of its row for K step s (the same K order for both
// operands).  Tile pairs (I, J >= I) run in row-major order and each needs
// one block of 16 rows (J's): the blocks are loaded kRing pairs ahead into a
// ring of registers (loads past the last pair re-read the last block, in
// bounds), and row block I's operand is the block of pair (I, I).
//
// Row a prunes column b when its integer distance RA[a] - K acc + CB[b] is
// at most b's threshold (alpha * d(a, b) <= d(p, b)), i.e. when
// RA[a] - K acc <= Tb[b] (K = 2 for L2, 1 for MIPS; RA, CB hold the norms and
// the uint8 offset terms); the tile's 16 x 16 bits go to the bit matrix as
// one 16-bit word per row.  Only words of blocks J >= I are written: bits
// of b before a are never read, since a is applied only when every earlier
// position is settled.
template <typename T, int METRIC, uint32_t NST>
__device__ __forceinline__ void gram_tiles(const PruneArgs& a, const uint32_t* Ci, const int32_t* RA,
                                           const int32_t* Tb, uint16_t* Pm16, uint32_t Mp) {
  constexpr bool U8 = sizeof(T) == 1 && T(-1) > T(0);
  constexpr uint32_t kFlip = U8 ? 0x80808080u : 0u;
  constexpr uint32_t kRing = 4;
  constexpr int32_t K = METRIC == L2 ? 2 : 1;
  const T* base = (const T*)a.base;
  const uint32_t lane = lane_id();
  const uint32_t g4 = lane >> 4, c16 = lane & 15u;
  const uint32_t nb = (Mp + 15u) / 16u, npairs = nb * (nb + 1u) / 2u;
  auto load_block = [&](uint32_t J, uint4* v) {
    const uint32_t Jc = J < nb ? J : nb - 1u;
    const uint32_t rb = 16u * Jc + c16 < Mp ? 16u * Jc + c16 : Mp - 1u;
    const uint4* rc = (const uint4*)(base + (uint64_t)Ci[rb] * a.ad) + g4 * NST;
#pragma unroll
    for (uint32_t st = 0; st < NST; st++) v[st] = rc[st];
  };
  auto frag = [&](const uint4& v) {
    return v4i_t{(int)(v.x ^ kFlip), (int)(v.y ^ kFlip), (int)(v.z ^ kFlip), (int)(v.w ^ kFlip)};
  };
  auto next = [&](uint32_t& I, uint32_t& J) {
    if (++J >= nb) J = ++I;
  };
  uint4 ring[kRing][NST];
  uint32_t Il = 0, Jl = 0;  // the next pair to load
#pragma unroll
  for (uint32_t s = 0; s < kRing; s++) {
    load_block(Jl, ring[s]);
    next(Il, Jl);
  }
  uint32_t I = 0, J = 0;  // the next pair to multiply
  v4i_t af[NST];
  int32_t ra[4];
  for (uint32_t k0 = 0; k0 < npairs; k0 += kRing) {
#pragma unroll
    for (uint32_t s = 0; s < kRing; s++) {
      if (k0 + s < npairs) {
        if (J == I) {
#pragma unroll
          for (uint32_t st = 0; st < NST; st++) af[st] = frag(ring[s][st]);
#pragma unroll
          for (uint32_t r = 0; r < 4; r++) ra[r] = RA[16u * I + 4u * g4 + r];
        }
        v4i_t acc = v4i_t{0, 0, 0, 0};
#pragma unroll
        for (uint32_t st = 0; st < NST; st++)
          acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[st], frag(ring[s][st]), acc, 0, 0, 0);
        const int32_t tb = Tb[16u * J + c16];
        uint64_t m[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) m[r] = ballot(ra[r] - K * acc[r] <= tb);
        if (lane < 16u) {  // lane L writes row 16 I + L: bits 16 (L / 4) .. of m[L % 4]
          const uint32_t rr = lane & 3u;
          const uint64_t mm = rr == 0u ? m[0] : (rr == 1u ? m[1] : (rr == 2u ? m[2] : m[3]));
          Pm16[(16u * I + lane) * (kGramMax / 16u) + J] = (uint16_t)(mm >> (16u * (lane >> 2)));
        }
        next(I, J);
      }
      load_block(Jl, ring[s]);
      next(Il, Jl);
    }
  }
}

template <typename T, int METRIC>
__device__ __forceinline__ void prune_task_gram(const PruneArgs& a, uint32_t t, uint32_t p, uint32_t* grow,
                                                uint64_t a0, uint64_t alen, uint32_t gdeg, T* pv, uint32_t* sel,
                                                float* Cd, uint32_t* Ci, unsigned char* gram) {
  using D = Dist<T, METRIC>;
  constexpr bool U8 = sizeof(T) == 1 && T(-1) > T(0);
  constexpr uint32_t kSlots = kGramMax / kWave;
  const uint32_t lane = lane_id();
  const uint32_t row_bytes = a.ad * (uint32_t)sizeof(T);
  const uint32_t nchunks = row_bytes >> 4;
  const T* base = (const T*)a.base;
  const uint32_t M = (uint32_t)(alen + gdeg), M4 = (M + 3u) & ~3u;
  int32_t* Nrm = (int32_t*)gram;  // norms, then the row terms RA
  int32_t* Sm = Nrm + kGramMax;
  uint32_t* Pm = (uint32_t*)(Sm + kGramMax);
  int32_t* Tb = (int32_t*)Cd;  // the column thresholds replace the distances
  // until the ranks are known, the bit matrix's space holds each candidate's
  // key, norm, sum and a tie counter by list position
  uint64_t* Key = (uint64_t*)Pm;
  int32_t* NrmL = (int32_t*)(Key + kGramMax);
  int32_t* SmL = NrmL + kGramMax;
  uint32_t* Tie = (uint32_t*)(SmL + kGramMax);
  stage_row(pv, base + (uint64_t)p * a.ad, row_bytes);
  for (uint32_t j = lane; j < M; j += kWave) Ci[j] = j < alen ? a.a_ids[a0 + j] : grow[1 + (j - (uint32_t)alen)];
  wave_sync();
  const typename D::Q ps = D::prep(pv, nchunks);
  // one pass over each candidate row: distance to p (unless given), norm and
  // (uint8) byte sum; candidates equal to p get (inf, kPruned), last in the
  // order, never taken
  for (uint32_t j = lane; j < M4; j += kWave) {
    float dj = __builtin_inff();
    uint32_t ij = kPruned;
    int32_t nn = 0, ss = 0;
    uint64_t key = ~0ull;  // padding: after every candidate
    if (j < M) {
      ij = Ci[j];
      if (ij == p) {
        ij = kPruned;
      } else {
        const float dc = int_row_terms<T, METRIC>((const uint4*)(base + (uint64_t)ij * a.ad), pv, ps, nchunks, nn, ss);
        dj = (j < alen && a.a_dists) ? a.a_dists[a0 + j] : dc;
      }
      key = prune_key(dj, ij);
      Cd[j] = dj;
      Ci[j] = ij;
    }
    Key[j] = key;
    NrmL[j] = nn;
    SmL[j] = ss;
    Tie[j] = 0u;
  }
  wave_sync();
  // sorted position = number of smaller keys; copies of one candidate (the
  // same id, row and distance, so interchangeable) take consecutive positions
  // in any order
  uint64_t ku[kSlots];
  uint32_t rk[kSlots];
#pragma unroll
  for (uint32_t u = 0; u < kSlots; u++) {
    const uint32_t j = lane + u * kWave;
    ku[u] = j < M ? Key[j] : ~0ull;
    rk[u] = 0;
  }
  const uint32_t nslots = (M + kWave - 1u) / kWave;
  if (nslots <= 1) count_below_keys<1>(Key, M4, ku, rk);
  else if (nslots == 2) count_below_keys<2>(Key, M4, ku, rk);
  else if (nslots == 3) count_below_keys<3>(Key, M4, ku, rk);
  else count_below_keys<4>(Key, M4, ku, rk);
  float dv[kSlots];
  uint32_t iv[kSlots];
  int32_t nv[kSlots], sv[kSlots];
#pragma unroll
  for (uint32_t u = 0; u < kSlots; u++) {
    const uint32_t j = lane + u * kWave;
    if (j < M) {
      rk[u] += atomicAdd(&Tie[rk[u]], 1u);
      dv[u] = Cd[j];
      iv[u] = Ci[j];
      nv[u] = NrmL[j];
      sv[u] = SmL[j];
    }
  }
  wave_sync();
#pragma unroll
  for (uint32_t u = 0; u < kSlots; u++) {
    if (lane + u * kWave < M) {
      Cd[rk[u]] = dv[u];
      Ci[rk[u]] = iv[u];
      Nrm[rk[u]] = nv[u];
      Sm[rk[u]] = sv[u];
    }
  }
  wave_sync();
  uint32_t Mp = 0;  // candidates other than p
  for (uint32_t b0 = 0; b0 < M; b0 += kWave) Mp += popc(ballot(b0 + lane < M && Ci[b0 + lane] != kPruned));
  // row terms RA and column thresholds Tb (d(a, b) = RA[a] - K acc + CB[b],
  // acc the dot product of the offset rows; Tb = threshold(d(p, b)) - CB[b]);
  // columns past the candidates never prune
  const uint32_t nb = (Mp + 15u) / 16u;
  for (uint32_t j = lane; j < 16u * nb; j += kWave) {
    int32_t ra = 0, tb = INT32_MIN;
    if (j < Mp) {
      const int32_t nn = Nrm[j], ss = Sm[j];
      const int32_t off = U8 ? 128 * ss - 16384 * (int32_t)a.ad : 0;  // dot = acc + off[a] + off[b] + 16384 ad
      int32_t cb;
      if (METRIC == L2) {  // d = na + nb - 2 dot
        ra = nn - 2 * off - (U8 ? 2 * 16384 * (int32_t)a.ad : 0);
        cb = nn - 2 * off;
      } else {  // d = -dot
        ra = -off - (U8 ? 16384 * (int32_t)a.ad : 0);
        cb = -off;
      }
      tb = prune_threshold(a.alpha, (double)Cd[j]) - cb;
    }
    Nrm[j] = ra;
    Tb[j] = tb;
  }
  wave_sync();
  if (Mp > 0) {
    const uint32_t nst = a.ad / 64u;
    uint16_t* Pm16 = (uint16_t*)Pm;
    if (nst == 1) gram_tiles<T, METRIC, 1>(a, Ci, Nrm, Tb, Pm16, Mp);
    else if (nst == 2) gram_tiles<T, METRIC, 2>(a, Ci, Nrm, Tb, Pm16, Mp);
    else if (nst == 3) gram_tiles<T, METRIC, 3>(a, Ci, Nrm, Tb, Pm16, Mp);
    else gram_tiles<T, METRIC, 4>(a, Ci, Nrm, Tb, Pm16, Mp);
  }
  wave_sync();

  // greedy sweep over the bit rows (index.h:84-106): lane w < kGramWords holds
  // live word w; a step selects the lowest live position and drops every
  // position its row marks.  Selections come in increasing position order:
  // they are kept as a bit mask and written out in that order afterwards.
  uint32_t live = 0, chosen = 0;
  if (lane < kGramWords) {
    const uint32_t lo = lane * 32u;
    live = Mp >= lo + 32u ? 0xFFFFFFFFu : (Mp > lo ? (1u << (Mp - lo)) - 1u : 0u);
  }
  uint32_t nsel = 0;
  for (;;) {
    const uint64_t nz = ballot(lane < kGramWords && live != 0u);
    if (!nz) break;
    const uint32_t w = (uint32_t)first_lane(nz);
    const uint32_t bit = 1u << (uint32_t)__builtin_ctz(readlane_u32(live, (int)w));
    const uint32_t at = w * 32u + (uint32_t)__builtin_ctz(bit);
    if (lane == w) chosen |= bit;
    if (++nsel >= a.Rsel) break;
    if (lane < kGramWords) {
      live &= ~Pm[at * kGramWords + lane];
      if (lane == w) live &= ~bit;
    }
  }
  uint32_t cw[kGramWords], before[kGramWords];
  uint32_t run = 0;
#pragma unroll
  for (uint32_t w = 0; w < kGramWords; w++) {
    cw[w] = readlane_u32(chosen, (int)w);
    before[w] = run;
    run += popc(cw[w]);
  }
#pragma unroll
  for (uint32_t u = 0; u < kSlots; u++) {
    const uint32_t j = lane + u * kWave;
    const uint32_t word = lane < 32u ? cw[2u * u] : cw[2u * u + 1u];
    const uint32_t pre = lane < 32u ? before[2u * u] : before[2u * u + 1u];
    if (j < Mp && ((word >> (j & 31u)) & 1u)) sel[pre + popc(word & ((1u << (j & 31u)) - 1u))] = Ci[j];
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

template <typename T, int METRIC>
__global__ __launch_bounds__(256) void k_prune(PruneArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t lane = lane_id();
  const uint32_t wave = bcast0(threadIdx.x >> 6);  // scalar: per-wave pointers stay in SGPRs
  unsigned char* my = smem + wave * a.lds_per_wave;
  T* pv = (T*)my;                                   // p's row
  T* sv = (T*)(my + ((a.ad * sizeof(T) + 15u) & ~15u));  // p*'s row
  uint32_t* sel = (uint32_t*)(my + 2u * ((a.ad * sizeof(T) + 15u) & ~15u));
  float* Cd = (float*)(my + a.off_cand);
  uint32_t* Ci = (uint32_t*)(Cd + a.ccap);
  uint16_t* Cs = (uint16_t*)(my + a.off_cand + 8u * a.ccap);
  T* rows = (T*)(my + a.off_rows);
  const uint32_t cap = a.ccap;
  uint64_t ntasks = a.np;
  if (a.np_dev && a.np_dev[0] < ntasks) ntasks = a.np_dev[0];

  for (;;) {
    const uint32_t t = fetch_task(a.work);
    if (t >= ntasks) break;
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
    } else if (M64 > cap) {
      if (lane == 0) {
        const uint32_t s = atomicAdd(a.nbig, 1u);
        if (a.big_off) {
          const uint64_t need = M64;  // the task's candidates, nothing more
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
    } else if (sizeof(T) == 1 && M64 <= a.gram_max && a.alpha > 0.0) {
      prune_task_gram<T, METRIC>(a, t, p, grow, a0, alen, gdeg, pv, sel, Cd, Ci, my + a.off_gram);
    } else if (M64 <= a.rrows) {
      prune_task<T, METRIC, true>(a, t, p, grow, a0, alen, gdeg, pv, sv, sel, Cd, Ci, Cs, rows);
    } else if (sizeof(T) == 4 && a.nat_split > 1u) {
      if constexpr (sizeof(T) == 4) prune_task<T, METRIC, false, true>(a, t, p, grow, a0, alen, gdeg, pv, sv, sel, Cd, Ci, Cs, rows);
    } else {
      prune_task<T, METRIC, false>(a, t, p, grow, a0, alen, gdeg, pv, sv, sel, Cd, Ci, Cs, rows);
    }
    wave_sync();
  }
}

// The queued (hub) tasks, one per workgroup: all waves of the block share a
// task.  Its candidates and their distances to p go to global scratch (a
// slot of gcap entries per block, or the task's slice of the bump pool), and
// the greedy sweep (index.h:76-106) then runs over windows of the list in
// (distance, id) order: a window is every remaining candidate whose distance
// key is at most a bound chosen so that at most kWin of them qualify (a
// radix select over the keys), gathered into LDS.  Inside a window the sweep
// is the plain one: take the live minimum by (distance, id), prune the
// window's other live candidates against it, compact the survivors.  Every
// candidate left in the list comes after the whole window in sorted order,
// so when a window runs out before Rsel selections the rest of the list is
// pruned against the window's selections and compacted, and the next window
// is taken from what remains.  A hub list of 10^5 candidates is swept over
// its first few thousand only, instead of every greedy step reading all of
// it.  A list of at most kWin candidates, and one where a single distance
// value has more than kWin candidates, is swept in place in scratch.  One
// thread per candidate row (spreading float32 rows over lanes measured
// slower here: a step's rounds multiply with the lanes per row).  The same
// arithmetic as prune_task, so the selections are identical.
constexpr uint32_t kWin = 2048, kHistBins = 2048;

// The ordered key of a distance: the bits sign-folded (zero taken as +0.0),
// so the unsigned order of keys is the float order.
__device__ __forceinline__ uint32_t dist_key(float d) {
  uint32_t b = __float_as_uint(d);
  if ((b << 1) == 0u) b = 0u;
  return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}

template <typename T, int METRIC>
__global__ __launch_bounds__(1024) void k_prune_big(PruneArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using D = Dist<T, METRIC>;
  const uint32_t lane = lane_id();
  const uint32_t wave = bcast0(threadIdx.x >> 6);
  const uint32_t tid = threadIdx.x, nth = blockDim.x, W = nth >> 6;
  const uint32_t row_bytes = a.ad * (uint32_t)sizeof(T);
  const uint32_t nchunks = row_bytes >> 4;
  const uint32_t rowa = (row_bytes + 15u) & ~15u;
  T* pv = (T*)smem;                     // p's row
  T* sv = (T*)(smem + rowa);            // p*'s row
  uint32_t* sel = (uint32_t*)(smem + 2u * rowa);
  uint32_t* wcnt = (uint32_t*)(smem + 2u * rowa + ((a.Rsel * 4u + 15u) & ~15u));  // per wave
  float* wd = (float*)(wcnt + kMaxBigWaves);
  uint32_t* wi = (uint32_t*)(wd + kMaxBigWaves);
  uint32_t* wp = wi + kMaxBigWaves;
  uint32_t* task = wp + kMaxBigWaves;   // 4 words: task, selected bin, residual rank
  float* Wd = (float*)(task + 4);       // the window (LDS)
  uint32_t* Wi = (uint32_t*)(Wd + kWin);
  uint32_t* hist = Wi + kWin;
  const T* base = (const T*)a.base;
  const uint64_t ntasks = a.nbig[0];

  float bd = __builtin_inff();
  uint32_t bi = kPruned, bp = kPruned;  // this lane's smallest (distance, id, position)
  auto offer = [&](float d, uint32_t i, uint32_t pos) {
    if (d < bd || (d == bd && (i < bi || (i == bi && pos < bp)))) { bd = d; bi = i; bp = pos; }
  };
  // the block's smallest (distance, id, position), on every lane
  auto block_min = [&]() {
    for (int o = 1; o < (int)kWave; o <<= 1) {
      const float od = __shfl_xor(bd, o);
      const uint32_t oi = __shfl_xor(bi, o), op = __shfl_xor(bp, o);
      offer(od, oi, op);
    }
    if (lane == 0) { wd[wave] = bd; wi[wave] = bi; wp[wave] = bp; }
    __syncthreads();
    for (uint32_t w = 0; w < W; w++) offer(wd[w], wi[w], wp[w]);
    bp = bcast0(bp);
    bi = bcast0(bi);
    __syncthreads();  // the slots are read before any later write
  };
  // one compaction round: threads with keep hold (c, cd); survivors go to
  // Ci / Cd from position `out` on, in thread order; returns the round's
  // survivors.  Every entry of the round was read before the barrier, and a
  // survivor's new position is at most its old one (in place) or the
  // destination is another array, so no unread entry is overwritten.
  auto compact = [&](float* Cd, uint32_t* Ci, bool keep, uint32_t c, float cd, uint32_t out) {
    const uint64_t km = ballot(keep);
    if (lane == 0) wcnt[wave] = popc(km);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < W; w++) {
      const uint32_t n = wcnt[w];
      before += w < wave ? n : 0u;
      total += n;
    }
    if (keep) {
      const uint32_t o = out + before + mbcnt(km);
      Ci[o] = c;
      Cd[o] = cd;
      offer(cd, c, o);
    }
    __syncthreads();  // stores and counts done before the next round
    return total;
  };
  // exclusive prefix of v over the block's threads; total on every thread
  auto block_scan = [&](uint32_t v, uint32_t& total) {
    uint32_t x = v;
    for (int o = 1; o < (int)kWave; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if ((int)lane >= o) x += y;
    }
    if (lane == kWave - 1u) wcnt[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < W; w++) {
      const uint32_t n = wcnt[w];
      before += w < wave ? n : 0u;
      total += n;
    }
    __syncthreads();
    return before + x - v;
  };

  for (;;) {
    if (tid == 0) task[0] = atomicAdd(a.work, 1u);
    __syncthreads();
    const uint32_t tq = bcast0(task[0]);
    __syncthreads();
    if (tq >= ntasks) break;
    const uint32_t t = a.big_tasks[tq];
    if (t == kNoTask) continue;
    const uint64_t off = a.big_off ? a.big_off[tq] : (uint64_t)blockIdx.x * a.gcap;
    float* Cd = a.gscratch_d + off;
    uint32_t* Ci = a.gscratch_i + off;
    const uint32_t p = a.pts[t];
    uint32_t* grow = a.graph + (uint64_t)p * a.R1;
    uint64_t a0, alen;
    if (a.a_off) { a0 = a.a_off[t]; alen = a.a_off[t + 1] - a0; }
    else { a0 = (uint64_t)t * a.a_stride; alen = a.a_len[t]; }
    const uint32_t deg = grow[0];
    const uint32_t gdeg = a.add ? deg : 0u;
    const uint32_t M = (uint32_t)(alen + gdeg);

    // ---- candidates: A (given or computed distances) then G[p]; copies of
    // p dropped
    if (wave == 0) stage_row(pv, base + (uint64_t)p * a.ad, row_bytes);
    __syncthreads();
    const typename D::Q ps = D::prep(pv, nchunks);
    bd = __builtin_inff();
    bi = bp = kPruned;
    uint32_t live = 0;
    for (uint32_t b0 = 0; b0 < M; b0 += nth) {
      const uint32_t j = b0 + tid;
      uint32_t c = 0;
      float cd = 0.f;
      if (j < M) {
        c = j < alen ? a.a_ids[a0 + j] : grow[1 + (j - (uint32_t)alen)];
        cd = (j < alen && a.a_dists) ? a.a_dists[a0 + j] : D::dist(base + (uint64_t)c * a.ad, pv, ps, a.d, nchunks);
      }
      live += compact(Cd, Ci, j < M && c != p, c, cd, live);
    }

    uint32_t nsel = 0;
    for (bool done = false; !done;) {
      // ---- the next window's bound: one below the (kWin + 1)-th smallest
      // remaining key (a radix select over 11 + 11 + 10 bits); at most kWin
      // remaining are swept in place
      uint32_t tau = 0xFFFFFFFFu, rank = kWin + 1u, prefix = 0, pmask = 0;
      bool last = live <= kWin, in_place = last;
      for (uint32_t pass = 0; pass < 3 && !last; pass++) {
        const uint32_t shift = pass == 0 ? 21u : (pass == 1 ? 10u : 0u);
        const uint32_t nbins = pass == 2 ? 1024u : 2048u;
        for (uint32_t i = tid; i < nbins; i += nth) hist[i] = 0u;
        __syncthreads();
        for (uint32_t j = tid; j < live; j += nth) {
          const uint32_t k = dist_key(Cd[j]);
          if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & (nbins - 1u)], 1u);
        }
        __syncthreads();
        const uint32_t per = nbins / nth ? nbins / nth : 1u;  // bins per thread
        const uint32_t b0 = tid * per;
        uint32_t own = 0;
        for (uint32_t i = 0; i < per; i++) own += b0 + i < nbins ? hist[b0 + i] : 0u;
        uint32_t total = 0;
        const uint32_t excl = block_scan(own, total);
        if (excl < rank && rank <= excl + own) {  // this thread's bins hold the rank-th key
          uint32_t cum = excl, b = b0;
          while (cum + hist[b] < rank) cum += hist[b++];
          task[1] = b;
          task[2] = rank - cum;
        }
        __syncthreads();
        prefix |= bcast0(task[1]) << shift;
        pmask |= (nbins - 1u) << shift;
        rank = bcast0(task[2]);
        __syncthreads();
      }
      if (!last) {
        // kWin + 1 - rank remaining keys lie below the selected one
        if (rank == kWin + 1u) in_place = last = true;  // ties beyond kWin: sweep the rest in place
        else tau = prefix - 1u;
      }
      // ---- gather the window (in place: the list as it stands, whose last
      // compaction offered every entry)
      float* Gd = in_place ? Cd : Wd;
      uint32_t* Gi = in_place ? Ci : Wi;
      uint32_t wl = live;
      if (!in_place) {
        bd = __builtin_inff();
        bi = bp = kPruned;
        wl = 0;
        for (uint32_t b0 = 0; b0 < live; b0 += nth) {
          const uint32_t j = b0 + tid;
          bool keep = false;
          uint32_t c = 0;
          float cd = 0.f;
          if (j < live) {
            c = Ci[j];
            cd = Cd[j];
            keep = dist_key(cd) <= tau;
          }
          wl += compact(Gd, Gi, keep, c, cd, wl);
        }
      }
      block_min();
      const uint32_t sel0 = nsel;

      // ---- the sweep inside the window
      while (bp != kPruned) {
        const uint32_t pstar = bi, at = bp;
        if (tid == 0) sel[nsel] = pstar;
        nsel++;
        if (nsel >= a.Rsel) break;  // the reference's remaining prunes cannot change the result
        if (wave == 0) stage_row(sv, base + (uint64_t)pstar * a.ad, row_bytes);
        __syncthreads();
        const typename D::Q ss = D::prep(sv, nchunks);
        bd = __builtin_inff();
        bi = bp = kPruned;
        uint32_t kept = 0;
        for (uint32_t b0 = 0; b0 < wl; b0 += nth) {
          const uint32_t j = b0 + tid;
          bool keep = false;
          uint32_t c = 0;
          float cd = 0.f;
          if (j < wl && j != at) {
            c = Gi[j];
            cd = Gd[j];
            const float dsp = D::dist(base + (uint64_t)c * a.ad, sv, ss, a.d, nchunks);
            keep = !(a.alpha * (double)dsp <= (double)cd);
          }
          kept += compact(Gd, Gi, keep, c, cd, kept);
        }
        wl = kept;
        block_min();  // its barriers also order the reads of sv before the next stage
      }
      done = last || nsel >= a.Rsel;
      if (!done) {
        // ---- the rest of the list (keys above the window's bound) pruned
        // against the window's selections, one compacting pass each, so a
        // candidate pruned early is not tested again
        __syncthreads();  // sel written
        bd = __builtin_inff();
        bi = bp = kPruned;
        uint32_t kept = 0;
        for (uint32_t b0 = 0; b0 < live; b0 += nth) {  // the window's entries leave
          const uint32_t j = b0 + tid;
          bool keep = false;
          uint32_t c = 0;
          float cd = 0.f;
          if (j < live) {
            c = Ci[j];
            cd = Cd[j];
            keep = dist_key(cd) > tau;
          }
          kept += compact(Cd, Ci, keep, c, cd, kept);
        }
        live = kept;
        for (uint32_t i = sel0; i < nsel && live > 0; i++) {
          if (wave == 0) stage_row(sv, base + (uint64_t)sel[i] * a.ad, row_bytes);
          __syncthreads();
          const typename D::Q ss = D::prep(sv, nchunks);
          bd = __builtin_inff();
          bi = bp = kPruned;
          kept = 0;
          for (uint32_t b0 = 0; b0 < live; b0 += nth) {
            const uint32_t j = b0 + tid;
            bool keep = false;
            uint32_t c = 0;
            float cd = 0.f;
            if (j < live) {
              c = Ci[j];
              cd = Cd[j];
              const float dsc = D::dist(base + (uint64_t)c * a.ad, sv, ss, a.d, nchunks);
              keep = !(a.alpha * (double)dsc <= (double)cd);
            }
            kept += compact(Cd, Ci, keep, c, cd, kept);
          }
          live = kept;
        }
      }
      __syncthreads();  // sel and the window are read before the next gather writes
    }
    __syncthreads();
    if (a.write_graph) {
      for (uint32_t j = tid; j < nsel; j += nth) grow[1 + j] = sel[j];
      __syncthreads();
      if (tid == 0) grow[0] = nsel;
    } else {
      for (uint32_t j = tid; j < nsel; j += nth) a.out_ids[(uint64_t)t * a.Rsel + j] = sel[j];
      if (tid == 0) a.out_len[t] = nsel;
    }
    __syncthreads();
  }
}

uint32_t prune_big_lds_bytes(uint32_t row_bytes, uint32_t Rsel) {
  return 2u * ((row_bytes + 15u) & ~15u) + ((Rsel * 4u + 15u) & ~15u) + 4u * kMaxBigWaves * 4u + 16u +
         8u * kWin + 4u * kHistBins;
}

template <typename T, int M>
static void set_attrs_t(int lds) {
  (void)hipFuncSetAttribute((const void*)&k_prune<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void*)&k_prune_big<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

uint32_t prune_gram_lds_bytes() { return gram_lds_bytes(); }

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
    hipLaunchKernelGGL((k_prune_big<T, M>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s, a);
  else
    hipLaunchKernelGGL((k_prune<T, M>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s, a);
  return hipGetLastError();
}

hipError_t launch_prune(int dtype, int metric, const PruneArgs& a, const Launch& l, hipStream_t s) {
  if (dtype == U8) return metric == L2 ? launch_t<uint8_t, L2>(a, l, s) : launch_t<uint8_t, MIPS>(a, l, s);
  if (dtype == I8) return metric == L2 ? launch_t<int8_t, L2>(a, l, s) : launch_t<int8_t, MIPS>(a, l, s);
  return metric == L2 ? launch_t<float, L2>(a, l, s) : launch_t<float, MIPS>(a, l, s);
}

}  // namespace ddg
