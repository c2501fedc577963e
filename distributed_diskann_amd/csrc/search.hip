// Batched greedy beam search (beamSearch.h:50-184) — one wavefront per
// query / inserted point, all tasks of a batch in one persistent launch fed
// by a device work counter.
//
// Per-wave LDS:  [query row][visited filter 2^bits u32][frontier A][frontier B]
//                [kept ids][candidates][candidate ranks][slot buckets, 1 KB]
// With GTAB the visited filter lives in global memory instead, one slice of
// 2^bits words per resident wave (large beams: the LDS then holds only a few
// KB per wave, and many more waves fit on a CU).
//
// Exactness (ids, distance bits, dist_cmps and |visited| equal to the
// reference's):
//  * the visited filter is the reference's lossy table (same size, same
//    hash64_2 slot, check-then-overwrite in neighbor-list order); lanes that
//    share a slot within one neighbor list resolve as the sequential loop
//    would: a lane sees the id written by the closest lower lane with the same
//    slot, and only the highest such lane writes;
//  * a neighbor equal to p.id() is skipped without touching the table;
//  * candidates are kept iff !(dist >= cutoff), cutoff = (float)INT_MAX while
//    the frontier is not full, else the frontier's last distance;
//  * the frontier / candidate merge is the multiset std::set_union: a
//    candidate equal to k frontier entries is emitted only for its copies
//    beyond k;
//  * the trim is min(beam, size) then, for k > 0 on a metric space, the
//    upper_bound against (0, (float)(cut * d_k));
//  * set_difference(frontier, visited) is tracked with a visited flag on
//    each frontier entry (bit 31 of the id).  A visited vertex that left the
//    frontier can never re-enter it (its distance is >= the cutoff, or above
//    cut * d_k once trimmed, and both only shrink), so the flag is exact for
//    cut >= 1 (the host rejects cut < 1 with k > 0).
#include <type_traits>

#include "device.hpp"
#include "internal.hpp"
#include "trav.hpp"

namespace ddg {

constexpr uint32_t kBuckets = 1024;  // slot buckets for the lane-conflict check
constexpr int kRowsLane = 0, kRowsStage = 1, kRowsSplit = 2;
// split rows whose lanes hold 5, 6 or 7 chunks (kRowsSplit: 8), each its own
// instantiation so the sum's chain walks exactly the lane's share
constexpr int kRowsSplit5 = 5, kRowsSplit6 = 6, kRowsSplit7 = 7;
constexpr bool split_rows(int R) { return R == kRowsSplit || R >= kRowsSplit5; }
constexpr uint32_t split_share(int R) { return R == kRowsSplit ? kSplitChunks : (uint32_t)R; }

// Frontier and candidate entries are (distance bits, id) pairs; bit 31 of a
// frontier id marks a visited entry.  Order: (distance, id) of
// beamSearch.h:59-61, ignoring the flag.
__device__ __forceinline__ bool less_e(uint2 a, uint2 b) {
  return less_di(__uint_as_float(a.x), a.y & kIdMask, __uint_as_float(b.x), b.y & kIdMask);
}
__device__ __forceinline__ bool same_e(uint2 a, uint2 b) {
  return (__uint_as_float(a.x) == __uint_as_float(b.x)) & ((a.y & kIdMask) == (b.y & kIdMask));
}

// A candidate of one expansion is kept in LDS as a 64-bit key, id in the low
// word and the distance's bits in the high word, sign-folded so that the
// unsigned order of the key is the (distance, id) order: one 64-bit compare
// per pair.  The fold is exact and is undone when a candidate enters the
// frontier.  It orders -0.0 below +0.0, which the float order calls equal;
// a search never holds both (L2 distances are sums of squares, +0.0 when
// zero; inner-product distances are negated sums that start at +0.0, so
// every zero is -0.0).
__device__ __forceinline__ uint2 cand_key(float d, uint32_t id) {
  const uint32_t b = __float_as_uint(d);
  return make_uint2(id, b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u));
}
__device__ __forceinline__ uint2 cand_entry(uint2 k) {  // back to (distance bits, id)
  const uint32_t o = k.y;
  return make_uint2(o ^ ((o & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu), k.x);
}
__device__ __forceinline__ uint64_t key64(uint2 k) { return ((uint64_t)k.y << 32) | k.x; }

// Number of entries of the sorted list F[0, n) strictly below c: binary
// lifting over powers of two, the same number of steps on every lane (n is
// wave-uniform).
__device__ __forceinline__ uint32_t count_below(const uint2* F, uint32_t n, uint2 c) {
  uint32_t lo = 0;
  for (uint32_t step = n ? 1u << (31 - __builtin_clz(n)) : 0u; step; step >>= 1) {
    const uint32_t probe = lo + step;
    const uint2 f = F[(probe <= n ? probe : n) - 1u];
    lo = ((probe <= n) & less_e(f, c)) ? probe : lo;
  }
  return lo;
}

// count_below for n < 256: eight fixed steps, no loop and no branch.
__device__ __forceinline__ uint32_t count_below_255(const uint2* F, uint32_t n, uint2 c) {
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t step = 128; step; step >>= 1) {
    const uint32_t probe = lo + step;
    const uint2 f = F[(probe <= n ? probe : n) - 1u];
    lo = ((probe <= n) & less_e(f, c)) ? probe : lo;
  }
  return lo;
}

// Number of entries of the ascending list U[0, n) not above x.
__device__ __forceinline__ uint32_t count_not_above(const uint32_t* U, uint32_t n, uint32_t x) {
  uint32_t lo = 0;
  for (uint32_t step = n ? 1u << (31 - __builtin_clz(n)) : 0u; step; step >>= 1) {
    const uint32_t probe = lo + step;
    const uint32_t u = U[(probe <= n ? probe : n) - 1u];
    lo = ((probe <= n) & (u <= x)) ? probe : lo;
  }
  return lo;
}

// Lanes of a prefetched adjacency row whose filter words are loaded ahead:
// those the next expansion's neighbor filter reads (below the row's degree
// and the degree limit).  Words past the degree hold stale ids from earlier
// rounds of the build; their random slots would each cost a line of the
// filter slice for nothing.
__device__ __forceinline__ uint32_t pf_m(const SearchArgs& a, uint32_t deg) {
  const uint32_t m = deg < a.degree_limit ? deg : a.degree_limit;
  return m < a.R ? m : a.R;
}

// STAGE: candidate rows are gathered into LDS before their distances (a
// launch choice for long rows; a separate instantiation, so the other
// kernels do not carry its registers).  The native kernels are compiled for
// at least 5 waves per SIMD (96 VGPRs; the compiler's own choice hoists
// every staged query chunk of a distance).  At 6 waves (80 VGPRs) they
// spill, and every spill reload is a memory load that waits for all loads in
// flight (they retire in order), the prefetched ones included: 10M x 128
// uint8 at Ls=140, 3.37 ms at 5 against 3.60 ms at 6; the 10M builds 9-14 %
// faster.  The BF16 kernels take 5 (96 VGPRs; 1.39 -> 1.23 ms on 100K x 20,
// the same as 4 at 10M); FP8 keeps the compiler's choice (no gain at 5).
// ROWS: how candidate rows reach their distances: kRowsLane (one lane per
// row), kRowsStage (gathered into LDS first) or kRowsSplit (native float32
// rows split over a.nat_split lanes, split_row_dist); each is its own
// instantiation, so no kernel carries another's registers.
template <typename T, int METRIC, int TRAV, bool GTAB, int ROWS>
__global__ __launch_bounds__(256, (TRAV == TRAV_NATIVE ? (split_rows(ROWS) ? 4 : 5) : (TRAV == TRAV_BF16 ? 5 : 1)))
void k_search(SearchArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using P = TravDist<T, METRIC, TRAV>;
  constexpr bool STAGE = ROWS == kRowsStage;
  const uint32_t lane = lane_id();
  // the wave's index in its block, as a scalar: every per-wave pointer below
  // is then wave-uniform (SGPRs), not per-lane
  const uint32_t wv = bcast0(threadIdx.x >> 6);
  unsigned char* my = smem + wv * a.lds_per_wave;
  unsigned char* ql = my;  // staged query
  uint32_t* table = GTAB ? a.gtable + ((uint64_t)blockIdx.x * (blockDim.x >> 6) + wv) * a.gtable_words
                        : (uint32_t*)(my + a.off_table);
  uint2* F0 = (uint2*)(my + a.off_f0);
  uint2* F1 = (uint2*)(my + a.off_f1);
  uint32_t* Kid = (uint32_t*)(my + a.off_keep);  // kept neighbor ids; later the candidates' ub list
  uint8_t* bkt = (uint8_t*)(my + a.off_bkt);
  uint2* Cn = (uint2*)(my + a.off_cand);   // candidates of one expansion, in neighbor order
  uint2* Ax = (uint2*)(my + a.off_aux);    // per candidate: (rank, entries of F not above it)
  uint32_t* Ub = Kid;
  uint32_t* wb = (uint32_t*)(my + a.off_wbits);  // written-slot bitmap (GTAB with wbits)
  // the bitmap is used by the native kernels only (the host sets wbits
  // only for them); the reduced-precision kernels compile without it
  const bool WB = GTAB && TRAV == TRAV_NATIVE && a.wbits != 0u;
  // Epoch-tagged global filter (no bitmap, a.gepoch set): the slice stores
  // id | epoch << id_bits, with a per-slice epoch advanced once per query,
  // so a query starts without clearing its slice: a word an earlier query
  // stored carries another epoch and equals no tagged id of this one, as an
  // empty slot would.  The slice is cleared only when the epoch wraps, and
  // its epoch is kept in a.gepoch across launches.
  const bool EP = GTAB && !WB && a.gepoch != nullptr;
  const uint32_t slice_ix = blockIdx.x * (blockDim.x >> 6) + wv;
  uint32_t epoch = EP ? bcast0(a.gepoch[slice_ix]) : 0u;

  const uint64_t R1 = (uint64_t)a.R + 1;
  const T* base = (const T*)a.base;

  for (;;) {
    const uint32_t t = fetch_task(a.work);
    if (t >= a.nq) break;
    const uint32_t q = a.order ? a.order[t] : t;  // the query this task serves

    // a per-query beam is clamped to [1, beam]: the frontier and the filter
    // were sized for the largest beam the host declared
    uint32_t beam = a.beam;
    if (a.per_query_beam) {
      const uint32_t bq = a.per_query_beam[q];
      beam = bq < 1u ? 1u : (bq < a.beam ? bq : a.beam);
    }
    const uint32_t mask = (1u << filter_bits(beam)) - 1u;
    uint32_t self_id;  // ids and query indices are below 2^32
    const T* qsrc;
    if (a.build_mode) {
      uint32_t pid = a.qids[q];
      self_id = pid;
      qsrc = base + (uint64_t)pid * a.ad;
    } else {
      self_id = a.qids ? a.qids[q] : q;
      qsrc = (const T*)a.queries + (uint64_t)q * a.ad;
    }
    stage_traversal_query<T, METRIC, TRAV>(a, ql, qsrc);
    if constexpr (GTAB) {
      // the slice's earlier stores complete first; then either the bitmap of
      // written slots is cleared (the slice is only read where a bit is set,
      // i.e. where this query stored) or the slice itself is
      wave_sync_global();
      if (WB) {
        for (uint32_t i = lane; i < (mask + 1u) / 32u; i += kWave) wb[i] = 0u;
      } else if (EP && epoch + 1u < a.epoch_end) {
        epoch++;
      } else {
        for (uint32_t i = lane; i < (mask + 1u) / 4u; i += kWave)
          ((uint4*)table)[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        wave_sync_global();
        epoch = 0;
      }
    } else {
      for (uint32_t i = lane; i <= mask; i += kWave) table[i] = 0xFFFFFFFFu;
    }
    wave_sync();
    const typename P::Q qs = P::prep(a, ql);
    const uint32_t etag = EP ? epoch << a.id_bits : 0u;

    // starting frontier, sorted by (distance, id).  A start id outside
    // [0, n) (device memory the host could not check) raises bit 1 of the
    // status word and the query runs with no start: an empty frontier
    uint32_t ns = a.nstarts;
    if (ballot(lane < ns && a.starts[lane] >= a.n)) {
      if (lane == 0) atomicOr(a.overflow, 2u);
      ns = 0;
    }
    {
      uint2 se = make_uint2(0u, 0u);
      if (lane < ns) {
        const uint32_t sid = a.starts[lane];
        se = make_uint2(__float_as_uint(P::dist(a, sid, ql, qs)), sid);
        Cn[lane] = se;
      }
      wave_sync();
      if (lane < ns) {
        uint32_t rank = 0;
        for (uint32_t t = 0; t < ns; t++) {
          const uint2 st = Cn[t];
          rank += (less_e(st, se) || (same_e(st, se) && t < lane)) ? 1u : 0u;
        }
        F0[rank] = se;
      }
      wave_sync();
    }
    uint32_t fsz = ns;
    uint32_t cur = 0;
    uint64_t cmps = ns;
    uint32_t nvis = 0;
    const uint32_t limit = a.limit < 0xFFFFFFFFull ? (uint32_t)a.limit : 0xFFFFFFFFu;
    int pos = fsz > 0 ? 0 : -1;
    bool trimmed = false;  // the frontier has been through one union + trim
    // Adjacency row (degree and first 64 neighbors) of the vertex predicted to
    // be expanded next.  It is requested after the neighbor filter for the
    // first unvisited frontier entry after the current one, and again after
    // the distances when a new candidate beats that entry; the prediction
    // (the smallest unvisited entry of frontier and new candidates) is the
    // vertex the next expansion takes whenever the search continues.  With a
    // global filter the table words of those neighbors are loaded ahead too,
    // after this expansion's table stores: no table store happens between that
    // load and the next expansion's filter, so they are the words the filter
    // would read.
    uint32_t pf_id = 0xFFFFFFFFu, pf_deg = 0, pf_nb = 0, pf_prev = 0xFFFFFFFFu, pf_slot = 0;
    bool pf_table = false;

    while (pos >= 0 && nvis < limit) {
      uint2* FR = cur ? F1 : F0;
      const uint2 ce = FR[pos];
      const float cdist = __uint_as_float(ce.x);
      const uint32_t cid = ce.y & kIdMask;
      const uint32_t* row = a.graph + (uint64_t)cid * R1;
      const bool use_pf_table = GTAB && pf_table && pf_id == cid;
      // the adjacency row is in pf_nb / pf_deg when the prediction holds;
      // otherwise it is loaded into those registers (one round trip: degree
      // and neighbor words together).  No copy is made of them here: a copy
      // would wait for every load in flight, the table words loaded ahead
      // included, instead of the filter's first use waiting for what it needs
      if (pf_id != cid) {
        pf_nb = lane < a.R ? row[1 + lane] : 0u;
        pf_deg = row[0];
      }
      const uint32_t nb0 = pf_nb;
      const uint32_t deg = pf_deg;
      wave_sync();
      if (lane == 0) {
        ((uint32_t*)FR)[2 * pos + 1] = cid | kFlag;
        if (a.vis_ids) {
          if (nvis < a.vcap) {
            a.vis_ids[(uint64_t)q * a.vcap + nvis] = cid;
            a.vis_dists[(uint64_t)q * a.vcap + nvis] = cdist;
          } else {
            a.overflow[0] = 1u;
          }
        }
      }
      nvis++;

      // the first unvisited frontier entry after pos (every entry before pos
      // is visited): the next vertex to expand unless a new candidate beats it
      int ui = -1;
      auto unvisited = [&](uint32_t b0) {
        const uint32_t i = b0 + lane;
        return ballot(i < fsz && !(FR[i].y & kFlag));
      };
      {  // the first 64 entries outside the loop (no loop header to wait at)
        const uint32_t b1 = (uint32_t)pos + 1u;
        const uint64_t um0 = unvisited(b1);
        if (um0) {
          ui = (int)(b1 + (uint32_t)first_lane(um0));
        } else {
          for (uint32_t b0 = b1 + kWave; b0 < fsz; b0 += kWave) {
            const uint64_t um = unvisited(b0);
            if (um) {
              ui = (int)(b0 + (uint32_t)first_lane(um));
              break;
            }
          }
        }
      }
      const uint2 ue = ui >= 0 ? FR[ui] : make_uint2(0u, 0u);
      pf_id = ui >= 0 ? (ue.y & kIdMask) : 0xFFFFFFFFu;
      pf_table = false;

      // ---- neighbor filter (beamSearch.h:125-131)
      const uint32_t m = deg < a.degree_limit ? deg : a.degree_limit;
      uint32_t nk = 0;
      auto filter_batch = [&](uint32_t b0) __attribute__((always_inline)) {
        const uint32_t j = b0 + lane;
        const bool valid = j < m;
        const uint32_t av = valid ? (b0 == 0 ? nb0 : row[1 + j]) : 0u;
        const bool act = valid && (av != self_id);
        const uint32_t tav = av | etag;  // the word this neighbor stores (tagged with the epoch)
        const bool pre = b0 == 0 && use_pf_table;  // slot and table word loaded ahead
        const uint32_t slot = pre ? pf_slot : (uint32_t)hash64_2((uint64_t)av) & mask;
        // table stores of the previous 64 neighbors by other lanes (those of
        // earlier expansions were ordered by the fence before the table
        // words were loaded ahead; every expansion after the first has one)
        if (GTAB && b0 > 0) wave_sync_global();
        uint32_t tv = 0xFFFFFFFFu;  // the table word before this neighbor list
        if (act) {
          if (pre) tv = pf_prev;
          else if (!WB || ((wb[slot >> 5] >> (slot & 31u)) & 1u)) tv = table[slot];
        }
        uint32_t prev = tv;
        bool has_next = false;
        // Lanes that share a slot share a bucket of `bkt`.  Every active lane
        // writes its lane id into its bucket; a lane that reads back another
        // id lost a race, and only the losers and their buckets' winners go
        // through the exact lane-order resolution below (usually nobody).
        const uint32_t bk = slot & (kBuckets - 1u);
        if (act) bkt[bk] = (uint8_t)lane;
        wave_sync();
        const uint32_t owner = act ? (uint32_t)bkt[bk] : lane;
        const uint64_t lost = ballot(owner != lane);
        if (lost) {
          uint64_t conf = lost, it = lost;
          while (it) {
            const int l = first_lane(it);
            it &= it - 1;
            conf |= 1ull << readlane_u32(owner, l);
          }
          it = conf;
          while (it) {
            const int jj = first_lane(it);
            it &= it - 1;
            const uint32_t sj = readlane_u32(slot, jj);
            const uint32_t aj = readlane_u32(tav, jj);
            if (act && sj == slot) {
              if ((uint32_t)jj < lane) prev = aj;
              else if ((uint32_t)jj > lane) has_next = true;
            }
          }
        }
        const bool keep = act && (prev != tav);
        wave_sync();
        if (act && !has_next && tv != tav) {  // the last writer of a slot; skip same-value stores
          table[slot] = tav;
          if (WB) atomicOr(&wb[slot >> 5], 1u << (slot & 31u));
        }
        const uint64_t km = ballot(keep);
        if (keep) Kid[nk + mbcnt(km)] = av;
        nk += popc(km);
        wave_sync();
      };
      // the first 64 neighbors outside the loop: the compiler waits for
      // every pending load at a loop header, and the first batch's table
      // words may still be in flight
      if (m > 0) filter_batch(0);
      for (uint32_t b0 = kWave; b0 < m; b0 += kWave) filter_batch(b0);
      cmps += nk;
      // the predicted vertex's adjacency row, requested with the candidate
      // rows below: the wait for those rows covers it too (requested earlier,
      // the compiler's waits in the filter would stall on it)
      if (pf_id != 0xFFFFFFFFu) {
        const uint32_t* prow = a.graph + (uint64_t)pf_id * R1;
        pf_nb = lane < a.R ? prow[1 + lane] : 0u;
        pf_deg = prow[0];
      }

      // ---- distances and cutoff (beamSearch.h:135-145)
      const float cutoff = (fsz < beam) ? 2147483648.0f : __uint_as_float(FR[fsz - 1].x);
      uint32_t nc = 0;
      if constexpr (STAGE) {
        // Long rows: the wave gathers stage_rows candidate rows into LDS
        // together (all lanes' loads in flight at once), then lane j computes
        // row j's distance from LDS with the same arithmetic.
        const uint32_t C = a.stage_chunks, CS = a.stage_stride;
        uint4* stg = (uint4*)(my + a.off_stage);
        for (uint32_t r0 = 0; r0 < nk; r0 += a.stage_rows) {
          const uint32_t nr = nk - r0 < a.stage_rows ? nk - r0 : a.stage_rows;
          const uint32_t total = nr * C;
          for (uint32_t e0 = 0; e0 < total; e0 += kChunkBatch * kWave) {
            // unpredicated loads (an element beyond the batch re-reads the
            // last one): a conditionally assigned array is kept in scratch
            uint4 xs[kChunkBatch];
#pragma unroll
            for (uint32_t u = 0; u < kChunkBatch; u++) {
              const uint32_t e = e0 + u * kWave + lane < total ? e0 + u * kWave + lane : total - 1u;
              const uint32_t r = e / C, c = e - r * C;
              xs[u] = P::row_ptr(a, Kid[r0 + r])[c];
            }
#pragma unroll
            for (uint32_t u = 0; u < kChunkBatch; u++) {
              const uint32_t e = e0 + u * kWave + lane;
              if (e < total) {
                const uint32_t r = e / C, c = e - r * C;
                stg[r * CS + c] = xs[u];
              }
            }
          }
          wave_sync();
          const bool v = lane < nr;
          uint32_t id = 0;
          float dd = 0.f;
          if (v) {
            id = Kid[r0 + lane];
            dd = P::dist_row(a, stg + lane * CS, ql, qs);
          }
          const bool pass = v && !(dd >= cutoff);
          const uint64_t pm = ballot(pass);
          if (pass) Cn[nc + mbcnt(pm)] = cand_key(dd, id);
          nc += popc(pm);
          wave_sync();  // the next batch overwrites the staged rows
        }
      } else if constexpr (TRAV != TRAV_NATIVE) {
        // a reduced row is split over trav_split lanes (one load batch each)
        const uint32_t S = a.trav_split;
        const uint32_t sub = lane & (S - 1u);
        const uint32_t step = kWave / S;
        auto batch = [&](uint32_t b0) {
          const uint32_t j = b0 + lane / S;
          const bool v = j < nk;
          const uint32_t id = v ? Kid[j] : 0u;
          const float dd = split_dist<T, METRIC, TRAV>(a, id, v, ql, sub);
          const bool pass = v && sub == 0u && !(dd >= cutoff);
          const uint64_t pm = ballot(pass);
          if (pass) Cn[nc + mbcnt(pm)] = cand_key(dd, id);
          nc += popc(pm);
        };
        // the first batch outside the loop: the compiler waits for every
        // pending load at a loop header, which would stall on the prefetch
        if (nk > 0) batch(0);
        for (uint32_t b0 = step; b0 < nk; b0 += step) batch(b0);
      } else if constexpr (split_rows(ROWS)) {
        // long float32 rows: one row over nat_split lanes, all its chunks in
        // flight at once, the running sum handed along the lanes in the
        // reference's order (split_row_dist)
        const uint32_t S = a.nat_split;
        const uint32_t sub = lane & (S - 1u);
        const uint32_t step = kWave / S;
        auto batch = [&](uint32_t b0) {
          const uint32_t j = b0 + lane / S;
          const bool v = j < nk;
          const uint32_t id = Kid[v ? j : b0];  // lanes past the list read a row of the batch
          const float dd = split_row_dist<METRIC, split_share(ROWS)>((const float*)a.base + (uint64_t)id * a.ad,
                                                                     (const float*)ql, a.ad >> 2, S, a.nat_cps, sub);
          const bool pass = v && sub == S - 1u && !(dd >= cutoff);
          const uint64_t pm = ballot(pass);
          if (pass) Cn[nc + mbcnt(pm)] = cand_key(dd, id);
          nc += popc(pm);
        };
        if (nk > 0) batch(0);  // outside the loop, as above
        for (uint32_t b0 = step; b0 < nk; b0 += step) batch(b0);
      } else {
        auto batch = [&](uint32_t b0) {
          const uint32_t j = b0 + lane;
          const bool v = j < nk;
          uint32_t id = 0;
          float dd = 0.f;
          if (v) {
            id = Kid[j];
            dd = P::dist(a, id, ql, qs);
          }
          const bool pass = v && !(dd >= cutoff);
          const uint64_t pm = ballot(pass);
          if (pass) Cn[nc + mbcnt(pm)] = cand_key(dd, id);
          nc += popc(pm);
        };
        if (nk > 0) batch(0);  // outside the loop, as above
        for (uint32_t b0 = kWave; b0 < nk; b0 += kWave) batch(b0);
      }
      wave_sync();

      // ---- order of the candidates (std::sort, beamSearch.h:148): the rank
      // of candidate j counts the candidates below it plus its equal copies at
      // lower indices; the smallest one (rank 0) is the prediction's rival
      bool have_min = false;
      uint2 cmin = make_uint2(0u, 0u);
      for (uint32_t cb = 0; cb < nc; cb += kWave) {
        const uint32_t j = cb + lane;
        const bool v = j < nc;
        uint2 cj = make_uint2(0u, 0u);
        uint32_t lt = 0, eqb = 0;
        if (v) {
          cj = Cn[j];
          const uint64_t kj = key64(cj);
          // four candidates per step, their LDS reads in flight together (the
          // array holds a multiple of 64 entries, so reads past nc stay in
          // it; those entries are not counted)
          for (uint32_t t0 = 0; t0 < nc; t0 += 4) {
            uint2 ct[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) ct[u] = Cn[t0 + u];
#pragma unroll
            for (uint32_t u = 0; u < 4; u++) {
              const uint32_t t = t0 + u;
              const uint64_t kt = key64(ct[u]);
              lt += ((t < nc) & (kt < kj)) ? 1u : 0u;
              eqb += ((t < j) & (kt == kj)) ? 1u : 0u;
            }
          }
          Ax[j].x = lt + eqb;
        }
        const uint64_t zm = ballot(v && lt + eqb == 0u);
        if (zm) {
          const int l = first_lane(zm);
          cmin = cand_entry(make_uint2(readlane_u32(cj.x, l), readlane_u32(cj.y, l)));
          have_min = true;
        }
      }
      // predicted next vertex: the smaller of the first unvisited entry and
      // the smallest candidate; its adjacency row is loaded during the merge
      const bool cand_next = have_min && (ui < 0 || less_e(cmin, ue));
      if (cand_next) {
        pf_id = cmin.y;
        const uint32_t* prow = a.graph + (uint64_t)pf_id * R1;
        pf_nb = lane < a.R ? prow[1 + lane] : 0u;
        pf_deg = prow[0];
      }
      if constexpr (GTAB) {
        // the head's row is the prediction: its table words are loaded
        // during the merge
        if (!cand_next && pf_id != 0xFFFFFFFFu) {
          wave_sync_global();  // this expansion's table stores by other lanes
          pf_slot = (uint32_t)hash64_2((uint64_t)pf_nb) & mask;
          pf_prev = (lane < pf_m(a, pf_deg) && (!WB || ((wb[pf_slot >> 5] >> (pf_slot & 31u)) & 1u)))
                        ? table[pf_slot]
                        : 0xFFFFFFFFu;
          pf_table = true;
        }
      }

      // ---- multiset union (beamSearch.h:151-154): a candidate equal to e
      // frontier entries is emitted only for its copies beyond e.  For each
      // candidate: ub = #frontier entries not above it.
      uint32_t ne = 0;
      for (uint32_t cb = 0; cb < nc; cb += kWave) {
        const uint32_t j = cb + lane;
        const bool v = j < nc;
        bool emit = false;
        if (v) {
          const uint2 kc = Cn[j];
          const uint2 cj = cand_entry(kc);
          const uint32_t rank = Ax[j].x;
          const uint32_t lb = fsz < 256u ? count_below_255(FR, fsz, cj) : count_below(FR, fsz, cj);
          uint32_t eq = 0;
          if (lb < fsz && same_e(FR[lb], cj)) {  // equal frontier entries (rare)
            eq = 1;
            while (lb + eq < fsz && same_e(FR[lb + eq], cj)) eq++;
          }
          // equal copies at lower indices: rank - #candidates strictly below
          uint32_t lt = 0;
          if (eq > 0)
            for (uint32_t t = 0; t < nc; t++) lt += (key64(Cn[t]) < key64(kc)) ? 1u : 0u;
          emit = eq == 0 || rank - lt >= eq;
          Ax[j] = make_uint2(emit ? rank : (rank | kFlag), lb + eq);
        }
        ne += popc(ballot(emit));
      }
      wave_sync();

      uint32_t nsz = fsz;
      uint32_t p_next = 0xFFFFFFFFu;  // the next expansion's position after the merge
      if (ne > 0) {
        // ---- merge F and E, keeping the first min(beam, |F|+|E|): emitted
        // candidate e (its position among the emitted ones: its rank, unless
        // a duplicate of a frontier entry was dropped) goes to e + ub_e,
        // frontier entry i to i + #{emitted e : ub_e <= i}.  Ub[e] = ub_e is
        // ascending in e.
        const uint32_t total = fsz + ne;
        nsz = total < beam ? total : beam;
        uint2* NF = cur ? F0 : F1;
        for (uint32_t cb = 0; cb < nc; cb += kWave) {
          const uint32_t j = cb + lane;
          bool em = false;
          uint32_t e = 0, ub = 0;
          if (j < nc) {
            const uint2 aj = Ax[j];
            em = !(aj.x & kFlag);
            e = aj.x;
            ub = aj.y;
            if (em && ne < nc) {
              e = 0;
              for (uint32_t t = 0; t < nc; t++) {
                const uint32_t rt = Ax[t].x;
                e += (!(rt & kFlag) && rt < aj.x) ? 1u : 0u;
              }
            }
          }
          if (em) {
            Ub[e] = ub;
            if (e + ub < nsz) NF[e + ub] = cand_entry(Cn[j]);
          }
        }
        wave_sync();
        for (uint32_t i = lane; i < fsz; i += kWave) {
          const uint32_t p = i + count_not_above(Ub, ne, i);
          if (p < nsz) NF[p] = FR[i];
        }
        // every emitted candidate is above the first unvisited entry or
        // below it; the smallest emitted one (e = 0) lands at ub_0
        const uint32_t ub0 = Ub[0];
        p_next = (ui >= 0 && (uint32_t)ui < ub0) ? (uint32_t)ui : ub0;
        wave_sync();
        cur ^= 1u;
      } else if (ui >= 0) {
        p_next = (uint32_t)ui;
      }
      if constexpr (GTAB) {
        if (cand_next) {  // a new candidate is the prediction: its row was loaded during the merge
          wave_sync_global();
          pf_slot = (uint32_t)hash64_2((uint64_t)pf_nb) & mask;
          pf_prev = (lane < pf_m(a, pf_deg) && (!WB || ((wb[pf_slot >> 5] >> (pf_slot & 31u)) & 1u)))
                        ? table[pf_slot]
                        : 0xFFFFFFFFu;
          pf_table = true;
        }
      }
      // ---- size limit and cut trim (beamSearch.h:157-167); an unchanged
      // frontier that went through them once stays as it is
      if (ne > 0 || !trimmed) {
        const uint2* F2 = cur ? F1 : F0;
        if (nsz > beam) nsz = beam;
        if (a.k > 0 && nsz > a.k && a.cut_trim) {
          const float thr = (float)(a.cut * (double)__uint_as_float(F2[a.k].x));
          auto over_mask = [&](uint32_t b0) {
            const uint32_t i = b0 + lane;
            bool over = false;
            if (i < nsz) {
              const uint2 e = F2[i];
              const float ed = __uint_as_float(e.x);
              over = (thr < ed) | ((thr == ed) & ((e.y & kIdMask) > 0u));
            }
            return ballot(over);
          };
          // the first 64 entries outside the loop (no loop header to wait at)
          uint32_t cut_at = nsz;
          const uint64_t om0 = over_mask(0);
          if (om0) {
            cut_at = (uint32_t)first_lane(om0);
          } else {
            for (uint32_t b0 = kWave; b0 < nsz; b0 += kWave) {
              const uint64_t om = over_mask(b0);
              if (om) {
                cut_at = b0 + (uint32_t)first_lane(om);
                break;
              }
            }
          }
          nsz = cut_at;
        }
        trimmed = true;
      }
      fsz = nsz;
      // ---- next unvisited frontier entry (set_difference, beamSearch.h:175-178)
      pos = p_next < fsz ? (int)p_next : -1;
    }

    if constexpr (TRAV != TRAV_NATIVE) {
      // FP32 re-rank of the first nr beam entries against the full rows, in
      // the reference's arithmetic, then the first k by (distance, id)
      const uint2* FR = cur ? F1 : F0;
      uint2* XR = cur ? F0 : F1;
      const uint32_t nr = (a.rerank == 0 || a.rerank > fsz) ? fsz : a.rerank;
      for (uint32_t j = lane; j < nr; j += kWave) {
        const uint32_t id = FR[j].y & kIdMask;
        const float dd = Dist<float, METRIC>::dist((const float*)a.base + (uint64_t)id * a.ad, (const float*)ql,
                                                   typename Dist<float, METRIC>::Q{0}, a.d, a.ad >> 2);
        XR[j] = make_uint2(__float_as_uint(dd), id);
      }
      wave_sync();
      for (uint32_t j = lane; j < nr; j += kWave) {
        const uint2 xj = XR[j];
        uint32_t rank = 0;
        for (uint32_t t = 0; t < nr; t++) {
          const uint2 xt = XR[t];
          rank += (less_e(xt, xj) || (same_e(xt, xj) && t < j)) ? 1u : 0u;
        }
        if (rank < a.k) {
          a.out_ids[(uint64_t)q * a.k + rank] = xj.y;
          if (a.out_dists) a.out_dists[(uint64_t)q * a.k + rank] = __uint_as_float(xj.x);
        }
      }
      for (uint32_t j = nr + lane; j < a.k; j += kWave) {
        a.out_ids[(uint64_t)q * a.k + j] = 0xFFFFFFFFu;
        if (a.out_dists) a.out_dists[(uint64_t)q * a.k + j] = __builtin_inff();
      }
    } else if (!a.build_mode) {
      const uint2* FR = cur ? F1 : F0;
      for (uint32_t j = lane; a.out_ids && j < a.k; j += kWave) {
        const bool ok = j < fsz;
        const uint2 e = ok ? FR[j] : make_uint2(0u, 0u);
        a.out_ids[(uint64_t)q * a.k + j] = ok ? (e.y & kIdMask) : 0xFFFFFFFFu;
        if (a.out_dists) a.out_dists[(uint64_t)q * a.k + j] = ok ? __uint_as_float(e.x) : __builtin_inff();
      }
      if (a.fr_ids) {  // beam_search's frontier (beamSearch.h:183)
        for (uint32_t j = lane; j < fsz; j += kWave) {
          const uint2 e = FR[j];
          a.fr_ids[(uint64_t)q * a.fr_stride + j] = e.y & kIdMask;
          a.fr_dists[(uint64_t)q * a.fr_stride + j] = __uint_as_float(e.x);
        }
        if (lane == 0) a.fr_len[q] = fsz;
      }
    }
    if (lane == 0) {
      if (a.dist_cmps) a.dist_cmps[q] = cmps;
      if (a.visited_count) a.visited_count[q] = (uint32_t)nvis;
    }
    wave_sync();
  }
  if (EP && lane_id() == 0) a.gepoch[slice_ix] = epoch;
}

// Distances of query q (wave q) to ids[q*m .. q*m+m), computed exactly as the
// search's traversal computes them; ids >= n give NaN.
template <typename T, int METRIC, int TRAV>
__global__ __launch_bounds__(256) void k_trav_dist(SearchArgs a, const uint32_t* ids, uint32_t m, float* out) {
  extern __shared__ __align__(16) unsigned char smem[];
  using P = TravDist<T, METRIC, TRAV>;
  const uint32_t wv = bcast0(threadIdx.x >> 6);
  unsigned char* ql = smem + wv * a.lds_per_wave;
  const uint64_t q = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  if (q >= a.nq) return;  // whole waves leave together
  stage_traversal_query<T, METRIC, TRAV>(a, ql, (const T*)a.queries + q * a.ad);
  wave_sync();
  const typename P::Q qs = P::prep(a, ql);
  for (uint32_t j = lane_id(); j < m; j += kWave) {
    const uint32_t id = ids[q * m + j];
    out[q * m + j] = id < a.n ? P::dist(a, id, ql, qs) : __builtin_nanf("");
  }
}

// Ordering keys of a launch's queries: the distance of query q to the start
// point in the traversal's own arithmetic, larger first (a query far from
// the start tends to need more expansions), as a 16-bit key that sorts
// ascending: the top half of the complemented order-preserving bits.  One
// wave per query; vals[q] = q.
template <typename T, int METRIC, int TRAV>
__global__ __launch_bounds__(256) void k_start_keys(SearchArgs a, uint32_t* keys, uint32_t* vals) {
  extern __shared__ __align__(16) unsigned char smem[];
  using P = TravDist<T, METRIC, TRAV>;
  const uint32_t wv = bcast0(threadIdx.x >> 6);
  unsigned char* ql = smem + wv * a.lds_per_wave;
  const uint64_t q = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  if (q >= a.nq) return;  // whole waves leave together
  stage_traversal_query<T, METRIC, TRAV>(a, ql, (const T*)a.queries + q * a.ad);
  wave_sync();
  const typename P::Q qs = P::prep(a, ql);
  const uint32_t sid = a.starts[0];
  const float d = sid < a.n ? P::dist(a, sid, ql, qs) : 0.f;
  if (lane_id() == 0) {
    const uint32_t b = __float_as_uint(d);
    const uint32_t ord = b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);  // ascending with d
    keys[q] = (~ord) >> 16;
    vals[q] = (uint32_t)q;
  }
}

template <typename T, int M, int TR>
static hipError_t launch_k(const SearchArgs& a, uint32_t* keys, uint32_t* vals, const Launch& l, hipStream_t s) {
  hipLaunchKernelGGL((k_start_keys<T, M, TR>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s, a,
                     keys, vals);
  return hipGetLastError();
}

hipError_t launch_start_keys(int dtype, int metric, int trav, const SearchArgs& a, uint32_t* keys, uint32_t* vals,
                             const Launch& l, hipStream_t s) {
  if (trav == TRAV_BF16)
    return metric == L2 ? launch_k<float, L2, TRAV_BF16>(a, keys, vals, l, s)
                        : launch_k<float, MIPS, TRAV_BF16>(a, keys, vals, l, s);
  if (trav == TRAV_FP8)
    return metric == L2 ? launch_k<float, L2, TRAV_FP8>(a, keys, vals, l, s)
                        : launch_k<float, MIPS, TRAV_FP8>(a, keys, vals, l, s);
  if (dtype == U8)
    return metric == L2 ? launch_k<uint8_t, L2, TRAV_NATIVE>(a, keys, vals, l, s)
                        : launch_k<uint8_t, MIPS, TRAV_NATIVE>(a, keys, vals, l, s);
  if (dtype == I8)
    return metric == L2 ? launch_k<int8_t, L2, TRAV_NATIVE>(a, keys, vals, l, s)
                        : launch_k<int8_t, MIPS, TRAV_NATIVE>(a, keys, vals, l, s);
  return metric == L2 ? launch_k<float, L2, TRAV_NATIVE>(a, keys, vals, l, s)
                      : launch_k<float, MIPS, TRAV_NATIVE>(a, keys, vals, l, s);
}

template <typename T, int M, int TR, bool G, int RW>
static hipError_t launch_g(const SearchArgs& a, const Launch& l, hipStream_t s) {
  hipLaunchKernelGGL((k_search<T, M, TR, G, RW>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s,
                     a);
  return hipGetLastError();
}

// split rows exist for native float32 only
template <typename T, int TR>
constexpr bool has_split() { return std::is_same<T, float>::value && TR == TRAV_NATIVE; }

template <typename T, int M, int TR>
static hipError_t launch_t(const SearchArgs& a, const Launch& l, hipStream_t s) {
  if (a.stage_rows)
    return a.gtable ? launch_g<T, M, TR, true, kRowsStage>(a, l, s) : launch_g<T, M, TR, false, kRowsStage>(a, l, s);
  if constexpr (has_split<T, TR>()) {
    if (a.nat_split > 1u) {
      switch (a.nat_cps) {
        case 5: return a.gtable ? launch_g<T, M, TR, true, kRowsSplit5>(a, l, s) : launch_g<T, M, TR, false, kRowsSplit5>(a, l, s);
        case 6: return a.gtable ? launch_g<T, M, TR, true, kRowsSplit6>(a, l, s) : launch_g<T, M, TR, false, kRowsSplit6>(a, l, s);
        case 7: return a.gtable ? launch_g<T, M, TR, true, kRowsSplit7>(a, l, s) : launch_g<T, M, TR, false, kRowsSplit7>(a, l, s);
        default: return a.gtable ? launch_g<T, M, TR, true, kRowsSplit>(a, l, s) : launch_g<T, M, TR, false, kRowsSplit>(a, l, s);
      }
    }
  }
  return a.gtable ? launch_g<T, M, TR, true, kRowsLane>(a, l, s) : launch_g<T, M, TR, false, kRowsLane>(a, l, s);
}

template <typename T, int M, int TR>
static hipError_t launch_d(const SearchArgs& a, const uint32_t* ids, uint32_t m, float* out, const Launch& l,
                           hipStream_t s) {
  hipLaunchKernelGGL((k_trav_dist<T, M, TR>), dim3(l.blocks), dim3(l.waves_per_block * kWave), l.lds_bytes, s, a,
                     ids, m, out);
  return hipGetLastError();
}

template <typename T, int M, int TR>
static void set_attrs_t(int lds) {
  const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
  (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsLane>, at, lds);
  (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsLane>, at, lds);
  (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsStage>, at, lds);
  (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsStage>, at, lds);
  if constexpr (has_split<T, TR>()) {
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsSplit>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsSplit>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsSplit5>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsSplit5>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsSplit6>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsSplit6>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, true, kRowsSplit7>, at, lds);
    (void)hipFuncSetAttribute((const void*)&k_search<T, M, TR, false, kRowsSplit7>, at, lds);
  }
  (void)hipFuncSetAttribute((const void*)&k_trav_dist<T, M, TR>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  (void)hipFuncSetAttribute((const void*)&k_start_keys<T, M, TR>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

// Raises the dynamic-LDS limit of every search kernel once (outside any
// stream capture); the launchers then only launch.
void search_set_attributes(int max_lds) {
  set_attrs_t<uint8_t, L2, TRAV_NATIVE>(max_lds);
  set_attrs_t<uint8_t, MIPS, TRAV_NATIVE>(max_lds);
  set_attrs_t<int8_t, L2, TRAV_NATIVE>(max_lds);
  set_attrs_t<int8_t, MIPS, TRAV_NATIVE>(max_lds);
  set_attrs_t<float, L2, TRAV_NATIVE>(max_lds);
  set_attrs_t<float, MIPS, TRAV_NATIVE>(max_lds);
  set_attrs_t<float, L2, TRAV_BF16>(max_lds);
  set_attrs_t<float, MIPS, TRAV_BF16>(max_lds);
  set_attrs_t<float, L2, TRAV_FP8>(max_lds);
  set_attrs_t<float, MIPS, TRAV_FP8>(max_lds);
}

hipError_t launch_search(int dtype, int metric, int trav, const SearchArgs& a, const Launch& l, hipStream_t s) {
  if (trav == TRAV_BF16)
    return metric == L2 ? launch_t<float, L2, TRAV_BF16>(a, l, s) : launch_t<float, MIPS, TRAV_BF16>(a, l, s);
  if (trav == TRAV_FP8)
    return metric == L2 ? launch_t<float, L2, TRAV_FP8>(a, l, s) : launch_t<float, MIPS, TRAV_FP8>(a, l, s);
  if (dtype == U8)
    return metric == L2 ? launch_t<uint8_t, L2, TRAV_NATIVE>(a, l, s) : launch_t<uint8_t, MIPS, TRAV_NATIVE>(a, l, s);
  if (dtype == I8)
    return metric == L2 ? launch_t<int8_t, L2, TRAV_NATIVE>(a, l, s) : launch_t<int8_t, MIPS, TRAV_NATIVE>(a, l, s);
  return metric == L2 ? launch_t<float, L2, TRAV_NATIVE>(a, l, s) : launch_t<float, MIPS, TRAV_NATIVE>(a, l, s);
}

hipError_t launch_trav_dist(int dtype, int metric, int trav, const SearchArgs& a, const uint32_t* ids, uint32_t m,
                            float* out, const Launch& l, hipStream_t s) {
  if (trav == TRAV_BF16)
    return metric == L2 ? launch_d<float, L2, TRAV_BF16>(a, ids, m, out, l, s)
                        : launch_d<float, MIPS, TRAV_BF16>(a, ids, m, out, l, s);
  if (trav == TRAV_FP8)
    return metric == L2 ? launch_d<float, L2, TRAV_FP8>(a, ids, m, out, l, s)
                        : launch_d<float, MIPS, TRAV_FP8>(a, ids, m, out, l, s);
  if (dtype == U8)
    return metric == L2 ? launch_d<uint8_t, L2, TRAV_NATIVE>(a, ids, m, out, l, s)
                        : launch_d<uint8_t, MIPS, TRAV_NATIVE>(a, ids, m, out, l, s);
  if (dtype == I8)
    return metric == L2 ? launch_d<int8_t, L2, TRAV_NATIVE>(a, ids, m, out, l, s)
                        : launch_d<int8_t, MIPS, TRAV_NATIVE>(a, ids, m, out, l, s);
  return metric == L2 ? launch_d<float, L2, TRAV_NATIVE>(a, ids, m, out, l, s)
                      : launch_d<float, MIPS, TRAV_NATIVE>(a, ids, m, out, l, s);
}

}  // namespace ddg
