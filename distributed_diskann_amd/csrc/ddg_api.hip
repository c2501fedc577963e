// C-ABI (include/ddg.h) over the HIP kernels: device memory ownership,
// launch sizing, the batch_insert driver and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ddg.h"
#include "device.hpp"
#include "internal.hpp"

using namespace ddg;

namespace {

constexpr uint32_t kLdsPerCU = 160u * 1024u;

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

struct Scratch {
  Buf b[48];
};

enum BufId {
  B_QUERIES, B_STARTS, B_QBEAM, B_QIDS, B_OUT_IDS, B_OUT_D, B_CMPS, B_VISN, B_VIS_IDS, B_VIS_D, B_FLAGS,
  B_INSERTS, B_NEW_IDS, B_NEW_LEN, B_OFFS, B_KEYS, B_VALS, B_KEYS2, B_VALS2, B_TMP, B_GFLAGS, B_GOFFS, B_GKEYS,
  B_GOFF, B_BIG, B_GSD, B_GSI, B_VC, B_PRUNE_OFF, B_PRUNE_IDS, B_PRUNE_D, B_GT_D, B_GTAB, B_BIGOFF, B_RPTS,
  B_FR_IDS, B_FR_D, B_FR_LEN, B_ROW_IDS, B_ROWS, B_GEPOCH, B_OKEYS, B_OVALS, B_OKEYS2, B_OVALS2, B_OTMP, B_COUNT
};
static_assert(B_COUNT <= 48, "scratch slots");

// Device words of B_FLAGS (u32 index): [0] search work counter, [1] visited
// list overflow, [2..3] the ordering pre-pass's work counter and status, [4] error status (bit 0 degree overflow, bit 1 prune scratch
// exhausted), [8..9] pair count (u64), [10..11] group count (u64), [12]
// largest group, [16] prune work counter, [17] long-list count, [18]
// long-list work counter, [20..21] long-list scratch handed out (u64),
// [24..25] a build round's batch size, 0 when its insert search overflowed
// (u64), [26..27] dist_cmps of the build's insert searches (u64).
constexpr size_t kFlagBytes = 256;
enum FlagWord : uint32_t {
  F_WORK = 0, F_OVERFLOW = 1, F_PREWORK = 2, F_PRESTATUS = 3, F_STATUS = 4, F_PAIRS = 8, F_GROUPS = 10,
  F_MAXG = 12, F_PWORK = 16, F_NBIG = 17,
  F_PWORK2 = 18, F_GUSED = 20, F_ROUNDB = 24, F_CMPS = 26
};

}  // namespace

struct ddg_index {
  int dtype = 0, metric = 0;
  uint64_t n = 0;
  uint32_t d = 0, ad = 0, R = 0, es = 1;
  void* rows = nullptr;
  uint32_t* graph = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // reduced traversal copies of a float32 index, indexed by ddg_trav
  void* trav_rows[3] = {nullptr, nullptr, nullptr};
  uint32_t trav_stride[3] = {0, 0, 0};
  float* trav_scale = nullptr;  // FP8: s_j at [0, d), 1/s_j at [d, 2d)
  // search launch shapes picked by the autotuner, per (traversal, filter bits)
  struct Tuned {
    uint32_t key;
    int gtab;
    uint32_t wpb, bpc, stage;  // stage: rows gathered into LDS per batch (0 = one lane per row)
  };
  std::vector<Tuned> tuned;
  const Tuned* cfg_override = nullptr;  // set while the autotuner times a shape
  uint32_t num_cu = 256;
  unsigned char* noalloc_scratch = nullptr;  // this GPU's copy of `scratch` (null: its address was not found)
  // what the global filter pool holds for epoch-tagged launches: the pool and
  // epoch buffers it was initialised in, its slice size and how many slices
  // are initialised (all 0xFFFFFFFF with epoch 0, or left by earlier
  // epoch-tagged launches of this layout); null after any other use
  const void* ep_pool = nullptr;
  const void* ep_epochs = nullptr;
  uint32_t ep_words = 0;
  uint64_t ep_slices = 0;
  std::string noalloc_err;
  Scratch s;
  std::string err;
  double prof[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

static std::string g_err;

// DDG_TRACE=1 prints each build phase to stderr after its kernels finished
static bool trace_on() {
  static int on = -1;
  if (on < 0) { const char* e = getenv("DDG_TRACE"); on = (e && e[0] == '1') ? 1 : 0; }
  return on == 1;
}
#define TRACE(...)                                                     \
  do {                                                                 \
    if (trace_on()) {                                                  \
      (void)hipStreamSynchronize(idx->stream);                         \
      fprintf(stderr, "[ddg] " __VA_ARGS__);                           \
      fputc('\n', stderr);                                             \
      fflush(stderr);                                                  \
    }                                                                  \
  } while (0)

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(idx, DDG_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

static ddg_status fail(ddg_index* idx, ddg_status st, const std::string& msg) {
  if (idx) idx->err = msg; else g_err = msg;
  return st;
}

static hipError_t ensure(ddg_index* idx, int id, size_t bytes) {
  Buf& b = idx->s.b[id];
  if (b.cap >= bytes && b.p) return hipSuccess;
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.cap = 0; }
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) { b.p = nullptr; return e; }
  b.cap = want;
  return hipSuccess;
}

template <class T>
static T* buf(ddg_index* idx, int id) { return (T*)idx->s.b[id].p; }

#define ENSURE(id, bytes)                                                                            \
  do {                                                                                               \
    hipError_t e_ = ensure(idx, (id), (bytes));                                                      \
    if (e_ != hipSuccess) return fail(idx, e_ == hipErrorOutOfMemory ? DDG_EOOM : DDG_EDEVICE,       \
                                      std::string("device allocation failed: ") + hipGetErrorString(e_)); \
  } while (0)

static uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

// ----------------------------------------------------------- launch sizing
struct SearchLayout {
  SearchArgs a;
  Launch l;
  bool ok;
};

// Element size and 16-byte chunks per row of a traversal copy.
static uint32_t trav_esize(const ddg_index* idx, int trav) {
  return trav == DDG_TRAV_BF16 ? 2u : trav == DDG_TRAV_FP8E4M3 ? 1u : idx->es;
}

// Row stride of a reduced copy: one 64-byte line for short rows, whole
// 128-byte lines otherwise.
static uint32_t trav_stride_for(const ddg_index* idx, int trav) {
  const uint32_t bytes = idx->d * trav_esize(idx, trav);
  return bytes <= 64 ? 64u : round_up(bytes, 128);
}

// Fills the traversal fields of `a` and the query staging sizes; returns the
// per-wave LDS bytes of the staged query (plus its FP8 factors).
// float32 rows of more than one load batch whose reference order walks whole
// groups of four (MIPS, or L2 with d % 4 == 0) are split over the fewest
// lanes (a power of two) that leave each at most kSplitChunks chunks: one
// round trip per row instead of one per batch of eight chunks (searches and
// prunes).  DDG_NAT_SPLIT=0 keeps one lane per row (for measurement; the
// results are the same).
static void nat_split_for(const ddg_index* idx, uint32_t& S_out, uint32_t& cps_out) {
  S_out = 0;
  cps_out = 0;
  const uint32_t chunks = idx->ad / 4;
  const char* ns = getenv("DDG_NAT_SPLIT");
  if (idx->dtype == DDG_F32 && chunks > kChunkBatch && (idx->metric == DDG_MIPS || idx->d % 4 == 0) &&
      !(ns && ns[0] == '0')) {
    uint32_t S = 2;
    while (S < kWave && (chunks + S - 1) / S > kSplitChunks) S *= 2;
    // rows beyond 64 lanes of kSplitChunks (d > 2048) keep one lane per row
    if ((chunks + S - 1) / S <= kSplitChunks) {
      S_out = S;
      cps_out = (chunks + S - 1) / S;
    }
  }
}

static uint32_t trav_layout(const ddg_index* idx, int trav, SearchArgs& a) {
  a.trav_split = 1;
  a.trav_cps = 0;
  if (trav == DDG_TRAV_NATIVE) {
    a.qrow_bytes = round_up(idx->ad * idx->es, 16);
    a.off_qaux = a.qrow_bytes;
    nat_split_for(idx, a.nat_split, a.nat_cps);
    return a.qrow_bytes;
  }
  const uint32_t esz = trav_esize(idx, trav);
  // whole rows of the copy (zero beyond d, like the staged query): every
  // distance walks complete 64-byte multiples, never a masked tail
  a.trav_stride = trav_stride_for(idx, trav);
  a.trav_chunks = a.trav_stride / 16;
  // a reduced row is split over the fewest lanes (1, 2, 4 or 8) that leave
  // each lane at most kLaneChunks chunks (rows above 512 bytes: as few as
  // possible over 8 lanes).  Expansions have few candidates on sparse graphs,
  // so narrow lane shares cost less arithmetic than full 64-lane batches:
  // 10M x 200 MIPS at Ls=250, FP8 split 4 ways 3.75 ms against 3.86 ms for
  // 2 ways, BF16 8 ways 4.42 ms against 4.67 ms for 4 ways.  (A fixed pairwise
  // tree of 8 blocks, which would let any expansion pick its own split, was
  // measured slower: 4.04 ms.)
  while (a.trav_split < 8 && (a.trav_chunks + a.trav_split - 1) / a.trav_split > kLaneChunks) a.trav_split *= 2;
  a.trav_cps = (a.trav_chunks + a.trav_split - 1) / a.trav_split;
  a.trav_rows = idx->trav_rows[trav];
  a.trav_scale = idx->trav_scale;
  const uint32_t qn = std::max<uint32_t>(idx->ad, a.trav_chunks * (16 / esz));
  a.qrow_bytes = round_up(qn * 4, 16);
  a.off_qaux = a.qrow_bytes;
  return trav == DDG_TRAV_FP8E4M3 ? 2 * a.qrow_bytes : a.qrow_bytes;
}

// Where the visited filter of a search lives: LDS up to kLdsTableMax bytes,
// global memory above (one slice per resident wave).  DDG_GTABLE=0 / 1
// forces LDS / global for every launch.
constexpr uint32_t kLdsTableMax = 8192;
static bool global_table(uint32_t bits) {
  static int force = -2;
  if (force == -2) {
    const char* e = getenv("DDG_GTABLE");
    force = (e && (e[0] == '0' || e[0] == '1')) ? e[0] - '0' : -1;
  }
  if (force >= 0) return force == 1;
  return (4u << bits) > kLdsTableMax;
}

static uint32_t tune_key(int trav, uint32_t bits) { return ((uint32_t)trav << 8) | bits; }

// Launch shapes the autotuner chose, for the whole process: any handle with
// the same dtype, metric, dimensions, degree bound, traversal mode and filter
// size (on a GPU with the same CU count) starts from them.
struct ProcessTuned {
  int dtype, metric;
  uint32_t d, ad, R, key, num_cu;
  ddg_index::Tuned cfg;
};
static std::mutex g_tune_mu;
static std::vector<ProcessTuned> g_tune_table;

static bool process_tuned(const ddg_index* idx, uint32_t key, ddg_index::Tuned* out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto& e : g_tune_table)
    if (e.dtype == idx->dtype && e.metric == idx->metric && e.d == idx->d && e.ad == idx->ad && e.R == idx->R &&
        e.key == key && e.num_cu == idx->num_cu) {
      *out = e.cfg;
      return true;
    }
  return false;
}

static void process_tune_store(const ddg_index* idx, const ddg_index::Tuned& cfg) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& e : g_tune_table)
    if (e.dtype == idx->dtype && e.metric == idx->metric && e.d == idx->d && e.ad == idx->ad && e.R == idx->R &&
        e.key == cfg.key && e.num_cu == idx->num_cu) {
      e.cfg = cfg;
      return;
    }
  g_tune_table.push_back(ProcessTuned{idx->dtype, idx->metric, idx->d, idx->ad, idx->R, cfg.key, idx->num_cu, cfg});
}

// Filter sizes (log2 slots) that keep a written-slot bitmap with a global
// filter (native traversal; the reduced-precision kernels are compiled
// without it): 2^12 .. 2^13 slots.  Larger filters with a bitmap were
// measured slower (10M x 200 MIPS, native at Ls=275, 2^15 slots: 9.36 ms
// against 8.72 ms without): the search waits on latency, not on the
// filter's traffic, and the bitmap costs LDS and occupancy.
// DDG_WBITS="lo-hi" (or "0" for none) overrides the range for measurement;
// it is read at every launch and results are the same either way.
static std::pair<uint32_t, uint32_t> wbits_range() {
  std::pair<uint32_t, uint32_t> v{12u, 13u};
  const char* e = getenv("DDG_WBITS");
  if (e) {
    unsigned lo = 0, hi = 0;
    if (sscanf(e, "%u-%u", &lo, &hi) == 2 && lo <= hi && hi <= 24) v = {lo, hi};
    else if (e[0] == '0') v = {1u, 0u};
  }
  return v;
}

// Fills the LDS layout and launch shape of a search; a.gtable_words > 0 asks
// the caller for a global filter pool of blocks * waves_per_block slices.
// The filter placement and block shape come from, in order: the autotuner's
// shape being timed, a tuned shape for (trav, filter bits), the defaults
// (filter in LDS up to 8 KB, up to 4 waves per block, as many blocks per CU
// as LDS and 32 waves allow).  Returns false when the shape does not fit.
static bool search_layout_cfg(const ddg_index* idx, uint32_t beam_max, uint32_t nstarts, int trav,
                              const ddg_index::Tuned* cfg, SearchArgs& a, Launch& l);

static bool search_layout(const ddg_index* idx, uint32_t beam_max, uint32_t nstarts, int trav, SearchArgs& a,
                          Launch& l, bool use_tuned = true, uint32_t* source = nullptr) {
  if (source) *source = 0;
  if (idx->cfg_override) return search_layout_cfg(idx, beam_max, nstarts, trav, idx->cfg_override, a, l);
  const uint32_t bits = filter_bits(beam_max);
  const ddg_index::Tuned* cfg = nullptr;
  uint32_t src = 0;
  ddg_index::Tuned pt{};
  if (use_tuned) {
    for (const auto& t : idx->tuned)
      if (t.key == tune_key(trav, bits)) { cfg = &t; src = 1; }
    if (!cfg && process_tuned(idx, tune_key(trav, bits), &pt)) { cfg = &pt; src = 2; }
  }
  // a shape tuned at another beam or start count of the same filter class may
  // not fit this frontier; the default shape is then used
  if (cfg && search_layout_cfg(idx, beam_max, nstarts, trav, cfg, a, l)) {
    if (source) *source = src;
    return true;
  }
  return search_layout_cfg(idx, beam_max, nstarts, trav, nullptr, a, l);
}

static bool search_layout_cfg(const ddg_index* idx, uint32_t beam_max, uint32_t nstarts, int trav,
                              const ddg_index::Tuned* cfg, SearchArgs& a, Launch& l) {
  a = SearchArgs{};
  l = Launch{};
  const uint32_t qrow = trav_layout(idx, trav, a);
  const uint32_t bits = filter_bits(beam_max);
  const bool gtab = cfg ? cfg->gtab != 0 : global_table(bits);
  const uint32_t table = gtab ? 0u : 4u << bits;
  a.gtable = nullptr;
  a.gtable_words = gtab ? 1u << bits : 0u;
  const uint32_t fcap = round_up(std::max(beam_max, nstarts), 16);
  const uint32_t rcap = round_up(std::max<uint32_t>(idx->R, 64), 64);
  uint32_t off = 0;
  off += qrow;
  a.off_table = off; off += table;
  a.off_f0 = off; off += fcap * 8;
  a.off_f1 = off; off += fcap * 8;
  a.off_keep = off; off += rcap * 4;
  a.off_cand = off; off += rcap * 8;
  a.off_aux = off; off += rcap * 8;
  a.off_bkt = off; off += 1024;
  // with a global filter, a bitmap of the slots the current query has
  // written (2^bits / 8 bytes): a never-written slot needs no global read,
  // and the slice needs no clearing per query (13 % less fetched and 29 %
  // less written at the 10M uint8 workload, Ls=140).  Below 2^12 slots a
  // query writes most slots anyway (measured 2 % slower at Ls=80).  The
  // range is wbits_range().
  const auto wr = wbits_range();
  a.wbits = gtab && bits >= wr.first && bits <= wr.second && trav == DDG_TRAV_NATIVE ? 1u : 0u;
  a.off_wbits = off;
  if (a.wbits) off += std::max<uint32_t>(16u, (1u << bits) / 8u);
  // rows of more than one load batch (8 chunks of 16 B) can be gathered into
  // LDS, cfg->stage rows at a time (at most 512 chunks, one batch per lane;
  // stride odd in chunks); the LDS this takes costs occupancy, so only a
  // tuned shape turns it on
  const uint32_t C = trav == DDG_TRAV_NATIVE ? (idx->ad * idx->es) / 16 : a.trav_chunks;
  a.stage_rows = 0;
  if (cfg && cfg->stage && C > 8) {
    a.stage_chunks = C;
    a.stage_stride = C % 2 ? C : C + 1;
    a.stage_rows = std::max<uint32_t>(1, std::min<uint32_t>(std::min<uint32_t>(64, 512 / C), cfg->stage));
    a.off_stage = round_up(off, 16);
    off = a.off_stage + a.stage_rows * a.stage_stride * 16;
  }
  a.lds_per_wave = round_up(off, 16);
  a.fcap = fcap;
  a.rcap = rcap;
  if (a.lds_per_wave > kLdsPerCU) return false;
  uint32_t wpb = std::min<uint32_t>(4, kLdsPerCU / a.lds_per_wave);
  if (cfg) wpb = cfg->wpb;
  const uint32_t per_block = wpb * a.lds_per_wave;
  if (per_block > kLdsPerCU) return false;
  uint32_t bpc = std::max<uint32_t>(1, std::min<uint32_t>(32 / wpb, kLdsPerCU / per_block));
  if (cfg) {
    if (cfg->bpc * per_block > kLdsPerCU || cfg->bpc * wpb > 32) return false;
    bpc = cfg->bpc;
  }
  l.waves_per_block = wpb;
  l.lds_bytes = per_block;
  l.blocks = idx->num_cu * bpc;
  return true;
}

// Bits an id below n takes (at least 1).
static uint32_t id_bits_for(uint64_t n) {
  uint32_t b = 1;
  while (b < 64 && (1ull << b) < n) b++;
  return b;
}

// The global filter pool of a launch (a.gtable_words > 0): one slice per
// resident wave.  Without the written-slot bitmap, and when ids leave at
// least two bits of a 32-bit word free, the slices are epoch-tagged: the
// pool and its per-slice epochs are initialised once per slice size, and
// later launches of that size continue their epochs (any other use of the
// pool, or a new allocation, starts over).
static ddg_status gtable_prepare(ddg_index* idx, SearchArgs& a, const Launch& l) {
  a.gepoch = nullptr;
  if (!a.gtable_words) return DDG_OK;
  const uint64_t slices = (uint64_t)l.blocks * l.waves_per_block;
  ENSURE(B_GTAB, (size_t)slices * a.gtable_words * 4);
  a.gtable = buf<uint32_t>(idx, B_GTAB);
  const uint32_t ib = id_bits_for(idx->n);
  const char* ep_env = getenv("DDG_EPOCH");  // DDG_EPOCH=0: clear per query, for measurement
  if (a.wbits || ib > 30 || (ep_env && ep_env[0] == '0')) {
    idx->ep_pool = nullptr;  // this launch stores untagged ids
    return DDG_OK;
  }
  ENSURE(B_GEPOCH, (size_t)slices * 4);
  uint32_t* ep = buf<uint32_t>(idx, B_GEPOCH);
  const bool same = idx->ep_pool == a.gtable && idx->ep_epochs == ep && idx->ep_words == a.gtable_words;
  const uint64_t from = same ? std::min(idx->ep_slices, slices) : 0;
  if (from < slices) {
    HIPCHK(launch_fill_u32(a.gtable + from * a.gtable_words, 0xFFFFFFFFu, (slices - from) * a.gtable_words,
                           idx->stream));
    HIPCHK(launch_fill_u32(ep + from, 0u, slices - from, idx->stream));
  }
  idx->ep_pool = a.gtable;
  idx->ep_epochs = ep;
  idx->ep_words = a.gtable_words;
  idx->ep_slices = std::max<uint64_t>(same ? idx->ep_slices : 0, slices);
  a.gepoch = ep;
  a.id_bits = ib;
  a.epoch_end = (1u << (32 - ib)) - 1u;  // epoch_end << ib | id is never 0xFFFFFFFF below it
  // DDG_EPOCH_END=<e> (e >= 2) shortens the epoch cycle, so that tests reach
  // the wrap-around clear with few queries; results are the same
  const char* ee = getenv("DDG_EPOCH_END");
  if (ee && atoi(ee) >= 2) a.epoch_end = std::min<uint32_t>(a.epoch_end, (uint32_t)atoi(ee));
  return DDG_OK;
}

#define ENSURE_GTABLE(a, l)                            \
  do {                                                 \
    ddg_status st_ = gtable_prepare(idx, (a), (l));    \
    if (st_ != DDG_OK) return st_;                     \
  } while (0)

// Launch tunables, read once from the environment (for measurement sweeps):
//   DDG_GTABLE=0|1       visited filter always in LDS / always global
//   DDG_PRUNE_LDS=<KB>   per-wave LDS budget of robustPrune (0: no row cache)
//   DDG_PRUNE_CCAP=<n>   candidates a prune keeps in LDS; longer lists take
//                        the global-scratch pass
//   DDG_PRUNE_BIG_WAVES=<w> waves sharing one task of the global-scratch pass
static int env_int(const char* name, int dflt, int lo, int hi) {
  const char* e = getenv(name);
  if (!e) return dflt;
  const int v = atoi(e);
  return (v < lo || v > hi) ? dflt : v;
}
// Defaults from a 10M x 128 uint8 build sweep: 12 KB per wave; forward
// prunes keep up to 512 candidates in LDS (visited lists), reverse prunes
// 128 or 256 (|S| + R for most targets; hub targets with thousands of new
// in-edges take the global-scratch pass instead of inflating every wave).
// Waves of a block of the queued-task pass (1, 2, 4, 8 or 16).  16 measured
// best over the 10M x 128 uint8, 1M x 100 int8 and 10M x 200 float32 builds
// with the reverse caps below (uint8 reverse 801 -> 779 ms against 8): the
// queued tasks are hubs whose greedy steps are a serial chain per block.
static uint32_t prune_big_waves() {
  static const int w = env_int("DDG_PRUNE_BIG_WAVES", 16, 1, (int)kMaxBigWaves);
  return (uint32_t)(w & (w - 1) ? 16 : w);
}
static uint32_t prune_wave_lds() {
  static const int kb = env_int("DDG_PRUNE_LDS", 12, 0, 160);
  return (uint32_t)kb * 1024u;
}
static uint32_t prune_ccap_max(uint32_t dflt) {
  static const int c = env_int("DDG_PRUNE_CCAP", 0, 64, 4096);
  return c ? (uint32_t)c : dflt;
}
// Reverse lists of up to 256 candidates stay in the wave's LDS when the rows
// prune from the Gram matrix (integer rows): 1M x 100 int8 reverse 74 -> 43
// ms, 10M x 128 uint8 826 -> 801 ms against 128; float32 rows keep 128
// (10M x 200 MIPS 300 -> 316 ms at 256).
constexpr uint32_t kForwardCcap = 512, kReverseCcap = 128, kReverseCcapGram = 256, kAnyCcap = 4096;
// DDG_PRUNE_GRAM=0 prunes integer rows pair by pair instead of from the
// Gram matrix (measurement only; the same selections)
static bool prune_gram_enabled() {
  static const int on = env_int("DDG_PRUNE_GRAM", 1, 0, 1);
  return on != 0;
}
// integer rows of at most 256 bytes, a multiple of 64: the Gram-matrix prune
// applies (its tiles split each row into four lane groups of whole 16-byte K
// slabs, which covers the row only when it is a multiple of 64 bytes)
static bool prune_gram_rows(const ddg_index* idx) {
  return idx->es == 1 && idx->ad <= 256 && idx->ad % 64 == 0 && prune_gram_enabled();
}

// Per-wave LDS of a prune launch: p's and p*'s rows, the selection, ccap
// candidates (distance, id, row slot), then a cache of rrows candidate rows
// sized to keep the wave within the prune LDS budget (no cache when fewer
// than 16 rows fit).  Cached rows are an odd number of 16-byte chunks apart.
static void prune_layout(const ddg_index* idx, uint32_t Rsel, uint32_t want_ccap, uint32_t ccap_max, PruneArgs& a,
                         Launch& l) {
  const uint32_t row = round_up(idx->ad * idx->es, 16);
  const uint32_t rs = (row / 16) % 2 == 0 ? row + 16 : row;
  const uint32_t budget = prune_wave_lds();
  a.rstride = rs / idx->es;
  uint32_t ccap = 64;
  while (ccap < want_ccap && ccap < prune_ccap_max(ccap_max)) ccap <<= 1;
  auto fixed = [&](uint32_t c) { return 2 * row + round_up(Rsel * 4, 16) + round_up(c * 10, 16); };
  while (ccap > 64 && fixed(ccap) * 2 > kLdsPerCU) ccap >>= 1;
  a.ccap = ccap;
  a.off_cand = 2 * row + round_up(Rsel * 4, 16);
  a.off_rows = fixed(ccap);
  // Gram rows: tasks of up to 256 candidates prune from an MFMA Gram matrix
  // (its LDS replaces the row cache)
  a.gram_max = 0;
  a.off_gram = fixed(ccap);
  if (prune_gram_rows(idx)) {
    a.gram_max = std::min<uint32_t>(ccap, 256);
    a.rrows = 0;
    a.lds_per_wave = round_up(fixed(ccap) + prune_gram_lds_bytes(), 16);
  } else {
    uint32_t rrows = fixed(ccap) < budget ? (budget - fixed(ccap)) / rs : 0u;
    rrows = std::min<uint32_t>(rrows, std::min<uint32_t>(ccap, round_up(std::max<uint32_t>(want_ccap, 16), 16)));
    a.rrows = rrows >= 16 ? rrows : 0u;
    a.lds_per_wave = round_up(fixed(ccap) + a.rrows * rs, 16);
  }
  // the block shape (1, 2 or 4 waves) that fits the most waves on a CU
  uint32_t wpb = 1, bpc = 1, best = 0;
  for (uint32_t w : {4u, 2u, 1u}) {
    if (w * a.lds_per_wave > kLdsPerCU) continue;
    const uint32_t b = std::min<uint32_t>(32 / w, kLdsPerCU / (w * a.lds_per_wave));
    if (w * b > best) { best = w * b; wpb = w; bpc = b; }
  }
  l.waves_per_block = wpb;
  l.lds_bytes = wpb * a.lds_per_wave;
  l.blocks = idx->num_cu * bpc;
}

// ---------------------------------------------------------------- timing
// Device time between construction and stop() (HIP events on the handle's
// stream; stop() synchronises).  A disabled timer records nothing, so it can
// sit in code that a stream capture records.
struct Timer {
  ddg_index* idx;
  double* acc;
  bool on;
  Timer(ddg_index* i, double* a, bool enabled = true) : idx(i), acc(a), on(enabled) {
    if (on) (void)hipEventRecord(idx->ev0, idx->stream);
  }
  void stop() {
    if (!on) return;
    (void)hipEventRecord(idx->ev1, idx->stream);
    (void)hipEventSynchronize(idx->ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, idx->ev0, idx->ev1);
    *acc += ms;
  }
};

// ------------------------------------------------------------- build parts
static uint32_t key_bits_for(uint64_t n) {
  uint32_t b = 1;
  while (b < 32 && (1ull << b) < n) b++;
  return ((b + 7) / 8) * 8;
}

// Beam search in build mode over pts[0..B) (device): visited lists to
// B_VIS_IDS / B_VIS_D with stride vcap, lengths to B_VISN, dist_cmps to
// B_CMPS, the start point read from B_STARTS.  insert_search_prepare sizes
// the buffers and picks the launch shape (it allocates, and may run the
// filter pool's set-up, so it runs before any recorded launch);
// insert_search_enqueue clears the work counter and the overflow word and
// launches, with no host synchronisation (so it can be recorded).
static ddg_status insert_search_prepare(ddg_index* idx, const uint32_t* pts, uint64_t B, uint32_t L, uint32_t vcap,
                                        SearchArgs& a, Launch& l) {
  ENSURE(B_VIS_IDS, (size_t)B * vcap * 4);
  ENSURE(B_VIS_D, (size_t)B * vcap * 4);
  ENSURE(B_VISN, (size_t)B * 4);
  ENSURE(B_CMPS, (size_t)B * 8);
  ENSURE(B_FLAGS, kFlagBytes);
  ENSURE(B_STARTS, 64);
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  a = SearchArgs{};
  l = Launch{};
  // DDG_BUILD_SHAPE="gtab,waves_per_block,blocks_per_cu,stage_rows" fixes
  // the insert search's launch shape, for measurement (results are the
  // same for every shape)
  ddg_index::Tuned bcfg{};
  const char* bshape = getenv("DDG_BUILD_SHAPE");
  const bool pinned = bshape && sscanf(bshape, "%d,%u,%u,%u", &bcfg.gtab, &bcfg.wpb, &bcfg.bpc, &bcfg.stage) == 4;
  if (pinned ? !search_layout_cfg(idx, L, 1, DDG_TRAV_NATIVE, &bcfg, a, l)
             : !search_layout(idx, L, 1, DDG_TRAV_NATIVE, a, l, false))
    return fail(idx, DDG_EINVAL, "build beam L too large for LDS");
  a.base = idx->rows;
  a.graph = idx->graph;
  a.qids = pts;
  a.starts = buf<uint32_t>(idx, B_STARTS);
  a.dist_cmps = buf<uint64_t>(idx, B_CMPS);
  a.visited_count = buf<uint32_t>(idx, B_VISN);
  a.vis_ids = buf<uint32_t>(idx, B_VIS_IDS);
  a.vis_dists = buf<float>(idx, B_VIS_D);
  a.overflow = flags + F_OVERFLOW;
  a.work = flags + F_WORK;
  a.n = idx->n;
  a.nq = B;
  a.limit = idx->n;
  a.cut = 0.0;
  a.d = idx->d;
  a.ad = idx->ad;
  a.R = idx->R;
  a.nstarts = 1;
  a.k = 0;
  a.beam = L;
  a.degree_limit = idx->R;
  a.vcap = vcap;
  a.build_mode = 1;
  a.cut_trim = 0;
  l.blocks = (uint32_t)std::min<uint64_t>(l.blocks, (B + l.waves_per_block - 1) / l.waves_per_block);
  ENSURE_GTABLE(a, l);
  return DDG_OK;
}

static hipError_t insert_search_enqueue(ddg_index* idx, const SearchArgs& a, const Launch& l) {
  hipError_t e = launch_fill_u32(buf<uint32_t>(idx, B_FLAGS) + F_WORK, 0u, 2, idx->stream);
  if (e != hipSuccess) return e;
  return launch_search(idx->dtype, idx->metric, DDG_TRAV_NATIVE, a, l, idx->stream);
}

// The build-mode search on its own (ddg_insert_search): a visited list that
// outgrows its slot reruns the search with room to spare (*vcap doubled).
static ddg_status run_insert_search(ddg_index* idx, const uint32_t* pts, uint64_t B, uint32_t L, uint32_t start,
                                    uint32_t* vcap) {
  for (int attempt = 0; attempt < 8; attempt++) {
    SearchArgs a{};
    Launch l{};
    ddg_status st = insert_search_prepare(idx, pts, B, L, *vcap, a, l);
    if (st != DDG_OK) return st;
    HIPCHK(hipMemcpyAsync(buf<uint32_t>(idx, B_STARTS), &start, 4, hipMemcpyHostToDevice, idx->stream));
    Timer t(idx, &idx->prof[0]);
    HIPCHK(insert_search_enqueue(idx, a, l));
    t.stop();
    uint32_t ov = 0;
    HIPCHK(hipMemcpy(&ov, buf<uint32_t>(idx, B_FLAGS) + F_OVERFLOW, 4, hipMemcpyDeviceToHost));
    if (ov & 2u) return fail(idx, DDG_ERANGE, "start point out of range");
    if (!ov) return DDG_OK;
    *vcap *= 2;
  }
  return fail(idx, DDG_EINVAL, "visited lists exceed the capacity limit");
}

// Long-list scratch a prune can need: a task takes the global path only when
// its M = |A| + deg candidates exceed ccap, and takes M entries.  With
// `total` candidates over np tasks and deg <= Radd, the sum is at most
// total + np Radd, and below total (1 + Radd / (ccap - Radd)) because each
// such task has |A| > ccap - Radd.
static uint64_t prune_pool(uint64_t total, uint64_t np, uint32_t Radd, uint32_t ccap, uint32_t max_cand) {
  if (max_cand <= ccap) return 0;
  uint64_t b = total + np * (uint64_t)Radd;
  if (ccap > Radd) b = std::min<uint64_t>(b, total + (total * Radd + (ccap - Radd) - 1) / (ccap - Radd));
  return b + 64;
}

// Layout, launch shapes and buffers of a prune over at most np tasks whose
// lists hold at most max_cand candidates (total in all).  Allocates, so it
// runs before any recorded (captured) launch.
struct PrunePlan {
  PruneArgs a;
  Launch l, lb;
};
// max_exact: max_cand bounds every list (per-wave scratch slots); otherwise
// max_cand only sizes the layout and long lists take bump-allocated slices.
static ddg_status prune_prepare(ddg_index* idx, PrunePlan& pp, uint64_t np, uint32_t max_cand, uint64_t total,
                                uint32_t ccap_max, bool max_exact) {
  PruneArgs& a = pp.a;
  prune_layout(idx, a.Rsel, max_cand, ccap_max, a, pp.l);
  const uint32_t big_waves = 4 * idx->num_cu;
  uint64_t pool = 0;
  a.gcap = 0;
  bool bump = false;
  if (max_cand > a.ccap) {
    if (max_exact) {
      uint64_t g = 64;
      while (g < max_cand) g <<= 1;
      a.gcap = (uint32_t)g;
      pool = g * big_waves;
    } else {
      bump = true;
      pool = prune_pool(total, np, a.add ? idx->R : 0u, a.ccap, max_cand);
    }
  }
  ENSURE(B_FLAGS, kFlagBytes);
  ENSURE(B_BIG, std::max<uint64_t>(np, 1) * 4);
  ENSURE(B_BIGOFF, std::max<uint64_t>(np, 1) * 8);
  ENSURE(B_GSD, std::max<uint64_t>(pool, 64) * 4);
  ENSURE(B_GSI, std::max<uint64_t>(pool, 64) * 4);
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  a.base = idx->rows;
  a.graph = idx->graph;
  a.work = flags + F_PWORK;
  a.nbig = flags + F_NBIG;
  a.status = flags + F_STATUS;
  a.gused = (uint64_t*)(flags + F_GUSED);
  a.big_tasks = buf<uint32_t>(idx, B_BIG);
  a.big_off = bump ? buf<uint64_t>(idx, B_BIGOFF) : nullptr;
  a.gscratch_d = buf<float>(idx, B_GSD);
  a.gscratch_i = buf<uint32_t>(idx, B_GSI);
  a.gpool = pool;
  a.np = np;
  a.d = idx->d;
  a.ad = idx->ad;
  a.R1 = idx->R + 1;
  nat_split_for(idx, a.nat_split, a.nat_cps);
  pp.l.blocks = (uint32_t)std::min<uint64_t>(pp.l.blocks, (np + pp.l.waves_per_block - 1) / pp.l.waves_per_block);
  // the queued tasks: one block of prune_big_waves() waves per task, 16
  // waves per CU (at most big_waves blocks: one slot of gcap entries each)
  pp.lb.waves_per_block = prune_big_waves();
  pp.lb.blocks = std::min<uint32_t>(idx->num_cu * (16 / pp.lb.waves_per_block), big_waves);
  pp.lb.lds_bytes = round_up(prune_big_lds_bytes(idx->ad * idx->es, a.Rsel), 16);
  return DDG_OK;
}

// Both launches of a prepared prune, no host synchronisation.
static hipError_t prune_enqueue(ddg_index* idx, const PrunePlan& pp) {
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  hipError_t e = launch_fill_u32(flags + F_PWORK, 0u, F_GUSED + 2 - F_PWORK, idx->stream);
  if (e != hipSuccess) return e;
  PruneArgs a = pp.a;
  a.big_pass = 0;
  e = launch_prune(idx->dtype, idx->metric, a, pp.l, idx->stream);
  if (e != hipSuccess) return e;
  a.big_pass = 1;
  a.work = flags + F_PWORK2;
  a.lds_per_wave = pp.lb.lds_bytes;
  a.rrows = 0;
  return launch_prune(idx->dtype, idx->metric, a, pp.lb, idx->stream);
}

static ddg_status check_status(ddg_index* idx) {
  uint32_t st = 0;
  HIPCHK(hipMemcpy(&st, buf<uint32_t>(idx, B_FLAGS) + F_STATUS, 4, hipMemcpyDeviceToHost));
  if (st & 1u) return fail(idx, DDG_EDEGREE, "adjacency list would exceed max degree");
  if (st & 2u) return fail(idx, DDG_EINVAL, "prune candidate list exceeds scratch capacity");
  return DDG_OK;
}

// robustPrune over np tasks (fixed-stride lists or CSR) with `total`
// candidates in all, timed into *acc, errors checked.
static ddg_status run_prune(ddg_index* idx, PruneArgs a, uint64_t np, uint32_t max_cand, uint64_t total, double* acc,
                            uint32_t ccap_max = kAnyCcap) {
  if (np == 0) return DDG_OK;
  PrunePlan pp{};
  pp.a = a;
  ddg_status st = prune_prepare(idx, pp, np, max_cand, total, ccap_max, true);
  if (st != DDG_OK) return st;
  HIPCHK(launch_fill_u32(buf<uint32_t>(idx, B_FLAGS) + F_STATUS, 0u, 1, idx->stream));
  double ms = 0.0;
  Timer t(idx, &ms);
  HIPCHK(prune_enqueue(idx, pp));
  t.stop();
  *acc += ms;
  TRACE("  prune: %llu tasks, ccap %u, rrows %u, %u B/wave, %u waves x %u blocks: %.2f ms",
        (unsigned long long)np, pp.a.ccap, pp.a.rrows, pp.a.lds_per_wave, pp.l.waves_per_block, pp.l.blocks, ms);
  return check_status(idx);
}

static void schedule(uint64_t m, double base, size_t max_batch, std::vector<std::pair<uint64_t, uint64_t>>& out) {
  // index.h:250-261 / bsp_index.hpp:238-246
  out.clear();
  size_t inc = 0, count = 0;
  while (count < m) {
    size_t lo, hi;
    if (std::pow(base, (double)inc) <= (double)max_batch) {
      lo = (size_t)std::pow(base, (double)inc) - 1;
      hi = std::min<size_t>((size_t)std::pow(base, (double)(inc + 1)), m) - 1;
      count = hi;
    } else {
      lo = count;
      hi = std::min<size_t>(count + max_batch, m);
      count += max_batch;
    }
    out.push_back({lo, hi});
    inc++;
  }
}

static size_t max_batch_size(double max_fraction, uint64_t count) {
  // index.h:233-236 / bsp_index.hpp:164-166
  size_t mb = std::min<size_t>((size_t)(max_fraction * (double)(float)count), 1000000ul);
  return mb == 0 ? (size_t)count : mb;
}

__global__ void k_acc_visited(uint32_t* vc, const uint32_t* pts, const uint32_t* visn, const uint64_t* B_dev) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *B_dev) vc[pts[i]] += visn[i];
}

// A build round's size as the kernels after its insert search see it: B, or
// 0 when a visited list outgrew its slot (or the start point was out of
// range), so that the rest of the round changes nothing and the host can
// rerun it with room to spare.
__global__ void k_round_gate(uint32_t* flags, uint64_t B) {
  if (threadIdx.x == 0) *(uint64_t*)(flags + F_ROUNDB) = flags[F_OVERFLOW] ? 0ull : B;
}

// dist_cmps of a round's insert searches added to the build's total.
__global__ void k_sum_cmps(const uint64_t* cmps, const uint64_t* B_dev, unsigned long long* total) {
  __shared__ unsigned long long part[4];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = i < *B_dev ? (unsigned long long)cmps[i] : 0ull;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

__global__ void k_max_group(const uint64_t* goff, const uint64_t* ngroups, uint32_t* out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < *ngroups) atomicMax(out, (uint32_t)(goff[g + 1] - goff[g]));
}

// knn_index::batch_insert over device ids ins[0..m)
// Everything of a build round after its insert search (index.h:281-323): the
// forward prune, the graph update and pair list, the radix sort and grouping
// of the pairs, the reverse append / prune.  Every size it needs beyond B is
// produced on the device (pair and group counts, long-list queue), so the
// whole sequence can be recorded once as a hipGraph and replayed for every
// round of the same B.  The round's points are in B_RPTS, its visited lists
// in B_VIS_* with stride vcap.
struct RoundShape {
  uint64_t B;
  uint32_t vcap, Rsel, kb;
  double alpha;
};

static ddg_status round_prepare(ddg_index* idx, const RoundShape& r, PrunePlan& fwd, PrunePlan& rev) {
  const uint64_t B = r.B, Pmax = r.B * r.Rsel;
  ENSURE(B_NEW_IDS, (size_t)B * r.Rsel * 4);
  ENSURE(B_NEW_LEN, (size_t)B * 4);
  ENSURE(B_OFFS, (size_t)(B + 1) * 8);
  ENSURE(B_KEYS, Pmax * 4);
  ENSURE(B_VALS, Pmax * 4);
  ENSURE(B_KEYS2, Pmax * 4);
  ENSURE(B_VALS2, Pmax * 4);
  ENSURE(B_TMP, std::max(scan_tmp_bytes(std::max(B, Pmax)), radix_tmp_bytes(Pmax)) + 64);
  ENSURE(B_GFLAGS, Pmax * 4);
  ENSURE(B_GOFFS, Pmax * 8);
  ENSURE(B_GKEYS, Pmax * 4);
  ENSURE(B_GOFF, (Pmax + 1) * 8);
  // forward prune (index.h:281): visited lists with their distances, plus G[p]
  fwd.a = PruneArgs{};
  fwd.a.pts = buf<uint32_t>(idx, B_RPTS);
  fwd.a.a_ids = buf<uint32_t>(idx, B_VIS_IDS);
  fwd.a.a_dists = buf<float>(idx, B_VIS_D);
  fwd.a.a_len = buf<uint32_t>(idx, B_VISN);
  fwd.a.a_stride = r.vcap;
  fwd.a.out_ids = buf<uint32_t>(idx, B_NEW_IDS);
  fwd.a.out_len = buf<uint32_t>(idx, B_NEW_LEN);
  fwd.a.alpha = r.alpha;
  fwd.a.Rsel = r.Rsel;
  fwd.a.add = 1;
  fwd.a.np_dev = (const uint64_t*)(buf<uint32_t>(idx, B_FLAGS) + F_ROUNDB);
  ddg_status st = prune_prepare(idx, fwd, B, r.vcap + idx->R, B * (uint64_t)r.vcap, kForwardCcap, true);
  if (st != DDG_OK) return st;
  // reverse append / prune per target (index.h:309-322); the cap keeps at
  // least 2R so the long-list pool bound holds
  rev.a = PruneArgs{};
  rev.a.pts = buf<uint32_t>(idx, B_GKEYS);
  rev.a.a_off = buf<uint64_t>(idx, B_GOFF);
  rev.a.alpha = r.alpha;
  rev.a.Rsel = r.Rsel;
  rev.a.add = 1;
  rev.a.write_graph = 1;
  rev.a.append_if_fits = 1;
  uint32_t rcap = prune_gram_rows(idx) ? kReverseCcapGram : kReverseCcap;
  while (rcap < 2 * idx->R) rcap <<= 1;
  return prune_prepare(idx, rev, Pmax, rcap + 1 + idx->R, Pmax, rcap, false);
}

// Enqueues the whole round (index.h:270-323): the insert search, the gate
// that turns the rest into no-ops when a visited list overflowed, the
// visited-count and dist_cmps accumulation, then the tail.  No host
// synchronisation unless `timed` (then each part is timed into the profile,
// which synchronises).
static ddg_status round_enqueue(ddg_index* idx, const RoundShape& r, const SearchArgs& sa, const Launch& sl,
                                uint32_t* vc_dev, const PrunePlan& fwd0, const PrunePlan& rev0, bool timed) {
  const uint64_t B = r.B, Pmax = r.B * r.Rsel;
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  uint64_t* dpairs = (uint64_t*)(flags + F_PAIRS);
  uint64_t* dgroups = (uint64_t*)(flags + F_GROUPS);
  const uint64_t* dB = (const uint64_t*)(flags + F_ROUNDB);
  double t_search = 0.0, t_fwd = 0.0, t_edges = 0.0, t_rev = 0.0;
  {
    Timer t(idx, &t_search, timed);
    HIPCHK(insert_search_enqueue(idx, sa, sl));
    hipLaunchKernelGGL(k_round_gate, dim3(1), dim3(64), 0, idx->stream, flags, B);
    HIPCHK(hipGetLastError());
    const uint32_t nb = (uint32_t)((B + 255) / 256);
    if (vc_dev)
      hipLaunchKernelGGL(k_acc_visited, dim3(nb), dim3(256), 0, idx->stream, vc_dev, buf<uint32_t>(idx, B_RPTS),
                         buf<uint32_t>(idx, B_VISN), dB);
    hipLaunchKernelGGL(k_sum_cmps, dim3(nb), dim3(256), 0, idx->stream, buf<uint64_t>(idx, B_CMPS), dB,
                       (unsigned long long*)(flags + F_CMPS));
    HIPCHK(hipGetLastError());
    t.stop();
  }
  {
    Timer t(idx, &t_fwd, timed);
    HIPCHK(prune_enqueue(idx, fwd0));
    t.stop();
  }
  const bool in_alt = Pmax > 1 && ((r.kb + 7) / 8) % 2 == 1;
  uint32_t* skeys = in_alt ? buf<uint32_t>(idx, B_KEYS2) : buf<uint32_t>(idx, B_KEYS);
  uint32_t* svals = in_alt ? buf<uint32_t>(idx, B_VALS2) : buf<uint32_t>(idx, B_VALS);
  {
    Timer t(idx, &t_edges, timed);
    HIPCHK(launch_apply_rows(idx->graph, idx->R + 1, buf<uint32_t>(idx, B_RPTS), buf<uint32_t>(idx, B_NEW_IDS),
                             buf<uint32_t>(idx, B_NEW_LEN), r.Rsel, B, flags + F_STATUS, idx->stream, dB));
    HIPCHK(scan_exclusive_u32(buf<uint32_t>(idx, B_NEW_LEN), buf<uint64_t>(idx, B_OFFS), B, dpairs,
                              buf<void>(idx, B_TMP), idx->s.b[B_TMP].cap, idx->stream, dB));
    HIPCHK(launch_emit_pairs(buf<uint32_t>(idx, B_RPTS), buf<uint32_t>(idx, B_NEW_IDS), buf<uint32_t>(idx, B_NEW_LEN),
                             buf<uint64_t>(idx, B_OFFS), r.Rsel, B, buf<uint32_t>(idx, B_KEYS),
                             buf<uint32_t>(idx, B_VALS), idx->stream, dB));
    bool alt = false;
    HIPCHK(radix_sort_pairs(buf<uint32_t>(idx, B_KEYS), buf<uint32_t>(idx, B_VALS), buf<uint32_t>(idx, B_KEYS2),
                            buf<uint32_t>(idx, B_VALS2), Pmax, r.kb, buf<void>(idx, B_TMP), idx->s.b[B_TMP].cap,
                            idx->stream, &alt, dpairs));
    if (Pmax > 1 && alt != in_alt) return fail(idx, DDG_EDEVICE, "radix pass count mismatch");
    HIPCHK(launch_group_starts(skeys, Pmax, buf<uint32_t>(idx, B_GFLAGS), buf<uint64_t>(idx, B_GOFFS), dgroups,
                               buf<uint32_t>(idx, B_GKEYS), buf<uint64_t>(idx, B_GOFF), buf<void>(idx, B_TMP),
                               idx->s.b[B_TMP].cap, idx->stream, dpairs));
    t.stop();
  }
  {
    PrunePlan rev = rev0;
    rev.a.a_ids = svals;
    rev.a.np_dev = dgroups;
    Timer t(idx, &t_rev, timed);
    HIPCHK(prune_enqueue(idx, rev));
    t.stop();
  }
  if (timed) {
    idx->prof[0] += t_search;
    idx->prof[1] += t_fwd;
    idx->prof[2] += t_edges;
    idx->prof[3] += t_rev;
  }
  return DDG_OK;
}

// Build rounds, insert search included, run as recorded hipGraphs (one per
// batch size B), unless DDG_GRAPH=0 (then they run launch by launch with
// per-part timing).
static bool use_graphs() {
  return env_int("DDG_GRAPH", 1, 0, 1) == 1;  // read at every build (tests switch it)
}

struct GraphCache {
  std::vector<std::pair<uint64_t, hipGraphExec_t>> execs;
  uint32_t vcap = 0;
  ~GraphCache() { clear(); }
  void clear() {
    for (auto& e : execs) (void)hipGraphExecDestroy(e.second);
    execs.clear();
  }
};

static ddg_status batch_insert_dev(ddg_index* idx, const uint32_t* ins, uint64_t m, double alpha, uint32_t L,
                                   uint32_t Rsel, uint32_t start, double base, double max_fraction,
                                   uint64_t max_batch_base, uint32_t* vc_dev) {
  std::vector<std::pair<uint64_t, uint64_t>> batches;
  schedule(m, base, max_batch_size(max_fraction, max_batch_base), batches);
  // DDG_BUILD_VCAP=<v> starts the visited-list slots smaller (tests reach the
  // overflow rerun with it; results are the same)
  uint32_t vcap = std::max<uint32_t>(2 * L + 64, 128);
  const int vcap_env = env_int("DDG_BUILD_VCAP", 0, 1, 1 << 20);
  if (vcap_env > 0) vcap = (uint32_t)vcap_env;
  const uint32_t kb = key_bits_for(idx->n);
  uint64_t Bmax = 0;
  for (auto [lo, hi] : batches) Bmax = std::max<uint64_t>(Bmax, hi - lo);
  if (Bmax == 0) return DDG_OK;
  // every buffer at its largest size up front: recorded rounds keep their pointers
  ENSURE(B_FLAGS, kFlagBytes);
  ENSURE(B_RPTS, Bmax * 4);
  ENSURE(B_STARTS, 64);
  auto prepare_max = [&]() -> ddg_status {
    SearchArgs a{};
    Launch l{};
    ddg_status st = insert_search_prepare(idx, buf<uint32_t>(idx, B_RPTS), Bmax, L, vcap, a, l);
    if (st != DDG_OK) return st;
    PrunePlan f{}, r{};
    return round_prepare(idx, RoundShape{Bmax, vcap, Rsel, kb, alpha}, f, r);
  };
  ddg_status pst = prepare_max();
  if (pst != DDG_OK) return pst;
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  HIPCHK(hipMemcpyAsync(buf<uint32_t>(idx, B_STARTS), &start, 4, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_fill_u32(flags + F_STATUS, 0u, 1, idx->stream));
  HIPCHK(launch_fill_u32(flags + F_CMPS, 0u, 2, idx->stream));
  const bool graphs = use_graphs() && !trace_on();
  GraphCache cache;
  for (size_t bi = 0; bi < batches.size(); bi++) {
    const uint64_t lo = batches[bi].first, Bn = batches[bi].second - lo;
    if (Bn == 0) continue;
    TRACE("batch [%llu, %llu)", (unsigned long long)lo, (unsigned long long)(lo + Bn));
    HIPCHK(hipMemcpyAsync(buf<uint32_t>(idx, B_RPTS), ins + lo, Bn * 4, hipMemcpyDeviceToDevice, idx->stream));
    // the whole round as one recorded graph per batch size (or launch by
    // launch with DDG_GRAPH=0); a visited list that outgrew its slot leaves
    // the graph unchanged, and the round is run again with slots twice as
    // long
    for (int attempt = 0;; attempt++) {
      if (attempt == 8) return fail(idx, DDG_EINVAL, "visited lists exceed the capacity limit");
      const RoundShape rs{Bn, vcap, Rsel, kb, alpha};
      SearchArgs sa{};
      Launch sl{};
      PrunePlan fwd{}, rev{};
      ddg_status st = insert_search_prepare(idx, buf<uint32_t>(idx, B_RPTS), Bn, L, vcap, sa, sl);
      if (st != DDG_OK) return st;
      st = round_prepare(idx, rs, fwd, rev);
      if (st != DDG_OK) return st;
      if (!graphs) {
        st = round_enqueue(idx, rs, sa, sl, vc_dev, fwd, rev, true);
        if (st != DDG_OK) return st;
      } else {
        hipGraphExec_t exec = nullptr;
        for (auto& e : cache.execs)
          if (e.first == Bn) exec = e.second;
        if (!exec) {
          hipGraph_t g = nullptr;
          HIPCHK(hipStreamBeginCapture(idx->stream, hipStreamCaptureModeThreadLocal));
          st = round_enqueue(idx, rs, sa, sl, vc_dev, fwd, rev, false);
          hipError_t ce = hipStreamEndCapture(idx->stream, &g);
          if (st != DDG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
          }
          if (ce != hipSuccess) return fail(idx, DDG_EDEVICE, std::string("round capture: ") + hipGetErrorString(ce));
          ce = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (ce != hipSuccess)
            return fail(idx, DDG_EDEVICE, std::string("round instantiate: ") + hipGetErrorString(ce));
          cache.execs.push_back({Bn, exec});
        }
        Timer t(idx, &idx->prof[2]);
        HIPCHK(hipGraphLaunch(exec, idx->stream));
        t.stop();
      }
      // the round's overflow and error words (this read waits for the round)
      uint32_t w[8];
      HIPCHK(hipMemcpy(w, flags, sizeof(w), hipMemcpyDeviceToHost));
      if (w[F_OVERFLOW] & 2u) return fail(idx, DDG_ERANGE, "start point out of range");
      if (w[F_STATUS] & 1u) return fail(idx, DDG_EDEGREE, "adjacency list would exceed max degree");
      if (w[F_STATUS] & 2u) return fail(idx, DDG_EINVAL, "prune candidate list exceeds scratch capacity");
      if (!(w[F_OVERFLOW] & 1u)) break;
      // every buffer again at its largest size for the longer slots; the
      // recorded rounds point at the old ones
      vcap *= 2;
      cache.clear();
      st = prepare_max();
      if (st != DDG_OK) return st;
      TRACE("  visited lists overflowed: rerunning the round with vcap %u", vcap);
    }
    TRACE("  round done (vcap %u)", vcap);
  }
  HIPCHK(hipStreamSynchronize(idx->stream));
  ddg_status st = check_status(idx);
  if (st != DDG_OK) return st;
  uint64_t total_cmps = 0;
  HIPCHK(hipMemcpy(&total_cmps, flags + F_CMPS, 8, hipMemcpyDeviceToHost));
  idx->prof[7] += (double)total_cmps;
  return DDG_OK;
}

// Builds the reduced traversal copy `trav` of a float32 index once per handle.
static ddg_status prepare_trav(ddg_index* idx, int trav, double* seconds) {
  if (seconds) *seconds = 0.0;
  if (trav == DDG_TRAV_NATIVE) return DDG_OK;
  if (trav != DDG_TRAV_BF16 && trav != DDG_TRAV_FP8E4M3) return fail(idx, DDG_EINVAL, "unknown traversal mode");
  if (idx->dtype != DDG_F32) return fail(idx, DDG_EINVAL, "reduced-precision traversal needs a float32 index");
  if (idx->trav_rows[trav]) return DDG_OK;
  const uint32_t stride = trav_stride_for(idx, trav);
  void* rows = nullptr;
  hipError_t e = hipMalloc(&rows, idx->n * (size_t)stride);
  if (e != hipSuccess)
    return fail(idx, e == hipErrorOutOfMemory ? DDG_EOOM : DDG_EDEVICE,
                std::string("traversal copy allocation failed: ") + hipGetErrorString(e));
  double ms = 0.0;
  if (trav == DDG_TRAV_BF16) {
    Timer t(idx, &ms);
    e = launch_make_bf16((const float*)idx->rows, idx->n, idx->ad, idx->d, rows, stride, idx->stream);
    t.stop();
  } else {
    if (!idx->trav_scale) {
      e = hipMalloc((void**)&idx->trav_scale, 2 * (size_t)idx->d * 4);
      if (e != hipSuccess) {
        (void)hipFree(rows);
        idx->trav_scale = nullptr;
        return fail(idx, DDG_EOOM, "scale allocation failed");
      }
    }
    // per-dimension max |x|, scales on the host in IEEE float: s = m / 240, 1/s = 240 / m
    std::vector<uint32_t> mb(idx->d);
    std::vector<float> sc(2 * (size_t)idx->d);
    uint32_t* maxbits = (uint32_t*)idx->trav_scale;
    Timer t(idx, &ms);
    e = launch_fill_u32(maxbits, 0u, idx->d, idx->stream);
    if (e == hipSuccess) e = launch_col_absmax((const float*)idx->rows, idx->n, idx->ad, idx->d, maxbits, idx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(mb.data(), maxbits, idx->d * 4, hipMemcpyDeviceToHost, idx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(idx->stream);
    if (e == hipSuccess) {
      for (uint32_t j = 0; j < idx->d; j++) {
        float m;
        std::memcpy(&m, &mb[j], 4);
        const bool ok = std::isfinite(m) && m >= 1e-30f;
        sc[j] = ok ? m / 240.0f : 1.0f;
        sc[idx->d + j] = ok ? 240.0f / m : 1.0f;
      }
      e = hipMemcpyAsync(idx->trav_scale, sc.data(), sc.size() * 4, hipMemcpyHostToDevice, idx->stream);
    }
    if (e == hipSuccess)
      e = launch_make_fp8((const float*)idx->rows, idx->n, idx->ad, idx->d, idx->trav_scale + idx->d, rows, stride,
                          idx->stream);
    t.stop();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(idx->stream);
  if (e != hipSuccess) {
    (void)hipFree(rows);
    return fail(idx, DDG_EDEVICE, std::string("traversal copy: ") + hipGetErrorString(e));
  }
  idx->trav_rows[trav] = rows;
  idx->trav_stride[trav] = stride;
  if (seconds) *seconds = ms / 1000.0;
  return DDG_OK;
}

// ------------------------------------------------ allocation-free search
// Scratch of ddg_search_device_noalloc: the 256-byte status block (work
// counter, status word) and then the global visited-filter slices of beams
// above 90.  One copy of the array per GPU, held until the process ends and
// shared by every call on that GPU (calls that overlap on one GPU race).  In
// host code `scratch` names a host-side stand-in, never the device memory:
// only the address hipGetSymbolAddress returns, once per device with that
// device current, is passed to kernels and fills.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static bool g_scratch_done[kMaxDevices];

// This device's copy of `scratch` (looked up on the device's first handle).
static unsigned char* noalloc_scratch_of(int dev, hipError_t* err) {
  if (dev < 0 || dev >= kMaxDevices) { *err = hipErrorInvalidDevice; return nullptr; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_done[dev]) {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? (unsigned char*)p : nullptr;
    g_scratch_done[dev] = true;
  }
  *err = g_scratch_err[dev];
  return g_scratch_ptr[dev];
}

// ====================================================================== ABI
extern "C" {

const char* ddg_last_error(const ddg_index* idx) { return idx ? idx->err.c_str() : g_err.c_str(); }

ddg_status ddg_create(ddg_dtype dtype, ddg_metric metric, uint64_t n, uint32_t d, uint32_t aligned_d,
                      const void* rows, uint32_t R, ddg_index** out) {
  ddg_index* idx = nullptr;
  if (!out) return fail(idx, DDG_EINVAL, "out is NULL");
  *out = nullptr;
  if (dtype < DDG_U8 || dtype > DDG_F32 || metric < DDG_L2 || metric > DDG_MIPS)
    return fail(idx, DDG_EINVAL, "bad dtype/metric");
  if (n == 0 || d == 0 || aligned_d < d || !rows || R == 0) return fail(idx, DDG_EINVAL, "bad sizes");
  if (n >= (1ull << 31)) return fail(idx, DDG_EINVAL, "n must be below 2^31");
  const uint32_t es = dtype == DDG_F32 ? 4 : 1;
  if ((aligned_d * es) % 16 != 0) return fail(idx, DDG_EINVAL, "aligned_d * sizeof(T) must be a multiple of 16");
  if (d * es > 16384) return fail(idx, DDG_EINVAL, "rows above 16 KiB are not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(idx, DDG_EDEVICE, "no HIP device");
  idx = new ddg_index();
  idx->dtype = dtype;
  idx->metric = metric;
  idx->n = n;
  idx->d = d;
  idx->ad = aligned_d;
  idx->R = R;
  idx->es = es;
  int dev = 0;
  hipGetDevice(&dev);
  static bool attrs_set = false;  // per process: kernel attributes are process-wide
  if (!attrs_set) {
    search_set_attributes((int)kLdsPerCU);
    prune_set_attributes((int)kLdsPerCU);
    attrs_set = true;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    idx->num_cu = (uint32_t)prop.multiProcessorCount;
  {
    hipError_t se = hipSuccess;
    idx->noalloc_scratch = noalloc_scratch_of(dev, &se);
    if (!idx->noalloc_scratch) idx->noalloc_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
  }
  ddg_status st = DDG_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == DDG_OK) {
      st = e == hipErrorOutOfMemory ? DDG_EOOM : DDG_EDEVICE;
      idx->err = std::string(what) + ": " + hipGetErrorString(e);
    }
  };
  chk(hipStreamCreate(&idx->stream), "hipStreamCreate");
  chk(hipEventCreate(&idx->ev0), "hipEventCreate");
  chk(hipEventCreate(&idx->ev1), "hipEventCreate");
  const size_t row_bytes = (size_t)aligned_d * es;
  chk(hipMalloc(&idx->rows, n * row_bytes), "hipMalloc rows");
  chk(hipMalloc((void**)&idx->graph, n * (size_t)(R + 1) * 4), "hipMalloc graph");
  if (st == DDG_OK) chk(hipMemcpyAsync(idx->rows, rows, n * row_bytes, hipMemcpyHostToDevice, idx->stream), "H2D rows");
  if (st == DDG_OK) chk(launch_pad_rows(idx->rows, n, d * es, (uint32_t)row_bytes, idx->stream), "pad rows");
  if (st == DDG_OK) chk(hipMemsetAsync(idx->graph, 0, n * (size_t)(R + 1) * 4, idx->stream), "zero graph");
  if (st == DDG_OK) chk(hipStreamSynchronize(idx->stream), "sync");
  if (st != DDG_OK) {
    g_err = idx->err;
    ddg_destroy(idx);
    return st;
  }
  *out = idx;
  return DDG_OK;
}

void ddg_destroy(ddg_index* idx) {
  if (!idx) return;
  if (idx->stream) hipStreamSynchronize(idx->stream);
  for (auto& b : idx->s.b)
    if (b.p) hipFree(b.p);
  if (idx->rows) hipFree(idx->rows);
  if (idx->graph) hipFree(idx->graph);
  for (void* p : idx->trav_rows)
    if (p) hipFree(p);
  if (idx->trav_scale) hipFree(idx->trav_scale);
  if (idx->ev0) hipEventDestroy(idx->ev0);
  if (idx->ev1) hipEventDestroy(idx->ev1);
  if (idx->stream) hipStreamDestroy(idx->stream);
  delete idx;
}

uint64_t ddg_size(const ddg_index* idx) { return idx ? idx->n : 0; }
uint32_t ddg_max_degree(const ddg_index* idx) { return idx ? idx->R : 0; }
uint32_t ddg_dims(const ddg_index* idx) { return idx ? idx->d : 0; }
uint32_t ddg_aligned_dims(const ddg_index* idx) { return idx ? idx->ad : 0; }

ddg_status ddg_last_profile(const ddg_index* idx, double* out9) {
  if (!idx || !out9) return DDG_EINVAL;
  for (int i = 0; i < 9; i++) out9[i] = idx->prof[i];
  return DDG_OK;
}

ddg_status ddg_graph_set(ddg_index* idx, const uint32_t* rows) {
  if (!idx || !rows) return fail(idx, DDG_EINVAL, "NULL argument");
  for (uint64_t i = 0; i < idx->n; i++) {
    const uint32_t* r = rows + i * (uint64_t)(idx->R + 1);
    if (r[0] > idx->R) return fail(idx, DDG_EDEGREE, "graph row degree exceeds R");
    for (uint32_t j = 0; j < r[0]; j++)
      if (r[1 + j] >= idx->n) return fail(idx, DDG_ERANGE, "graph row holds an id >= n");
  }
  HIPCHK(hipMemcpyAsync(idx->graph, rows, idx->n * (size_t)(idx->R + 1) * 4, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  return DDG_OK;
}

ddg_status ddg_graph_get(ddg_index* idx, uint32_t* rows) {
  if (!idx || !rows) return fail(idx, DDG_EINVAL, "NULL argument");
  HIPCHK(hipMemcpyAsync(rows, idx->graph, idx->n * (size_t)(idx->R + 1) * 4, hipMemcpyDeviceToHost, idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  return DDG_OK;
}

static ddg_status build_common(ddg_index* idx, const ddg_build_params* bp, uint32_t* visited_counts,
                               double* build_seconds) {
  if (!bp) return fail(idx, DDG_EINVAL, "NULL params");
  if (bp->R == 0 || bp->R > idx->R) return fail(idx, DDG_EINVAL, "BuildParams.R must be in [1, graph R]");
  if (bp->L == 0) return fail(idx, DDG_EINVAL, "BuildParams.L must be positive");
  if (bp->start_point >= idx->n) return fail(idx, DDG_ERANGE, "start point out of range");
  if (!(bp->base > 1.0)) return fail(idx, DDG_EINVAL, "base must exceed 1");
  for (double& p : idx->prof) p = 0;
  const uint64_t n = idx->n;
  uint64_t m = n;
  if (bp->bsp && bp->max_build) m = std::min<uint64_t>(bp->max_build, n);
  ENSURE(B_INSERTS, n * 4);
  std::vector<uint32_t> ids(n);
  for (uint64_t i = 0; i < n; i++) ids[i] = (uint32_t)i;
  HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_INSERTS), ids.data(), n * 4, hipMemcpyHostToDevice));
  uint32_t* vc = nullptr;
  if (visited_counts) {
    ENSURE(B_VC, n * 4);
    vc = buf<uint32_t>(idx, B_VC);
    HIPCHK(launch_fill_u32(vc, 0u, n, idx->stream));
  }
  HIPCHK(hipStreamSynchronize(idx->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, idx->stream));
  const uint64_t mbb = bp->bsp ? m : n;
  ddg_status st = DDG_OK;
  if (bp->two_pass && !bp->bsp)
    st = batch_insert_dev(idx, buf<uint32_t>(idx, B_INSERTS), m, 1.0, bp->L, bp->R, bp->start_point, bp->base,
                          bp->max_fraction, mbb, vc);
  if (st == DDG_OK)
    st = batch_insert_dev(idx, buf<uint32_t>(idx, B_INSERTS), m, bp->alpha, bp->L, bp->R, bp->start_point, bp->base,
                          bp->max_fraction, mbb, vc);
  if (st == DDG_OK && bp->final_sort) {
    TRACE("final row sort");
    Timer t(idx, &idx->prof[4]);
    HIPCHK(launch_sort_rows(idx->dtype, idx->metric, idx->rows, idx->graph, n, idx->R + 1, idx->d, idx->ad,
                            idx->stream));
    t.stop();
  }
  HIPCHK(hipEventRecord(e1, idx->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (build_seconds) *build_seconds = ms / 1000.0;
  if (st != DDG_OK) return st;
  if (visited_counts) HIPCHK(hipMemcpy(visited_counts, vc, n * 4, hipMemcpyDeviceToHost));
  return DDG_OK;
}

ddg_status ddg_build(ddg_index* idx, const ddg_build_params* bp, uint32_t* visited_counts, double* build_seconds) {
  if (!idx) return fail(idx, DDG_EINVAL, "NULL index");
  return build_common(idx, bp, visited_counts, build_seconds);
}

// The insertion order batch_insert draws for random_order (index.h:237-241):
// parlay::random_permutation<int>(m) with the default generator, i.e. the
// numerical-recipes hash (parlay/utilities.h:131-140) as the generator
// (random.h:33-45); below 8192 elements one Fisher-Yates pass from the back
// (random.h:52-60), above it a stable bucket sort of the positions by
// hash(i) & (2^bits - 1) with bits = (ceil(log2 m) - 7) / 2 (m < 2^27) or
// ceil(log2 m) - 17, each bucket then shuffled with its own forked state
// (random.h:62-98).
static uint64_t nr_hash(uint64_t u) {
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

static void fisher_yates(uint32_t* a, uint64_t n, uint64_t state) {
  for (uint64_t i = n; i-- > 1;) std::swap(a[i], a[nr_hash(i + state) % (i + 1)]);
}

static void insertion_permutation(uint64_t m, std::vector<uint32_t>& perm) {
  perm.resize(m);
  if (m < 8192) {
    for (uint64_t i = 0; i < m; i++) perm[i] = (uint32_t)i;
    fisher_yates(perm.data(), m, 0);
    return;
  }
  uint32_t lg = 0;
  for (uint64_t b = m - 1; b; b >>= 1) lg++;
  const uint32_t bits = m < (1ull << 27) ? (lg - 7) / 2 : lg - 17;
  const uint64_t nb = 1ull << bits, mask = nb - 1;
  std::vector<uint64_t> first(nb + 1, 0);
  for (uint64_t i = 0; i < m; i++) first[(nr_hash(i) & mask) + 1]++;
  for (uint64_t b = 0; b < nb; b++) first[b + 1] += first[b];
  std::vector<uint64_t> next(first.begin(), first.end() - 1);
  for (uint64_t i = 0; i < m; i++) perm[next[nr_hash(i) & mask]++] = (uint32_t)i;
  for (uint64_t b = 0; b < nb; b++) fisher_yates(perm.data() + first[b], first[b + 1] - first[b], nr_hash(nr_hash(b)));
}

ddg_status ddg_batch_insert(ddg_index* idx, const uint32_t* inserts, uint64_t m, double alpha, uint32_t L,
                            uint32_t start_point, int random_order, double base, double max_fraction) {
  if (!idx || (!inserts && m)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (random_order && m > (uint64_t)INT32_MAX) return fail(idx, DDG_EINVAL, "random_order needs m <= INT_MAX");
  if (L == 0 || !(base > 1.0)) return fail(idx, DDG_EINVAL, "bad L/base");
  if (start_point >= idx->n) return fail(idx, DDG_ERANGE, "start point out of range");
  for (uint64_t i = 0; i < m; i++)
    if (inserts[i] >= idx->n) return fail(idx, DDG_ERANGE, "ERROR: invalid point given to batch_insert");
  for (double& p : idx->prof) p = 0;
  ENSURE(B_INSERTS, std::max<uint64_t>(m, 1) * 4);
  if (random_order) {
    std::vector<uint32_t> perm, shuffled(m);
    insertion_permutation(m, perm);
    for (uint64_t i = 0; i < m; i++) shuffled[i] = inserts[perm[i]];
    HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_INSERTS), shuffled.data(), m * 4, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_INSERTS), inserts, m * 4, hipMemcpyHostToDevice));
  }
  return batch_insert_dev(idx, buf<uint32_t>(idx, B_INSERTS), m, alpha, L, idx->R, start_point, base, max_fraction,
                          idx->n, nullptr);
}

ddg_status ddg_sort_rows(ddg_index* idx) {
  if (!idx) return fail(idx, DDG_EINVAL, "NULL index");
  HIPCHK(launch_sort_rows(idx->dtype, idx->metric, idx->rows, idx->graph, idx->n, idx->R + 1, idx->d, idx->ad,
                          idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  return DDG_OK;
}

ddg_status ddg_robust_prune(ddg_index* idx, uint64_t np, const uint32_t* pts, const uint64_t* cand_off,
                            const uint32_t* cand_ids, const float* cand_dists, double alpha, int add, uint32_t R,
                            uint32_t* out_ids, uint32_t* out_len) {
  if (!idx || (np && (!pts || !cand_off || !out_ids || !out_len))) return fail(idx, DDG_EINVAL, "NULL argument");
  if (R == 0) return fail(idx, DDG_EINVAL, "R must be positive");
  if (np == 0) return DDG_OK;
  const uint64_t nc = cand_off[np];
  uint32_t maxc = 0;
  for (uint64_t t = 0; t < np; t++) {
    if (pts[t] >= idx->n) return fail(idx, DDG_ERANGE, "prune point out of range");
    if (cand_off[t + 1] < cand_off[t]) return fail(idx, DDG_EINVAL, "cand_off must be non-decreasing");
    maxc = std::max<uint32_t>(maxc, (uint32_t)(cand_off[t + 1] - cand_off[t]));
  }
  for (uint64_t i = 0; i < nc; i++)
    if (cand_ids[i] >= idx->n) return fail(idx, DDG_ERANGE, "candidate id out of range");
  ENSURE(B_QIDS, np * 4);
  ENSURE(B_PRUNE_OFF, (np + 1) * 8);
  ENSURE(B_PRUNE_IDS, std::max<uint64_t>(nc, 1) * 4);
  ENSURE(B_PRUNE_D, std::max<uint64_t>(nc, 1) * 4);
  ENSURE(B_NEW_IDS, np * R * 4);
  ENSURE(B_NEW_LEN, np * 4);
  HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_QIDS), pts, np * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(buf<uint64_t>(idx, B_PRUNE_OFF), cand_off, (np + 1) * 8, hipMemcpyHostToDevice));
  if (nc) HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_PRUNE_IDS), cand_ids, nc * 4, hipMemcpyHostToDevice));
  if (nc && cand_dists) HIPCHK(hipMemcpy(buf<float>(idx, B_PRUNE_D), cand_dists, nc * 4, hipMemcpyHostToDevice));
  PruneArgs a{};
  a.pts = buf<uint32_t>(idx, B_QIDS);
  a.a_ids = buf<uint32_t>(idx, B_PRUNE_IDS);
  a.a_dists = cand_dists ? buf<float>(idx, B_PRUNE_D) : nullptr;
  a.a_off = buf<uint64_t>(idx, B_PRUNE_OFF);
  a.out_ids = buf<uint32_t>(idx, B_NEW_IDS);
  a.out_len = buf<uint32_t>(idx, B_NEW_LEN);
  a.alpha = alpha;
  a.Rsel = R;
  a.add = add ? 1 : 0;
  for (double& p : idx->prof) p = 0;
  ddg_status st = run_prune(idx, a, np, maxc + (add ? idx->R : 0), nc, &idx->prof[1]);
  if (st != DDG_OK) return st;
  HIPCHK(hipMemcpy(out_ids, buf<uint32_t>(idx, B_NEW_IDS), np * R * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(out_len, buf<uint32_t>(idx, B_NEW_LEN), np * 4, hipMemcpyDeviceToHost));
  return DDG_OK;
}

// Outputs of a search launch beyond searchAll's: beam_search's whole final
// frontier and its visited list (visit order; sorted on the host).
struct SearchExtra {
  uint32_t* fr_ids = nullptr;
  float* fr_dists = nullptr;
  uint32_t* fr_len = nullptr;
  uint32_t fr_stride = 0;
  uint32_t* vis_ids = nullptr;
  float* vis_dists = nullptr;
  uint32_t vcap = 0;
  bool overflow = false;  // out: some visited list outgrew vcap
};

// The order a launch takes its queries in (a.order, task t -> query): the
// longest expected first, so that the launch's end is not set by a long
// query that started late.  With per-query beams the key is the beam (the
// work grows with it): 100K x 20 with beams {10..200}, 1.70 -> 1.40 ms.
// With a uniform beam of at least kPreMinBeam and more queries than resident
// waves, a pre-pass searches every query with beam kPreBeam for at most
// kPreLimit visits (k = 1), and the launch takes the queries in descending
// order of that search's best distance: a query whose neighborhood the short
// search finds far away tends to need many expansions (Spearman 0.49 against
// the visit count at Ls=250 on 10M x 200 MIPS FP8, where the distance to the
// start point gives 0.18).  Wall time per call, pre-pass, keys and sort
// included (main kernel alone in brackets), 10,000 queries: 10M x 200 MIPS
// FP8 at Ls=250 3.54 -> 3.36 ms (3.49 -> 2.97), BF16 at Ls=275 4.21 -> 4.15
// ms, FP32 at Ls=275 5.78 -> 5.80 ms; 10M x 128 uint8 at Ls=140 3.44 -> 3.31
// ms (3.40 -> 2.96).  Pre-pass, keys and sort cost about 0.35 ms whatever
// the beam, so shorter searches lose (FP8 at Ls=100 1.93 -> 2.05 ms, uint8
// at Ls=100 2.67 -> 2.71 ms): hence the threshold.
// DDG_ORDER=start orders uniform-beam launches by the distance to the start
// point instead (larger first; 10M x 200 MIPS FP8 unchanged), DDG_ORDER=0
// keeps the queries in their own order, DDG_ORDER_MIN_BEAM=<b> moves the
// pre-pass threshold.  Keys sort ascending and stably, so equal keys keep the
// queries' own order.  Results do not depend on the order.  Allocates: not
// for the allocation-free path.
constexpr uint32_t kPreBeam = 10, kPreLimit = 8, kPreMinBeam = 140;

__global__ void k_beam_order_keys(const uint32_t* beams, uint32_t beam_max, uint64_t nq, uint32_t* keys,
                                  uint32_t* vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint32_t b = beams[i];
  keys[i] = beam_max - (b < 1u ? 1u : (b < beam_max ? b : beam_max));
  vals[i] = (uint32_t)i;
}

// Keys of the pre-pass order: the best distance, larger first, as a 16-bit
// key that sorts ascending (the top half of the complemented
// order-preserving bits: sign, exponent and 7 mantissa bits; two sort
// passes instead of four).
__global__ void k_dist_order_keys(const float* d, uint64_t nq, uint32_t* keys, uint32_t* vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint32_t b = __float_as_uint(d[i]);
  keys[i] = (~(b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u))) >> 16;
  vals[i] = (uint32_t)i;
}

// The pre-pass search of the uniform-beam order: beam kPreBeam, k = 1, at
// most kPreLimit visits, its filter in LDS; best distances to dist (nq
// floats), ids to ids.  Its work counter and status word are F_PREWORK and
// F_PRESTATUS.  False when the short search's default shape would need a
// global filter (the order is then not taken).
static bool prepass_search(ddg_index* idx, int trav, const SearchArgs& a, uint32_t* ids, float* dist,
                           hipError_t* err) {
  SearchArgs p{};
  Launch pl{};
  *err = hipSuccess;
  if (!search_layout_cfg(idx, kPreBeam, a.nstarts, trav, nullptr, p, pl) || p.gtable_words) return false;
  p.base = a.base;
  p.graph = a.graph;
  p.queries = a.queries;
  p.qids = a.qids;
  p.starts = a.starts;
  p.out_ids = ids;
  p.out_dists = dist;
  p.work = a.work - F_WORK + F_PREWORK;
  p.overflow = a.work - F_WORK + F_PRESTATUS;
  p.n = a.n;
  p.nq = a.nq;
  p.limit = kPreLimit;
  p.cut = a.cut;
  p.d = a.d;
  p.ad = a.ad;
  p.R = a.R;
  p.nstarts = a.nstarts;
  p.k = 1;
  p.beam = kPreBeam;
  p.degree_limit = a.degree_limit;
  p.build_mode = 0;
  p.cut_trim = a.cut_trim;
  p.rerank = 1;
  pl.blocks = (uint32_t)std::min<uint64_t>(pl.blocks, (a.nq + pl.waves_per_block - 1) / pl.waves_per_block);
  *err = launch_fill_u32(p.work, 0u, 2, idx->stream);
  if (*err == hipSuccess) *err = launch_search(idx->dtype, idx->metric, trav, p, pl, idx->stream);
  return *err == hipSuccess;
}

static ddg_status order_queries(ddg_index* idx, int trav, const uint32_t* beams, uint32_t beam_max, uint64_t nq,
                                const Launch& l, SearchArgs& a) {
  a.order = nullptr;
  const char* e = getenv("DDG_ORDER");
  if (nq < 2 || (e && e[0] == '0')) return DDG_OK;
  const bool by_start = e && strcmp(e, "start") == 0;
  const bool more = nq > (uint64_t)l.blocks * l.waves_per_block;  // queries beyond the resident waves
  uint32_t min_beam = kPreMinBeam;
  const char* mb = getenv("DDG_ORDER_MIN_BEAM");
  if (mb && atoi(mb) > 0) min_beam = (uint32_t)atoi(mb);
  const bool by_pre = !beams && !by_start && more && a.beam >= min_beam && a.beam > kPreBeam;
  if (!beams && !(by_start && more) && !by_pre) return DDG_OK;
  ENSURE(B_OKEYS, nq * 4);
  ENSURE(B_OVALS, nq * 4);
  ENSURE(B_OKEYS2, nq * 4);
  ENSURE(B_OVALS2, nq * 4);
  ENSURE(B_OTMP, radix_tmp_bytes(nq) + 64);
  uint32_t kb = 16;
  if (beams) {
    hipLaunchKernelGGL(k_beam_order_keys, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, idx->stream, beams,
                       beam_max, nq, buf<uint32_t>(idx, B_OKEYS), buf<uint32_t>(idx, B_OVALS));
    HIPCHK(hipGetLastError());
    kb = 1;
    while (kb < 32 && (1ull << kb) <= beam_max) kb++;
  } else if (by_pre) {
    // the pre-pass writes into the sort's second buffers, which the sort
    // overwrites only after the keys are made from them
    hipError_t err;
    if (!prepass_search(idx, trav, a, buf<uint32_t>(idx, B_OKEYS2), buf<float>(idx, B_OVALS2), &err)) {
      HIPCHK(err);
      return DDG_OK;
    }
    hipLaunchKernelGGL(k_dist_order_keys, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, idx->stream,
                       buf<float>(idx, B_OVALS2), nq, buf<uint32_t>(idx, B_OKEYS), buf<uint32_t>(idx, B_OVALS));
    HIPCHK(hipGetLastError());
  } else {
    SearchArgs k = a;
    k.lds_per_wave = trav == DDG_TRAV_FP8E4M3 ? 2 * a.qrow_bytes : a.qrow_bytes;
    Launch kl{};
    kl.waves_per_block = 4;
    kl.lds_bytes = 4 * k.lds_per_wave;
    kl.blocks = (uint32_t)((nq + 3) / 4);
    HIPCHK(launch_start_keys(idx->dtype, idx->metric, trav, k, buf<uint32_t>(idx, B_OKEYS),
                             buf<uint32_t>(idx, B_OVALS), kl, idx->stream));
  }
  bool alt = false;
  HIPCHK(radix_sort_pairs(buf<uint32_t>(idx, B_OKEYS), buf<uint32_t>(idx, B_OVALS), buf<uint32_t>(idx, B_OKEYS2),
                          buf<uint32_t>(idx, B_OVALS2), nq, kb, buf<void>(idx, B_OTMP), idx->s.b[B_OTMP].cap,
                          idx->stream, &alt));
  a.order = alt ? buf<uint32_t>(idx, B_OVALS2) : buf<uint32_t>(idx, B_OVALS);
  return DDG_OK;
}

// One search launch over device-resident queries and outputs; the caller
// has checked every argument.
// `fixed` (nullable): a caller-provided scratch of fixed_bytes that takes the
// status words and the global filter slices instead of the handle's own
// buffers; the launch then allocates nothing.
static ddg_status search_launch(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                                const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                                const uint32_t* per_query_beam, uint32_t beam_max, uint32_t* out_ids,
                                float* out_dists, uint64_t* dist_cmps, uint32_t* visited, SearchExtra* ex,
                                unsigned char* fixed = nullptr, size_t fixed_bytes = 0) {
  SearchArgs a{};
  Launch l{};
  const int trav = (int)qp->trav;
  if (fixed) {
    if (trav != DDG_TRAV_NATIVE && !idx->trav_rows[trav])
      return fail(idx, DDG_EINVAL, "the reduced traversal copy does not exist: call ddg_prepare_traversal first");
  } else {
    ddg_status pst = prepare_trav(idx, trav, nullptr);
    if (pst != DDG_OK) return pst;
  }
  if (!search_layout(idx, beam_max, nstarts, trav, a, l))
    return fail(idx, DDG_EINVAL, "beam width too large for the LDS frontier");
  uint32_t* flags = nullptr;
  if (fixed) {
    flags = (uint32_t*)fixed;
  } else {
    ENSURE(B_FLAGS, kFlagBytes);
    flags = buf<uint32_t>(idx, B_FLAGS);
  }
  HIPCHK(launch_fill_u32(flags, 0u, 2, idx->stream));
  a.base = idx->rows;
  a.graph = idx->graph;
  a.queries = queries;
  a.qids = query_ids;
  a.starts = starts;
  a.per_query_beam = per_query_beam;
  a.out_ids = out_ids;
  a.out_dists = out_dists;
  a.dist_cmps = dist_cmps;
  a.visited_count = visited;
  a.work = flags;
  a.overflow = flags + 1;
  a.n = idx->n;
  a.nq = nq;
  a.limit = qp->limit ? qp->limit : idx->n;
  a.cut = qp->cut;
  a.d = idx->d;
  a.ad = idx->ad;
  a.R = idx->R;
  a.nstarts = nstarts;
  a.k = qp->k;
  a.beam = per_query_beam ? beam_max : qp->beam;  // per-query beams are clamped to beam_max
  a.degree_limit = qp->degree_limit ? std::min(qp->degree_limit, idx->R) : idx->R;
  a.build_mode = 0;
  a.cut_trim = idx->metric == DDG_L2 ? 1 : 0;
  a.rerank = qp->rerank;
  if (ex) {
    a.fr_ids = ex->fr_ids;
    a.fr_dists = ex->fr_dists;
    a.fr_len = ex->fr_len;
    a.fr_stride = ex->fr_stride;
    a.vis_ids = ex->vis_ids;
    a.vis_dists = ex->vis_dists;
    a.vcap = ex->vcap;
  }
  l.blocks = (uint32_t)std::min<uint64_t>(l.blocks, (nq + l.waves_per_block - 1) / l.waves_per_block);
  if (l.blocks == 0) return DDG_OK;
  if (fixed) {
    if (a.gtable_words) {
      // as many resident blocks as whole filter slices fit after the status
      // words; the launch shape does not change results
      const size_t slice = (size_t)a.gtable_words * 4, per_block = slice * l.waves_per_block;
      const size_t fit = (fixed_bytes - kFlagBytes) / per_block;
      if (fit == 0)
        return fail(idx, DDG_EINVAL, "beam too large for the allocation-free search: one block's visited filters "
                                     "exceed its scratch");
      l.blocks = (uint32_t)std::min<size_t>(l.blocks, fit);
      a.gtable = (uint32_t*)(fixed + kFlagBytes);
    }
  } else {
    ENSURE_GTABLE(a, l);
    ddg_status ost = order_queries(idx, trav, per_query_beam, beam_max, nq, l, a);
    if (ost != DDG_OK) return ost;
  }
  Timer t(idx, &idx->prof[0]);
  HIPCHK(launch_search(idx->dtype, idx->metric, trav, a, l, idx->stream));
  t.stop();
  uint32_t ov = 0;
  HIPCHK(hipMemcpy(&ov, flags + 1, 4, hipMemcpyDeviceToHost));
  if (ov & 2u) return fail(idx, DDG_ERANGE, "start point out of range");
  if (ex) ex->overflow = (ov & 1u) != 0;
  return DDG_OK;
}

// Search over device-resident queries/outputs (all pointers device memory).
ddg_status ddg_search_device(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                             const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                             const uint32_t* per_query_beam, uint32_t beam_max, uint32_t* out_ids, float* out_dists,
                             uint64_t* dist_cmps, uint32_t* visited) {
  // every host-side scalar is checked before anything is launched: the
  // kernel trusts nstarts (<= 64 seeds one lane each), the frontier capacity
  // (beam_max) and cut >= 1 (the exactness of the visited flag)
  if (!idx || !qp) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nq && (!queries || !out_ids || !starts)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nstarts == 0 || nstarts > 64) return fail(idx, DDG_EINVAL, "nstarts must be in [1, 64]");
  if (qp->trav != DDG_TRAV_NATIVE && qp->trav != DDG_TRAV_BF16 && qp->trav != DDG_TRAV_FP8E4M3)
    return fail(idx, DDG_EINVAL, "unknown traversal mode");
  if (qp->trav != DDG_TRAV_NATIVE && idx->dtype != DDG_F32)
    return fail(idx, DDG_EINVAL, "reduced-precision traversal needs a float32 index");
  if (qp->trav != DDG_TRAV_NATIVE && qp->rerank != 0 && qp->rerank < qp->k)
    return fail(idx, DDG_EINVAL, "rerank must be 0 (all) or at least k");
  if (beam_max == 0 || qp->k > beam_max)
    return fail(idx, DDG_EINVAL, "Error: beam search parameter Q smaller than k");  // beamSearch.h:327-331
  if (!per_query_beam && (qp->beam == 0 || qp->beam > beam_max || qp->k > qp->beam))
    return fail(idx, DDG_EINVAL, "beam must be in [k, beam_max]");
  if (qp->k > 0 && qp->cut < 1.0 && idx->metric == DDG_L2)
    return fail(idx, DDG_EINVAL, "cut below 1.0 is not supported");
  if (nq >= (1ull << 32)) return fail(idx, DDG_EINVAL, "nq must be below 2^32");
  return search_launch(idx, queries, nq, query_ids, starts, nstarts, qp, per_query_beam, beam_max, out_ids,
                       out_dists, dist_cmps, visited, nullptr);
}

ddg_status ddg_search_device_noalloc(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                                     const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                                     const uint32_t* per_query_beam, uint32_t beam_max, uint32_t* out_ids,
                                     float* out_dists, uint64_t* dist_cmps, uint32_t* visited) {
  if (!idx || !qp) return fail(idx, DDG_EINVAL, "NULL argument");
  if (!idx->noalloc_scratch) return fail(idx, DDG_EDEVICE, idx->noalloc_err);
  if (nq && (!queries || !out_ids || !starts)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nstarts == 0 || nstarts > 64) return fail(idx, DDG_EINVAL, "nstarts must be in [1, 64]");
  if (qp->trav != DDG_TRAV_NATIVE && qp->trav != DDG_TRAV_BF16 && qp->trav != DDG_TRAV_FP8E4M3)
    return fail(idx, DDG_EINVAL, "unknown traversal mode");
  if (qp->trav != DDG_TRAV_NATIVE && idx->dtype != DDG_F32)
    return fail(idx, DDG_EINVAL, "reduced-precision traversal needs a float32 index");
  if (qp->trav != DDG_TRAV_NATIVE && qp->rerank != 0 && qp->rerank < qp->k)
    return fail(idx, DDG_EINVAL, "rerank must be 0 (all) or at least k");
  if (beam_max == 0 || qp->k > beam_max)
    return fail(idx, DDG_EINVAL, "Error: beam search parameter Q smaller than k");  // beamSearch.h:327-331
  if (!per_query_beam && (qp->beam == 0 || qp->beam > beam_max || qp->k > qp->beam))
    return fail(idx, DDG_EINVAL, "beam must be in [k, beam_max]");
  if (qp->k > 0 && qp->cut < 1.0 && idx->metric == DDG_L2)
    return fail(idx, DDG_EINVAL, "cut below 1.0 is not supported");
  if (nq >= (1ull << 32)) return fail(idx, DDG_EINVAL, "nq must be below 2^32");
  return search_launch(idx, queries, nq, query_ids, starts, nstarts, qp, per_query_beam, beam_max, out_ids,
                       out_dists, dist_cmps, visited, nullptr, idx->noalloc_scratch, kNoAllocScratchBytes);
}

ddg_status ddg_beam_search(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                           const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp, uint32_t fcap,
                           uint32_t* frontier_ids, float* frontier_dists, uint32_t* frontier_len, uint32_t vcap,
                           uint32_t* visited_ids, float* visited_dists, uint32_t* visited_len, uint64_t* dist_cmps) {
  if (!idx || !qp || !starts) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nq && (!queries || !frontier_ids || !frontier_dists || !frontier_len || !visited_len))
    return fail(idx, DDG_EINVAL, "NULL argument");
  if (vcap && (!visited_ids || !visited_dists)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nstarts == 0 || nstarts > 64) return fail(idx, DDG_EINVAL, "nstarts must be in [1, 64]");
  if (qp->trav != DDG_TRAV_NATIVE) return fail(idx, DDG_EINVAL, "beam_search traverses the native rows only");
  if (qp->beam == 0) return fail(idx, DDG_EINVAL, "beam must be positive");
  if (fcap < std::max(qp->beam, nstarts)) return fail(idx, DDG_EINVAL, "fcap must be at least max(beam, nstarts)");
  if (qp->k > 0 && qp->cut < 1.0 && idx->metric == DDG_L2)
    return fail(idx, DDG_EINVAL, "cut below 1.0 is not supported");
  if (nq >= (1ull << 32)) return fail(idx, DDG_EINVAL, "nq must be below 2^32");
  for (uint32_t i = 0; i < nstarts; i++)
    if (starts[i] >= idx->n) return fail(idx, DDG_ERANGE, "start point out of range");
  if (nq == 0) return DDG_OK;
  for (double& p : idx->prof) p = 0;
  const size_t row_bytes = (size_t)idx->ad * idx->es;
  const uint32_t fs = std::max(qp->beam, nstarts);
  ENSURE(B_QUERIES, nq * row_bytes);
  ENSURE(B_STARTS, 256);
  ENSURE(B_CMPS, nq * 8);
  ENSURE(B_VISN, nq * 4);
  ENSURE(B_FR_IDS, nq * fs * 4);
  ENSURE(B_FR_D, nq * fs * 4);
  ENSURE(B_FR_LEN, nq * 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_QUERIES), queries, nq * row_bytes, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_pad_rows(buf<void>(idx, B_QUERIES), nq, idx->d * idx->es, (uint32_t)row_bytes, idx->stream));
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_STARTS), starts, nstarts * 4, hipMemcpyHostToDevice, idx->stream));
  uint32_t* dqid = nullptr;
  if (query_ids) {
    ENSURE(B_QIDS, nq * 4);
    dqid = buf<uint32_t>(idx, B_QIDS);
    HIPCHK(hipMemcpyAsync(dqid, query_ids, nq * 4, hipMemcpyHostToDevice, idx->stream));
  }
  // the visited lists are recorded into device slots of vdev entries; a list
  // that outgrows them reruns the batch with twice the room (the limit
  // bounds every list), so visited_len always holds the true lengths
  const uint64_t lim = qp->limit ? std::min<uint64_t>(qp->limit, idx->n) : idx->n;
  uint32_t vdev = (uint32_t)std::max<uint64_t>(std::min<uint64_t>(std::max<uint32_t>(vcap, 2 * qp->beam + 64), lim), 1);
  SearchExtra ex;
  for (;;) {
    ENSURE(B_VIS_IDS, nq * (size_t)vdev * 4);
    ENSURE(B_VIS_D, nq * (size_t)vdev * 4);
    ex.fr_ids = buf<uint32_t>(idx, B_FR_IDS);
    ex.fr_dists = buf<float>(idx, B_FR_D);
    ex.fr_len = buf<uint32_t>(idx, B_FR_LEN);
    ex.fr_stride = fs;
    ex.vis_ids = buf<uint32_t>(idx, B_VIS_IDS);
    ex.vis_dists = buf<float>(idx, B_VIS_D);
    ex.vcap = vdev;
    ddg_status st = search_launch(idx, buf<void>(idx, B_QUERIES), nq, dqid, buf<uint32_t>(idx, B_STARTS), nstarts,
                                  qp, nullptr, qp->beam, nullptr, nullptr, buf<uint64_t>(idx, B_CMPS),
                                  buf<uint32_t>(idx, B_VISN), &ex);
    if (st != DDG_OK) return st;
    if (!ex.overflow) break;
    if (vdev >= lim) return fail(idx, DDG_EDEVICE, "visited list longer than the visit limit");
    vdev = (uint32_t)std::min<uint64_t>(2ull * vdev, lim);
  }
  std::vector<uint32_t> fi((size_t)nq * fs), fl(nq), vi((size_t)nq * vdev), vl(nq);
  std::vector<float> fd((size_t)nq * fs), vd((size_t)nq * vdev);
  HIPCHK(hipMemcpy(fi.data(), buf<void>(idx, B_FR_IDS), fi.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(fd.data(), buf<void>(idx, B_FR_D), fd.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(fl.data(), buf<void>(idx, B_FR_LEN), nq * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vi.data(), buf<void>(idx, B_VIS_IDS), vi.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vd.data(), buf<void>(idx, B_VIS_D), vd.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vl.data(), buf<void>(idx, B_VISN), nq * 4, hipMemcpyDeviceToHost));
  if (dist_cmps) HIPCHK(hipMemcpy(dist_cmps, buf<void>(idx, B_CMPS), nq * 8, hipMemcpyDeviceToHost));
  bool short_cap = false;
  std::vector<std::pair<float, uint32_t>> tmp;
  for (uint64_t q = 0; q < nq; q++) {
    frontier_len[q] = fl[q];
    std::memcpy(frontier_ids + q * fcap, fi.data() + q * fs, fl[q] * 4);
    std::memcpy(frontier_dists + q * fcap, fd.data() + q * fs, fl[q] * 4);
    visited_len[q] = vl[q];
    if (vl[q] > vcap) { short_cap = true; continue; }
    // beam_search keeps visited sorted by (distance, id) (beamSearch.h:115-117)
    tmp.resize(vl[q]);
    for (uint32_t j = 0; j < vl[q]; j++) tmp[j] = {vd[q * vdev + j], vi[q * vdev + j]};
    std::sort(tmp.begin(), tmp.end(), [](const std::pair<float, uint32_t>& x, const std::pair<float, uint32_t>& y) {
      return x.first < y.first || (x.first == y.first && x.second < y.second);
    });
    for (uint32_t j = 0; j < vl[q]; j++) {
      visited_ids[q * vcap + j] = tmp[j].second;
      visited_dists[q * vcap + j] = tmp[j].first;
    }
  }
  if (short_cap) return fail(idx, DDG_EINVAL, "a visited list is longer than vcap (visited_len holds the lengths)");
  return DDG_OK;
}

ddg_status ddg_graph_set_rows(ddg_index* idx, const uint32_t* ids, uint64_t m, const uint32_t* rows) {
  if (!idx || (m && (!ids || !rows))) return fail(idx, DDG_EINVAL, "NULL argument");
  const uint32_t R1 = idx->R + 1;
  for (uint64_t i = 0; i < m; i++) {
    if (ids[i] >= idx->n) return fail(idx, DDG_ERANGE, "row id out of range");
    const uint32_t* r = rows + i * (uint64_t)R1;
    if (r[0] > idx->R) return fail(idx, DDG_EDEGREE, "graph row degree exceeds R");
    for (uint32_t j = 0; j < r[0]; j++)
      if (r[1 + j] >= idx->n) return fail(idx, DDG_ERANGE, "graph row holds an id >= n");
  }
  if (m == 0) return DDG_OK;
  ENSURE(B_ROW_IDS, m * 4);
  ENSURE(B_ROWS, m * (size_t)R1 * 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_ROW_IDS), ids, m * 4, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_ROWS), rows, m * (size_t)R1 * 4, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_copy_rows(idx->graph, R1, buf<uint32_t>(idx, B_ROW_IDS), m, buf<uint32_t>(idx, B_ROWS), 1,
                          idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  return DDG_OK;
}

ddg_status ddg_graph_get_rows(ddg_index* idx, const uint32_t* ids, uint64_t m, uint32_t* rows) {
  if (!idx || (m && (!ids || !rows))) return fail(idx, DDG_EINVAL, "NULL argument");
  for (uint64_t i = 0; i < m; i++)
    if (ids[i] >= idx->n) return fail(idx, DDG_ERANGE, "row id out of range");
  if (m == 0) return DDG_OK;
  const uint32_t R1 = idx->R + 1;
  ENSURE(B_ROW_IDS, m * 4);
  ENSURE(B_ROWS, m * (size_t)R1 * 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_ROW_IDS), ids, m * 4, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_copy_rows(idx->graph, R1, buf<uint32_t>(idx, B_ROW_IDS), m, buf<uint32_t>(idx, B_ROWS), 0,
                          idx->stream));
  HIPCHK(hipMemcpyAsync(rows, buf<void>(idx, B_ROWS), m * (size_t)R1 * 4, hipMemcpyDeviceToHost, idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  return DDG_OK;
}

ddg_status ddg_search(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                      const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                      const uint32_t* per_query_beam, uint32_t* out_ids, float* out_dists, uint64_t* dist_cmps,
                      uint32_t* visited) {
  if (!idx || !qp || (nq && (!queries || !out_ids)) || !starts) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nstarts == 0 || nstarts > 64) return fail(idx, DDG_EINVAL, "nstarts must be in [1, 64]");
  if (qp->trav != DDG_TRAV_NATIVE && qp->trav != DDG_TRAV_BF16 && qp->trav != DDG_TRAV_FP8E4M3)
    return fail(idx, DDG_EINVAL, "unknown traversal mode");
  if (qp->trav != DDG_TRAV_NATIVE && idx->dtype != DDG_F32)
    return fail(idx, DDG_EINVAL, "reduced-precision traversal needs a float32 index");
  if (qp->trav != DDG_TRAV_NATIVE && qp->rerank != 0 && qp->rerank < qp->k)
    return fail(idx, DDG_EINVAL, "rerank must be 0 (all) or at least k");
  for (uint32_t i = 0; i < nstarts; i++)
    if (starts[i] >= idx->n) return fail(idx, DDG_ERANGE, "start point out of range");
  uint32_t beam_max = qp->beam;
  if (per_query_beam) {
    beam_max = 0;
    for (uint64_t i = 0; i < nq; i++) {
      if (per_query_beam[i] < qp->k || per_query_beam[i] == 0)
        return fail(idx, DDG_EINVAL, "Error: beam search parameter Q smaller than k");
      beam_max = std::max(beam_max, per_query_beam[i]);
    }
  } else if (qp->k > qp->beam || qp->beam == 0) {
    return fail(idx, DDG_EINVAL, "Error: beam search parameter Q smaller than k");  // beamSearch.h:327-331
  }
  if (qp->k > 0 && qp->cut < 1.0 && idx->metric == DDG_L2)
    return fail(idx, DDG_EINVAL, "cut below 1.0 is not supported");
  for (double& p : idx->prof) p = 0;
  const size_t row_bytes = (size_t)idx->ad * idx->es;
  ENSURE(B_QUERIES, std::max<uint64_t>(nq, 1) * row_bytes);
  ENSURE(B_STARTS, 256);
  ENSURE(B_OUT_IDS, std::max<uint64_t>(nq * qp->k, 1) * 4);
  ENSURE(B_OUT_D, std::max<uint64_t>(nq * qp->k, 1) * 4);
  ENSURE(B_CMPS, std::max<uint64_t>(nq, 1) * 8);
  ENSURE(B_VISN, std::max<uint64_t>(nq, 1) * 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_QUERIES), queries, nq * row_bytes, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_pad_rows(buf<void>(idx, B_QUERIES), nq, idx->d * idx->es, (uint32_t)row_bytes, idx->stream));
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_STARTS), starts, nstarts * 4, hipMemcpyHostToDevice, idx->stream));
  uint32_t* dqb = nullptr;
  uint32_t* dqid = nullptr;
  if (per_query_beam) {
    ENSURE(B_QBEAM, nq * 4);
    dqb = buf<uint32_t>(idx, B_QBEAM);
    HIPCHK(hipMemcpyAsync(dqb, per_query_beam, nq * 4, hipMemcpyHostToDevice, idx->stream));
  }
  if (query_ids) {
    ENSURE(B_QIDS, nq * 4);
    dqid = buf<uint32_t>(idx, B_QIDS);
    HIPCHK(hipMemcpyAsync(dqid, query_ids, nq * 4, hipMemcpyHostToDevice, idx->stream));
  }
  ddg_status st = ddg_search_device(idx, buf<void>(idx, B_QUERIES), nq, dqid, buf<uint32_t>(idx, B_STARTS), nstarts, qp,
                                    dqb, beam_max, buf<uint32_t>(idx, B_OUT_IDS), buf<float>(idx, B_OUT_D),
                                    buf<uint64_t>(idx, B_CMPS), buf<uint32_t>(idx, B_VISN));
  if (st != DDG_OK) return st;
  HIPCHK(hipMemcpyAsync(out_ids, buf<void>(idx, B_OUT_IDS), nq * qp->k * 4, hipMemcpyDeviceToHost, idx->stream));
  if (out_dists)
    HIPCHK(hipMemcpyAsync(out_dists, buf<void>(idx, B_OUT_D), nq * qp->k * 4, hipMemcpyDeviceToHost, idx->stream));
  std::vector<uint64_t> cm(nq);
  std::vector<uint32_t> vs(nq);
  HIPCHK(hipMemcpyAsync(cm.data(), buf<void>(idx, B_CMPS), nq * 8, hipMemcpyDeviceToHost, idx->stream));
  HIPCHK(hipMemcpyAsync(vs.data(), buf<void>(idx, B_VISN), nq * 4, hipMemcpyDeviceToHost, idx->stream));
  HIPCHK(hipStreamSynchronize(idx->stream));
  double tc = 0, tv = 0;
  for (uint64_t i = 0; i < nq; i++) { tc += (double)cm[i]; tv += (double)vs[i]; }
  const double s_trav = trav_esize(idx, (int)qp->trav);
  idx->prof[6] = tc * idx->d * s_trav + tv * 4.0 * (1 + idx->R) + (double)nq * idx->d * idx->es;
  if (qp->trav != DDG_TRAV_NATIVE) {  // the FP32 re-rank reads up to min(rerank, beam) full rows per query
    const uint32_t nr = (qp->rerank == 0 || qp->rerank > beam_max) ? beam_max : qp->rerank;
    idx->prof[6] += (double)nq * nr * idx->d * 4.0;
  }
  idx->prof[7] = tc;
  idx->prof[8] = tv;
  if (dist_cmps) std::memcpy(dist_cmps, cm.data(), nq * 8);
  if (visited) std::memcpy(visited, vs.data(), nq * 4);
  return DDG_OK;
}

ddg_status ddg_autotune_search(ddg_index* idx, const void* queries, uint64_t nq, const ddg_query_params* qp,
                               uint32_t* out4, double* ms2) {
  if (!idx || !qp || (nq && !queries)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nq == 0 || qp->beam == 0 || qp->k > qp->beam) return fail(idx, DDG_EINVAL, "bad query set or beam");
  ddg_status st = prepare_trav(idx, (int)qp->trav, nullptr);
  if (st != DDG_OK) return st;
  const size_t row_bytes = (size_t)idx->ad * idx->es;
  ENSURE(B_QUERIES, nq * row_bytes);
  ENSURE(B_STARTS, 256);
  ENSURE(B_OUT_IDS, nq * qp->k * 4 + 4);
  ENSURE(B_VC, nq * qp->k * 4 + 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_QUERIES), queries, nq * row_bytes, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_pad_rows(buf<void>(idx, B_QUERIES), nq, idx->d * idx->es, (uint32_t)row_bytes, idx->stream));
  HIPCHK(launch_fill_u32(buf<uint32_t>(idx, B_STARTS), 0u, 1, idx->stream));
  const uint32_t bits = filter_bits(qp->beam);
  const uint32_t key = tune_key((int)qp->trav, bits);
  // drop an earlier choice so the default shape is the reference point
  for (size_t i = 0; i < idx->tuned.size(); i++)
    if (idx->tuned[i].key == key) { idx->tuned.erase(idx->tuned.begin() + (long)i); break; }
  // a warm-up launch, then `reps` timed ones: their median
  auto time_shape = [&](const ddg_index::Tuned* cfg, uint32_t* out_ids, double* ms, int reps = 1) -> ddg_status {
    idx->cfg_override = cfg;
    std::vector<double> ts;
    for (int rep = 0; rep <= reps; rep++) {
      const double before = idx->prof[0];
      ddg_status s2 = ddg_search_device(idx, buf<void>(idx, B_QUERIES), nq, nullptr, buf<uint32_t>(idx, B_STARTS), 1,
                                        qp, nullptr, qp->beam, out_ids, nullptr, nullptr, nullptr);
      if (s2 != DDG_OK) { idx->cfg_override = nullptr; return s2; }
      if (rep > 0) ts.push_back(idx->prof[0] - before);
    }
    idx->cfg_override = nullptr;
    std::sort(ts.begin(), ts.end());
    *ms = ts[ts.size() / 2];
    return DDG_OK;
  };
  double base_ms = 0.0;
  st = time_shape(nullptr, buf<uint32_t>(idx, B_OUT_IDS), &base_ms);
  if (st != DDG_OK) return st;
  std::vector<uint32_t> ref(nq * qp->k), got(nq * qp->k);
  HIPCHK(hipMemcpy(ref.data(), buf<void>(idx, B_OUT_IDS), ref.size() * 4, hipMemcpyDeviceToHost));
  ddg_index::Tuned best{key, global_table(bits) ? 1 : 0, 0, 0, 0};
  double best_ms = base_ms;
  uint32_t tried = 0;
  std::vector<std::pair<double, ddg_index::Tuned>> timed;  // every shape's single timing
  for (int gtab = 0; gtab <= 1; gtab++) {
    if (!gtab && (4u << bits) > 64u * 1024u) continue;  // LDS filters above 64 KB leave too few waves
    for (uint32_t wpb : {1u, 2u, 4u})
      for (uint32_t bpc : {1u, 2u, 3u, 4u, 6u, 8u}) {
        if (wpb * bpc > 32) continue;
        ddg_index::Tuned c{key, gtab, wpb, bpc, 0};
        SearchArgs a{};
        Launch l{};
        idx->cfg_override = &c;
        const bool fits = search_layout(idx, qp->beam, 1, (int)qp->trav, a, l);
        idx->cfg_override = nullptr;
        if (!fits) continue;
        double ms = 0.0;
        st = time_shape(&c, buf<uint32_t>(idx, B_VC), &ms);
        if (st != DDG_OK) return st;
        tried++;
        HIPCHK(hipMemcpy(got.data(), buf<void>(idx, B_VC), got.size() * 4, hipMemcpyDeviceToHost));
        if (got != ref) return fail(idx, DDG_EDEVICE, "autotuner: a launch shape changed the search result");
        timed.push_back({ms, c});
        if (ms < best_ms) { best_ms = ms; best = c; }
      }
  }
  // one launch each leaves a few percent of noise between close shapes: the
  // three fastest and the default are timed again over five launches, and
  // the smallest median wins
  if (!timed.empty()) {
    std::stable_sort(timed.begin(), timed.end(),
                     [](const std::pair<double, ddg_index::Tuned>& x, const std::pair<double, ddg_index::Tuned>& y) {
                       return x.first < y.first;
                     });
    st = time_shape(nullptr, buf<uint32_t>(idx, B_VC), &best_ms, 5);
    if (st != DDG_OK) return st;
    base_ms = best_ms;
    best = ddg_index::Tuned{key, global_table(bits) ? 1 : 0, 0, 0, 0};
    for (size_t i = 0; i < timed.size() && i < 3; i++) {
      double ms = 0.0;
      st = time_shape(&timed[i].second, buf<uint32_t>(idx, B_VC), &ms, 5);
      if (st != DDG_OK) return st;
      if (ms < best_ms) { best_ms = ms; best = timed[i].second; }
    }
  }
  // long rows: with the best shape, try gathering 2 / 4 / 8 / as many rows as
  // one load batch holds into LDS per round trip
  const uint32_t C = qp->trav == DDG_TRAV_NATIVE ? (idx->ad * idx->es) / 16 : (idx->d * trav_esize(idx, (int)qp->trav) + 15) / 16;
  if (C > 8 && best.wpb) {
    const uint32_t smax = std::max<uint32_t>(1, std::min<uint32_t>(64, 512 / C));
    std::vector<uint32_t> stages;
    for (uint32_t sr : {2u, 4u, 8u, smax})
      if (sr <= smax && std::find(stages.begin(), stages.end(), sr) == stages.end()) stages.push_back(sr);
    for (uint32_t sr : stages) {
      ddg_index::Tuned c = best;
      c.stage = sr;
      SearchArgs a{};
      Launch l{};
      idx->cfg_override = &c;
      const bool fits = search_layout(idx, qp->beam, 1, (int)qp->trav, a, l);
      idx->cfg_override = nullptr;
      if (!fits) continue;
      double ms = 0.0;
      st = time_shape(&c, buf<uint32_t>(idx, B_VC), &ms, 3);
      if (st != DDG_OK) return st;
      tried++;
      HIPCHK(hipMemcpy(got.data(), buf<void>(idx, B_VC), got.size() * 4, hipMemcpyDeviceToHost));
      if (got != ref) return fail(idx, DDG_EDEVICE, "autotuner: a launch shape changed the search result");
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
  }
  if (best.wpb) {
    idx->tuned.push_back(best);
    process_tune_store(idx, best);
  }
  if (out4) {
    SearchArgs a{};
    Launch l{};
    search_layout(idx, qp->beam, 1, (int)qp->trav, a, l);
    out4[0] = (a.gtable_words ? 1u : 0u) | (a.stage_rows << 8);
    out4[1] = l.waves_per_block;
    out4[2] = l.blocks / idx->num_cu;
    out4[3] = tried;
  }
  if (ms2) { ms2[0] = base_ms; ms2[1] = best_ms; }
  return DDG_OK;
}

ddg_status ddg_search_shape(ddg_index* idx, const ddg_query_params* qp, uint32_t* out4) {
  if (!idx || !qp || !out4) return fail(idx, DDG_EINVAL, "NULL argument");
  if (qp->beam == 0) return fail(idx, DDG_EINVAL, "bad beam");
  SearchArgs a{};
  Launch l{};
  uint32_t src = 0;
  if (!search_layout(idx, qp->beam, 1, (int)qp->trav, a, l, true, &src))
    return fail(idx, DDG_EINVAL, "beam width too large for the LDS frontier");
  out4[0] = (a.gtable_words ? 1u : 0u) | (a.stage_rows << 8);
  out4[1] = l.waves_per_block;
  out4[2] = l.blocks / idx->num_cu;
  out4[3] = src;
  return DDG_OK;
}

ddg_status ddg_prepare_traversal(ddg_index* idx, ddg_trav trav, double* seconds) {
  if (!idx) return fail(idx, DDG_EINVAL, "NULL argument");
  return prepare_trav(idx, (int)trav, seconds);
}

ddg_status ddg_distances(ddg_index* idx, ddg_trav trav, const void* queries, uint64_t nq, const uint32_t* ids,
                         uint32_t m, float* out) {
  if (!idx || (nq && m && (!queries || !ids || !out))) return fail(idx, DDG_EINVAL, "NULL argument");
  if (trav != DDG_TRAV_NATIVE && trav != DDG_TRAV_BF16 && trav != DDG_TRAV_FP8E4M3)
    return fail(idx, DDG_EINVAL, "unknown traversal mode");
  if (nq == 0 || m == 0) return DDG_OK;
  if (nq > (1ull << 31)) return fail(idx, DDG_EINVAL, "too many queries");
  for (uint64_t i = 0; i < nq * m; i++)
    if (ids[i] >= idx->n) return fail(idx, DDG_ERANGE, "point id out of range");
  ddg_status st = prepare_trav(idx, (int)trav, nullptr);
  if (st != DDG_OK) return st;
  SearchArgs a{};
  const uint32_t qbytes = trav_layout(idx, (int)trav, a);
  const size_t row_bytes = (size_t)idx->ad * idx->es;
  ENSURE(B_QUERIES, nq * row_bytes);
  ENSURE(B_VIS_IDS, nq * (size_t)m * 4);
  ENSURE(B_VIS_D, nq * (size_t)m * 4);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_QUERIES), queries, nq * row_bytes, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_VIS_IDS), ids, nq * (size_t)m * 4, hipMemcpyHostToDevice, idx->stream));
  a.base = idx->rows;
  a.queries = buf<void>(idx, B_QUERIES);
  a.n = idx->n;
  a.nq = nq;
  a.d = idx->d;
  a.ad = idx->ad;
  a.lds_per_wave = qbytes;
  Launch l{};
  l.waves_per_block = 4;
  l.lds_bytes = 4 * qbytes;
  l.blocks = (uint32_t)((nq + 3) / 4);
  Timer t(idx, &idx->prof[0]);
  HIPCHK(launch_trav_dist(idx->dtype, idx->metric, (int)trav, a, buf<uint32_t>(idx, B_VIS_IDS), m,
                          buf<float>(idx, B_VIS_D), l, idx->stream));
  t.stop();
  HIPCHK(hipMemcpy(out, buf<void>(idx, B_VIS_D), nq * (size_t)m * 4, hipMemcpyDeviceToHost));
  return DDG_OK;
}

ddg_status ddg_insert_search(ddg_index* idx, const uint32_t* ids, uint64_t m, uint32_t L, uint32_t start_point,
                             uint32_t vcap, uint32_t* vis_ids, float* vis_dists, uint32_t* vis_len,
                             uint64_t* dist_cmps) {
  if (!idx || (m && (!ids || !vis_ids || !vis_dists || !vis_len))) return fail(idx, DDG_EINVAL, "NULL argument");
  if (L == 0 || vcap == 0) return fail(idx, DDG_EINVAL, "bad L/vcap");
  if (start_point >= idx->n) return fail(idx, DDG_ERANGE, "start point out of range");
  for (uint64_t i = 0; i < m; i++)
    if (ids[i] >= idx->n) return fail(idx, DDG_ERANGE, "point id out of range");
  if (m == 0) return DDG_OK;
  for (double& p : idx->prof) p = 0;
  ENSURE(B_INSERTS, m * 4);
  HIPCHK(hipMemcpy(buf<uint32_t>(idx, B_INSERTS), ids, m * 4, hipMemcpyHostToDevice));
  uint32_t cap = vcap;
  ddg_status st = run_insert_search(idx, buf<uint32_t>(idx, B_INSERTS), m, L, start_point, &cap);
  if (st != DDG_OK) return st;
  std::vector<uint32_t> vi((size_t)m * cap);
  std::vector<float> vd((size_t)m * cap);
  HIPCHK(hipMemcpy(vi.data(), buf<void>(idx, B_VIS_IDS), (size_t)m * cap * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vd.data(), buf<void>(idx, B_VIS_D), (size_t)m * cap * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(vis_len, buf<void>(idx, B_VISN), m * 4, hipMemcpyDeviceToHost));
  if (dist_cmps) HIPCHK(hipMemcpy(dist_cmps, buf<void>(idx, B_CMPS), m * 8, hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < m; i++) {
    if (vis_len[i] > vcap) return fail(idx, DDG_EINVAL, "visited list longer than vcap");
    std::memcpy(vis_ids + i * vcap, vi.data() + i * cap, vis_len[i] * 4);
    std::memcpy(vis_dists + i * vcap, vd.data() + i * cap, vis_len[i] * 4);
  }
  return DDG_OK;
}

// Test hook (not part of include/ddg.h): device bytes of the handle's own
// scratch buffers (the allocation-free entry point must leave them at zero).
uint64_t ddg_internal_buffer_bytes(const ddg_index* idx) {
  uint64_t t = 0;
  if (idx)
    for (const auto& b : idx->s.b) t += b.p ? b.cap : 0;
  return t;
}

// Test hook (not part of include/ddg.h): records a search launch shape for
// (trav, beam) as the autotuner would, so tests can run every shape.
ddg_status ddg_internal_set_shape(ddg_index* idx, int trav, uint32_t beam, int gtab, uint32_t wpb, uint32_t bpc,
                                  uint32_t stage) {
  if (!idx || wpb == 0 || bpc == 0 || wpb > 4 || wpb * bpc > 32) return fail(idx, DDG_EINVAL, "bad shape");
  const uint32_t key = tune_key(trav, filter_bits(beam));
  for (size_t i = 0; i < idx->tuned.size(); i++)
    if (idx->tuned[i].key == key) { idx->tuned.erase(idx->tuned.begin() + (long)i); break; }
  idx->tuned.push_back(ddg_index::Tuned{key, gtab, wpb, bpc, stage});
  return DDG_OK;
}

// Test hook (not part of include/ddg.h): one prune launch in any of the
// modes the build uses — fixed-stride candidate lists (a_off == NULL) or CSR,
// forward (out_ids / out_len) or reverse (write_graph, append_if_fits).
ddg_status ddg_internal_prune(ddg_index* idx, uint64_t np, const uint32_t* pts, const uint32_t* a_ids,
                              const float* a_dists, const uint32_t* a_len, uint32_t a_stride,
                              const uint64_t* a_off, uint64_t total_ids, double alpha, uint32_t Rsel, int add,
                              int write_graph, int append_if_fits, uint32_t max_cand, uint32_t* out_ids,
                              uint32_t* out_len) {
  if (!idx || np == 0) return fail(idx, DDG_EINVAL, "bad arguments");
  ENSURE(B_QIDS, np * 4);
  ENSURE(B_PRUNE_IDS, std::max<uint64_t>(total_ids, 1) * 4);
  ENSURE(B_PRUNE_D, std::max<uint64_t>(total_ids, 1) * 4);
  ENSURE(B_PRUNE_OFF, (np + 1) * 8);
  ENSURE(B_VISN, np * 4);
  ENSURE(B_NEW_IDS, np * Rsel * 4);
  ENSURE(B_NEW_LEN, np * 4);
  HIPCHK(hipMemcpy(buf<void>(idx, B_QIDS), pts, np * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(buf<void>(idx, B_PRUNE_IDS), a_ids, total_ids * 4, hipMemcpyHostToDevice));
  if (a_dists) HIPCHK(hipMemcpy(buf<void>(idx, B_PRUNE_D), a_dists, total_ids * 4, hipMemcpyHostToDevice));
  if (a_off) HIPCHK(hipMemcpy(buf<void>(idx, B_PRUNE_OFF), a_off, (np + 1) * 8, hipMemcpyHostToDevice));
  if (a_len) HIPCHK(hipMemcpy(buf<void>(idx, B_VISN), a_len, np * 4, hipMemcpyHostToDevice));
  PruneArgs a{};
  a.pts = buf<uint32_t>(idx, B_QIDS);
  a.a_ids = buf<uint32_t>(idx, B_PRUNE_IDS);
  a.a_dists = a_dists ? buf<float>(idx, B_PRUNE_D) : nullptr;
  a.a_off = a_off ? buf<uint64_t>(idx, B_PRUNE_OFF) : nullptr;
  a.a_len = buf<uint32_t>(idx, B_VISN);
  a.a_stride = a_stride;
  a.out_ids = buf<uint32_t>(idx, B_NEW_IDS);
  a.out_len = buf<uint32_t>(idx, B_NEW_LEN);
  a.alpha = alpha;
  a.Rsel = Rsel;
  a.add = add ? 1 : 0;
  a.write_graph = write_graph ? 1 : 0;
  a.append_if_fits = append_if_fits ? 1 : 0;
  uint64_t total = 0;
  if (a_off) total = a_off[np] - a_off[0];
  else for (uint64_t t = 0; t < np; t++) total += a_len[t];
  ddg_status st = run_prune(idx, a, np, max_cand, total, &idx->prof[1]);
  if (st != DDG_OK) return st;
  if (!write_graph) {
    HIPCHK(hipMemcpy(out_ids, buf<void>(idx, B_NEW_IDS), np * Rsel * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out_len, buf<void>(idx, B_NEW_LEN), np * 4, hipMemcpyDeviceToHost));
  }
  return DDG_OK;
}

// Test hook (not part of include/ddg.h): the reverse-edge grouping passes
// alone.  Sorts n (key, value) pairs stably by key on the device and returns
// the sorted pairs, the group keys / offsets and the largest group.
ddg_status ddg_internal_group_pairs(ddg_index* idx, const uint32_t* keys, const uint32_t* vals, uint64_t n,
                                    uint32_t* out_keys, uint32_t* out_vals, uint64_t* ngroups,
                                    uint32_t* group_keys, uint64_t* group_off, uint32_t* maxg) {
  if (!idx || n == 0) return fail(idx, DDG_EINVAL, "bad arguments");
  for (uint64_t i = 0; i < n; i++)
    if (keys[i] >= idx->n) return fail(idx, DDG_ERANGE, "key out of range");
  ENSURE(B_KEYS, n * 4);
  ENSURE(B_VALS, n * 4);
  ENSURE(B_KEYS2, n * 4);
  ENSURE(B_VALS2, n * 4);
  ENSURE(B_TMP, std::max(scan_tmp_bytes(n), radix_tmp_bytes(n)) + 64);
  ENSURE(B_GFLAGS, n * 4);
  ENSURE(B_GOFFS, n * 8);
  ENSURE(B_GKEYS, n * 4);
  ENSURE(B_GOFF, (n + 1) * 8);
  ENSURE(B_FLAGS, kFlagBytes);
  uint32_t* flags = buf<uint32_t>(idx, B_FLAGS);
  HIPCHK(hipMemcpy(buf<void>(idx, B_KEYS), keys, n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(buf<void>(idx, B_VALS), vals, n * 4, hipMemcpyHostToDevice));
  bool in_alt = false;
  HIPCHK(radix_sort_pairs(buf<uint32_t>(idx, B_KEYS), buf<uint32_t>(idx, B_VALS), buf<uint32_t>(idx, B_KEYS2),
                          buf<uint32_t>(idx, B_VALS2), n, key_bits_for(idx->n), buf<void>(idx, B_TMP),
                          idx->s.b[B_TMP].cap, idx->stream, &in_alt));
  uint32_t* skeys = in_alt ? buf<uint32_t>(idx, B_KEYS2) : buf<uint32_t>(idx, B_KEYS);
  uint32_t* svals = in_alt ? buf<uint32_t>(idx, B_VALS2) : buf<uint32_t>(idx, B_VALS);
  uint64_t* dng = (uint64_t*)(flags + 10);
  HIPCHK(launch_group_starts(skeys, n, buf<uint32_t>(idx, B_GFLAGS), buf<uint64_t>(idx, B_GOFFS), dng,
                             buf<uint32_t>(idx, B_GKEYS), buf<uint64_t>(idx, B_GOFF), buf<void>(idx, B_TMP),
                             idx->s.b[B_TMP].cap, idx->stream));
  HIPCHK(launch_fill_u32(flags + 12, 0u, 1, idx->stream));
  hipLaunchKernelGGL(k_max_group, dim3((n + 255) / 256), dim3(256), 0, idx->stream, buf<uint64_t>(idx, B_GOFF), dng,
                     flags + 12);
  HIPCHK(hipStreamSynchronize(idx->stream));
  HIPCHK(hipMemcpy(out_keys, skeys, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(out_vals, svals, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ngroups, dng, 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(group_keys, buf<void>(idx, B_GKEYS), *ngroups * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(group_off, buf<void>(idx, B_GOFF), (*ngroups + 1) * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(maxg, flags + 12, 4, hipMemcpyDeviceToHost));
  return DDG_OK;
}

ddg_status ddg_quantized_mips_distances(ddg_dtype code_dtype, uint32_t d, float max_coord, float min_coord,
                                        const void* queries, int queries_quantized, uint64_t nq, uint32_t q_stride,
                                        const void* codes, uint64_t n, uint32_t c_stride, float* out) {
  ddg_index* idx = nullptr;
  if (code_dtype != DDG_U8 && code_dtype != DDG_I8) return fail(idx, DDG_EINVAL, "codes must be uint8 or int8");
  if (d == 0 || q_stride < d || c_stride < d) return fail(idx, DDG_EINVAL, "bad dimension or stride");
  if (nq && n && (!queries || !codes || !out)) return fail(idx, DDG_EINVAL, "NULL argument");
  if (nq == 0 || n == 0) return DDG_OK;
  if (nq > 65535) return fail(idx, DDG_EINVAL, "at most 65535 queries per call");
  if (d > 16384) return fail(idx, DDG_EINVAL, "d above 16384 is not supported");
  // mips_point.h:114-119: maxval = (float)(T)(2^(8 sizeof T) - 1), in float
  const float maxval = code_dtype == DDG_U8 ? (float)(uint8_t)255u : (float)(int8_t)(uint8_t)255u;
  const float delta = max_coord - min_coord;
  const float mult = delta / maxval;
  const size_t qbytes = nq * (size_t)q_stride * (queries_quantized ? 1 : 4), cbytes = n * (size_t)c_stride;
  void *dq = nullptr, *dc = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&dq, qbytes);
  if (e == hipSuccess) e = hipMalloc(&dc, cbytes);
  if (e == hipSuccess) e = hipMalloc(&dout, nq * n * 4);
  if (e == hipSuccess) e = hipMemcpy(dq, queries, qbytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dc, codes, cbytes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_quantized_mips(code_dtype, dq, queries_quantized ? 1 : 0, nq, q_stride, dc, n, c_stride, d, mult,
                              min_coord, (float*)dout, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out, dout, nq * n * 4, hipMemcpyDeviceToHost);
  if (dq) (void)hipFree(dq);
  if (dc) (void)hipFree(dc);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess)
    return fail(idx, e == hipErrorOutOfMemory ? DDG_EOOM : DDG_EDEVICE,
                std::string("quantized inner products: ") + hipGetErrorString(e));
  return DDG_OK;
}

ddg_status ddg_ground_truth(ddg_index* idx, const void* queries, uint64_t nq, uint32_t k, uint32_t* ids,
                            double* dists) {
  if (!idx || (nq && (!queries || !ids))) return fail(idx, DDG_EINVAL, "NULL argument");
  if (k == 0 || k > 100 || k > idx->n) return fail(idx, DDG_EINVAL, "k must be in [1, min(100, n)]");
  if (nq == 0) return DDG_OK;
  for (double& p : idx->prof) p = 0;
  const size_t row_bytes = (size_t)idx->ad * idx->es;
  ENSURE(B_QUERIES, nq * row_bytes);
  ENSURE(B_OUT_IDS, nq * k * 4);
  ENSURE(B_GT_D, nq * k * 8);
  const size_t tmp_bytes = ground_truth_tmp_bytes(idx->dtype, idx->n, nq, idx->d, k);
  ENSURE(B_TMP, tmp_bytes);
  HIPCHK(hipMemcpyAsync(buf<void>(idx, B_QUERIES), queries, nq * row_bytes, hipMemcpyHostToDevice, idx->stream));
  HIPCHK(launch_pad_rows(buf<void>(idx, B_QUERIES), nq, idx->d * idx->es, (uint32_t)row_bytes, idx->stream));
  Timer t(idx, &idx->prof[5]);
  HIPCHK(launch_ground_truth(idx->dtype, idx->metric, idx->rows, idx->n, buf<void>(idx, B_QUERIES), nq, idx->d,
                             idx->ad, k, buf<uint32_t>(idx, B_OUT_IDS), buf<double>(idx, B_GT_D),
                             buf<void>(idx, B_TMP), tmp_bytes, idx->stream));
  t.stop();
  HIPCHK(hipMemcpy(ids, buf<void>(idx, B_OUT_IDS), nq * k * 4, hipMemcpyDeviceToHost));
  if (dists) HIPCHK(hipMemcpy(dists, buf<void>(idx, B_GT_D), nq * k * 8, hipMemcpyDeviceToHost));
  return DDG_OK;
}

}  // extern "C"
