// Internal interfaces between the C-ABI layer (ddg_api.hip) and the kernel
// launchers.  Not installed; the public surface is include/ddg.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace ddg {

// ------------------------------------------------------------------ search

// A lane's share of a reduced (BF16 / FP8) row in a search: at most
// kLaneChunks 16-byte chunks for rows up to 512 bytes (split over up to 8
// lanes).
constexpr uint32_t kLaneChunks = 4;

struct SearchArgs {
  const void* base;              // n x ad rows
  const uint32_t* graph;         // n x (R+1)
  const void* queries;           // nq x ad rows (query mode)
  const uint32_t* qids;          // build mode: point id per task; query mode: p.id() (nullable)
  const uint32_t* starts;
  const uint32_t* per_query_beam;  // nullable
  const uint32_t* order;         // nullable: task t serves query order[t] (longest expected first)
  uint32_t* out_ids;             // nq x k (query mode)
  float* out_dists;              // nullable
  uint64_t* dist_cmps;           // nullable, nq
  uint32_t* visited_count;       // nullable, nq
  uint32_t* vis_ids;             // nullable (set in build mode): nq x vcap visited (id, dist), visit order
  float* vis_dists;
  uint32_t* fr_ids;              // nullable: the whole final frontier, nq x fr_stride (native traversal)
  float* fr_dists;
  uint32_t* fr_len;
  uint32_t fr_stride;            // >= max(beam, nstarts): a search that never expands keeps all its starts
  uint32_t* overflow;            // build mode: set to 1 when a list exceeds vcap
  uint32_t* work;                // work-queue counter, zero at launch
  uint64_t n, nq, limit;
  double cut;
  uint32_t d, ad, R, nstarts, k, beam, degree_limit, vcap;
  uint32_t build_mode, cut_trim;
  // reduced-precision traversal (float32 indices): the copy, its row stride
  // in bytes and 16-byte chunks per row, FP8 per-dimension scales, and the
  // number of final beam entries re-ranked in FP32 (0 = all)
  const void* trav_rows;
  const float* trav_scale;
  uint32_t trav_stride, trav_chunks, rerank;
  uint32_t trav_split, trav_cps;  // lanes per reduced row and chunks per lane (split_dist)
  // native float32 rows of more than kChunkBatch chunks: lanes per row (0 or 1:
  // one lane per row) and chunks per lane (split_row_dist)
  uint32_t nat_split, nat_cps;
  // visited filter in global memory (null: in LDS), gtable_words per
  // resident wave, indexed by block * waves_per_block + wave
  uint32_t* gtable;
  uint32_t gtable_words;
  // epoch-tagged slices (null: each query clears its slice): one epoch word
  // per slice; stored words are id | epoch << id_bits, epoch < epoch_end
  uint32_t* gepoch;
  uint32_t id_bits, epoch_end;
  // rows longer than one load batch are gathered into LDS cooperatively:
  // stage_rows rows of stage_chunks 16-byte chunks, stage_stride apart
  // (chunks), at off_stage; stage_rows = 0 keeps one lane per row
  uint32_t stage_rows, stage_chunks, stage_stride, off_stage;
  // per-wave LDS layout (bytes); the staged query takes qrow_bytes at offset
  // 0, its FP8 factors another qrow_bytes at off_qaux
  uint32_t lds_per_wave, qrow_bytes, off_qaux, off_table, off_f0, off_f1, off_keep, off_cand, off_aux, off_bkt;
  uint32_t off_wbits, wbits;  // global filter: LDS bitmap of the slots written by the current query (wbits != 0)
  uint32_t fcap, rcap;
};

// ------------------------------------------------------------------- prune
struct PruneArgs {
  const void* base;
  uint32_t* graph;             // n x (R1); read for add, written when write_graph
  const uint32_t* pts;         // task -> point id
  const uint32_t* a_ids;       // candidate ids
  const float* a_dists;        // nullable: computed when absent
  const uint64_t* a_off;       // CSR offsets (np + 1) or null
  const uint32_t* a_len;       // fixed-stride lists: length per task (when a_off is null)
  uint32_t* out_ids;           // np x Rsel (when !write_graph)
  uint32_t* out_len;           // np
  uint32_t* work;
  uint32_t* big_tasks;         // tasks whose candidate list exceeds the LDS capacity
  uint32_t* nbig;
  uint32_t* status;            // [0]: degree overflow
  float* gscratch_d;           // global candidate arrays of long-list tasks (bump-allocated)
  uint32_t* gscratch_i;
  uint64_t* big_off;           // scratch offset of each queued long-list task
  uint64_t* gused;             // scratch entries handed out so far
  uint64_t gpool;              // scratch entries available
  const uint64_t* np_dev;      // nullable: task count on the device (np then bounds it)
  uint64_t np;
  double alpha;
  uint32_t a_stride, d, ad, R1, Rsel, ccap, add, write_graph, append_if_fits;
  uint32_t lds_per_wave, off_cand, gcap, big_pass;
  // LDS row cache: tasks with at most rrows candidates stage their rows at
  // off_rows, rstride elements apart
  uint32_t rrows, off_rows, rstride;
  // integer rows: tasks with at most gram_max candidates (<= 256) prune from
  // an MFMA Gram matrix; its norms, sums and prune bit matrix live at off_gram
  uint32_t gram_max, off_gram;
  // float32 rows longer than one load batch: each distance split over
  // nat_split lanes of nat_cps chunks (split_row_dist; 0 = one lane per row)
  uint32_t nat_split, nat_cps;
};

struct Launch {
  uint32_t blocks, waves_per_block, lds_bytes;
};

// launchers (return hipSuccess or an error code)
hipError_t launch_search(int dtype, int metric, int trav, const SearchArgs& a, const Launch& l, hipStream_t s);
// distances of query q to ids[q*m .. q*m+m) as the traversal computes them;
// one wave per query, `a` carries the rows, copies and LDS layout
hipError_t launch_trav_dist(int dtype, int metric, int trav, const SearchArgs& a, const uint32_t* ids, uint32_t m,
                            float* out, const Launch& l, hipStream_t s);
// ordering keys of a launch's queries (distance to the start point, larger
// first, 16 bits), vals[q] = q; one wave per query, `a` carries the queries,
// the start point, the copies and the staged-query LDS layout
hipError_t launch_start_keys(int dtype, int metric, int trav, const SearchArgs& a, uint32_t* keys, uint32_t* vals,
                             const Launch& l, hipStream_t s);
// reduced traversal copies of float32 rows (trav.hip)
hipError_t launch_col_absmax(const float* rows, uint64_t n, uint32_t ad, uint32_t d, uint32_t* maxbits,
                             hipStream_t s);
hipError_t launch_make_bf16(const float* rows, uint64_t n, uint32_t ad, uint32_t d, void* out, uint32_t stride,
                            hipStream_t s);
hipError_t launch_make_fp8(const float* rows, uint64_t n, uint32_t ad, uint32_t d, const float* inv_scale, void* out,
                           uint32_t stride, hipStream_t s);
hipError_t launch_prune(int dtype, int metric, const PruneArgs& a, const Launch& l, hipStream_t s);
// raise the dynamic-LDS limit of every search / prune kernel (once per process)
void search_set_attributes(int max_lds);
void prune_set_attributes(int max_lds);
// LDS bytes the Gram-matrix prune path needs per wave
uint32_t prune_gram_lds_bytes();
// waves of a k_prune_big block (at most), and its LDS bytes
constexpr uint32_t kMaxBigWaves = 16;
uint32_t prune_big_lds_bytes(uint32_t row_bytes, uint32_t Rsel);

// edges.hip
// B_dev (nullable): the batch size the kernels honour, min(*B_dev, B)
hipError_t launch_apply_rows(uint32_t* graph, uint32_t R1, const uint32_t* pts, const uint32_t* new_ids,
                             const uint32_t* new_len, uint32_t Rsel, uint64_t B, uint32_t* status, hipStream_t s,
                             const uint64_t* B_dev = nullptr);
// Launchers taking (n, n_dev): n sizes the grid; when n_dev is not null the
// kernels process min(*n_dev, n) elements (the count is produced on the
// device by an earlier kernel of the same round).
hipError_t scan_exclusive_u32(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* tmp,
                              size_t tmp_bytes, hipStream_t s, const uint64_t* n_dev = nullptr);
size_t scan_tmp_bytes(uint64_t n);
hipError_t launch_emit_pairs(const uint32_t* pts, const uint32_t* new_ids, const uint32_t* new_len,
                             const uint64_t* offs, uint32_t Rsel, uint64_t B, uint32_t* keys, uint32_t* vals,
                             hipStream_t s, const uint64_t* B_dev = nullptr);
size_t radix_tmp_bytes(uint64_t n);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint64_t n,
                            uint32_t key_bits, void* tmp, size_t tmp_bytes, hipStream_t s, bool* result_in_alt,
                            const uint64_t* n_dev = nullptr);
hipError_t launch_group_starts(const uint32_t* keys, uint64_t n, uint32_t* flags, uint64_t* offs,
                               uint64_t* ngroups_dev, uint32_t* group_keys, uint64_t* group_off, void* tmp,
                               size_t tmp_bytes, hipStream_t s, const uint64_t* n_dev = nullptr);
hipError_t launch_sort_rows(int dtype, int metric, const void* base, uint32_t* graph, uint64_t n, uint32_t R1,
                            uint32_t d, uint32_t ad, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// graph rows ids[0..m) to (to_graph = 0) or from (to_graph = 1) the packed m x R1 array rows
hipError_t launch_copy_rows(uint32_t* graph, uint32_t R1, const uint32_t* ids, uint64_t m, uint32_t* rows,
                            int to_graph, hipStream_t s);
hipError_t launch_pad_rows(void* rows, uint64_t n, uint32_t d_bytes, uint32_t row_bytes, hipStream_t s);

// quant.hip: Quantized_Mips_Point distances of nq queries (float, or codes when
// qq) to n code rows, out nq x n; strides in elements
hipError_t launch_quantized_mips(int dtype, const void* queries, int qq, uint64_t nq, uint32_t qstride,
                                 const void* codes, uint64_t n, uint32_t cstride, uint32_t d, float mult, float minc,
                                 float* out, hipStream_t s);

// dense.hip
hipError_t launch_ground_truth(int dtype, int metric, const void* base, uint64_t n, const void* queries,
                               uint64_t nq, uint32_t d, uint32_t ad, uint32_t k, uint32_t* ids, double* dists,
                               void* tmp, size_t tmp_bytes, hipStream_t s);
size_t ground_truth_tmp_bytes(int dtype, uint64_t n, uint64_t nq, uint32_t d, uint32_t k);

}  // namespace ddg
