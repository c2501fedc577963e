/* ddg.h — C-ABI of the MI355X-native Vamana/DiskANN build and search.
 *
 * One handle (ddg_index) owns the device copy of one PointRange (vectors) and
 * one Graph (adjacency rows), plus one HIP stream and the build scratch.  The
 * memory layouts are the reference's, byte for byte:
 *   vectors  n rows of aligned_d elements (point_range.h:45-50, 80); the
 *            library zeroes the padding of its device copy
 *   graph    n rows of (R+1) u32: [degree, nbr_0 .. nbr_{R-1}] (graph.h:126-128)
 *
 * Every call is blocking and stream-synchronous; a handle is not thread-safe
 * (use one handle per host thread).  The caller owns every host buffer, the
 * library copies in and out.  No C++ exception crosses this boundary: each
 * call returns a ddg_status and ddg_last_error() describes the last failure.
 * The reference has no status codes — it prints and abort()s (graph.h:55-90,
 * index.h:217-223, beamSearch.h:278-282); include/ddg/knn_index.hpp maps a
 * non-OK status back to that print + abort().
 */
#ifndef DDG_H
#define DDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddg_index ddg_index;

typedef enum {
  DDG_OK = 0,
  DDG_EINVAL = 1,  /* bad argument / unsupported parameter combination      */
  DDG_ERANGE = 2,  /* a point id outside [0, n) (index.h:217-223)           */
  DDG_EDEGREE = 3, /* an adjacency list would exceed R (graph.h:62-90)      */
  DDG_EOOM = 4,    /* device allocation failed                             */
  DDG_EDEVICE = 5, /* HIP runtime error or missing GPU                      */
  DDG_EIO = 6      /* short or corrupt input                                */
} ddg_status;

typedef enum { DDG_U8 = 0, DDG_I8 = 1, DDG_F32 = 2 } ddg_dtype;

/* Euclidian_Point (euclidian_point.h:68-107, is_metric) and Mips_Point
 * (mips_point.h:68-109, not metric: no cut trimming in search). */
typedef enum { DDG_L2 = 0, DDG_MIPS = 1 } ddg_metric;

/* Vector copy the search traverses with:
 * NATIVE  = the index's own rows, the reference's arithmetic (any dtype);
 * BF16    = a BF16 copy of float32 rows (round to nearest even), FP32 query
 *           and FP32 accumulation;
 * FP8E4M3 = an FP8-e4m3 copy of float32 rows with one scale per dimension,
 *           x_ij ~ s_j * e_ij, s_j = max_i |x_ij| / 240, FP32 accumulation.
 * Both reduced modes end with an FP32 re-rank of the final beam against the
 * full rows (the reference's own distance), so returned distances are exact
 * FP32 distances. */
typedef enum { DDG_TRAV_NATIVE = 0, DDG_TRAV_BF16 = 1, DDG_TRAV_FP8E4M3 = 2 } ddg_trav;

/* BuildParams (types.h:103-142) + the batch_insert schedule knobs
 * (index.h:213-216) + the BSP keys of para.json (dataset.hpp:54-62). */
typedef struct {
  uint32_t R, L;
  double alpha;
  int two_pass;         /* index.h:131: a first pass with alpha = 1.0        */
  uint32_t start_point; /* 0 for knn_index::build_index; start_point_index in BSP */
  double base;          /* prefix-doubling base, 2 in the reference          */
  double max_fraction;  /* 1.0 for build_index; xita in BSP                  */
  uint64_t max_build;   /* BSP "max": insert only the first max points (0 = all) */
  int final_sort;       /* index.h:133-136 per-row sort (build_index only)   */
  int bsp;              /* 1: batch sizing over m (bsp_index.hpp:164), 0: over n */
} ddg_build_params;

/* QueryParams (types.h:145-156) plus the traversal precision. */
typedef struct {
  uint32_t k, beam;
  double cut;
  uint64_t limit;        /* max visited vertices (0 = n)                     */
  uint32_t degree_limit; /* 0 = R                                            */
  ddg_trav trav;
  uint32_t rerank;       /* reduced trav: beam entries re-ranked in FP32
                            (0 = the whole final beam, otherwise >= k)      */
} ddg_query_params;

/* Replaces PointRange(char*) + Graph(R, n) (point_range.h:61-97,
 * graph.h:126-128; diskann.cpp:48-52): copies n rows of aligned_d elements
 * (aligned_d >= d, any padding content) and allocates a zeroed graph. */
ddg_status ddg_create(ddg_dtype dtype, ddg_metric metric, uint64_t n, uint32_t d, uint32_t aligned_d,
                      const void* rows, uint32_t R, ddg_index** out);
void ddg_destroy(ddg_index* idx);
const char* ddg_last_error(const ddg_index* idx); /* idx may be NULL */

/* Graph(char*) / Graph::save in memory (graph.h:130-200). */
ddg_status ddg_graph_set(ddg_index* idx, const uint32_t* rows /* n*(R+1) */);
ddg_status ddg_graph_get(ddg_index* idx, uint32_t* rows /* n*(R+1) */);

/* knn_index::build_index (index.h:125-137) when bsp == 0, or one rank's
 * BSP_index::insert at P = 1 (bsp_index.hpp:141-434) when bsp == 1.
 * visited_counts (nullable, n entries) receives BuildStats.visited
 * (index.h:280).  build_seconds (nullable) receives the device time of the
 * build, vectors already resident. */
ddg_status ddg_build(ddg_index* idx, const ddg_build_params* bp, uint32_t* visited_counts,
                     double* build_seconds);

/* knn_index::batch_insert (index.h:213-341).  random_order != 0 inserts in
 * the order of parlay::random_permutation<int>(m) (index.h:237-241), which is
 * deterministic (default generator state); m must then be <= INT_MAX. */
ddg_status ddg_batch_insert(ddg_index* idx, const uint32_t* inserts, uint64_t m, double alpha, uint32_t L,
                            uint32_t start_point, int random_order, double base, double max_fraction);

/* knn_index::robustPrune (index.h:63-123), batched: task t prunes point
 * pts[t] over candidates cand_ids[cand_off[t] .. cand_off[t+1]) (with their
 * distances to pts[t] when cand_dists != NULL, computed otherwise), plus
 * G[pts[t]] when add != 0.  Writes up to R ids per task into out_ids
 * (np*R) and the count into out_len.  Candidates are ordered by
 * (distance, id). */
ddg_status ddg_robust_prune(ddg_index* idx, uint64_t np, const uint32_t* pts, const uint64_t* cand_off,
                            const uint32_t* cand_ids, const float* cand_dists, double alpha, int add, uint32_t R,
                            uint32_t* out_ids, uint32_t* out_len);

/* beam_search / searchAll (beamSearch.h:50-184, 322-346) for nq queries in
 * one persistent launch.  query_ids gives p.id() for the self-skip of
 * beamSearch.h:128 (NULL = the query's row index, as searchAll does).
 * per_query_beam (nullable) gives each query its own beam width.  Outputs
 * frontier[0:k] ids and distances, dist_cmps and |visited| per query
 * (nullable).  The launch takes its queries longest expected first: by beam
 * when beams differ, and, for a uniform beam of 140 or more with more queries
 * than resident wavefronts, by the best distance of a short pre-pass search
 * (beam 10, at most 8 visits) run in the same call.  Outputs do not depend
 * on that order; environment DDG_ORDER=0 turns it off. */
ddg_status ddg_search(ddg_index* idx, const void* queries /* nq*aligned_d */, uint64_t nq,
                      const uint32_t* query_ids, const uint32_t* starts, uint32_t nstarts,
                      const ddg_query_params* qp, const uint32_t* per_query_beam, uint32_t* out_ids,
                      float* out_dists, uint64_t* dist_cmps, uint32_t* visited);

/* ddg_search over device memory: queries, query_ids, starts,
 * per_query_beam and every output are device pointers (e.g. torch tensors on
 * the same GPU); beam_max bounds per_query_beam (or is at least qp->beam).
 * The host scalars are checked as ddg_search checks them (nstarts in
 * [1, 64], 0 < k <= beam <= beam_max, cut >= 1 on L2 when k > 0, the
 * traversal mode against the dtype) before any launch.  Device contents the
 * host cannot see: a per-query beam is clamped to [1, beam_max], and a start
 * id >= n fails the call with DDG_ERANGE (that query's outputs are then
 * empty).  The call returns once the search kernel has finished. */
ddg_status ddg_search_device(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                             const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                             const uint32_t* per_query_beam, uint32_t beam_max, uint32_t* out_ids, float* out_dists,
                             uint64_t* dist_cmps, uint32_t* visited);

/* ddg_search_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as ddg_search_device; the search's status words and
 * its global visited-filter slices (beams above 90) live in a fixed 64 MiB
 * device array that the library holds per GPU until the process ends, so
 * large beams run with as many resident waves as their filter slices fit
 * there, and a beam whose one block of filters exceeds it is rejected
 * (DDG_EINVAL).  A reduced traversal copy must exist already
 * (ddg_prepare_traversal).  Every call on a GPU shares that array: calls
 * that overlap on one GPU race. */
ddg_status ddg_search_device_noalloc(ddg_index* idx, const void* queries, uint64_t nq, const uint32_t* query_ids,
                                     const uint32_t* starts, uint32_t nstarts, const ddg_query_params* qp,
                                     const uint32_t* per_query_beam, uint32_t beam_max, uint32_t* out_ids,
                                     float* out_dists, uint64_t* dist_cmps, uint32_t* visited);

/* beam_search (beamSearch.h:41-54 signature, 50-184 algorithm) for nq points
 * in one launch, with everything the reference returns,
 * ((frontier, visited), dist_cmps): query i's final frontier (frontier_len[i]
 * <= max(beam, nstarts) (id, distance) pairs sorted by (distance, id)) at
 * frontier_ids / frontier_dists + i*fcap, its visited list sorted by
 * (distance, id) (beamSearch.h:115-117) at visited_ids / visited_dists +
 * i*vcap, and dist_cmps[i] (nullable).  query_ids gives p.id() for the
 * self-skip of beamSearch.h:128 (NULL = the query's row index); qp->k only
 * drives the cut trim (k = 0 is the build mode of index.h:277); native
 * traversal only.  fcap >= max(beam, nstarts).  visited_len[i] always holds
 * the true length; a list longer than vcap fails the call with DDG_EINVAL
 * (the caller can retry with a larger vcap). */
ddg_status ddg_beam_search(ddg_index* idx, const void* queries /* nq*aligned_d */, uint64_t nq,
                           const uint32_t* query_ids, const uint32_t* starts, uint32_t nstarts,
                           const ddg_query_params* qp, uint32_t fcap, uint32_t* frontier_ids, float* frontier_dists,
                           uint32_t* frontier_len, uint32_t vcap, uint32_t* visited_ids, float* visited_dists,
                           uint32_t* visited_len, uint64_t* dist_cmps);

/* Graph rows by id (graph.h:43-117: G[i] reads and G[i].update_neighbors,
 * batched): rows is m packed rows of R+1 words [degree, nbrs...].  set
 * checks every degree <= R and every id < n before writing. */
ddg_status ddg_graph_set_rows(ddg_index* idx, const uint32_t* ids, uint64_t m, const uint32_t* rows);
ddg_status ddg_graph_get_rows(ddg_index* idx, const uint32_t* ids, uint64_t m, uint32_t* rows);

/* Build-mode traces: beam_search(Points[ids[i]], G, Points, start,
 * QueryParams(0, L, 0, n, R)) as batch_insert calls it (index.h:277-279).
 * Writes |visited| and the visited (id, distance) lists, vcap per point, in
 * visit order, plus dist_cmps. */
ddg_status ddg_insert_search(ddg_index* idx, const uint32_t* ids, uint64_t m, uint32_t L, uint32_t start_point,
                             uint32_t vcap, uint32_t* vis_ids, float* vis_dists, uint32_t* vis_len,
                             uint64_t* dist_cmps);

/* Brute-force top-k (k <= 100, the width of the reference's ground-truth
 * files) of nq queries over the index's vectors: FP64 accumulation for
 * float32 (MFMA tiles), exact integer for uint8/int8, ties by id.  dists may
 * be NULL. */
ddg_status ddg_ground_truth(ddg_index* idx, const void* queries /* nq*aligned_d */, uint64_t nq, uint32_t k,
                            uint32_t* ids, double* dists);

/* Quantized_Mips_Point<T>::distance (mips_point.h:111-187) of nq queries to
 * n code rows (T = uint8 or int8 by code_dtype), in the reference's
 * arithmetic: code c decodes to (float)c * mult + min_coord with
 * mult = (max_coord - min_coord) / maxval, maxval = (float)(T)255 (255 for
 * uint8; the cast makes it -1 for int8); the products with the query (float
 * values, or codes decoded the same way when queries_quantized != 0) are
 * added in dimension order; the sum is returned unnegated, as the reference
 * returns it.  Strides in elements; out is nq x n (row-major); nq <= 65535.
 * Host buffers; runs on the current device.  The reference's drivers never
 * instantiate this point type, and no search here traverses with it. */
ddg_status ddg_quantized_mips_distances(ddg_dtype code_dtype, uint32_t d, float max_coord, float min_coord,
                                        const void* queries, int queries_quantized, uint64_t nq, uint32_t q_stride,
                                        const void* codes, uint64_t n, uint32_t c_stride, float* out);

/* index.h:133-136: sort every adjacency list by (distance to its owner, id). */
ddg_status ddg_sort_rows(ddg_index* idx);

/* Builds the reduced traversal copy of a float32 index (BF16 or FP8E4M3; a
 * no-op for NATIVE or when it exists).  ddg_search builds it on first use;
 * calling this first keeps the conversion out of a timed search.  seconds
 * (nullable) receives the device time of the conversion.  No reference
 * counterpart: the reference traverses its float32 rows only. */
ddg_status ddg_prepare_traversal(ddg_index* idx, ddg_trav trav, double* seconds);

/* Autotuner (no reference counterpart): times the search launch shapes that
 * give the same results -- visited filter in LDS or in global memory, 1/2/4
 * waves per block, 1..8 resident blocks per CU -- on these queries
 * (nq*aligned_d, host memory) with qp, checks that every shape returns the
 * same ids, and keeps the fastest for this handle's later searches with the
 * same traversal mode and filter size (each shape timed over one launch
 * after a warm-up; the three fastest and the default again over five, the
 * smallest median winning).  The filter size and frontier width
 * are fixed by beamSearch.h:64-73 and the beam, so they are not tuned.
 * For rows longer than 128 bytes it also tries gathering 2, 4, 8 or as
 * many candidate rows as one load batch holds into LDS per round trip.
 * The choice is also kept in a table of the process, keyed by (dtype,
 * metric, d, aligned d, R, traversal mode, filter size, CU count), which
 * every handle with the same key uses until it tunes for itself.
 * out4 (nullable): [filter in global memory (bit 0) | staged rows << 8,
 * waves per block, blocks per CU, shapes tried]; ms2 (nullable): [default
 * shape ms, chosen shape ms]. */
ddg_status ddg_autotune_search(ddg_index* idx, const void* queries, uint64_t nq, const ddg_query_params* qp,
                               uint32_t* out4, double* ms2);

/* The launch shape a search of this handle with qp (beam qp->beam, one
 * start point) would use now, in ddg_autotune_search's out4 layout except
 * out4[3]: where the shape comes from (0 the defaults, 1 this handle's own
 * tuning, 2 the process table).  No reference counterpart: it lets a caller
 * (and the tests) see what the autotuner left behind. */
ddg_status ddg_search_shape(ddg_index* idx, const ddg_query_params* qp, uint32_t* out4);

/* Point::distance (euclidian_point.h:44-65, mips_point.h:44-66) batched, in
 * the arithmetic of traversal mode trav: out[i*m + j] = distance of query i
 * (nq*aligned_d rows) to point ids[i*m + j].  NATIVE gives the reference's
 * bits; BF16 / FP8E4M3 give the values the reduced traversal ranks by. */
ddg_status ddg_distances(ddg_index* idx, ddg_trav trav, const void* queries, uint64_t nq, const uint32_t* ids,
                         uint32_t m, float* out);

/* Size / layout queries. */
uint64_t ddg_size(const ddg_index* idx);
uint32_t ddg_max_degree(const ddg_index* idx);
uint32_t ddg_dims(const ddg_index* idx);
uint32_t ddg_aligned_dims(const ddg_index* idx);

/* Kernel timing of the last call (milliseconds of device time per kernel
 * family, measured with HIP events on the handle's stream), for the
 * roofline report: [0] search, [1] forward prune, [2] edges (pairs, sort,
 * groups), [3] reverse append/prune, [4] row sort, [5] dense.  Also the
 * counters of the last ddg_search: [6] algorithmic bytes
 * = sum over queries of (dist_cmps * d * s + |visited| * 4 * (1 + R) + d * s),
 * with s the element size of the traversed copy, [7] total dist_cmps,
 * [8] total |visited|. */
ddg_status ddg_last_profile(const ddg_index* idx, double* out9);

#ifdef __cplusplus
}
#endif
#endif /* DDG_H */
