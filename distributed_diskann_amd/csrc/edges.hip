// Reverse-edge insertion (index.h:287-323) as device passes over HBM:
//   apply rows     G[p_b] = new_out[b]                  (update_neighbors, graph.h:72-81)
//   scan + emit    (target, source) pairs in (b, j) order (index.h:290-297)
//   radix sort     stable LSD on the target, so each target's sources stay
//                  in batch order — parlay::group_by_key's order
//   group starts   segment boundaries of equal targets
// followed by the prune kernel in append_if_fits mode (index.h:309-322).
// Also: the final per-row sort (index.h:133-136) and small fill helpers.
#include "device.hpp"
#include "internal.hpp"

namespace ddg {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kTile = kBlock * kItems;  // 4096 elements per block

static inline uint32_t blocks_for(uint64_t n, uint64_t per) { return (uint32_t)((n + per - 1) / per); }

// Sizes given as (n, n_dev): n bounds the grid, n_dev (nullable) holds the
// count the kernel honours, so a launch can be recorded before its input
// size is known (graph capture of a build round).
__device__ __forceinline__ uint64_t dev_count(uint64_t n, const uint64_t* n_dev) {
  if (!n_dev) return n;
  const uint64_t v = *n_dev;
  return v < n ? v : n;
}

// ------------------------------------------------------------- small helpers
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// Graph rows ids[r] <-> packed row r of `rows` (R1 words each), one thread per
// word: a batch of G[i].update_neighbors / G[i] reads (graph.h:72-81).
__global__ void k_copy_rows(uint32_t* graph, uint32_t R1, const uint32_t* ids, uint64_t m, uint32_t* rows,
                            int to_graph) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * R1) return;
  const uint64_t r = i / R1, c = i - r * R1;
  uint32_t* g = graph + (uint64_t)ids[r] * R1 + c;
  if (to_graph) *g = rows[i];
  else rows[i] = *g;
}

hipError_t launch_copy_rows(uint32_t* graph, uint32_t R1, const uint32_t* ids, uint64_t m, uint32_t* rows,
                            int to_graph, hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_rows, dim3(blocks_for(m * R1, 256)), dim3(256), 0, s, graph, R1, ids, m, rows,
                     to_graph);
  return hipGetLastError();
}

// zero bytes [d_bytes, row_bytes) of every row
__global__ void k_pad_rows(unsigned char* rows, uint64_t n, uint32_t d_bytes, uint32_t row_bytes) {
  const uint32_t pad = row_bytes - d_bytes;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * pad) return;
  const uint64_t r = i / pad, c = i % pad;
  rows[r * row_bytes + d_bytes + c] = 0;
}

hipError_t launch_pad_rows(void* rows, uint64_t n, uint32_t d_bytes, uint32_t row_bytes, hipStream_t s) {
  if (row_bytes == d_bytes || n == 0) return hipSuccess;
  const uint64_t total = n * (row_bytes - d_bytes);
  hipLaunchKernelGGL(k_pad_rows, dim3(blocks_for(total, 256)), dim3(256), 0, s, (unsigned char*)rows, n, d_bytes,
                     row_bytes);
  return hipGetLastError();
}

// --------------------------------------------------------- apply forward rows
__global__ void k_apply_rows(uint32_t* graph, uint32_t R1, const uint32_t* pts, const uint32_t* new_ids,
                             const uint32_t* new_len, uint32_t Rsel, uint64_t B_max, const uint64_t* B_dev,
                             uint32_t* status) {
  const uint64_t b = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= dev_count(B_max, B_dev)) return;
  const uint32_t lane = lane_id();
  const uint32_t len = new_len[b];
  if (len > R1 - 1) {
    if (lane == 0) atomicOr(status, 1u);
    return;
  }
  uint32_t* row = graph + (uint64_t)pts[b] * R1;
  for (uint32_t j = lane; j < len; j += kWave) row[1 + j] = new_ids[b * Rsel + j];
  if (lane == 0) row[0] = len;
}

hipError_t launch_apply_rows(uint32_t* graph, uint32_t R1, const uint32_t* pts, const uint32_t* new_ids,
                             const uint32_t* new_len, uint32_t Rsel, uint64_t B, uint32_t* status, hipStream_t s,
                             const uint64_t* B_dev) {
  if (B == 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_rows, dim3(blocks_for(B, 4)), dim3(256), 0, s, graph, R1, pts, new_ids, new_len, Rsel,
                     B, B_dev, status);
  return hipGetLastError();
}

// --------------------------------------------------------------------- scan
// Exclusive scan of u32 into u64: per-block sums, a single-block scan of the
// sums, then the per-block scan with its offset.
__device__ __forceinline__ uint64_t block_scan_excl(uint64_t v, uint64_t* sh, uint64_t* total) {
  // wave-level inclusive scan, then across the 4 waves of the block
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x >> 6;
  uint64_t x = v;
  for (uint32_t o = 1; o < kWave; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) sh[w] = x;
  __syncthreads();
  uint64_t woff = 0, tot = 0;
  for (uint32_t i = 0; i < (blockDim.x >> 6); i++) {
    if (i < w) woff += sh[i];
    tot += sh[i];
  }
  __syncthreads();
  *total = tot;
  return woff + x - v;
}


__global__ void k_scan_reduce(const uint32_t* in, uint64_t n_max, const uint64_t* n_dev, uint64_t* bsum) {
  __shared__ uint64_t sh[4];
  const uint64_t n = dev_count(n_max, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint64_t s = 0;
  for (uint32_t k = 0; k < kItems; k++) {
    const uint64_t i = base + (uint64_t)threadIdx.x * kItems + k;
    if (i < n) s += in[i];
  }
  uint64_t tot;
  block_scan_excl(s, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_sums(uint64_t* bsum, uint64_t nb, uint64_t* total) {
  __shared__ uint64_t sh[4];
  uint64_t carry = 0;
  for (uint64_t c = 0; c < nb; c += kBlock) {
    const uint64_t i = c + threadIdx.x;
    const uint64_t v = i < nb ? bsum[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_scan_excl(v, sh, &tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_down(const uint32_t* in, uint64_t n_max, const uint64_t* n_dev, const uint64_t* bsum,
                            uint64_t* out) {
  __shared__ uint64_t sh[4];
  const uint64_t n = dev_count(n_max, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint64_t s = 0;
  for (uint32_t k = 0; k < kItems; k++) {
    const uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint64_t tot;
  uint64_t ex = block_scan_excl(s, sh, &tot) + bsum[blockIdx.x];
  for (uint32_t k = 0; k < kItems; k++) {
    const uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

size_t scan_tmp_bytes(uint64_t n) { return ((n + kTile - 1) / kTile + 1) * sizeof(uint64_t); }

hipError_t scan_exclusive_u32(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* tmp,
                              size_t tmp_bytes, hipStream_t s, const uint64_t* n_dev) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = blocks_for(n, kTile);
  if (tmp_bytes < scan_tmp_bytes(n)) return hipErrorInvalidValue;
  uint64_t* bsum = (uint64_t*)tmp;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, n_dev, bsum);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, bsum, (uint64_t)nb, total);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kBlock), 0, s, in, n, n_dev, bsum, out);
  return hipGetLastError();
}

// --------------------------------------------------------------- emit pairs
__global__ void k_emit_pairs(const uint32_t* pts, const uint32_t* new_ids, const uint32_t* new_len,
                             const uint64_t* offs, uint32_t Rsel, uint64_t B_max, const uint64_t* B_dev, uint32_t* keys,
                             uint32_t* vals) {
  const uint64_t b = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= dev_count(B_max, B_dev)) return;
  const uint32_t lane = lane_id();
  const uint32_t len = new_len[b];
  const uint64_t o = offs[b];
  const uint32_t src = pts[b];
  for (uint32_t j = lane; j < len; j += kWave) {
    keys[o + j] = new_ids[b * Rsel + j];
    vals[o + j] = src;
  }
}

hipError_t launch_emit_pairs(const uint32_t* pts, const uint32_t* new_ids, const uint32_t* new_len,
                             const uint64_t* offs, uint32_t Rsel, uint64_t B, uint32_t* keys, uint32_t* vals,
                             hipStream_t s, const uint64_t* B_dev) {
  if (B == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit_pairs, dim3(blocks_for(B, 4)), dim3(256), 0, s, pts, new_ids, new_len, offs, Rsel, B,
                     B_dev, keys, vals);
  return hipGetLastError();
}

// --------------------------------------------------------------- radix sort
// Stable LSD radix sort of (key, value) pairs, 8-bit digits.  A block owns a
// tile of 4096 consecutive elements: each of its 4 waves walks 1024 of them
// in order, 64 at a time, so the rank of an element among equal digits is
// (global digit offset of the block) + (same-digit elements of earlier waves)
// + (earlier steps of this wave) + (lower lanes of this step).
constexpr uint32_t kRadix = 256;
constexpr uint32_t kWaveItems = kTile / 4;  // per wave

__device__ __forceinline__ uint64_t match_digit(uint32_t dig, bool valid) {
  uint64_t m = ballot(valid);
  for (uint32_t b = 0; b < 8; b++) {
    const bool bit = (dig >> b) & 1u;
    const uint64_t bb = ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, uint64_t n_max, const uint64_t* n_dev,
                                                    uint32_t shift, uint32_t* hist, uint32_t nblocks) {
  __shared__ uint32_t h[kRadix];
  const uint64_t n = dev_count(n_max, n_dev);
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  for (uint32_t k = 0; k < kItems; k++) {
    const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* keys, const uint32_t* vals, uint64_t n_max,
                                                       const uint64_t* n_dev, uint32_t shift, const uint64_t* hoff,
                                                       uint32_t nblocks, uint32_t* okeys, uint32_t* ovals) {
  __shared__ uint32_t cnt[4][kRadix];
  const uint64_t n = dev_count(n_max, n_dev);
  __shared__ uint64_t wbase[4][kRadix];
  const uint32_t lane = lane_id();
  const uint32_t w = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < 4 * kRadix; i += kBlock) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t wstart = (uint64_t)blockIdx.x * kTile + (uint64_t)w * kWaveItems;
  // pass 1: per-wave digit counts
  for (uint32_t st = 0; st < kWaveItems; st += kWave) {
    const uint64_t i = wstart + st + lane;
    const bool v = i < n;
    const uint32_t dig = v ? (keys[i] >> shift) & 0xFFu : 0u;
    const uint64_t m = match_digit(dig, v);
    // the highest lane of each digit group adds the group size
    if (v && (m >> lane) == 1ull) cnt[w][dig] += popc(m);
    wave_sync();
  }
  __syncthreads();
  // per-digit base of each wave: global offset + counts of earlier waves
  {
    const uint32_t dig = threadIdx.x;
    uint64_t acc = hoff[(uint64_t)dig * nblocks + blockIdx.x];
    for (uint32_t ww = 0; ww < 4; ww++) {
      wbase[ww][dig] = acc;
      acc += cnt[ww][dig];
    }
  }
  __syncthreads();
  // pass 2: scatter
  for (uint32_t st = 0; st < kWaveItems; st += kWave) {
    const uint64_t i = wstart + st + lane;
    const bool v = i < n;
    uint32_t key = 0, val = 0;
    if (v) { key = keys[i]; val = vals[i]; }
    const uint32_t dig = (key >> shift) & 0xFFu;
    const uint64_t m = match_digit(dig, v);
    if (v) {
      const uint64_t dst = wbase[w][dig] + mbcnt(m);
      okeys[dst] = key;
      ovals[dst] = val;
    }
    wave_sync();
    if (v && (m >> lane) == 1ull) wbase[w][dig] += popc(m);
    wave_sync();
  }
}

size_t radix_tmp_bytes(uint64_t n) {
  const uint64_t nb = (n + kTile - 1) / kTile;
  const uint64_t h = nb * kRadix;
  return h * sizeof(uint32_t) + h * sizeof(uint64_t) + scan_tmp_bytes(h) + 256;
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint64_t n,
                            uint32_t key_bits, void* tmp, size_t tmp_bytes, hipStream_t s, bool* result_in_alt,
                            const uint64_t* n_dev) {
  *result_in_alt = false;
  if (n <= 1) return hipSuccess;
  if (tmp_bytes < radix_tmp_bytes(n)) return hipErrorInvalidValue;
  const uint32_t nb = blocks_for(n, kTile);
  const uint64_t h = (uint64_t)nb * kRadix;
  uint32_t* hist = (uint32_t*)tmp;
  uint64_t* hoff = (uint64_t*)((char*)tmp + ((h * sizeof(uint32_t) + 255) & ~(size_t)255));
  void* stmp = (char*)hoff + h * sizeof(uint64_t);
  size_t stmp_bytes = tmp_bytes - ((char*)stmp - (char*)tmp);
  uint32_t *ik = keys, *iv = vals, *ok = keys_alt, *ov = vals_alt;
  bool alt = false;
  for (uint32_t shift = 0; shift < key_bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, s, ik, n, n_dev, shift, hist, nb);
    hipError_t e = scan_exclusive_u32(hist, hoff, h, nullptr, stmp, stmp_bytes, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, s, ik, iv, n, n_dev, shift, hoff, nb, ok, ov);
    uint32_t* t;
    t = ik; ik = ok; ok = t;
    t = iv; iv = ov; ov = t;
    alt = !alt;
  }
  *result_in_alt = alt;
  return hipGetLastError();
}

// ------------------------------------------------------------- group starts
__global__ void k_group_flags(const uint32_t* keys, uint64_t n_max, const uint64_t* n_dev, uint32_t* flags) {
  const uint64_t n = dev_count(n_max, n_dev);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_group_write(const uint32_t* keys, const uint32_t* flags, const uint64_t* offs, uint64_t n_max,
                              const uint64_t* n_dev, uint32_t* group_keys, uint64_t* group_off,
                              const uint64_t* ngroups) {
  const uint64_t n = dev_count(n_max, n_dev);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) {
    group_keys[offs[i]] = keys[i];
    group_off[offs[i]] = i;
  }
  if (i == 0) group_off[*ngroups] = n;
}

hipError_t launch_group_starts(const uint32_t* keys, uint64_t n, uint32_t* flags, uint64_t* offs,
                               uint64_t* ngroups_dev, uint32_t* group_keys, uint64_t* group_off, void* tmp,
                               size_t tmp_bytes, hipStream_t s, const uint64_t* n_dev) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_flags, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, n_dev, flags);
  hipError_t e = scan_exclusive_u32(flags, offs, n, ngroups_dev, tmp, tmp_bytes, s, n_dev);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_group_write, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, flags, offs, n, n_dev,
                     group_keys, group_off, ngroups_dev);
  return hipGetLastError();
}

// ----------------------------------------------------------------- row sort
// index.h:133-136: one wave per vertex, rank sort of its list by
// (distance to the owner, id).
template <typename T, int METRIC>
__global__ __launch_bounds__(256) void k_sort_rows(const T* base, uint32_t* graph, uint64_t n, uint32_t R1,
                                                   uint32_t d, uint32_t ad) {
  extern __shared__ __align__(16) unsigned char smem[];
  using D = Dist<T, METRIC>;
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t row_bytes = ad * (uint32_t)sizeof(T);
  const uint32_t rb16 = (row_bytes + 15u) & ~15u;
  const uint32_t per_wave = rb16 + 8u * (R1 - 1);
  unsigned char* my = smem + w * per_wave;
  T* pv = (T*)my;
  float* sd = (float*)(my + rb16);
  uint32_t* si = (uint32_t*)(sd + (R1 - 1));
  const uint32_t nchunks = row_bytes >> 4;
  for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x >> 6) + w; v < n;
       v += (uint64_t)gridDim.x * (blockDim.x >> 6)) {
    uint32_t* row = graph + v * R1;
    const uint32_t deg = row[0];
    stage_row(pv, base + v * ad, row_bytes);
    wave_sync();
    const typename D::Q qs = D::prep(pv, nchunks);
    for (uint32_t j = lane; j < deg; j += kWave) {
      const uint32_t id = row[1 + j];
      sd[j] = D::dist(base + (uint64_t)id * ad, pv, qs, d, nchunks);
      si[j] = id;
    }
    wave_sync();
    for (uint32_t j = lane; j < deg; j += kWave) {
      const float dj = sd[j];
      const uint32_t ij = si[j];
      uint32_t rank = 0;
      for (uint32_t t = 0; t < deg; t++) {
        const float dt = sd[t];
        const uint32_t it = si[t];
        rank += (less_di(dt, it, dj, ij) || (dt == dj && it == ij && t < j)) ? 1u : 0u;
      }
      row[1 + rank] = ij;
    }
    wave_sync();
  }
}

template <typename T, int M>
static hipError_t sort_rows_t(const void* base, uint32_t* graph, uint64_t n, uint32_t R1, uint32_t d, uint32_t ad,
                              hipStream_t s) {
  const uint32_t rb16 = (ad * (uint32_t)sizeof(T) + 15u) & ~15u;
  const uint32_t lds = 4u * (rb16 + 8u * (R1 - 1));
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 3) / 4, 8192);
  hipLaunchKernelGGL((k_sort_rows<T, M>), dim3(blocks), dim3(256), lds, s, (const T*)base, graph, n, R1, d, ad);
  return hipGetLastError();
}

hipError_t launch_sort_rows(int dtype, int metric, const void* base, uint32_t* graph, uint64_t n, uint32_t R1,
                            uint32_t d, uint32_t ad, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (dtype == U8)
    return metric == L2 ? sort_rows_t<uint8_t, L2>(base, graph, n, R1, d, ad, s)
                        : sort_rows_t<uint8_t, MIPS>(base, graph, n, R1, d, ad, s);
  if (dtype == I8)
    return metric == L2 ? sort_rows_t<int8_t, L2>(base, graph, n, R1, d, ad, s)
                        : sort_rows_t<int8_t, MIPS>(base, graph, n, R1, d, ad, s);
  return metric == L2 ? sort_rows_t<float, L2>(base, graph, n, R1, d, ad, s)
                      : sort_rows_t<float, MIPS>(base, graph, n, R1, d, ad, s);
}

}  // namespace ddg
