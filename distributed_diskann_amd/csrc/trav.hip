// Reduced traversal copies of float32 rows: BF16 (round to nearest even) and
// FP8-e4m3 with one scale per dimension.  Each copy row is `stride` bytes,
// elements beyond d are zero.
#include "device.hpp"
#include "internal.hpp"

namespace ddg {

// max_i |x_ij| per dimension j, as the bit pattern of a non-negative float
// (which orders like the value).  Block b walks a stripe of rows; thread t
// covers dimensions t, t + blockDim.x, ...; one atomicMax per (block, dim).
__global__ __launch_bounds__(256) void k_col_absmax(const float* rows, uint64_t n, uint32_t ad, uint32_t d,
                                                    uint32_t* maxbits) {
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * per;
  const uint64_t r1 = r0 + per < n ? r0 + per : n;
  for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
    uint32_t m = 0;
#pragma unroll 8
    for (uint64_t r = r0; r < r1; r++) {
      const uint32_t b = __float_as_uint(rows[r * ad + j]) & 0x7FFFFFFFu;
      m = b > m ? b : m;
    }
    if (r0 < r1) atomicMax(&maxbits[j], m);
  }
}

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// one thread per 16-byte output chunk (8 elements)
__global__ __launch_bounds__(256) void k_make_bf16(const float* rows, uint64_t n, uint32_t ad, uint32_t d,
                                                   uint4* out, uint32_t stride) {
  const uint32_t cpr = stride >> 4;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cpr) return;
  const uint64_t r = t / cpr;
  const uint32_t c = (uint32_t)(t - r * cpr);
  const float* x = rows + r * ad;
  uint32_t w[4];
  for (uint32_t i = 0; i < 4; i++) {
    const uint32_t j = c * 8u + 2u * i;
    const uint32_t lo = j < d ? bf16_rne(x[j]) : 0u;
    const uint32_t hi = j + 1 < d ? bf16_rne(x[j + 1]) : 0u;
    w[i] = lo | (hi << 16);
  }
  out[t] = make_uint4(w[0], w[1], w[2], w[3]);
}

// one thread per 16-byte output chunk (16 codes); code j = fp8(x_j * inv_j),
// round to nearest even, byte order = element order
__global__ __launch_bounds__(256) void k_make_fp8(const float* rows, uint64_t n, uint32_t ad, uint32_t d,
                                                  const float* inv, uint4* out, uint32_t stride) {
  const uint32_t cpr = stride >> 4;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cpr) return;
  const uint64_t r = t / cpr;
  const uint32_t c = (uint32_t)(t - r * cpr);
  const float* x = rows + r * ad;
  uint32_t w[4];
  for (uint32_t i = 0; i < 4; i++) {
    float v[4];
    for (uint32_t e = 0; e < 4; e++) {
      const uint32_t j = c * 16u + 4u * i + e;
      v[e] = j < d ? x[j] * inv[j] : 0.f;
    }
    int p = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
    p = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], p, true);
    w[i] = (uint32_t)p;
  }
  out[t] = make_uint4(w[0], w[1], w[2], w[3]);
}

static uint32_t grid_for(uint64_t threads) { return (uint32_t)((threads + 255) / 256); }

hipError_t launch_col_absmax(const float* rows, uint64_t n, uint32_t ad, uint32_t d, uint32_t* maxbits,
                             hipStream_t s) {
  const uint32_t blocks = (uint32_t)(n < 2048 ? n : 2048);
  hipLaunchKernelGGL(k_col_absmax, dim3(blocks), dim3(256), 0, s, rows, n, ad, d, maxbits);
  return hipGetLastError();
}

hipError_t launch_make_bf16(const float* rows, uint64_t n, uint32_t ad, uint32_t d, void* out, uint32_t stride,
                            hipStream_t s) {
  const uint64_t t = n * (stride >> 4);
  if (t == 0) return hipSuccess;
  hipLaunchKernelGGL(k_make_bf16, dim3(grid_for(t)), dim3(256), 0, s, rows, n, ad, d, (uint4*)out, stride);
  return hipGetLastError();
}

hipError_t launch_make_fp8(const float* rows, uint64_t n, uint32_t ad, uint32_t d, const float* inv_scale, void* out,
                           uint32_t stride, hipStream_t s) {
  const uint64_t t = n * (stride >> 4);
  if (t == 0) return hipSuccess;
  hipLaunchKernelGGL(k_make_fp8, dim3(grid_for(t)), dim3(256), 0, s, rows, n, ad, d, inv_scale, (uint4*)out, stride);
  return hipGetLastError();
}

}  // namespace ddg
