// Quantized_Mips_Point::distance (mips_point.h:111-187) for every (query,
// code row) pair, in the reference's arithmetic: code c decodes to
// (float)c * mult + min_coord (a product, then a sum: the library is compiled
// with -ffp-contract=off), the products with the query are added in dimension
// order, and the sum is returned as it is (not negated).  A float query is a
// Mips_Point<float> query; a quantized query is decoded the same way.
#include "device.hpp"
#include "internal.hpp"

namespace ddg {

// One block column per query: the query's values (decoded when QQ) in LDS,
// one thread per code row.
template <typename T, bool QQ>
__global__ __launch_bounds__(256) void k_qmips(const void* queries, uint32_t qstride, const T* codes,
                                               uint32_t cstride, uint64_t n, uint32_t d, float mult, float minc,
                                               float* out) {
  extern __shared__ float qv[];
  const uint64_t q = blockIdx.y;
  for (uint32_t j = threadIdx.x; j < d; j += blockDim.x) {
    if (QQ) qv[j] = (float)((const T*)queries)[q * qstride + j] * mult + minc;
    else qv[j] = ((const float*)queries)[q * qstride + j];
  }
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* p = codes + i * cstride;
  float acc = 0.f;
  for (uint32_t j = 0; j < d; j++) {
    const float dc = (float)p[j] * mult + minc;
    acc = acc + dc * qv[j];
  }
  out[q * n + i] = acc;
}

template <typename T>
static hipError_t qmips_t(const void* queries, int qq, uint64_t nq, uint32_t qstride, const void* codes, uint64_t n,
                          uint32_t cstride, uint32_t d, float mult, float minc, float* out, hipStream_t s) {
  const dim3 grid((uint32_t)((n + 255) / 256), (uint32_t)nq);
  if (qq)
    hipLaunchKernelGGL((k_qmips<T, true>), grid, dim3(256), d * 4, s, queries, qstride, (const T*)codes, cstride, n,
                       d, mult, minc, out);
  else
    hipLaunchKernelGGL((k_qmips<T, false>), grid, dim3(256), d * 4, s, queries, qstride, (const T*)codes, cstride,
                       n, d, mult, minc, out);
  return hipGetLastError();
}

hipError_t launch_quantized_mips(int dtype, const void* queries, int qq, uint64_t nq, uint32_t qstride,
                                 const void* codes, uint64_t n, uint32_t cstride, uint32_t d, float mult, float minc,
                                 float* out, hipStream_t s) {
  if (dtype == U8) return qmips_t<uint8_t>(queries, qq, nq, qstride, codes, n, cstride, d, mult, minc, out, s);
  if (dtype == I8) return qmips_t<int8_t>(queries, qq, nq, qstride, codes, n, cstride, d, mult, minc, out, s);
  return hipErrorInvalidValue;
}

}  // namespace ddg
