// mx_common.h -- operand types and BF16 conversions shared by the
// mixed-precision MFMA kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lk {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));


// two values per conversion (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ bf16x8 to_bf16(const float* v) {
  bf16x8 r;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const bf16x2 h = __builtin_convertvector((f32x2){v[2 * p], v[2 * p + 1]}, bf16x2);
    r[2 * p] = h[0];
    r[2 * p + 1] = h[1];
  }
  return r;
}

// x = hi + lo + O(2^-16 |x|): hi = bf16(x), lo = bf16(x - hi); x - hi is exact
// in FP32, and hi widens by a shift of the packed bits
__device__ __forceinline__ void split_bf16(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const float x0 = v[2 * p], x1 = v[2 * p + 1];
    const bf16x2 h = __builtin_convertvector((f32x2){x0, x1}, bf16x2);
    const uint32_t hu = __builtin_bit_cast(uint32_t, h);
    const float r0 = x0 - __uint_as_float(hu << 16);
    const float r1 = x1 - __uint_as_float(hu & 0xffff0000u);
    const bf16x2 l = __builtin_convertvector((f32x2){r0, r1}, bf16x2);
    hi[2 * p] = h[0];
    hi[2 * p + 1] = h[1];
    lo[2 * p] = l[0];
    lo[2 * p + 1] = l[1];
  }
}

}  // namespace lk
