// eigen_common.h -- wave-level helpers shared by the eigen kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace lk {

// sum over the K lanes of a K-thread block (K = 32 or 64)
template <int K>
__device__ __forceinline__ double wsumK(double v) {
#pragma unroll
  for (int o = K / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, K);
  return v;
}

// value of lane `lane` (a constant or wave-uniform index) in every lane
__device__ __forceinline__ double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// 64-bit DPP move of lane (row_shr / row_bcast) into lanes the row mask
// enables, 0 elsewhere
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// sum over the 64 lanes of a wave, returned wave-uniform: row_shr 1, 2, 4, 8
// leave each 16-lane row's sum in its lane 15, row_bcast 15 and 31 carry
// them on to lane 63
__device__ __forceinline__ double wave_sum_dpp(double x) {
  x += dpp_d<0x111, 0xf>(x);
  x += dpp_d<0x112, 0xf>(x);
  x += dpp_d<0x114, 0xf>(x);
  x += dpp_d<0x118, 0xf>(x);
  x += dpp_d<0x142, 0xa>(x);
  x += dpp_d<0x143, 0xc>(x);
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
}

// sum over each 32-lane half of a wave, left in the half's last lane (31 or
// 63; the other lanes hold partial sums): row_shr 1, 2, 4, 8 leave each
// 16-lane row's sum in its lane 15, row_bcast 15 adds row 0 into row 1 and
// row 2 into row 3.  The whole wave must be active.
__device__ __forceinline__ double half_sum_dpp(double x) {
  x += dpp_d<0x111, 0xf>(x);
  x += dpp_d<0x112, 0xf>(x);
  x += dpp_d<0x114, 0xf>(x);
  x += dpp_d<0x118, 0xf>(x);
  x += dpp_d<0x142, 0xa>(x);
  return x;
}

}  // namespace lk
