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

}  // namespace lk
