// kernels_ingest.hip -- the steps either side of the local analysis:
//
//  k_state_from_members  letkf_state_init (src/letkf_state.F90:326-334): from
//                        the members in file order (gridpoint fastest, then
//                        slab, then member) to state_ij(k, ns, ij) background
//                        perturbations, the mean SUM(.,1)/ens_size and the
//                        spread SQRT(DOT_PRODUCT(x', x')/(ens_size - 1)), in
//                        FP32 with the reference's order of operations
//  k_members_from_state  the analysis back to file order (the layout change of
//                        letkf_state_write_ens, :877-955) and the mean and
//                        spread (letkf_state_write_meansprd, :357-451)
//  k_obs_hx_split        letkf_obs_read (src/letkf_obs.F90:459-466):
//                        obs_hx_mean = SUM(obs_hx, 1)/ens_size and the
//                        perturbations obs_hx - obs_hx_mean, FP32
//  k_obs_compact         the QC compaction of :479-489 (qc > 0 dropped, order
//                        kept) as a gather through the selected indices
//  k_f64_to_f32          the FP32 view of the analysis (state_ij's REAL kind)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "letkf_internal.h"

namespace lk {

#pragma clang fp contract(off)

// one thread per (gridpoint, slab); members[ij + nij * (s + ns * m)]
__global__ void k_state_from_members(int64_t nij, int64_t ns, int k,
                                     const float* __restrict__ members,
                                     float* __restrict__ state, float* __restrict__ mean,
                                     float* __restrict__ sprd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nij * ns) return;
  const int64_t s = e / nij, ij = e - s * nij;
  const int64_t stride = nij * ns;
  const float* x = members + ij + nij * s;
  float sum = 0.0f;
  for (int m = 0; m < k; ++m) sum = sum + x[m * stride];
  const float mu = sum / (float)k;
  float* out = state + (int64_t)k * (s + ns * ij);
  float dot = 0.0f;
  for (int m = 0; m < k; ++m) {
    const float d = x[m * stride] - mu;
    out[m] = d;
    dot = dot + d * d;
  }
  mean[s + ns * ij] = mu;
  sprd[s + ns * ij] = sqrtf(dot / (float)(k - 1));
}

// analysis (FP64 on the device) to FP32 file order
__global__ void k_members_from_state(int64_t nij, int64_t ns, int k,
                                     const double* __restrict__ ana,
                                     const double* __restrict__ amean,
                                     const double* __restrict__ asprd,
                                     float* __restrict__ members, float* __restrict__ mean,
                                     float* __restrict__ sprd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nij * ns) return;
  const int64_t s = e / nij, ij = e - s * nij;
  const int64_t stride = nij * ns;
  const double* in = ana + (int64_t)k * (s + ns * ij);
  if (members)
    for (int m = 0; m < k; ++m) members[ij + nij * s + m * stride] = (float)in[m];
  if (mean) mean[ij + nij * s] = (float)amean[s + ns * ij];
  if (sprd) sprd[ij + nij * s] = (float)asprd[s + ns * ij];
}

// one thread per observation: hx (k, n) full values in, mean and
// perturbations out; flag = qc <= 0 (kept)
__global__ void k_obs_hx_split(int64_t n, int k, const float* __restrict__ hx,
                               const int32_t* __restrict__ qc, float* __restrict__ pert,
                               float* __restrict__ hmean, uint8_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* x = hx + (int64_t)k * i;
  float sum = 0.0f;
  for (int m = 0; m < k; ++m) sum = sum + x[m];
  const float mu = sum / (float)k;
  float* p = pert + (int64_t)k * i;
  for (int m = 0; m < k; ++m) p[m] = x[m] - mu;
  hmean[i] = mu;
  keep[i] = qc[i] <= 0 ? 1 : 0;
}

// gather the kept observations (sel: their indices in input order); the
// field pointers travel by value
__global__ void k_obs_compact(int64_t nsel, int k, const int32_t* __restrict__ sel,
                              ObsFields in, ObsFields out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsel) return;
  const int64_t i = sel[j];
#pragma unroll
  for (int f = 0; f < ObsFields::kF; ++f) out.f[f][j] = in.f[f][i];
#pragma unroll
  for (int f = 0; f < ObsFields::kI; ++f) out.i[f][j] = in.i[f][i];
  for (int m = 0; m < k; ++m) out.hx[(int64_t)k * j + m] = in.hx[(int64_t)k * i + m];
}

// FP64 -> FP32 (round to nearest), grid-stride over n values
__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ in, float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (float)in[i];
}

void launch_f64_to_f32(hipStream_t st, int64_t n, const double* in, float* out) {
  if (n <= 0) return;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 8192);
  k_f64_to_f32<<<(unsigned)nb, 256, 0, st>>>(n, in, out);
}

void launch_state_from_members(hipStream_t st, int64_t nij, int64_t ns, int k,
                               const float* members, float* state, float* mean, float* sprd) {
  const int64_t tot = nij * ns;
  if (tot <= 0) return;
  k_state_from_members<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(nij, ns, k, members,
                                                                     state, mean, sprd);
}

void launch_members_from_state(hipStream_t st, int64_t nij, int64_t ns, int k,
                               const double* ana, const double* amean, const double* asprd,
                               float* members, float* mean, float* sprd) {
  const int64_t tot = nij * ns;
  if (tot <= 0) return;
  k_members_from_state<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(
      nij, ns, k, ana, amean, asprd, members, mean, sprd);
}

void launch_obs_hx_split(hipStream_t st, int64_t n, int k, const float* hx, const int32_t* qc,
                         float* pert, float* hmean, uint8_t* keep) {
  if (n <= 0) return;
  k_obs_hx_split<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, k, hx, qc, pert, hmean, keep);
}

void launch_obs_compact(hipStream_t st, int64_t nsel, int k, const int32_t* sel,
                        const ObsFields& in, const ObsFields& out) {
  if (nsel <= 0) return;
  k_obs_compact<<<(unsigned)((nsel + 255) / 256), 256, 0, st>>>(nsel, k, sel, in, out);
}

}  // namespace lk
